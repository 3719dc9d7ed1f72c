"""The C ABI library loads and exports every declared symbol (CPU, no GPU calls)."""
import ctypes
import os
import subprocess

import pytest

import framework_amd as fa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_function():
    L = fa.lib()
    names = fa.declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert missing == []


def test_exported_symbols_have_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", fa.mdmc.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    for n in fa.declared_functions():
        assert n in exported, n


def test_status_strings():
    L = fa.lib()
    assert L.mdmc_status_string(0) == b"ok"
    assert L.mdmc_status_string(-4) == b"output capacity too small"
    assert L.mdmc_status_string(12345) == b"unknown status"


def test_smoothing_needs_a_context():
    raw = ctypes.CDLL(fa.mdmc.LIB_PATH)
    null = ctypes.c_void_p(None)
    assert raw.mdmc_smooth_device(null, null, ctypes.c_uint64(0), null, ctypes.c_uint64(0), 1, null, null) == -1


def test_null_arguments_are_rejected():
    raw = ctypes.CDLL(fa.mdmc.LIB_PATH)  # untyped handle: plain null pointers
    null = ctypes.c_void_p(None)
    assert raw.mdmc_run(null, null, null, null, ctypes.c_float(0.0), 0, null) == -1
    assert raw.mdmc_ctx_create(null) == -1
    assert raw.mdmc_count_device(null, null, null, ctypes.c_float(0.0), null, null) == -1


def test_cpp_shim_header_compiles_and_links(tmp_path):
    """include/dmc.hpp (the dmc::run_dmc drop-in) builds with the host compiler
    and links against the library (the program is not run: no GPU here)."""
    src = os.path.join(ROOT, "tests", "golden", "shim_compile_check.cpp")
    main = tmp_path / "main.cpp"
    main.write_text("int shim_compile_check();\nint main() { return 0; }\n")
    libdir = os.path.dirname(fa.mdmc.LIB_PATH)
    subprocess.run(["g++", "-std=c++14", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(main), src,
                    "-L", libdir, "-lmdmc", "-Wl,-rpath-link,/opt/rocm/lib", "-o", str(tmp_path / "shim")],
                   check=True)


def test_obj_writers_agree(tmp_path):
    """mdmc_write_obj over quads and the shim's dump_obj over the triangles the
    shim's run_dmc makes from them give the same bytes; faces are 1-based."""
    import numpy as np

    rng = np.random.default_rng(3)
    v = (rng.standard_normal((257, 3)) * [1.0, 300.0, 1e-3]).astype(np.float32)
    v[5] = [0.1, 1e7, -0.0]
    v[9, 1] = np.float32(np.nan)
    q = rng.integers(0, 257, (300, 4)).astype(np.int32)
    (tmp_path / "v.f32").write_bytes(v.tobytes())
    (tmp_path / "q.i32").write_bytes(q.tobytes())
    src = os.path.join(ROOT, "tests", "golden", "obj_check.cpp")
    libdir = os.path.dirname(fa.mdmc.LIB_PATH)
    exe = tmp_path / "obj_check"
    subprocess.run(["g++", "-std=c++14", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), src, "-L",
                    libdir, "-lmdmc", "-Wl,-rpath," + libdir, "-Wl,-rpath-link,/opt/rocm/lib", "-o", str(exe)],
                   check=True)
    subprocess.run([str(exe), str(tmp_path / "v.f32"), str(tmp_path / "q.i32"), str(tmp_path / "a.obj"),
                    str(tmp_path / "b.obj")], check=True)
    a, b = (tmp_path / "a.obj").read_bytes(), (tmp_path / "b.obj").read_bytes()
    assert a == b
    lines = a.decode().splitlines()
    assert len(lines) == 257 + 600
    assert lines[0].startswith("v ") and lines[257] == "f %d %d %d" % (q[0, 0] + 1, q[0, 1] + 1, q[0, 2] + 1)
    fa.write_obj(str(tmp_path / "c.obj"), v, q)
    assert (tmp_path / "c.obj").read_bytes() == a


def test_obj_writer_rejects_out_of_range_quads(tmp_path):
    """A face naming a vertex outside the array is an error, and no file is made."""
    import numpy as np

    v = np.zeros((4, 3), np.float32)
    for bad in (4, -1):
        q = np.array([[0, 1, 2, bad]], np.int32)
        out = tmp_path / ("bad%d.obj" % bad)
        with pytest.raises(fa.mdmc.MDMCError, match="outside the vertex array"):
            fa.write_obj(str(out), v, q)
        assert not out.exists()
    fa.write_obj(str(tmp_path / "ok.obj"), v, np.array([[0, 1, 2, 3]], np.int32))
    assert (tmp_path / "ok.obj").read_text().count("\nf ") == 2
