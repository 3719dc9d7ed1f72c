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


@pytest.mark.parametrize("nv,nq", [(257, 300), (70001, 50003)])
def test_obj_writers_agree(tmp_path, nv, nq):
    """mdmc_write_obj over quads and the shim's dump_obj over the triangles the
    shim's run_dmc makes from them give the same bytes; faces are 1-based.  The
    larger mesh spans several of the chunks mdmc_write_obj formats on separate
    threads (16384 vertices or quads each), so their order is checked too."""
    import numpy as np

    rng = np.random.default_rng(3)
    v = (rng.standard_normal((nv, 3)) * [1.0, 300.0, 1e-3]).astype(np.float32)
    v[5] = [0.1, 1e7, -0.0]
    v[9, 1] = np.float32(np.nan)
    q = rng.integers(0, nv, (nq, 4)).astype(np.int32)
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
    assert len(lines) == nv + 2 * nq
    assert lines[0].startswith("v ") and lines[nv] == "f %d %d %d" % (q[0, 0] + 1, q[0, 1] + 1, q[0, 2] + 1)
    assert lines[-1] == "f %d %d %d" % (q[-1, 2] + 1, q[-1, 3] + 1, q[-1, 0] + 1)
    fa.write_obj(str(tmp_path / "c.obj"), v, q)
    assert (tmp_path / "c.obj").read_bytes() == a


def test_obj_number_text_matches_printf(tmp_path):
    """The OBJ writer formats numbers itself (mdmc_fmt.h): its "%g" of a
    float and its "%u" are byte for byte snprintf's, over every 997th float
    bit pattern, dense sweeps around powers of ten and decade carries, and
    random values (tests/golden/fmt_check.cpp)."""
    exe = tmp_path / "fmt_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "framework_amd", "csrc"),
                    os.path.join(ROOT, "tests", "golden", "fmt_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300).stdout
    assert out.strip() == "OK", out


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


def test_shim_utilities(tmp_path):
    """include/dmc.hpp's constants, typedefs, Array3D iterators, index
    helpers and argmin/argmax (tests/golden/shim_utils_check.cpp)."""
    src = os.path.join(ROOT, "tests", "golden", "shim_utils_check.cpp")
    # with the header's own vector types, and with HIP's (a caller that
    # includes HIP's headers, as the reference's callers include CUDA's)
    for i, extra in enumerate(([], ["-D__HIP_PLATFORM_AMD__", "-isystem", "/opt/rocm/include"])):
        exe = tmp_path / f"shim_utils_check{i}"
        subprocess.run(["g++", "-std=c++14", "-O1", "-Wall", "-Werror", *extra, "-I", os.path.join(ROOT, "include"),
                        src, "-o", str(exe)], check=True)
        out = subprocess.run([str(exe)], capture_output=True, text=True)
        assert out.returncode == 0 and out.stdout.strip() == "OK", out.stdout + out.stderr


def test_harness_builds(tmp_path):
    """The test_dmc-style program (11-pointer setup, run_dmc, dump_obj,
    11-pointer cleanup) compiles against the shim and links (run on the GPU
    by tests/test_gpu_dropin.py)."""
    exe = tmp_path / "dmc_harness"
    subprocess.run(harness_build_command(str(exe)), check=True)
    assert exe.exists()


def harness_build_command(exe: str) -> list:
    libdir = os.path.dirname(fa.mdmc.LIB_PATH)
    return ["g++", "-std=c++14", "-O2", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
            "-isystem", "/opt/rocm/include", os.path.join(ROOT, "tests", "golden", "dmc_harness.cpp"), "-L", libdir, "-lmdmc",
            "-L", "/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath," + libdir, "-Wl,-rpath,/opt/rocm/lib", "-o", exe]


# The reference's small tables (DMC_CUDA/DMC_CUDA/src/dmc.cu:622-661), as data.
REF_OPPOSITE_FACE = [5, 3, 4, 1, 2, 0]                       # :622
REF_FACE_TO_CHECK_DIR = [5, 3, 0, 2, 1, 4]                   # :631-632 (NEG_Z, NEG_Y, POS_X, POS_Y, NEG_X, POS_Z)
REF_EDGE_BELONGED_VOXEL = [0x60 | 10, 0x20 | 9, 0x20 | 10, 0xA0 | 9, 0xC0 | 6, 0x40 | 6, 0xFF, 0x80 | 6,
                           0x40 | 10, 0xFF, 0xFF, 0x80 | 9]  # :635-649
REF_CIRCULAR_EDGE = [6, 7, 4, 5, 9, 1, 3, 11, 10, 8, 0, 2]   # :653-658
REF_LOCAL_EDGES = [6, 9, 10]                                 # :661
OFFSETS = [0, 3072, 6144, 6400, 6656, 6692, 6728, 6734, 6740, 6752, 6764, 6784]


def reference_tables_host() -> list:
    """mdmc_reference_luts_host split into the 11 tables of mdmc_device_luts."""
    buf = (ctypes.c_uint8 * 6784)()
    assert fa.lib().mdmc_reference_luts_host(buf) == 0
    b = bytes(buf)
    return [b[OFFSETS[i]:OFFSETS[i + 1]] for i in range(11)]


def check_reference_tables(t: list) -> None:
    """The 11 tables of the 11-pointer setup against the reference: SHA-256
    pins of its case tables, its small tables as data."""
    import hashlib

    import numpy as np

    from tests.test_luts import PINS

    def sha(x):
        return hashlib.sha256(bytes(x)).hexdigest()

    assert sha(t[0]) == PINS["edge_label_v0"] and sha(t[1]) == PINS["edge_label_v1"]
    assert sha(t[2]) == PINS["num_vertex_v0"] and sha(t[3]) == PINS["num_vertex_v1"]
    amb = np.frombuffer(t[4], np.uint8)
    assert sha(t[4]) == PINS["amb_sorted"] and sha(t[5]) == PINS["amb_faces_sorted"]
    assert (np.diff(amb.astype(int)) > 0).all()
    assert list(t[6]) == REF_OPPOSITE_FACE
    assert list(t[7]) == REF_FACE_TO_CHECK_DIR
    assert list(t[8]) == REF_EDGE_BELONGED_VOXEL
    assert list(t[9]) == REF_CIRCULAR_EDGE
    assert list(t[10][:3]) == REF_LOCAL_EDGES


def test_reference_format_tables_on_host():
    check_reference_tables(reference_tables_host())


def test_table_cleanup_rejects_foreign_pointers():
    """The 11-pointer cleanup frees only what the setup handed out (CPU: no
    table was handed out, so any pointers are foreign)."""
    raw = ctypes.CDLL(fa.mdmc.LIB_PATH)
    ptrs = (ctypes.c_void_p * 11)(*[0x1000 + i for i in range(11)])
    assert raw.mdmc_cleanup_device_luts(ptrs) in (-1, -2)  # invalid, or no device to ask
