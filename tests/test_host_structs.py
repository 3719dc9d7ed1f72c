"""Host-side checks: ctypes mirrors of the C ABI structs match the C layout,
and the device-memory pool's bookkeeping (CPU, no GPU calls)."""
import ctypes
import os
import subprocess

import framework_amd as fa
from framework_amd import mdmc, server

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _c_layout(tmp_path):
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "golden", "abi_layout.c"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return {k: int(v) for k, v in (ln.split() for ln in out.splitlines())}


def test_ctypes_structs_match_c_layout(tmp_path):
    c = _c_layout(tmp_path)
    pairs = {"mdmc_counts": mdmc.Counts, "mdmc_mesh": mdmc.Mesh, "mdmc_volume": mdmc.Volume,
             "mdmc_server_config": server.ServerConfig, "mdmc_request": server.Request,
             "mdmc_result": server.Result, "mdmc_server_stats": server.ServerStats}
    for name, cls in pairs.items():
        assert ctypes.sizeof(cls) == c[name], name
    for key, val in c.items():
        if "." not in key:
            continue
        name, field = key.split(".")
        cls = {"mdmc_volume": mdmc.Volume, "mdmc_request": server.Request, "mdmc_result": server.Result,
               "mdmc_server_stats": server.ServerStats, "mdmc_gen_spec": None}.get(name)
        if cls is not None:
            assert getattr(cls, field).offset == val, key


def test_pool_bookkeeping(tmp_path):
    exe = tmp_path / "pool"
    subprocess.run(["g++", "-std=c++14", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "framework_amd", "csrc"),
                    os.path.join(ROOT, "tests", "golden", "pool_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120).stdout
    assert out.strip() == "OK", out


def test_batch_scratch_query_without_gpu():
    """mdmc_batch_scratch_bytes only reads dims (no device call)."""
    L = fa.lib()
    L.mdmc_batch_scratch_bytes.argtypes = [ctypes.POINTER(mdmc.Volume), ctypes.c_uint32,
                                           ctypes.POINTER(ctypes.c_uint64)]
    vols = (mdmc.Volume * 3)()
    for i, d in enumerate([(64, 64, 64), (1, 5, 5), (300, 20, 10)]):
        vols[i].dims[0], vols[i].dims[1], vols[i].dims[2] = d
    b = ctypes.c_uint64()
    assert L.mdmc_batch_scratch_bytes(vols, 3, ctypes.byref(b)) == 0
    # the record slots dominate: 8 bytes per slot.  The x-tile width with the
    # fewest padded cells is 256 here (63 -> 256, 299 -> 2 x 256 beats 512 and
    # 512), so the 300-wide volume spans two x-tiles
    segs = 63 * 63 + 2 * 19 * 9
    assert b.value >= segs * 256 * 8
    assert b.value < (63 * 63 + 19 * 9) * 512 * 8
    one = (mdmc.Volume * 1)()
    one[0].dims[0], one[0].dims[1], one[0].dims[2] = 1, 1, 1
    assert L.mdmc_batch_scratch_bytes(one, 1, ctypes.byref(b)) == 0 and b.value == 0


def test_single_scratch_query_without_gpu():
    """mdmc_scratch_bytes is host arithmetic: the record slots (8 bytes per
    slot, one x-tile of at least 256 slots per (y, z) row of cells) bound it
    from below, an empty grid needs none, a null argument is refused."""
    L = fa.lib()
    L.mdmc_scratch_bytes.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint64)]
    b = ctypes.c_uint64()
    dims = (ctypes.c_uint32 * 3)(200, 150, 100)
    assert L.mdmc_scratch_bytes(dims, ctypes.byref(b)) == 0
    assert b.value >= 149 * 99 * 256 * 8
    assert b.value < 149 * 99 * 1024 * 8
    big = ctypes.c_uint64()
    assert L.mdmc_scratch_bytes((ctypes.c_uint32 * 3)(400, 150, 100), ctypes.byref(big)) == 0
    assert big.value > b.value
    assert L.mdmc_scratch_bytes((ctypes.c_uint32 * 3)(1, 9, 9), ctypes.byref(b)) == 0 and b.value == 0
    assert L.mdmc_scratch_bytes(None, ctypes.byref(b)) != 0


def test_classify_tile_map_covers_every_tile_once(tmp_path):
    """The plane-tile classify's workgroup -> tile map (mdmc_tilemap.h), over
    3000 random grid shapes: every tile exactly once, XCD shares within one
    tile of each other (tests/golden/tile_map_check.cpp)."""
    exe = tmp_path / "tile_map_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "framework_amd", "csrc"),
                    os.path.join(ROOT, "tests", "golden", "tile_map_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120).stdout
    assert out.strip() == "OK", out


def test_emit_tickets_take_every_chunk_once(tmp_path):
    """The persistent emit's chunk tickets (mdmc_tilemap.h), simulated on the
    CPU in k_emit's control flow with the workgroups advanced in random order
    over 4000 random chunk counts, grids and wave counts: every chunk exactly
    once, none past the end, every workgroup stops, bounded counters
    (tests/golden/emit_ticket_check.cpp)."""
    exe = tmp_path / "emit_ticket_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "framework_amd", "csrc"),
                    os.path.join(ROOT, "tests", "golden", "emit_ticket_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300).stdout
    assert out.strip() == "OK", out
