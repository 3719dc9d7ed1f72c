"""The C++ drop-in (include/dmc.hpp) run on the GPU the way the reference's
own harness runs (test_dmc, DMC_CUDA/DMC_CUDA/src/dmc.cu:2248-2306): the
11-pointer table setup, run_dmc with 15 smoothing iterations on a 21^3 sphere
over [-5,5]^3 at iso 4.1, dump_obj, the 11-pointer cleanup.  Its vertices and
triangles must equal the oracle's (vertices bitwise, triangles exactly), its
OBJ file must equal the bytes mdmc_write_obj writes for the oracle's mesh, and
the tables it read back from the device must be the reference's."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import oracle
from tests.test_abi import OFFSETS, check_reference_tables, harness_build_command

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("harness") / "dmc_harness")
    subprocess.run(harness_build_command(exe), check=True)
    return exe


def _run(harness, out_dir, *args):
    print(f"dmc_harness {out_dir} {' '.join(map(str, args))}", flush=True)
    res = subprocess.run([harness, str(out_dir), *map(str, args)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    n = int(round(float(args[0]))) + 1 if args else 21
    field = np.fromfile(out_dir / "field.f32", np.float32).reshape(n, n, n)
    verts = np.fromfile(out_dir / "vertices.f32", np.float32).reshape(-1, 3)
    tris = np.fromfile(out_dir / "triangles.u32", np.uint32).reshape(-1, 3)
    return field, verts, tris


@pytest.mark.parametrize("args", [(), (48, 3.3, 0), (33, 2.0, 4)])
def test_reference_harness_sequence(harness, tmp_path, args):
    import framework_amd as fa

    field, verts, tris = _run(harness, tmp_path, *args)
    iso = float(np.float32(args[1])) if args else float(np.float32(4.1))
    num_smooth = args[2] if args else 15
    ref = oracle.run(field, iso, (-5.0, -5.0, -5.0), (5.0, 5.0, 5.0), num_smooth=num_smooth)
    assert ref["num_quads"] > 0
    assert verts.shape == ref["vertices"].shape
    same = (verts.view(np.uint32) == ref["vertices"].view(np.uint32)) | (np.isnan(verts) & np.isnan(ref["vertices"]))
    assert same.all(), f"{(~same).sum()} vertex components differ"
    assert np.array_equal(tris, fa.triangles_from_quads(ref["quads"]))
    # the (q0,q1,q2),(q2,q3,q0) split of dmc.cu:2039-2045, stated directly
    q = ref["quads"].astype(np.uint32)
    assert np.array_equal(tris[0::2], q[:, [0, 1, 2]]) and np.array_equal(tris[1::2], q[:, [2, 3, 0]])
    fa.write_obj(str(tmp_path / "oracle.obj"), ref["vertices"], ref["quads"])
    assert (tmp_path / "sphere.obj").read_bytes() == (tmp_path / "oracle.obj").read_bytes()
    tb = (tmp_path / "tables.bin").read_bytes()
    assert len(tb) == OFFSETS[-1] - 17  # the 11 tables without the block's padding
    check_reference_tables([tb[OFFSETS[i]:OFFSETS[i + 1]] for i in range(11)])
