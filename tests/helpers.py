"""Shared checks for the parity tests."""
from __future__ import annotations

import numpy as np

import oracle


def assert_same_mesh(gpu_verts: np.ndarray, gpu_quads: np.ndarray, ref: dict, tol: float = 1e-5,
                     bitwise: bool = True) -> None:
    """Topology bit-exact; vertex positions bitwise equal (or within `tol`)."""
    assert gpu_quads.shape == ref["quads"].shape, (gpu_quads.shape, ref["quads"].shape)
    if not np.array_equal(gpu_quads, ref["quads"]):
        bad = np.nonzero((gpu_quads != ref["quads"]).any(axis=1))[0]
        raise AssertionError(f"{bad.size} quads differ; first {bad[:5]}: gpu {gpu_quads[bad[:3]]} ref {ref['quads'][bad[:3]]}")
    assert gpu_verts.shape == ref["vertices"].shape
    rv = ref["vertices"]
    nan_g, nan_r = np.isnan(gpu_verts), np.isnan(rv)
    assert np.array_equal(nan_g, nan_r), "NaN pattern differs"
    fin = ~nan_r
    err = np.abs(gpu_verts[fin].astype(np.float64) - rv[fin].astype(np.float64))
    assert err.size == 0 or err.max() <= tol, f"max vertex error {err.max()}"
    if bitwise:
        same = gpu_verts.view(np.uint32) == rv.view(np.uint32)
        # NaN payloads may differ between the CPU and GPU for NaN inputs
        same |= nan_g & nan_r
        assert same.all(), f"{(~same).sum()} vertex components differ bitwise (max err {err.max() if err.size else 0})"


def assert_same_cells(cells: dict, ref: dict) -> None:
    assert np.array_equal(cells["cell_index"], ref["cell_index"])
    assert np.array_equal(cells["cfg"], ref["cfg"])
    assert np.array_equal(cells["lut2"], (ref["info"] >> 7) & 1)
    assert np.array_equal(cells["vertex_begin"], ref["vertex_begin"])


def check_against_oracle(ctx, field_t, iso: float, xyz_min=None, xyz_max=None, check_cells: bool = True):
    """Run the HIP path and the oracle on the same float bits; return the oracle result."""
    import torch

    verts, quads, cnt = ctx.extract(field_t, iso, xyz_min, xyz_max)
    field_np = field_t.cpu().numpy()
    ref = oracle.run(field_np, iso, xyz_min, xyz_max)
    assert ref["num_anomalies"] == 0
    assert cnt["num_active"] == ref["num_active"], (cnt, ref["num_active"])
    assert cnt["num_vertices"] == ref["num_vertices"]
    assert cnt["num_quads"] == ref["num_quads"]
    assert_same_mesh(verts.cpu().numpy(), quads.cpu().numpy(), ref)
    if check_cells and ref["num_active"]:
        assert_same_cells(ctx.cells(ref["num_active"]), ref)
    torch.cuda.synchronize()
    return ref
