"""Smoothing (num_smooth > 0) on the GPU against the oracle, bitwise."""
from __future__ import annotations

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    torch.cuda.set_device(0)
    return torch


@pytest.fixture(scope="module")
def fa_mod(torch_cuda):
    import framework_amd as fa

    fa.setup()
    return fa


def _fields():
    return [
        ("sphere", oracle.field_sphere((48, 45, 50), (23.3, 21.9, 24.6), 17.2), 0.0, None),
        ("gyroid", oracle.field_gyroid((40, 36, 44), 9.0, 7, 0.3), 0.02, None),
        ("blobs", oracle.field_blobs((50, 44, 38), 3, 9, sigma=(3.0, 9.0)), 0.1, None),
        ("box", oracle.field_sphere((33, 30, 35), (16.0, 14.2, 17.9), 11.1), 0.0,
         ((-5.0, -2.5, 1.0), (5.0, 7.5, 3.0))),
        # rows longer than one x-tile: cells at the start of a tile find their
        # -x neighbours in the previous tile's segment
        ("wide_gyroid", oracle.field_gyroid((300, 20, 18), 7.0, 11, 0.3), 0.0, None),
        ("wide_blobs", oracle.field_blobs((520, 16, 14), 6, 32, sigma=(4.0, 10.0)), 0.1, None),
        ("wide_blobs2", oracle.field_blobs((520, 16, 14), 7, 28, sigma=(3.0, 12.0)), 0.05, None),
    ]


def _same(v, q, ref):
    assert np.array_equal(q, ref["quads"])
    assert np.array_equal(v.view(np.uint32), ref["vertices"].view(np.uint32))


@pytest.mark.parametrize("n", [1, 2, 3])
def test_device_smoothing_matches_oracle(fa_mod, torch_cuda, n):
    torch = torch_cuda
    ctx = fa_mod.Context()
    for name, f, iso, box in _fields():
        mn, mx = box if box else (None, None)
        ref = oracle.run(f, iso, mn, mx, num_smooth=n)
        v, q, c = ctx.extract(torch.from_numpy(f).cuda(), iso, mn, mx, num_smooth=n)
        assert c["status"] == 0
        _same(v.cpu().numpy(), q.cpu().numpy(), ref)
    ctx.close()


def test_smoothing_after_combined_extract_and_host_entry(fa_mod, torch_cuda):
    torch = torch_cuda
    name, f, iso, _ = _fields()[1]
    ref = oracle.run(f, iso, num_smooth=2)
    ctx = fa_mod.Context()
    ft = torch.from_numpy(f).cuda()
    verts = torch.empty((ref["num_vertices"], 3), device="cuda")
    quads = torch.empty((ref["num_quads"], 4), dtype=torch.int32, device="cuda")
    ctx.extract_into(ft, iso, verts, quads)
    ctx.smooth(verts, quads, 2)
    torch.cuda.synchronize()
    assert ctx.read_counts()["status"] == 0
    _same(verts.cpu().numpy(), quads.cpu().numpy(), ref)
    v, q = fa_mod.run_host(f, iso, num_smooth=2)
    _same(v, q, ref)
    ctx.close()


def test_smoothing_without_an_emit_is_rejected(fa_mod, torch_cuda):
    torch = torch_cuda
    ctx = fa_mod.Context()
    v = torch.zeros((4, 3), device="cuda")
    q = torch.zeros((1, 4), dtype=torch.int32, device="cuda")
    with pytest.raises(fa_mod.MDMCError):
        ctx.smooth(v, q, 1)
    ctx.close()
