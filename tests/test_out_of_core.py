"""Out-of-core extraction (z-slabs through the device): identical to the
in-core host entry at every slab size (GPU); slab sizing (CPU)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle


def _budget_for(fa, dims, S):
    """Smallest budget whose slabs hold at least S cell planes."""
    lo, hi = 1 << 12, 1 << 34
    while lo < hi:
        mid = (lo + hi) // 2
        try:
            got = fa.out_of_core_slab_planes(dims, mid)
        except fa.MDMCError:
            got = 0
        if got >= S:
            hi = mid
        else:
            lo = mid + 1
    return lo


def test_slab_sizing_grows_with_the_budget():
    import framework_amd as fa

    dims = (70, 60, 90)
    prev = 0
    for b in [1 << 20, 1 << 21, 1 << 22, 1 << 23, 1 << 24, 1 << 25]:
        S = fa.out_of_core_slab_planes(dims, b)
        assert prev <= S <= dims[2] - 1
        prev = S
    assert fa.out_of_core_slab_planes(dims, 1 << 40) == dims[2] - 1
    with pytest.raises(fa.MDMCError):
        fa.out_of_core_slab_planes(dims, 1 << 10)
    assert fa.out_of_core_slab_planes((5, 5, 1), 1 << 20) == 0
    b = _budget_for(fa, dims, 7)
    assert fa.out_of_core_slab_planes(dims, b) >= 7 and fa.out_of_core_slab_planes(dims, b - 1) < 7


@pytest.mark.gpu
def test_out_of_core_equals_in_core():
    import torch

    import framework_amd as fa

    torch.cuda.set_device(0)
    fa.setup()
    cases = [
        (oracle.field_gyroid((70, 60, 90), 11.0, 5, 0.3), 0.0, None),
        (oracle.field_blobs((64, 48, 100), 2, 12, sigma=(3.0, 12.0)), 0.1, None),
        (oracle.field_sphere((41, 45, 50), (20.5, 22.0, 24.9), 15.3), 0.0, ((-5.0, -2.5, 1.0), (5.0, 7.5, 3.0))),
    ]
    for f, iso, box in cases:
        Z, Y, X = f.shape
        mn, mx = box if box else (None, None)
        v0, q0 = fa.run_host(f, iso, mn, mx)
        for S in (1, 2, 5, 17, Z - 1):
            budget = _budget_for(fa, (X, Y, Z), S)
            assert fa.out_of_core_slab_planes((X, Y, Z), budget) >= S
            v, q = fa.run_out_of_core(f, iso, budget, mn, mx)
            assert np.array_equal(q, q0), (f.shape, S)
            assert np.array_equal(v.view(np.uint32), v0.view(np.uint32)), (f.shape, S)
    # against the oracle directly once
    f, iso, _ = cases[0]
    ref = oracle.run(f, iso)
    v, q = fa.run_out_of_core(f, iso, _budget_for(fa, f.shape[::-1], 3))
    assert np.array_equal(q, ref["quads"]) and np.array_equal(v.view(np.uint32), ref["vertices"].view(np.uint32))


@pytest.mark.gpu
def test_out_of_core_repeated_calls_reuse_buffers():
    """The slab buffers stay with the library between calls: the same call
    twice, a single-slab call (no second field slab), a larger grid, then the
    first call again, each identical to the in-core mesh."""
    import torch

    import framework_amd as fa

    torch.cuda.set_device(0)
    fa.setup()
    small = oracle.field_blobs((64, 48, 100), 2, 12, sigma=(3.0, 12.0))
    big = oracle.field_gyroid((90, 70, 120), 13.0, 7, 0.3)
    ref = {id(f): fa.run_host(f, 0.1) for f in (small, big)}
    for f, S in ((small, 4), (small, 4), (small, small.shape[0] - 1), (big, 6), (small, 4)):
        Z, Y, X = f.shape
        v, q = fa.run_out_of_core(f, 0.1, _budget_for(fa, (X, Y, Z), S))
        v0, q0 = ref[id(f)]
        assert np.array_equal(q, q0), (f.shape, S)
        assert np.array_equal(v.view(np.uint32), v0.view(np.uint32)), (f.shape, S)


@pytest.mark.gpu
def test_out_of_core_keeps_to_its_budget():
    """The out-of-core entry has its own device buffers: after a call (they
    stay allocated), the device memory it holds is at most the budget plus one
    slab's mesh, also when an in-core extraction of the same grid ran just
    before it and when a call with a larger budget came first."""
    import torch

    import framework_amd as fa

    torch.cuda.set_device(0)
    fa.setup()
    f = oracle.field_gyroid((96, 80, 128), 12.0, 3, 0.3)
    Z, Y, X = f.shape
    v0, q0 = fa.run_host(f, 0.0)
    mesh = 12 * len(v0) + 16 * len(q0)  # the whole mesh bounds one slab's mesh
    slack = 8 << 20                      # allocation granularity, counts record
    for S in (60, 7, 19):
        budget = _budget_for(fa, (X, Y, Z), S)
        fa.cleanup()
        fa.setup()
        fa.run_host(f, 0.0)              # the in-core entry's buffers exist ...
        free0 = torch.cuda.mem_get_info()[0]
        v, q = fa.run_out_of_core(f, 0.0, budget)
        held = free0 - torch.cuda.mem_get_info()[0]
        print(f"slab planes {S} budget {budget} held {held}", flush=True)
        assert held <= budget + mesh + slack, (S, budget, held)
        assert np.array_equal(q, q0) and np.array_equal(v.view(np.uint32), v0.view(np.uint32))
    # a smaller budget right after a larger one gives the larger buffers back
    big, small = _budget_for(fa, (X, Y, Z), 60), _budget_for(fa, (X, Y, Z), 5)
    fa.cleanup()
    fa.setup()
    free0 = torch.cuda.mem_get_info()[0]
    fa.run_out_of_core(f, 0.0, big)
    fa.run_out_of_core(f, 0.0, small)
    held = free0 - torch.cuda.mem_get_info()[0]
    assert held <= small + mesh + slack, (big, small, held)
