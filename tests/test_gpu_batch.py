"""Batched extraction (many volumes in one launch sequence) against the
single-volume path and the oracle (GPU).

Each volume of a batch must give exactly the mesh the single-volume entry
gives for it alone: same counts, quads and vertex bits.  The single-volume
path is itself checked against the oracle, and the small volumes here are
also compared with the oracle directly.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
from tests.helpers import assert_same_mesh

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


def _mixed_batch(fa, torch, n, seed, max_dim=96):
    """Volumes of uneven dims (log-uniform per axis), alternating blob sums and
    gyroid+noise, isovalues in [-0.2, 0.2]."""
    rng = np.random.default_rng(seed)
    gen = fa.Context()
    fields, isos = [], []
    for i in range(n):
        dims = tuple(int(round(np.exp(rng.uniform(np.log(8), np.log(max_dim))))) for _ in range(3))
        if i % 2 == 0:
            f = gen.field_blobs(dims, seed=seed * 1000 + i, nblobs=int(rng.integers(8, 33)), sigma=(2.0, 10.0))
        else:
            f = gen.field_gyroid(dims, period=float(rng.uniform(6, 14)), seed=seed * 1000 + i, noise_amp=0.3)
        fields.append(f)
        isos.append(float(rng.uniform(-0.2, 0.2)))
    torch.cuda.synchronize()
    gen.close()
    return fields, isos


def _single(ctx, f, iso):
    v, q, c = ctx.extract(f, iso)
    return v.cpu().numpy(), q.cpu().numpy(), c


def test_batch_matches_single_volume(fa_mod, torch_cuda):
    fields, isos = _mixed_batch(fa_mod, torch_cuda, 24, seed=3)
    bctx, sctx = fa_mod.Context(), fa_mod.Context()
    res = bctx.batch_extract(fields, isos)
    for i, (f, iso) in enumerate(zip(fields, isos)):
        v, q, c = _single(sctx, f, iso)
        bv, bq, bc = res[i]
        assert (bc["num_active"], bc["num_vertices"], bc["num_quads"]) == \
            (c["num_active"], c["num_vertices"], c["num_quads"]), i
        assert np.array_equal(bq.cpu().numpy(), q), i
        assert np.array_equal(bv.cpu().numpy().view(np.uint32), v.view(np.uint32)), i
    bctx.close()
    sctx.close()


def test_batch_with_volumes_spanning_several_x_tiles(fa_mod, torch_cuda):
    """Narrow volumes next to wide ones: the batch takes 256-cell x-tiles and
    the wide volumes span 2-3 of them, while single extractions of the same
    fields use one 512- or 1024-cell x-tile.  Results must not differ."""
    gen = fa_mod.Context()
    dims_list = [(600, 20, 15), (40, 40, 40), (300, 30, 20), (520, 9, 9), (257, 12, 11), (33, 60, 7)]
    fields = [gen.field_gyroid(d, period=9.0, seed=70 + i, noise_amp=0.3) if i % 2 else
              gen.field_blobs(d, seed=80 + i, nblobs=10, sigma=(2.0, 8.0)) for i, d in enumerate(dims_list)]
    torch_cuda.cuda.synchronize()
    isos = [0.05, -0.1, 0.0, 0.1, 0.02, -0.05]
    bctx, sctx = fa_mod.Context(), fa_mod.Context()
    res = bctx.batch_extract(fields, isos)
    for i, (f, iso) in enumerate(zip(fields, isos)):
        v, q, c = _single(sctx, f, iso)
        bv, bq, bc = res[i]
        assert (bc["num_active"], bc["num_vertices"], bc["num_quads"]) == \
            (c["num_active"], c["num_vertices"], c["num_quads"]), dims_list[i]
        assert c["num_quads"] > 0, dims_list[i]
        assert np.array_equal(bq.cpu().numpy(), q), dims_list[i]
        assert np.array_equal(bv.cpu().numpy().view(np.uint32), v.view(np.uint32)), dims_list[i]
    bctx.close()
    sctx.close()
    gen.close()


def test_batch_against_oracle(fa_mod, torch_cuda):
    fields, isos = _mixed_batch(fa_mod, torch_cuda, 6, seed=11, max_dim=48)
    ctx = fa_mod.Context()
    res = ctx.batch_extract(fields, isos)
    for i, (f, iso) in enumerate(zip(fields, isos)):
        ref = oracle.run(f.cpu().numpy(), iso)
        v, q, c = res[i]
        assert c["num_vertices"] == ref["num_vertices"] and c["num_quads"] == ref["num_quads"]
        assert_same_mesh(v.cpu().numpy(), q.cpu().numpy(), ref)
    ctx.close()


def test_batch_with_empty_and_flat_volumes(fa_mod, torch_cuda):
    """Degenerate dims and volumes without a surface inside a batch."""
    torch = torch_cuda
    ctx = fa_mod.Context()
    a = ctx.field_blobs((40, 33, 27), seed=1, nblobs=10, sigma=(3.0, 8.0))
    flat = torch.full((20, 21, 22), 5.0, device="cuda")
    tiny = torch.zeros((1, 7, 7), device="cuda")
    b = ctx.field_gyroid((37, 41, 29), period=9.0, seed=2, noise_amp=0.2)
    torch.cuda.synchronize()
    fields, isos = [tiny, a, flat, tiny, b, tiny], [0.0, 0.1, 0.0, 0.0, 0.0, 0.0]
    res = ctx.batch_extract(fields, isos)
    sctx = fa_mod.Context()
    for i, (f, iso) in enumerate(zip(fields, isos)):
        v, q, c = _single(sctx, f, iso)
        assert res[i][2]["num_vertices"] == c["num_vertices"] and res[i][2]["num_quads"] == c["num_quads"]
        assert np.array_equal(res[i][1].cpu().numpy(), q)
        assert np.array_equal(res[i][0].cpu().numpy().view(np.uint32), v.view(np.uint32))
    assert res[2][2]["num_vertices"] == 0 and res[0][2]["num_vertices"] == 0
    # all volumes empty
    res = ctx.batch_extract([tiny, flat], [0.0, 0.0])
    assert all(r[2]["num_vertices"] == 0 for r in res)
    assert all(r[0].shape == (0, 3) and r[1].shape == (0, 4) for r in res)
    # no volumes at all
    assert ctx.batch_extract([], []) == []
    ctx.close()
    sctx.close()


def test_batch_capacity_overflow_is_flagged_per_volume(fa_mod, torch_cuda):
    torch = torch_cuda
    fields, isos = _mixed_batch(fa_mod, torch, 4, seed=5, max_dim=40)
    ctx = fa_mod.Context()
    exact = ctx.batch_extract(fields, isos)
    outs = []
    for i, (v, q, c) in enumerate(exact):
        nv = c["num_vertices"] - (3 if i == 1 else 0)
        outs.append((torch.empty((max(nv, 0), 3), device="cuda"), torch.empty((c["num_quads"], 4), dtype=torch.int32,
                                                                              device="cuda")))
    counts = ctx.batch_counts_tensor(len(fields))
    ctx.batch_extract_into(fields, isos, outs, counts)
    torch.cuda.synchronize()
    st = counts.cpu().numpy().view(np.uint32)[:, 3]
    assert st[1] & 1, st
    assert st[0] == 0 and st[2] == 0 and st[3] == 0, st
    ctx.close()


def test_batch_graph_replay(fa_mod, torch_cuda):
    """The batch sequence captured in a HIP graph, replayed several times with
    eager work in between; every replay equals the eager result."""
    torch = torch_cuda
    fields, isos = _mixed_batch(fa_mod, torch, 8, seed=7, max_dim=64)
    ctx = fa_mod.Context()
    exact = ctx.batch_extract(fields, isos)
    outs = [(torch.empty_like(v), torch.empty_like(q)) for v, q, _ in exact]
    counts = ctx.batch_counts_tensor(len(fields))
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        ctx.batch_extract_into(fields, isos, outs, counts, stream=stream)  # sizes the buffers
    stream.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        ctx.batch_extract_into(fields, isos, outs, counts, stream=torch.cuda.current_stream())
    for rep in range(3):
        for v, q in outs:
            v.fill_(-1.0)
            q.fill_(-1)
        counts.fill_(-1)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        c = counts.cpu().numpy().view(np.uint32)
        for i, (v, q, cnt) in enumerate(exact):
            assert c[i, 1] == cnt["num_vertices"] and c[i, 2] == cnt["num_quads"] and c[i, 3] == 0, (rep, i)
            assert torch.equal(outs[i][1], q), (rep, i)
            assert torch.equal(outs[i][0].view(torch.int32), v.view(torch.int32)), (rep, i)
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays
    # the captured context's buffers belong to the graph: eager calls on it
    # are refused (they would rewrite its staging tables and scratch), and the
    # graph still replays correctly afterwards
    with pytest.raises(fa_mod.MDMCError, match="captured graph"):
        ctx.batch_extract_into(fields, isos, outs, counts)
    with pytest.raises(fa_mod.MDMCError, match="captured graph"):
        ctx.count(fields[0], isos[0])
    g.replay()
    torch.cuda.synchronize()
    c = counts.cpu().numpy().view(np.uint32)
    for i, (v, q, cnt) in enumerate(exact):
        assert c[i, 1] == cnt["num_vertices"] and c[i, 3] == 0, i
        assert torch.equal(outs[i][1], q), i
    ctx.close()


def test_zmarch_and_plane_tile_kernels_agree(fa_mod, torch_cuda):
    """The two classify kernels produce identical meshes (row lengths with and
    without 16-byte row loads)."""
    torch = torch_cuda
    ctx = fa_mod.Context()
    for dims in [(131, 40, 30), (132, 40, 30), (300, 20, 17), (700, 9, 9)]:
        f = ctx.field_gyroid(dims, period=9.0, seed=4, noise_amp=0.5)
        torch.cuda.synchronize()
        v0, q0, _ = ctx.extract(f, 0.02)
        os.environ["MDMC_CLASSIFY"] = "zmarch"
        try:
            v1, q1, _ = ctx.extract(f, 0.02)
        finally:
            del os.environ["MDMC_CLASSIFY"]
        assert torch.equal(q0, q1), dims
        assert torch.equal(v0.view(torch.int32), v1.view(torch.int32)), dims
    ctx.close()


def test_batch_generator_matches_single_generators(fa_mod, torch_cuda):
    torch = torch_cuda
    ctx = fa_mod.Context()
    specs = [{"dims": (43, 43, 43), "kind": "gyroid", "seed": 15843, "period": 15.7, "noise_amp": 0.1},
             {"dims": (32, 30, 31), "kind": "blobs", "seed": 9, "nblobs": 12, "sigma": (2.0, 10.0)},
             {"dims": (7, 90, 5), "kind": "gyroid", "seed": 1, "period": 6.0, "noise_amp": 0.5}]
    for rep in range(3):  # staging reuse across calls
        outs = ctx.batch_generate(specs[rep:] + specs[:rep])
        torch.cuda.synchronize()
        for sp, o in zip(specs[rep:] + specs[:rep], outs):
            if sp["kind"] == "blobs":
                ref = ctx.field_blobs(sp["dims"], sp["seed"], sp["nblobs"], sp["sigma"])
            else:
                ref = ctx.field_gyroid(sp["dims"], sp["period"], sp["seed"], sp["noise_amp"])
            torch.cuda.synchronize()
            assert torch.equal(o.view(torch.int32), ref.view(torch.int32)), (rep, sp)
    ctx.close()


def _regenerate_cpu(spec):
    """The CPU generator's field for a batch spec (same definition, same bits)."""
    if spec["kind"] == "blobs":
        return oracle.field_blobs(spec["dims"], spec["seed"], spec["nblobs"], spec.get("sigma", (8.0, 32.0)),
                                  spec.get("amp", (-1.0, 1.0)))
    return oracle.field_gyroid(spec["dims"], spec.get("period", 24.0), spec["seed"], spec.get("noise_amp", 0.05))


def test_batched_workload_at_size(fa_mod, torch_cuda):
    """The batched workload itself (256 volumes, axes log-uniform in
    [32, 256], blob sums and gyroid+noise alternating, isovalues U[-0.2, 0.2];
    seed 1 as in the benchmark): every volume of the eager batch equals its
    single extraction bitwise and the closed-form active-cell and quad counts
    of its field; three HIP-graph replays equal the eager batch; 8 volumes
    equal the oracle on fields regenerated on the CPU."""
    import bench
    from tests.test_gpu_parity import _closed_form_counts

    torch = torch_cuda
    specs, isos = bench.batched_workload(256, seed=1)
    ctx, sctx = fa_mod.Context(), fa_mod.Context()
    fields = ctx.batch_generate(specs)
    torch.cuda.synchronize()
    res = ctx.batch_extract(fields, isos)
    for i, (f, iso) in enumerate(zip(fields, isos)):
        v, q, c = sctx.extract(f, iso)
        bv, bq, bc = res[i]
        assert (bc["num_active"], bc["num_vertices"], bc["num_quads"]) == \
            (c["num_active"], c["num_vertices"], c["num_quads"]), i
        assert torch.equal(bq, q) and torch.equal(bv.view(torch.int32), v.view(torch.int32)), i
        A, Q = _closed_form_counts(torch, f, iso)
        assert (c["num_active"], c["num_quads"]) == (A, Q), (i, specs[i]["dims"])
    print(f"batch of {len(fields)}: {sum(r[2]['num_vertices'] for r in res)} vertices", flush=True)
    # graph replays against the eager batch
    outs = [(torch.empty_like(v), torch.empty_like(q)) for v, q, _ in res]
    counts = ctx.batch_counts_tensor(len(fields))
    gctx = fa_mod.Context()  # a context of its own for the graph
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        gctx.batch_extract_into(fields, isos, outs, counts, stream=stream)
    stream.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        gctx.batch_extract_into(fields, isos, outs, counts, stream=torch.cuda.current_stream())
    for rep in range(3):
        for v, q in outs:
            v.fill_(-1.0)
            q.fill_(-1)
        counts.fill_(-1)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        c = counts.cpu().numpy().view(np.uint32)
        assert (c[:, 3] == 0).all(), rep
        for i, (v, q, cnt) in enumerate(res):
            assert c[i, 1] == cnt["num_vertices"] and c[i, 2] == cnt["num_quads"], (rep, i)
            assert torch.equal(outs[i][1], q) and torch.equal(outs[i][0].view(torch.int32), v.view(torch.int32)), \
                (rep, i)
    # the oracle on 8 volumes, both field kinds, small and large
    sizes = np.array([np.prod(s["dims"]) for s in specs])
    pick = list(np.argsort(sizes)[[0, 1, 2, 3, 128, 129, 254, 255]])
    for i in pick:
        fnp = _regenerate_cpu(specs[i])
        assert np.array_equal(fnp.view(np.uint32), fields[i].cpu().numpy().view(np.uint32)), i
        ref = oracle.run(fnp, isos[i])
        v, q, _ = res[i]
        assert_same_mesh(v.cpu().numpy(), q.cpu().numpy(), ref)
    del g
    ctx.close()
    sctx.close()
    gctx.close()
