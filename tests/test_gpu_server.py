"""The request loop (GPU): results alone and in a mix are identical, admission
by memory footprint, no request is lost under a tight budget."""
from __future__ import annotations

import numpy as np
import pytest

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


def _requests(n, seed, dmin=24, dmax=96):
    from framework_amd.server import make_request

    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        d = int(round(np.exp(rng.uniform(np.log(dmin), np.log(dmax)))))
        iso = float(rng.uniform(-0.2, 0.2))
        if i % 2 == 0:
            reqs.append(make_request((d, d, d), iso, "blobs", seed=seed * 7919 + i, nblobs=int(rng.integers(8, 33)),
                                     sigma=(2.0, 10.0), tag=i))
        else:
            reqs.append(make_request((d, d, d), iso, "gyroid", seed=seed * 7919 + i, period=float(rng.uniform(8, 16)),
                                     noise_amp=0.1, tag=i))
    return reqs


def _single(fa, r):
    """The same request through the single-volume entry points."""
    ctx = fa.Context()
    X, Y, Z = r.dims[0], r.dims[1], r.dims[2]
    if r.field_kind == 1:
        f = ctx.field_blobs((X, Y, Z), r.seed, r.nblobs, (r.sigma_lo, r.sigma_hi), (r.amp_lo, r.amp_hi))
    else:
        f = ctx.field_gyroid((X, Y, Z), r.period, r.seed, r.noise_amp)
    v, q, c = ctx.extract(f, r.iso, tuple(r.xyz_min), tuple(r.xyz_max))
    out = v.cpu().numpy(), q.cpu().numpy()
    ctx.close()
    return out


def test_result_alone_equals_result_in_mix(fa_mod, torch_cuda):
    from framework_amd.server import Server, run_stream

    reqs = _requests(40, seed=2)
    probe = [5, 17, 30]
    alone = {}
    srv = Server(2 << 30, max_batch=8, num_streams=2)
    for k in probe:
        res = run_stream(srv, [reqs[k]], np.zeros(1), keep={0})
        alone[k] = res["meshes"][0]
    srv.close()
    srv = Server(2 << 30, max_batch=8, num_streams=3)
    arrivals = np.cumsum(np.random.default_rng(0).exponential(2e-4, len(reqs)))
    res = run_stream(srv, reqs, arrivals, keep=set(probe))
    st = srv.stats()
    srv.close()
    assert (res["status"] == 0).all()
    assert st["batches"] >= 2 and st["completed"] == len(reqs)
    for k in probe:
        va, qa = alone[k]
        vm, qm = res["meshes"][k]
        vs, qs = _single(fa_mod, reqs[k])
        assert np.array_equal(qa, qm) and np.array_equal(qa, qs), k
        assert np.array_equal(va.view(np.uint32), vm.view(np.uint32)), k
        assert np.array_equal(va.view(np.uint32), vs.view(np.uint32)), k


@pytest.mark.parametrize("solo_cells", [None, 30 ** 3, 1])
def test_every_result_in_a_stream_equals_its_single_extraction(fa_mod, torch_cuda, solo_cells):
    """All results of a continuously batched stream (arrivals while batches
    run, uneven sizes, both field kinds) against the single-volume path; with
    the default routing (every request here is below the size that runs
    alone), with the larger requests running alone between batches, and with
    every request alone."""
    from framework_amd.server import OPT_SOLO_CELLS, Server, run_stream

    reqs = _requests(24, seed=9, dmin=16, dmax=72)
    srv = Server(1 << 30, max_batch=5, num_streams=3)
    if solo_cells is not None:
        srv.set_options({OPT_SOLO_CELLS: solo_cells})
    arrivals = np.cumsum(np.random.default_rng(1).exponential(1e-4, len(reqs)))
    res = run_stream(srv, reqs, arrivals, keep=set(range(len(reqs))))
    st = srv.stats()
    srv.close()
    assert (res["status"] == 0).all() and st["completed"] == len(reqs) and st["batches"] >= 3
    for k, r in enumerate(reqs):
        vs, qs = _single(fa_mod, r)
        vm, qm = res["meshes"][k]
        assert np.array_equal(qm, qs), k
        assert np.array_equal(vm.view(np.uint32), vs.view(np.uint32)), k


@pytest.mark.parametrize("solo_cells", [None, 40 ** 3])
def test_admission_under_a_tight_budget(fa_mod, torch_cuda, solo_cells):
    """Requests larger than the budget are rejected; everything else finishes
    with the peak memory within the budget (also with the larger requests
    running alone, whose scratch is the single-volume one)."""
    from framework_amd.server import OPT_SOLO_CELLS, Server, make_request, run_stream

    budget = 96 << 20
    reqs = _requests(30, seed=4, dmin=24, dmax=80)
    reqs.insert(7, make_request((400, 400, 400), 0.0, "gyroid", seed=1, tag=99))  # 256 MB field
    srv = Server(budget, max_batch=6, num_streams=2)
    if solo_cells is not None:
        srv.set_options({OPT_SOLO_CELLS: solo_cells})
    res = run_stream(srv, reqs, np.zeros(len(reqs)))
    st = srv.stats()
    srv.close()
    assert res["status"][7] == -8  # MDMC_ERR_REJECTED
    ok = np.delete(res["status"], 7)
    assert (ok == 0).all(), ok
    assert st["memory_peak"] <= budget
    assert st["completed"] == len(reqs)


def test_device_field_requests(fa_mod, torch_cuda):
    """Caller-owned fields are extracted in place."""
    from framework_amd.server import Server, make_request, run_stream

    ctx = fa_mod.Context()
    fields = [ctx.field_blobs((40 + 3 * i, 33, 29), seed=i, nblobs=9, sigma=(3.0, 9.0)) for i in range(5)]
    torch_cuda.cuda.synchronize()
    reqs = [make_request((f.shape[2], f.shape[1], f.shape[0]), 0.05, field=f, tag=i) for i, f in enumerate(fields)]
    srv = Server(512 << 20, max_batch=4, num_streams=2)
    res = run_stream(srv, reqs, np.zeros(len(reqs)), keep=set(range(5)))
    srv.close()
    for i, f in enumerate(fields):
        v, q, c = ctx.extract(f, 0.05)
        vm, qm = res["meshes"][i]
        assert np.array_equal(q.cpu().numpy(), qm)
        assert np.array_equal(v.cpu().numpy().view(np.uint32), vm.view(np.uint32))
    ctx.close()


def test_clients_on_several_threads(fa_mod, torch_cuda):
    """Four host threads submit while the caller's thread waits on poll and a
    fifth keeps changing launch options and the size that runs alone (options
    wait for a lane's launcher to finish with its context): the server's
    threads drive the lanes, every request comes back once, and the results
    equal their single extractions; a wait with nothing in flight returns at
    once."""
    import threading
    import time

    from framework_amd.mdmc import OPT_BATCH_PLANES, OPT_EMIT_BLOCKS_PER_CU
    from framework_amd.server import OPT_SOLO_CELLS, Server

    reqs = _requests(32, seed=13, dmin=16, dmax=64)
    srv = Server(1 << 30, max_batch=4, num_streams=3)
    ids = [None] * len(reqs)
    stop = threading.Event()

    def client(k0):
        for k in range(k0, len(reqs), 4):
            ids[k] = srv.submit(reqs[k])

    def tuner():
        i = 0
        while not stop.is_set():
            srv.set_options({OPT_EMIT_BLOCKS_PER_CU: (0, 4, 16)[i % 3], OPT_BATCH_PLANES: (0, 1, 2)[i % 3],
                             OPT_SOLO_CELLS: (0, 30 ** 3)[i % 2]})
            i += 1
            time.sleep(2e-4)

    threads = [threading.Thread(target=client, args=(t,)) for t in range(4)]
    opt_thread = threading.Thread(target=tuner)
    opt_thread.start()
    for t in threads:
        t.start()
    got, meshes = {}, {}
    while len(got) < len(reqs):
        for r in srv.poll(wait=True):
            assert r["id"] not in got
            got[r["id"]] = r
            if r["status"] == 0 and r["tag"] % 5 == 0:
                meshes[r["tag"]] = srv.copy_result(r)
            srv.release(r["id"])
    for t in threads:
        t.join()
    stop.set()
    opt_thread.join()
    assert sorted(got) == sorted(ids)
    assert all(r["status"] == 0 for r in got.values())
    assert srv.poll(wait=True) == []
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs) and st["queued"] == 0 and st["in_flight"] == 0
    for k, (vm, qm) in meshes.items():
        vs, qs = _single(fa_mod, reqs[k])
        assert np.array_equal(qm, qs), k
        assert np.array_equal(vm.view(np.uint32), vs.view(np.uint32)), k


def _regenerate_cpu(r):
    import oracle

    dims = (r.dims[0], r.dims[1], r.dims[2])
    if r.field_kind == 1:
        return oracle.field_blobs(dims, r.seed, r.nblobs, (r.sigma_lo, r.sigma_hi), (r.amp_lo, r.amp_hi))
    return oracle.field_gyroid(dims, r.period, r.seed, r.noise_amp)


def test_served_workload_stream(fa_mod, torch_cuda):
    """The served workload's distribution at size: 240 requests of the
    benchmark's stream (seed 5: cubic dims log-uniform in [48, 384], blob or
    gyroid fields from per-request seeds, isovalues U[-0.2, 0.2]) as a Poisson
    stream through a 64 GiB server.  Every result equals its single
    extraction bitwise (compared on the device); 12 results, the smallest and
    a spread of sizes up to the largest, equal the oracle on fields
    regenerated on the CPU; none is rejected; the peak stays in the budget."""
    import bench
    import oracle
    from framework_amd.server import Server, run_stream
    from tests.helpers import assert_same_mesh

    torch = torch_cuda
    reqs = bench._served_requests(240, seed=5)
    d = np.array([r.dims[0] for r in reqs])
    order = np.argsort(d)
    to_oracle = set(order[:8].tolist()) | set(order[[80, 160, 200, len(reqs) - 1]].tolist())
    assert d.min() >= 48 and d.max() <= 384 and d.max() > 300
    budget = 64 << 30
    srv = Server(budget, max_batch=32, num_streams=8)
    gen, sctx = fa_mod.Context(), fa_mod.Context()
    oracle.set_threads(8)
    seen = {"single": 0, "oracle": 0}

    def check(k, res):
        assert res["status"] == 0, (k, res["status"])
        r = reqs[k]
        X, Y, Z = r.dims[0], r.dims[1], r.dims[2]
        if r.field_kind == 1:
            f = gen.field_blobs((X, Y, Z), r.seed, r.nblobs, (r.sigma_lo, r.sigma_hi), (r.amp_lo, r.amp_hi))
        else:
            f = gen.field_gyroid((X, Y, Z), r.period, r.seed, r.noise_amp)
        v, q, c = sctx.extract(f, r.iso, tuple(r.xyz_min), tuple(r.xyz_max))
        sv, sq = srv.result_tensors(res)
        assert (res["num_active"], res["num_vertices"], res["num_quads"]) == \
            (c["num_active"], c["num_vertices"], c["num_quads"]), k
        assert torch.equal(sq, q) and torch.equal(sv.view(torch.int32), v.view(torch.int32)), k
        seen["single"] += 1
        if seen["single"] % 40 == 0:
            print(f"  {seen['single']} results checked", flush=True)
        if k in to_oracle:
            fnp = _regenerate_cpu(r)
            assert np.array_equal(fnp.view(np.uint32), f.cpu().numpy().view(np.uint32)), k
            ref = oracle.run(fnp, r.iso, tuple(r.xyz_min), tuple(r.xyz_max))
            assert_same_mesh(sv.cpu().numpy(), sq.cpu().numpy(), ref)
            seen["oracle"] += 1

    arrivals = np.cumsum(np.random.default_rng(6).exponential(1.0 / 4000.0, len(reqs)))
    res = run_stream(srv, reqs, arrivals, on_result=check)
    st = srv.stats()
    srv.close()
    oracle.set_threads(1)
    gen.close()
    sctx.close()
    print(f"served {len(reqs)}: batches {st['batches']} peak {st['memory_peak'] / 1e9:.2f} GB", flush=True)
    assert (res["status"] == 0).all() and st["rejected"] == 0
    assert seen == {"single": len(reqs), "oracle": len(to_oracle)}
    assert 0 < st["memory_peak"] <= budget and st["completed"] == len(reqs)
