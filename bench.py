#!/usr/bin/env python3
"""Benchmark: plain Dual Marching Cubes on one 512^3 FP32 field.

Workload (one "step" = one full extraction, field resident in HBM):
  512^3 points, sum of 32 Gaussian blobs (SplitMix64 seed 0; centres uniform in
  the grid, sigma U[8,32] voxels, amplitude U[-1,1]), isovalue 0.1; output is
  the reference vertex layout (iso vertices + owned edge points, FP32 xyz) and
  int32x4 quads.  Each step runs classify -> scan -> emit on one stream.

Prints ONE JSON line on stdout.  `value` is Gvoxels/s (field samples per
second) over the timed steps, which go round-robin to two contexts on two
streams (--streams), so that one extraction's short scan and emit overlap
the next one's classify; the one-stream time per extraction is reported
beside it (config.ms_per_extraction_one_stream).  `roofline` is for the dominant kernel (longest average launch),
timed live with HIP events the library records around each launch on its
stream.  `cpu_baseline` is the C oracle on the same field, on up to 16 threads
(and on one thread, `cpu_baseline.single_thread`).

Four more legs go into the same line:
  "host_path": the drop-in host entry (mdmc_run, replacing run_dmc) on the
      same 512^3 field, host field in and host mesh out, so PCIe transfers
      are included (never the headline value).
  "dense": the surface-dense case, 1024^3 gyroid (period 24) + noise 0.05,
      iso 0: ms per extraction, Gvoxels/s and the pipeline's algorithmic
      bytes (field + vertices + quads) against the HBM peak.
  "batched": 256 volumes, dims per axis log-uniform in [32, 256], alternating
      blob sums (8-32 blobs) and gyroid+noise, isovalues U[-0.2, 0.2]; the
      plain per-volume loop (count, read sizes, allocate, emit, per volume)
      against one batched launch sequence and its HIP-graph replay.
  "served": a Poisson stream of requests (cubic dims log-uniform in
      [48, 384], blob or gyroid fields generated from per-request seeds,
      isovalues U[-0.2, 0.2]) through the request loop (8 lanes, batches of
      up to 32) under a 64 GB budget at 0.8x its measured capacity:
      volumes/s, p50/p99 latency, and the plain per-volume loop on the same
      requests.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
DIMS = (512, 512, 512)
SEED, NBLOBS, ISO = 0, 32, 0.1


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--streams", type=int, default=2,
                    help="contexts/streams of the throughput figure beside the headline, whose steps go "
                         "strictly one after another on one stream")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget for the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="default launch shapes (no autotuner sweep)")
    ap.add_argument("--options", default="",
                    help="fixed single-volume launch options instead of the autotuner, e.g. "
                         "'pt_planes=3,emit_blocks_per_cu=16' (names as in framework_amd.autotune.OPTION_NAMES)")
    ap.add_argument("--skip-batched", action="store_true")
    ap.add_argument("--skip-served", action="store_true")
    ap.add_argument("--skip-dense", action="store_true")
    ap.add_argument("--skip-host", action="store_true")
    ap.add_argument("--dense-only", action="store_true", help="only the surface-dense leg (prints its JSON)")
    ap.add_argument("--served-requests", type=int, default=10000)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 512 MiB device-to-device copy --steps times (a kernel of known "
                         "HBM traffic, used to scale rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the extraction path has no multi-GPU form here)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import framework_amd as fa

    torch.cuda.set_device(0)
    if args.copy_probe:
        src = torch.ones(DIMS, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {src.numel() * 4} bytes x {args.steps}")
        return 0
    fa.setup()
    if args.dense_only:
        print(json.dumps(bench_dense(fa, torch, steps=args.steps)), flush=True)
        return 0
    ctx = fa.Context()
    N = DIMS[0] * DIMS[1] * DIMS[2]
    log("generating field")
    field = ctx.field_blobs(DIMS, seed=SEED, nblobs=NBLOBS, sigma=(8.0, 32.0), amp=(-1.0, 1.0))
    torch.cuda.synchronize()
    ctx.count(field, ISO)
    torch.cuda.synchronize()
    c = ctx.read_counts()
    A, V, Q = c["num_active"], c["num_vertices"], c["num_quads"]
    log(f"active={A} vertices={V} quads={Q}")
    verts = torch.empty((V, 3), dtype=torch.float32, device="cuda")
    quads = torch.empty((Q, 4), dtype=torch.int32, device="cuda")
    counts = torch.zeros(4, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    tuned = None
    from framework_amd import autotune

    if args.options:
        by_name = {v: k for k, v in autotune.OPTION_NAMES.items()}
        opts = {by_name[kv.split("=")[0].strip()]: int(kv.split("=")[1]) for kv in args.options.split(",")}
        autotune.apply(ctx, opts)
        tuned = {"options": opts}
        log(f"launch options {opts} (fixed)")
    elif not args.no_tune:
        tuned = autotune.tune_single(ctx, field, ISO)
        log(f"launch options {tuned['options']} ({tuned['ms']:.4f} ms)")

    for _ in range(args.warmup):
        ctx.extract_into(field, ISO, verts, quads, counts=counts)
    torch.cuda.synchronize()
    st = counts.cpu().numpy().view(np.uint32)
    if st[3] != 0 or st[1] != V or st[2] != Q:
        log(f"bad status after warmup: {st}")
        return 1

    # Timed steps (no per-kernel events inside the timed region).  Steps go
    # round-robin to `streams` contexts on their own streams, as a service
    # extracting one volume after another would run them: one extraction's
    # scan and emit (short, latency-bound launches) overlap the next one's
    # classify.  Every step is a full extraction into its own outputs.
    nst = max(1, args.streams)
    lanes = [(ctx, stream, verts, quads, counts)]
    for _ in range(nst - 1):
        c2 = fa.Context()
        if tuned is not None:
            autotune.apply(c2, tuned["options"])
        lanes.append((c2, torch.cuda.Stream(), torch.empty_like(verts), torch.empty_like(quads),
                      torch.zeros_like(counts)))

    def timed(n_lanes: int) -> float:
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0.record(stream)
        for ln in lanes[1:n_lanes]:
            ln[1].wait_stream(stream)
        for i in range(args.steps):
            cx, st, vv, qq, cc = lanes[i % n_lanes]
            cx.extract_into(field, ISO, vv, qq, counts=cc, stream=st)
        for ln in lanes[1:n_lanes]:
            stream.wait_stream(ln[1])
        t1.record(stream)
        torch.cuda.synchronize()
        return t0.elapsed_time(t1) / args.steps

    for i in range(2 * nst):  # warm every lane
        cx, st, vv, qq, cc = lanes[i % nst]
        cx.extract_into(field, ISO, vv, qq, counts=cc, stream=st)
    torch.cuda.synchronize()
    # the headline: extractions strictly one after another on one stream;
    # beside it the throughput of `streams` streams taking steps round-robin
    ms_step = timed(1)
    ms_multi = timed(nst) if nst > 1 else ms_step
    for cx, st, vv, qq, cc in lanes:
        stc = cc.cpu().numpy().view(np.uint32)
        if stc[3] != 0 or stc[1] != V or stc[2] != Q:
            log(f"bad status after timed steps: {stc}")
            return 1
    for cx, *_ in lanes[1:]:
        cx.close()
    # the same steps again with HIP events around every kernel launch (on the
    # launching stream) for the per-kernel averages the roofline uses
    ctx.profile(True)
    ctx.profile_read()
    for _ in range(args.steps):
        ctx.extract_into(field, ISO, verts, quads, counts=counts)
    torch.cuda.synchronize()
    prof = ctx.profile_read()
    ctx.profile(False)
    log(f"ms/step={ms_step:.4f} (one stream; {nst} streams {ms_multi:.4f}) kernels={prof}")

    algo = {"classify": 4.0 * N, "scan": 0.0, "emit": 12.0 * V + 16.0 * Q}
    per_kernel = {k: (ms / n if n else 0.0) for k, (ms, n) in prof.items() if k != "other"}
    dom = max(per_kernel, key=per_kernel.get)
    dom_ms = per_kernel[dom]
    achieved = algo[dom] / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else 0.0
    B = 4.0 * N + 12.0 * V + 16.0 * Q
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "dmc_blobs512_hbm_traffic.json")
    if os.path.exists(tfile):
        # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
        # (scaled by a copy of known size) on this workload
        with open(tfile) as fh:
            tk = json.load(fh)["kernels"].get(dom)
        if tk:
            traffic = tk["read_bytes"] + tk["write_bytes"]
    pipe_gbs = B / (ms_step * 1e-3) / 1e9

    cpu = None
    fp64_error = None
    if not args.no_cpu:
        import oracle

        fnp = field.cpu().numpy()

        def time_oracle(threads: int, budget: float) -> tuple:
            oracle.set_threads(threads)
            times = []
            tstart = time.perf_counter()
            while True:
                t = time.perf_counter()
                r = oracle.run(fnp, ISO)
                times.append(time.perf_counter() - t)
                if time.perf_counter() - tstart > budget or len(times) >= 5:
                    break
            assert r["num_quads"] == Q and r["num_vertices"] == V
            return sorted(times)[len(times) // 2], len(times)

        # the per-cell loops on every core of this process's CPU share: the
        # affinity set, within the OpenMP thread limit the host sets for the
        # process (the GPU's share of a shared host), and on one thread
        host = cpu_host_info()
        nthr = host["threads_used"]
        log(f"CPU baseline (oracle, {nthr} threads and 1 thread) on {host['model']}")
        tmul, nmul = time_oracle(nthr, args.cpu_seconds / 2)
        tone, none = time_oracle(1, args.cpu_seconds / 2)
        oracle.set_threads(1)
        cpu = {"value": N / tmul / 1e9, "unit": "Gvoxels/s", "cores": nthr, "kind": "port",
               "sample": f"full 512^3 blob field, {nmul} run(s), median {tmul:.3f} s (C oracle restating "
                         f"dmc::run_dmc, OpenMP over cell planes and active cells, {nthr} threads, on the host "
                         f"of the GPU that ran the steps)",
               "host": host,
               "single_thread": {"value": N / tone / 1e9, "cores": 1,
                                 "sample": f"same field, {none} run(s), median {tone:.3f} s"}}
        # vertex error of the FP32 path against the oracle's FP64 twin (the
        # same algorithm with every coordinate and interpolation in double)
        oracle.set_threads(nthr)
        r64 = oracle.run(fnp, ISO, fp64=True)
        oracle.set_threads(1)
        vg = verts.cpu().numpy().astype(np.float64)
        fin = np.isfinite(vg) & np.isfinite(r64["vertices64"])
        err = np.abs(vg - r64["vertices64"])[fin]
        same = np.array_equal(verts.cpu().numpy().view(np.uint32), r64["vertices"].view(np.uint32))
        fp64_error = {"max_voxel_units": float(err.max()), "mean_voxel_units": float(err.mean()),
                      "vertices": int(len(vg)), "fp32_bitwise_equal_to_oracle": bool(same),
                      "workload": "512^3 blobs (the headline field), voxel units (box 0..511)"}
        log(f"FP64 twin: {fp64_error}")

    # smoothing on the headline mesh (after the FP64 check, which reads the
    # unsmoothed vertices)
    smooth = bench_smooth(ctx, torch, field, ISO, verts, quads, counts, V)
    smooth["workload"] = "512^3 blobs (the headline field)"
    log(f"smoothing: {smooth}")
    host = None if args.skip_host else bench_host(fa, field.cpu().numpy(), V, Q)
    dense = None if args.skip_dense else bench_dense(fa, torch)
    batched = None if args.skip_batched else bench_batched(fa, torch, tune=not args.no_tune)
    served = None if args.skip_served else bench_served(
        fa, torch, args.served_requests, options=(batched or {}).get("launch_option_ids"))

    out = {
        "metric": "Gvoxels/s + %HBM peak at 512^3; served volumes/s with p50/p99 latency",
        "value": N / (ms_step * 1e-3) / 1e9,
        "unit": "Gvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain DMC, 512^3 FP32 sum of 32 Gaussian blobs (seed 0, sigma U[8,32], "
                               "amp U[-1,1]), iso 0.1, field resident in HBM",
                   "dims": list(DIMS), "iso": ISO, "active_cells": A, "vertices": V, "quads": Q,
                   "streams": 1, "ms_per_extraction_one_stream": ms_step,
                   "throughput_streams": {"streams": nst, "ms_per_extraction": ms_multi,
                                          "gvoxels_per_s": N / (ms_multi * 1e-3) / 1e9},
                   "launch_options": None if tuned is None else
                   {autotune.OPTION_NAMES[k]: v for k, v in tuned["options"].items()}},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "avg_launch_ms": dom_ms, "algorithmic_bytes": algo[dom]},
        "pipeline": {"algorithmic_bytes": B, "achieved_gbs": pipe_gbs, "frac": pipe_gbs / HBM_PEAK_GBS,
                     "kernel_avg_ms": per_kernel},
        "cpu_baseline": cpu,
        "vertex_error_vs_fp64": fp64_error,
        "smoothing": smooth,
        "host_path": host,
        "dense": dense,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


def cpu_host_info() -> dict:
    """The host CPU: model, hardware threads, this process's affinity set and
    the threads the CPU baseline uses (the affinity set, capped by
    OMP_NUM_THREADS when the host sets it for the process)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    used = affinity
    if omp.isdigit() and int(omp) > 0:
        used = min(used, int(omp))
    return {"model": model, "hardware_concurrency": os.cpu_count(), "affinity_cpus": affinity,
            "omp_num_threads": omp or None, "threads_used": max(1, used)}


def bench_host(fa, fnp, V: int, Q: int, reps: int = 5) -> dict:
    """The drop-in host entry (mdmc_run, the run_dmc replacement) on the same
    512^3 field: host field in, host mesh out, PCIe transfers included."""
    import numpy as np

    fa.run_host(fnp, ISO)
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        v, q = fa.run_host(fnp, ISO)
        times.append(time.perf_counter() - t)
    assert len(v) == V and len(q) == Q
    ms = float(np.median(times)) * 1e3
    N = fnp.size
    # the same through the out-of-core entry with a 256 MiB device budget
    budget = 256 << 20
    S = fa.out_of_core_slab_planes(fnp.shape[::-1], budget)
    fa.run_out_of_core(fnp, ISO, budget)
    t_ooc = []
    for _ in range(3):
        t = time.perf_counter()
        vo, qo = fa.run_out_of_core(fnp, ISO, budget)
        t_ooc.append(time.perf_counter() - t)
    assert np.array_equal(qo, q) and np.array_equal(vo.view(np.uint32), v.view(np.uint32))
    ms_ooc = float(np.median(t_ooc)) * 1e3
    # the OBJ writer (mdmc_write_obj, dump_obj's format) on that mesh
    import tempfile

    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "mesh.obj")
        fa.write_obj(path, v, q)
        t_obj = []
        for _ in range(3):
            t = time.perf_counter()
            fa.write_obj(path, v, q)
            t_obj.append(time.perf_counter() - t)
        obj_bytes = os.path.getsize(path)
    ms_obj = float(np.median(t_obj)) * 1e3
    out = {"entry": "mdmc_run (host field in, host mesh out)", "ms_per_extraction": ms,
           "gvoxels_per_s": N / (ms * 1e-3) / 1e9, "h2d_bytes": 4 * N, "d2h_bytes": 12 * V + 16 * Q,
           "transfer_gbs": (4 * N + 12 * V + 16 * Q) / (ms * 1e-3) / 1e9,
           "out_of_core": {"device_budget_bytes": budget, "slab_cell_planes": S, "ms_per_extraction": ms_ooc,
                           "gvoxels_per_s": N / (ms_ooc * 1e-3) / 1e9, "identical_to_in_core": True},
           "obj_writer": {"bytes": obj_bytes, "ms": ms_obj, "mb_per_s": obj_bytes / (ms_obj * 1e-3) / 1e6,
                          "vertices": V, "triangles": 2 * Q}}
    log(f"host path: {out}")
    return out


def bench_smooth(ctx, torch, field, iso: float, verts, quads, counts, V: int, iters: int = 3) -> dict:
    """Vertex smoothing (num_smooth, mdmc_smooth_device) on the mesh of one
    extraction: `iters` iterations (each k_smooth then k_reaverage) timed with
    HIP events on the launching stream.  Algorithmic bytes per iteration: every
    vertex read once and written once, 24 V."""
    ctx.extract_into(field, iso, verts, quads, counts=counts)
    stream = torch.cuda.current_stream()
    ctx.smooth(verts, quads, 1, counts=counts)  # warm
    ctx.extract_into(field, iso, verts, quads, counts=counts)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    ctx.smooth(verts, quads, iters, counts=counts)
    t1.record(stream)
    torch.cuda.synchronize()
    assert int(counts[3]) == 0
    ms = t0.elapsed_time(t1) / iters
    return {"iterations": iters, "ms_per_iteration": ms, "algorithmic_bytes_per_iteration": 24.0 * V,
            "achieved_gbs": 24.0 * V / (ms * 1e-3) / 1e9}


def bench_dense(fa, torch, n: int = 1024, steps: int = 5) -> dict:
    """The surface-dense case: n^3 gyroid (period 24) + noise 0.05, iso 0."""
    ctx = fa.Context()
    log(f"dense: generating {n}^3 gyroid")
    f = ctx.field_gyroid((n, n, n), period=24.0, seed=0, noise_amp=0.05)
    torch.cuda.synchronize()
    ctx.count(f, 0.0)
    torch.cuda.synchronize()
    c = ctx.read_counts()
    A, V, Q = c["num_active"], c["num_vertices"], c["num_quads"]
    verts = torch.empty((V, 3), dtype=torch.float32, device="cuda")
    quads = torch.empty((Q, 4), dtype=torch.int32, device="cuda")
    counts = torch.zeros(4, dtype=torch.int32, device="cuda")
    ctx.extract_into(f, 0.0, verts, quads, counts=counts)
    stream = torch.cuda.current_stream()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record(stream)
    for _ in range(steps):
        ctx.extract_into(f, 0.0, verts, quads, counts=counts)
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / steps
    ctx.profile(True)
    ctx.profile_read()
    ctx.extract_into(f, 0.0, verts, quads, counts=counts)
    torch.cuda.synchronize()
    prof = ctx.profile_read()
    ctx.profile(False)
    st = counts.cpu().numpy()
    assert st[3] == 0 and st[1] == V and st[2] == Q
    N = n ** 3
    B = 4.0 * N + 12.0 * V + 16.0 * Q
    out = {"workload": f"{n}^3 FP32 gyroid, period 24, noise U(-0.05, 0.05), iso 0", "active_cells": A,
           "vertices": V, "quads": Q, "ms_per_extraction": ms, "gvoxels_per_s": N / (ms * 1e-3) / 1e9,
           "algorithmic_bytes": B, "achieved_gbs": B / (ms * 1e-3) / 1e9,
           "hbm_frac": B / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "kernel_ms": {k: v[0] / max(v[1], 1) for k, v in prof.items() if k != "other"}}
    out["smoothing"] = bench_smooth(ctx, torch, f, 0.0, verts, quads, counts, V)
    log(f"dense: {out}")
    ctx.close()
    del f, verts, quads
    torch.cuda.empty_cache()
    return out


def _log_uniform_int(rng, lo, hi):
    import numpy as np

    return int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))


def batched_workload(nvol: int = 256, seed: int = 1):
    """The batched workload: nvol volumes with dims drawn per axis
    log-uniform in [32, 256], alternating blob sums (8-32 blobs, sigma
    U[4,16]) and gyroid (period 24) + noise 0.05, isovalues U[-0.2, 0.2].
    Returns (generator specs, isovalues)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    specs, isos = [], []
    for i in range(nvol):
        dims = tuple(_log_uniform_int(rng, 32, 256) for _ in range(3))
        if i % 2 == 0:
            specs.append({"dims": dims, "kind": "blobs", "seed": 1000 + i, "nblobs": int(rng.integers(8, 33)),
                          "sigma": (4.0, 16.0)})
        else:
            specs.append({"dims": dims, "kind": "gyroid", "seed": 1000 + i, "period": 24.0, "noise_amp": 0.05})
        isos.append(float(rng.uniform(-0.2, 0.2)))
    return specs, isos


def bench_batched(fa, torch, nvol: int = 256, reps: int = 5, tune: bool = True) -> dict:
    """256 uneven volumes: plain per-volume loop vs one batched launch sequence
    (launch shapes chosen by the autotuner unless tune is False)."""
    import numpy as np

    specs, isos = batched_workload(nvol, seed=1)
    ctx = fa.Context()
    log("batched: generating 256 fields")
    fields = ctx.batch_generate(specs)
    torch.cuda.synchronize()
    points = sum(int(f.numel()) for f in fields)
    # plain loop: the single-volume protocol per volume
    sctx = fa.Context()
    for f, iso in zip(fields[:4], isos[:4]):
        sctx.extract(f, iso)
    torch.cuda.synchronize()
    t = time.perf_counter()
    V = Q = 0
    for f, iso in zip(fields, isos):
        v, q, c = sctx.extract(f, iso)
        V += c["num_vertices"]
        Q += c["num_quads"]
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t
    # batched: count all, one sync, allocate, emit all
    tuned = None
    if tune:
        from framework_amd import autotune

        tuned = autotune.tune_batch(ctx, fields, isos)
        log(f"batched: launch options {tuned['options']} ({tuned['ms']:.3f} ms)")
    ctx.batch_extract(fields, isos)
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        res = ctx.batch_extract(fields, isos)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    t_batch = sorted(times)[len(times) // 2]
    assert sum(r[2]["num_vertices"] for r in res) == V and sum(r[2]["num_quads"] for r in res) == Q
    # graph replay of count + emit into fixed outputs
    outs = [(r[0], r[1]) for r in res]
    counts = ctx.batch_counts_tensor(nvol)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        ctx.batch_extract_into(fields, isos, outs, counts, stream=stream)
    stream.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        ctx.batch_extract_into(fields, isos, outs, counts, stream=torch.cuda.current_stream())
    g.replay()
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        g.replay()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    t_graph = sorted(times)[len(times) // 2]
    st = counts.cpu().numpy().view(np.uint32)
    assert (st[:, 3] == 0).all() and int(st[:, 1].sum()) == V
    ctx.close()
    sctx.close()
    out = {"volumes": nvol, "points": points, "vertices": V, "quads": Q,
           "plain_loop_volumes_per_s": nvol / t_plain, "batched_volumes_per_s": nvol / t_batch,
           "graph_volumes_per_s": nvol / t_graph, "speedup_batched": t_plain / t_batch,
           "speedup_graph": t_plain / t_graph, "batched_gvoxels_per_s": points / t_batch / 1e9,
           "graph_gvoxels_per_s": points / t_graph / 1e9,
           "launch_options": None if tuned is None else
           {autotune.OPTION_NAMES[k]: v for k, v in tuned["options"].items()},
           "launch_option_ids": None if tuned is None else dict(tuned["options"])}
    log(f"batched: {out}")
    return out


def _served_requests(n: int, seed: int):
    import numpy as np
    from framework_amd.server import make_request

    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        d = _log_uniform_int(rng, 48, 384)
        iso = float(rng.uniform(-0.2, 0.2))
        if rng.random() < 0.5:
            reqs.append(make_request((d, d, d), iso, "blobs", seed=seed * 1_000_003 + i,
                                     nblobs=int(rng.integers(8, 33)), sigma=(4.0, 16.0), tag=i))
        else:
            reqs.append(make_request((d, d, d), iso, "gyroid", seed=seed * 1_000_003 + i, period=24.0,
                                     noise_amp=0.05, tag=i))
    return reqs


def bench_served(fa, torch, nreq: int, budget: int = 64 << 30, options: dict | None = None) -> dict:
    """Poisson stream through the request loop at 0.8x its measured capacity
    (lanes use the batched leg's tuned launch options when given)."""
    import numpy as np
    from framework_amd.server import Server, run_plain_loop, run_stream

    reqs = _served_requests(nreq, seed=5)
    srv = Server(budget, max_batch=32, num_streams=8)
    if options:
        srv.set_options(options)
    # capacity: a saturating burst (everything arrives at once) on a sample
    ncap = min(1000, nreq)
    log(f"served: capacity burst of {ncap}")
    run_stream(srv, reqs[: min(50, nreq)], np.zeros(min(50, nreq)))  # warm the lanes
    cap = run_stream(srv, reqs[:ncap], np.zeros(ncap))
    capacity = ncap / cap["wall"]
    rate = 0.8 * capacity
    arrivals = np.cumsum(np.random.default_rng(6).exponential(1.0 / rate, nreq))
    log(f"served: {nreq} requests at {rate:.0f}/s")
    res = run_stream(srv, reqs, arrivals)
    st = srv.stats()
    srv.close()
    ok = res["status"] == 0
    lat = np.sort(res["latency"][ok])
    # latency histogram (ms): counts per bin [edges[i], edges[i+1]), the last open
    hist_edges = [0.0, 0.1, 0.2, 0.5, 1.0, 2.0, 5.0, 10.0, 20.0, 50.0]
    hist_counts = np.histogram(lat * 1e3, bins=hist_edges + [float("inf")])[0].tolist()
    # plain loop on a sample of the same requests
    nplain = min(300, nreq)
    ctx = fa.Context()
    run_plain_loop(ctx, reqs[:5])
    plain = run_plain_loop(ctx, reqs[:nplain])
    ctx.close()
    plain_vps = nplain / plain["wall"]
    out = {"requests": nreq, "completed_ok": int(ok.sum()), "rejected": int((res["status"] == -8).sum()),
           "offered_rate": rate, "capacity_volumes_per_s": capacity,
           "volumes_per_s": nreq / res["wall"], "p50_ms": float(np.percentile(lat, 50) * 1e3),
           "p99_ms": float(np.percentile(lat, 99) * 1e3),
           "latency_hist_ms": {"edges": hist_edges, "counts": hist_counts}, "batches": st["batches"],
           "memory_peak_gb": st["memory_peak"] / 1e9, "budget_gb": budget / 1e9,
           "plain_loop_volumes_per_s": plain_vps, "plain_loop_sample": nplain,
           "plain_loop_p50_ms": float(np.percentile(plain["latency"], 50) * 1e3),
           "speedup_vs_plain_loop": capacity / plain_vps,
           "speedup_stream_vs_plain_loop": (nreq / res["wall"]) / plain_vps}
    log(f"served: {out}")
    return out


if __name__ == "__main__":
    sys.exit(main())
