"""Launch-shape autotuner for the extraction kernels.

The tunable launch shapes are context options (MDMC_OPT_* in include/mdmc.h):

  single volume  MDMC_OPT_PT_ROWS              cell rows per classify tile (7, 15)
                 MDMC_OPT_PT_PLANES            cell planes per classify tile
                 MDMC_OPT_PT_BAND              columns per XCD group
                 MDMC_OPT_EMIT_BLOCKS_PER_CU   persistent emit grid
  batches        MDMC_OPT_BATCH_PLANES         cell planes per batch classify tile
                 MDMC_OPT_CLASSIFY_BLOCKS_PER_CU  persistent classify grid (0: a workgroup per tile)
                 MDMC_OPT_EMIT_BLOCKS_PER_CU   emit grid

None of them changes a result, only how the work is laid over the chip, so
the tuner may pick any of them; it checks the counts of every candidate
against the default's anyway.  The search is coordinate descent: one option
at a time over its candidate values, the others held at the best so far,
each candidate timed with HIP events on the launching stream (median of
`reps` runs).  Choices are cached in a JSON file keyed by device name and a
workload signature, so a later process applies them without sweeping.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

from .mdmc import (OPT_BATCH_PLANES, OPT_CLASSIFY_BLOCKS_PER_CU, OPT_EMIT_BLOCKS_PER_CU, OPT_PT_BAND,
                   OPT_PT_PLANES, OPT_PT_ROWS, MDMCError)

SINGLE_SPACE = {OPT_PT_ROWS: [0, 15], OPT_PT_PLANES: [0, 1, 2, 3], OPT_PT_BAND: [0, 2, 4],
                OPT_EMIT_BLOCKS_PER_CU: [0, 16]}
BATCH_SPACE = {OPT_BATCH_PLANES: [0, 1, 2, 4], OPT_CLASSIFY_BLOCKS_PER_CU: [0, 4, 8, 12],
               OPT_EMIT_BLOCKS_PER_CU: [0, 8, 32, 256]}
OPTION_NAMES = {OPT_BATCH_PLANES: "batch_planes", OPT_CLASSIFY_BLOCKS_PER_CU: "classify_blocks_per_cu",
                OPT_EMIT_BLOCKS_PER_CU: "emit_blocks_per_cu", OPT_PT_PLANES: "pt_planes", OPT_PT_BAND: "pt_band",
                OPT_PT_ROWS: "pt_rows"}


def default_cache_path() -> str:
    return os.environ.get("MDMC_TUNING_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "mdmc", "tuning.json"))


def _device_name() -> str:
    import torch

    return torch.cuda.get_device_name(torch.cuda.current_device())


def single_signature(dims_xyz) -> str:
    return "single:%dx%dx%d" % tuple(int(d) for d in dims_xyz)


def batch_signature(dims_list) -> str:
    h = hashlib.sha256(np.asarray(sorted(tuple(int(d) for d in dims) for dims in dims_list),
                                  dtype=np.int64).tobytes()).hexdigest()[:16]
    return f"batch:{len(dims_list)}:{h}"


def apply(ctx, options: dict) -> None:
    """Set options ({option id: value}) on a context; 0 restores a default."""
    for opt, val in options.items():
        ctx.set_option(int(opt), int(val))


def _time(torch, run, reps: int) -> float:
    stream = torch.cuda.current_stream()
    times = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        run()
        b.record(stream)
        b.synchronize()
        times.append(a.elapsed_time(b))
    return float(np.median(times))


def _descend(torch, ctx, space: dict, run, check, reps: int) -> tuple[dict, list]:
    best = {opt: 0 for opt in space}
    apply(ctx, best)
    run()
    torch.cuda.synchronize()
    ref = check()
    best_ms = _time(torch, run, reps)
    log = [({OPTION_NAMES[o]: v for o, v in best.items()}, best_ms)]
    for opt, values in space.items():
        for val in values:
            if val == best[opt]:
                continue
            cand = dict(best)
            cand[opt] = val
            apply(ctx, cand)
            run()
            torch.cuda.synchronize()
            if check() != ref:
                raise MDMCError(-7, f"autotune: option {OPTION_NAMES[opt]}={val} changed the counts")
            ms = _time(torch, run, reps)
            log.append(({OPTION_NAMES[o]: v for o, v in cand.items()}, ms))
            if ms < best_ms:
                best, best_ms = cand, ms
    # confirm against the defaults, alternating the two so that drifting
    # clocks hit both alike: a choice must win by more than 1% to stand
    default = {opt: 0 for opt in space}
    if best != default:
        t_best, t_def = [], []
        for _ in range(3):
            apply(ctx, default)
            t_def.append(_time(torch, run, reps))
            apply(ctx, best)
            t_best.append(_time(torch, run, reps))
        if float(np.median(t_best)) > 0.99 * float(np.median(t_def)):
            best = default
        log.append(({"confirm": "default" if best == default else "tuned"}, float(np.median(t_best))))
    apply(ctx, best)
    return best, log


def tune_single(ctx, field, iso: float, reps: int = 10, space: dict | None = None) -> dict:
    """Tunes the single-volume path on one resident field (a (Z, Y, X) float32
    CUDA tensor).  Returns {"options", "ms", "log"}; the options are left set
    on ctx."""
    import torch

    ctx.count(field, iso)
    torch.cuda.synchronize()
    c = ctx.read_counts()
    verts = torch.empty((max(c["num_vertices"], 1), 3), device="cuda")
    quads = torch.empty((max(c["num_quads"], 1), 4), dtype=torch.int32, device="cuda")
    counts = ctx.counts_tensor()

    def run():
        ctx.extract_into(field, iso, verts, quads, counts=counts)

    def check():
        return tuple(counts.cpu().numpy().view(np.uint32).tolist())

    best, log = _descend(torch, ctx, space or SINGLE_SPACE, run, check, reps)
    return {"options": best, "ms": min(ms for _, ms in log), "log": log}


def tune_batch(ctx, fields, isos, reps: int = 5, space: dict | None = None) -> dict:
    """Tunes the batch path on resident fields (outputs sized by one exact
    count).  Returns {"options", "ms", "log"}; the options are left set on ctx."""
    import torch

    exact = ctx.batch_extract(fields, isos)
    outs = [(v, q) for v, q, _ in exact]
    counts = ctx.batch_counts_tensor(len(fields))

    def run():
        ctx.batch_extract_into(fields, isos, outs, counts)

    def check():
        return counts.cpu().numpy().tobytes()

    best, log = _descend(torch, ctx, space or BATCH_SPACE, run, check, reps)
    return {"options": best, "ms": min(ms for _, ms in log), "log": log}


def load(path: str | None = None) -> dict:
    path = path or default_cache_path()
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save(entries: dict, path: str | None = None) -> None:
    path = path or default_cache_path()
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(entries, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def cached_or_tune(ctx, signature: str, tune, path: str | None = None) -> dict:
    """Applies the cached options for (device, signature), or runs `tune()`
    (which returns a tune_* result), caches and applies its choice."""
    key = f"{_device_name()}|{signature}"
    entries = load(path)
    if key in entries:
        opts = {int(k): int(v) for k, v in entries[key]["options"].items()}
        apply(ctx, opts)
        return {"options": opts, "ms": entries[key].get("ms"), "cached": True}
    res = tune()
    entries[key] = {"options": {str(k): int(v) for k, v in res["options"].items()}, "ms": res["ms"]}
    save(entries, path)
    res["cached"] = False
    return res


__all__ = ["tune_single", "tune_batch", "apply", "load", "save", "cached_or_tune", "single_signature",
           "batch_signature", "default_cache_path", "SINGLE_SPACE", "BATCH_SPACE", "OPTION_NAMES"]
