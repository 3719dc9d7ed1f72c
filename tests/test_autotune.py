"""Launch-shape options and the autotuner: every option value gives the same
mesh as the defaults (GPU), and the tuning cache round-trips (CPU)."""
from __future__ import annotations

import numpy as np
import pytest


class _FakeCtx:
    def __init__(self):
        self.opts = {}

    def set_option(self, opt, val):
        self.opts[opt] = val


def test_tuning_cache_roundtrip(tmp_path, monkeypatch):
    from framework_amd import autotune

    monkeypatch.setattr(autotune, "_device_name", lambda: "test-device")
    path = str(tmp_path / "tuning.json")
    calls = []

    def tune():
        calls.append(1)
        return {"options": {autotune.OPT_PT_BAND: 2, autotune.OPT_PT_PLANES: 3}, "ms": 0.5, "log": []}

    ctx = _FakeCtx()
    sig = autotune.single_signature((64, 32, 16))
    r1 = autotune.cached_or_tune(ctx, sig, tune, path)
    assert not r1["cached"] and calls == [1]
    ctx2 = _FakeCtx()
    r2 = autotune.cached_or_tune(ctx2, sig, tune, path)
    assert r2["cached"] and calls == [1]
    assert ctx2.opts == {autotune.OPT_PT_BAND: 2, autotune.OPT_PT_PLANES: 3}
    # another workload misses the cache
    autotune.cached_or_tune(_FakeCtx(), autotune.single_signature((64, 32, 17)), tune, path)
    assert calls == [1, 1]
    assert autotune.batch_signature([(1, 2, 3), (4, 5, 6)]) == autotune.batch_signature([(4, 5, 6), (1, 2, 3)])


def test_set_option_rejects_a_null_context():
    import framework_amd as fa

    lib = fa.lib()
    assert lib.mdmc_ctx_set_option(None, 0, 1) != 0


@pytest.mark.gpu
def test_every_single_volume_option_keeps_the_mesh():
    import torch

    import framework_amd as fa
    from framework_amd import autotune

    torch.cuda.set_device(0)
    fa.setup()
    gen = fa.Context()
    fields = [gen.field_gyroid((300, 45, 29), period=11.0, seed=3, noise_amp=0.2),   # 512-cell x-tiles
              gen.field_blobs((131, 70, 40), seed=5, nblobs=12, sigma=(4.0, 12.0)),  # 256-cell x-tiles
              gen.field_blobs((700, 20, 12), seed=6, nblobs=9, sigma=(4.0, 12.0))]   # two 512-cell x-tiles
    torch.cuda.synchronize()
    for f in fields:
        ref_ctx = fa.Context()
        v0, q0, c0 = ref_ctx.extract(f, 0.02)
        ref_ctx.close()
        for opt, values in autotune.SINGLE_SPACE.items():
            for val in values:
                ctx = fa.Context()
                ctx.set_option(opt, val)
                v, q, c = ctx.extract(f, 0.02)
                assert c == c0, (opt, val)
                assert torch.equal(q, q0), (opt, val)
                assert torch.equal(v.view(torch.int32), v0.view(torch.int32)), (opt, val)
                ctx.close()
    gen.close()


@pytest.mark.gpu
def test_tuners_pick_options_that_keep_results():
    import torch

    import framework_amd as fa
    from framework_amd import autotune

    torch.cuda.set_device(0)
    fa.setup()
    ctx = fa.Context()
    f = ctx.field_gyroid((128, 96, 64), period=13.0, seed=1, noise_amp=0.1)
    torch.cuda.synchronize()
    v0, q0, _ = fa.Context().extract(f, 0.0)
    res = autotune.tune_single(ctx, f, 0.0, reps=3)
    assert set(res["options"]) == set(autotune.SINGLE_SPACE) and len(res["log"]) >= 2
    v, q, _ = ctx.extract(f, 0.0)
    assert torch.equal(q, q0) and torch.equal(v.view(torch.int32), v0.view(torch.int32))

    rng = np.random.default_rng(2)
    specs = [{"dims": tuple(int(d) for d in rng.integers(20, 90, 3)), "kind": "blobs" if i % 2 else "gyroid",
              "seed": i, "nblobs": 10, "sigma": (3.0, 9.0), "period": 10.0, "noise_amp": 0.1} for i in range(12)]
    fields = ctx.batch_generate(specs)
    isos = [0.0] * len(fields)
    ref = fa.Context().batch_extract(fields, isos)
    bres = autotune.tune_batch(ctx, fields, isos, reps=2)
    assert set(bres["options"]) == set(autotune.BATCH_SPACE)
    out = ctx.batch_extract(fields, isos)
    for (va, qa, ca), (vb, qb, cb) in zip(ref, out):
        assert ca == cb and torch.equal(qa, qb) and torch.equal(va.view(torch.int32), vb.view(torch.int32))
    ctx.close()
