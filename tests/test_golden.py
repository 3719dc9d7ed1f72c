"""Committed fixtures (tests/golden/golden.json, written by make_golden.py):
the oracle (CPU) and the HIP path (GPU) both reproduce the recorded counts and
SHA-256 digests of quads + vertex bits."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "golden.json")))


def field_for(name):
    import tests.golden.make_golden as mg

    for n, f, iso in mg.cases():
        if n == name:
            return f, iso
    raise KeyError(name)


def digest(quads, verts):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(quads, dtype=np.int32).tobytes())
    h.update(np.ascontiguousarray(verts, dtype=np.float32).view(np.uint32).tobytes())
    return h.hexdigest()


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_oracle_reproduces_fixture(name):
    f, iso = field_for(name)
    r = oracle.run(f, iso)
    g = GOLDEN[name]
    assert (r["num_active"], r["num_vertices"], r["num_quads"]) == (g["num_active"], g["num_vertices"], g["num_quads"])
    assert digest(r["quads"], r["vertices"]) == g["sha256"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_gpu_reproduces_fixture(name):
    import torch

    import framework_amd as fa

    f, iso = field_for(name)
    fa.setup()
    ctx = fa.Context()
    verts, quads, cnt = ctx.extract(torch.from_numpy(f).cuda(), iso)
    g = GOLDEN[name]
    assert (cnt["num_active"], cnt["num_vertices"], cnt["num_quads"]) == (g["num_active"], g["num_vertices"],
                                                                           g["num_quads"])
    assert digest(quads.cpu().numpy(), verts.cpu().numpy()) == g["sha256"]
    ctx.close()
