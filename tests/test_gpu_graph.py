"""The single-volume extraction captured in a HIP graph (GPU): replays equal
the eager extraction bitwise, with eager work of other contexts and torch
between the replays, on a sparse and a surface-dense field; eager calls on
the captured context are refused."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    import framework_amd as fa

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    torch.cuda.set_device(0)
    fa.setup()
    return torch, fa


def _replays(torch, fa, field, iso, box, nrep=3):
    eager = fa.Context()
    v0, q0, c0 = eager.extract(field, iso, box[0], box[1])
    gctx = fa.Context()
    verts, quads = torch.empty_like(v0), torch.empty_like(q0)
    counts = torch.zeros(4, dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):  # sizes the context's scratch outside the capture
        gctx.extract_into(field, iso, verts, quads, xyz_min=box[0], xyz_max=box[1], counts=counts, stream=s)
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        gctx.extract_into(field, iso, verts, quads, xyz_min=box[0], xyz_max=box[1], counts=counts,
                          stream=torch.cuda.current_stream())
    other = torch.rand(257, 263, 91, device="cuda") - 0.5
    for rep in range(nrep):
        print(f"graph replay {rep} of {tuple(field.shape)}", flush=True)
        verts.fill_(-1.0)
        quads.fill_(-1)
        counts.fill_(-1)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        c = counts.cpu().numpy().view(np.uint32)
        assert c[3] == 0 and (c[0], c[1], c[2]) == (c0["num_active"], c0["num_vertices"], c0["num_quads"]), rep
        assert torch.equal(quads, q0) and torch.equal(verts.view(torch.int32), v0.view(torch.int32)), rep
        # eager work between replays: torch, and another context on the same field
        _ = (other * 2.0).sum().item()
        v1, q1, _ = eager.extract(field, iso, box[0], box[1])
        assert torch.equal(q1, q0)
    with pytest.raises(fa.MDMCError, match="captured graph"):
        gctx.extract_into(field, iso, verts, quads, xyz_min=box[0], xyz_max=box[1], counts=counts)
    del g
    gctx.close()
    eager.close()


def test_graph_replays_sparse(env):
    torch, fa = env
    ctx = fa.Context()
    f = ctx.field_blobs((300, 280, 260), seed=9, nblobs=20, sigma=(6.0, 20.0))
    torch.cuda.synchronize()
    _replays(torch, fa, f, 0.1, ((-1.0, 0.0, 2.0), (4.0, 3.0, 9.0)))
    ctx.close()


def test_graph_replays_dense(env):
    torch, fa = env
    ctx = fa.Context()
    f = ctx.field_gyroid((256, 200, 180), period=11.0, seed=3, noise_amp=0.05)
    torch.cuda.synchronize()
    _replays(torch, fa, f, 0.0, ((0.0, 0.0, 0.0), (255.0, 199.0, 179.0)))
    ctx.close()
