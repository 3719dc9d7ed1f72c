"""The no-allocation entry (mdmc_extract_device_noalloc), called through the C
ABI, against the allocating entry it varies (mdmc_extract_device): the same
counts, quads and vertex bits, on two calls in a row (the second on a
smaller grid, so stale contents of the shared scratch would show), and a grid
whose scratch does not fit the fixed array is refused before any launch."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_VP = ctypes.c_void_p
_U3 = ctypes.c_uint32 * 3
_F3 = ctypes.c_float * 3


@pytest.fixture(scope="module")
def env():
    import torch

    import framework_amd as fa

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    torch.cuda.set_device(0)
    fa.setup()
    L = fa.lib()
    L.mdmc_extract_device_noalloc.argtypes = [_VP, _U3, _F3, _F3, ctypes.c_float, _VP, ctypes.c_uint64, _VP,
                                              ctypes.c_uint64, _VP, _VP]
    L.mdmc_extract_device_noalloc.restype = ctypes.c_int
    return torch, fa, L


def _both(torch, fa, L, field, iso, box):
    """(counts, vertices, quads) from the allocating entry on a context and
    from the no-allocation entry, same field, same outputs' capacities."""
    Z, Y, X = field.shape
    ctx = fa.Context()
    ctx.count(field, iso)
    torch.cuda.synchronize()
    c = ctx.read_counts()
    V, Q = c["num_vertices"], c["num_quads"]
    out = []
    for noalloc in (False, True):
        verts = torch.full((V, 3), -7.0, device="cuda")
        quads = torch.full((Q, 4), -7, dtype=torch.int32, device="cuda")
        counts = torch.full((4,), -1, dtype=torch.int32, device="cuda")
        stream = torch.cuda.current_stream()
        if noalloc:
            print(f"noalloc call dims {(X, Y, Z)} iso {iso}", flush=True)
            rc = L.mdmc_extract_device_noalloc(_VP(field.data_ptr()), _U3(X, Y, Z), _F3(*box[0]), _F3(*box[1]),
                                               ctypes.c_float(iso), _VP(verts.data_ptr()), V, _VP(quads.data_ptr()),
                                               Q, _VP(counts.data_ptr()), _VP(stream.cuda_stream))
            assert rc == 0, fa.lib().mdmc_last_error()
        else:
            ctx.extract_into(field, iso, verts, quads, xyz_min=box[0], xyz_max=box[1], counts=counts)
        torch.cuda.synchronize()
        out.append((counts.cpu().numpy().view(np.uint32).copy(), verts, quads))
    ctx.close()
    return out


def test_noalloc_matches_the_allocating_entry_twice(env):
    torch, fa, L = env
    gen = fa.Context()
    big = gen.field_gyroid((200, 190, 170), period=13.0, seed=3, noise_amp=0.3)
    small = gen.field_blobs((90, 71, 63), seed=4, nblobs=12, sigma=(3.0, 10.0))
    torch.cuda.synchronize()
    cases = [(big, 0.05, ((0.0, 0.0, 0.0), (199.0, 189.0, 169.0))),
             (small, 0.1, ((-2.0, 1.0, 0.5), (3.0, 4.0, 9.0)))]
    for field, iso, box in cases:
        (c0, v0, q0), (c1, v1, q1) = _both(torch, fa, L, field, iso, box)
        assert c0[3] == 0 and c1[3] == 0, (c0, c1)
        assert np.array_equal(c0, c1), (c0, c1)
        assert c1[2] > 0
        assert torch.equal(q0, q1)
        assert torch.equal(v0.view(torch.int32), v1.view(torch.int32))
    gen.close()


def test_noalloc_refuses_a_grid_beyond_its_scratch(env):
    torch, fa, L = env
    X = Y = Z = 400
    field = torch.zeros((Z, Y, X), dtype=torch.float32, device="cuda")
    counts = torch.zeros(4, dtype=torch.int32, device="cuda")
    rc = L.mdmc_extract_device_noalloc(_VP(field.data_ptr()), _U3(X, Y, Z), _F3(0, 0, 0), _F3(1, 1, 1),
                                       ctypes.c_float(0.5), _VP(None), 0, _VP(None), 0, _VP(counts.data_ptr()),
                                       _VP(torch.cuda.current_stream().cuda_stream))
    assert rc == -4  # MDMC_ERR_CAPACITY
    assert b"no-allocation entry has" in L.mdmc_last_error()


def test_noalloc_under_graph_capture(env):
    """The callers the entry is for: a HIP graph captured around it (no
    allocation may happen inside a capture), replayed with other work between
    replays; every replay equals the allocating entry's result."""
    torch, fa, L = env
    gen = fa.Context()
    field = gen.field_gyroid((150, 130, 110), period=10.0, seed=8, noise_amp=0.1)
    torch.cuda.synchronize()
    v0, q0, c0 = gen.extract(field, 0.02)
    Z, Y, X = field.shape
    verts, quads = torch.empty_like(v0), torch.empty_like(q0)
    counts = torch.zeros(4, dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        rc = L.mdmc_extract_device_noalloc(_VP(field.data_ptr()), _U3(X, Y, Z), _F3(0, 0, 0),
                                           _F3(X - 1, Y - 1, Z - 1), ctypes.c_float(0.02), _VP(verts.data_ptr()),
                                           verts.shape[0], _VP(quads.data_ptr()), quads.shape[0],
                                           _VP(counts.data_ptr()), _VP(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, L.mdmc_last_error()
    for rep in range(3):
        print(f"noalloc graph replay {rep}", flush=True)
        verts.fill_(-1.0)
        quads.fill_(-1)
        counts.fill_(-1)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        c = counts.cpu().numpy().view(np.uint32)
        assert c[3] == 0 and (c[1], c[2]) == (c0["num_vertices"], c0["num_quads"]), rep
        assert torch.equal(quads, q0) and torch.equal(verts.view(torch.int32), v0.view(torch.int32)), rep
        _ = gen.extract(field, 0.1)  # other work on the GPU between replays
    del g
    gen.close()
