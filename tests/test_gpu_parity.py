"""HIP path vs the CPU oracle on identical float bits (GPU).

Bar: case indices, per-cell vertex counts/offsets and quads bit-exact; vertex
positions bitwise equal to the oracle's FP32 sequence (which implies the
1e-5 voxel-unit tolerance of the brief).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests.helpers import assert_same_mesh, check_against_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    torch.cuda.set_device(0)
    return torch


@pytest.fixture(scope="module")
def ctx(torch_cuda):
    import framework_amd as fa

    fa.setup()
    c = fa.Context()
    yield c
    c.close()


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def test_sphere_64(ctx, torch_cuda):
    """Config 0: 64^3 signed distance to a sphere r=20 at (32,32,32), iso 0."""
    f = oracle.field_sphere((64, 64, 64), (32.0, 32.0, 32.0), 20.0)
    ref = check_against_oracle(ctx, _dev(torch_cuda, f), 0.0)
    assert (ref["num_active"], ref["num_vertices"], ref["num_quads"]) == (7472, 14942, 7470)
    assert ref["num_tie_edges"] == 72


@pytest.mark.parametrize("dims,seed,nblobs,iso", [
    ((96, 80, 72), 1, 12, 0.1),
    ((130, 70, 33), 2, 8, 0.05),
    ((64, 64, 64), 3, 20, -0.1),
    ((200, 17, 9), 4, 6, 0.0),
])
def test_blobs_small(ctx, torch_cuda, dims, seed, nblobs, iso):
    f = ctx.field_blobs(dims, seed=seed, nblobs=nblobs, sigma=(3.0, 12.0))
    torch_cuda.cuda.synchronize()
    check_against_oracle(ctx, f, iso)


@pytest.mark.parametrize("dims,iso", [((96, 96, 96), 0.0), ((77, 50, 41), 0.15), ((33, 129, 20), -0.2)])
def test_gyroid_noise(ctx, torch_cuda, dims, iso):
    """Strongly perturbed gyroid: dense output and adjacent 2B/3B cells."""
    f = ctx.field_gyroid(dims, period=8.0, seed=5, noise_amp=0.8)
    torch_cuda.cuda.synchronize()
    ref = check_against_oracle(ctx, f, iso)
    assert ((ref["info"] >> 7) & 1).sum() > 0, "expected some variant-1 cells"


@pytest.mark.parametrize("dims", [(260, 9, 11), (1028, 5, 6), (516, 13, 7), (256, 8, 8), (772, 3, 4)])
def test_wide_rows_vectorised(ctx, torch_cuda, dims):
    """Rows that are a multiple of 4 points and >= 256 cells take the 16-byte
    load kernel: partial chunks, a second x-tile, rows ending mid-chunk."""
    f = ctx.field_gyroid(dims, period=7.0, seed=13, noise_amp=0.6)
    torch_cuda.cuda.synchronize()
    check_against_oracle(ctx, f, 0.05)


def test_misaligned_field_pointer(ctx, torch_cuda):
    """A field pointer that is not 16-byte aligned uses the scalar kernel."""
    torch = torch_cuda
    dims = (520, 6, 5)
    X, Y, Z = dims
    buf = torch.empty(X * Y * Z + 1, dtype=torch.float32, device="cuda")
    f = ctx.field_gyroid(dims, period=9.0, seed=14, noise_amp=0.4, out=buf[1:].view(Z, Y, X))
    torch.cuda.synchronize()
    check_against_oracle(ctx, f, 0.0)


def test_ragged_x_tiles(ctx, torch_cuda):
    """Rows longer than one 1024-cell x-tile (two segments per row)."""
    f = ctx.field_gyroid((1100, 9, 7), period=10.0, seed=9, noise_amp=0.2)
    torch_cuda.cuda.synchronize()
    check_against_oracle(ctx, f, 0.0)


@pytest.mark.parametrize("dims", [(2, 2, 2), (3, 2, 2), (2, 5, 3), (65, 2, 2), (64, 3, 2), (129, 4, 5)])
def test_tiny_grids(ctx, torch_cuda, dims):
    rng = np.random.default_rng(sum(dims))
    X, Y, Z = dims
    f = rng.standard_normal((Z, Y, X)).astype(np.float32)
    check_against_oracle(ctx, _dev(torch_cuda, f), 0.0)


def test_random_many_shapes(ctx, torch_cuda):
    rng = np.random.default_rng(1234)
    for _ in range(12):
        X, Y, Z = (int(v) for v in rng.integers(2, 90, size=3))
        f = rng.standard_normal((Z, Y, X)).astype(np.float32)
        check_against_oracle(ctx, _dev(torch_cuda, f), float(rng.uniform(-0.5, 0.5)))


def test_empty_surface(ctx, torch_cuda):
    """No active cell -> empty mesh (the reference exits here, dmc.cu:2113)."""
    f = np.full((20, 30, 40), 3.0, dtype=np.float32)
    verts, quads, cnt = ctx.extract(_dev(torch_cuda, f), 0.0)
    assert cnt["num_active"] == 0 and verts.shape == (0, 3) and quads.shape == (0, 4)


def test_degenerate_dims(ctx, torch_cuda):
    for dims in [(1, 5, 5), (5, 1, 5), (5, 5, 1)]:
        X, Y, Z = dims
        f = np.zeros((Z, Y, X), dtype=np.float32)
        verts, quads, cnt = ctx.extract(_dev(torch_cuda, f), 0.5)
        assert cnt["num_active"] == 0 and quads.shape == (0, 4)


def test_exact_ties(ctx, torch_cuda):
    """Integer field with integer iso: many corners exactly at iso."""
    rng = np.random.default_rng(7)
    f = rng.integers(-2, 3, size=(40, 37, 53)).astype(np.float32)
    ref = check_against_oracle(ctx, _dev(torch_cuda, f), 0.0)
    assert ref["num_tie_edges"] > 0


def test_nan_corners(ctx, torch_cuda):
    rng = np.random.default_rng(8)
    f = rng.standard_normal((30, 31, 45)).astype(np.float32)
    f.reshape(-1)[rng.choice(f.size, 200, replace=False)] = np.nan
    check_against_oracle(ctx, _dev(torch_cuda, f), 0.1)


def test_extreme_values(ctx, torch_cuda):
    """Magnitudes near the float range and infinities: differences overflow to
    inf, interpolation parameters become 0, nan or inf, the same IEEE results
    on both sides."""
    rng = np.random.default_rng(10)
    f = (rng.standard_normal((26, 30, 34)) * 3e38).astype(np.float32)
    flat = f.reshape(-1)
    flat[rng.choice(f.size, 150, replace=False)] = np.inf
    flat[rng.choice(f.size, 150, replace=False)] = -np.inf
    flat[rng.choice(f.size, 100, replace=False)] = np.float32(3.4e38)
    check_against_oracle(ctx, _dev(torch_cuda, f), 0.0)
    check_against_oracle(ctx, _dev(torch_cuda, f), 1e38)


def test_denormal_values(ctx, torch_cuda):
    rng = np.random.default_rng(9)
    f = (rng.standard_normal((20, 24, 28)) * 1e-40).astype(np.float32)
    check_against_oracle(ctx, _dev(torch_cuda, f), 0.0)


def test_world_box(ctx, torch_cuda):
    """Non-unit coordinate box (dmc.cu:2149 xyz_range)."""
    f = oracle.field_sphere((41, 45, 50), (20.5, 22.0, 24.9), 15.3)
    check_against_oracle(ctx, _dev(torch_cuda, f), 0.0, xyz_min=(-5.0, -2.5, 1.0), xyz_max=(5.0, 7.5, 3.0))


def test_world_box_through_both_coordinate_paths(ctx, torch_cuda):
    """The combined extract fills the coordinate tables during its count; a
    separate count + emit fills them in the emit.  Alternating boxes and paths
    on one context gives the oracle's vertices every time."""
    torch = torch_cuda
    f = oracle.field_sphere((37, 29, 44), (18.0, 14.5, 21.0), 11.7)
    ft = _dev(torch, f)
    boxes = [((-5.0, -2.5, 1.0), (5.0, 7.5, 3.0)), ((0.0, 0.0, 0.0), (36.0, 28.0, 43.0)),
             ((2.0, -1.0, -3.0), (2.5, 9.0, 0.0))]
    for rep in range(2):
        for mn, mx in boxes:
            ref = oracle.run(f, 0.0, mn, mx)
            V, Q = ref["num_vertices"], ref["num_quads"]
            verts = torch.empty((V, 3), device="cuda")
            quads = torch.empty((Q, 4), dtype=torch.int32, device="cuda")
            if rep == 0:
                ctx.extract_into(ft, 0.0, verts, quads, xyz_min=mn, xyz_max=mx)
            else:
                ctx.count(ft, 0.0)
                ctx.emit(ft, verts, quads, xyz_min=mn, xyz_max=mx)
            torch.cuda.synchronize()
            assert_same_mesh(verts.cpu().numpy(), quads.cpu().numpy(), ref)
    # count with one box (combined call), then emit the same count with another box
    mn, mx = boxes[0]
    ref = oracle.run(f, 0.0, *boxes[2])
    verts = torch.empty((ref["num_vertices"], 3), device="cuda")
    quads = torch.empty((ref["num_quads"], 4), dtype=torch.int32, device="cuda")
    ctx.extract_into(ft, 0.0, verts, quads, xyz_min=mn, xyz_max=mx)
    ctx.emit(ft, verts, quads, xyz_min=boxes[2][0], xyz_max=boxes[2][1])
    torch.cuda.synchronize()
    assert_same_mesh(verts.cpu().numpy(), quads.cpu().numpy(), ref)
    # ... and back to the count's box: the tables must be rewritten again
    ctx.emit(ft, verts, quads, xyz_min=mn, xyz_max=mx)
    torch.cuda.synchronize()
    assert_same_mesh(verts.cpu().numpy(), quads.cpu().numpy(), oracle.run(f, 0.0, mn, mx))


@pytest.mark.parametrize("emit_per_cu", [0, 1, 3, 1000])
def test_repeated_emits_of_one_count(ctx, torch_cuda, emit_per_cu):
    """The emit takes chunks by ticket from counters that the count zeroes;
    an emit that follows another without a count between zeroes them first.
    Three emits of one count equal the oracle on a surface-dense field with
    many more chunks than the emit's first tickets cover (with one workgroup
    per CU most chunks come through the counters), at grids of 8, 1, 3 and
    1000 workgroups per CU (the last capped by the chunk capacity)."""
    torch = torch_cuda
    from framework_amd.mdmc import OPT_EMIT_BLOCKS_PER_CU

    dims = (160, 150, 140)
    ft = ctx.field_gyroid(dims, period=7.0, seed=4, noise_amp=0.05)
    torch.cuda.synchronize()
    f = ft.cpu().numpy()
    mn, mx = (0.0, 0.0, 0.0), tuple(float(d - 1) for d in dims)
    ref = oracle.run(f, 0.0, mn, mx)
    assert ref["num_active"] > 64 * 8 * 1024  # more chunks than the first tickets of 256 CUs cover
    ctx.set_option(OPT_EMIT_BLOCKS_PER_CU, emit_per_cu)
    try:
        ctx.count(ft, 0.0)
        for rep in range(3):
            verts = torch.full((ref["num_vertices"], 3), -1.0, device="cuda")
            quads = torch.full((ref["num_quads"], 4), -1, dtype=torch.int32, device="cuda")
            ctx.emit(ft, verts, quads, xyz_min=mn, xyz_max=mx)
            torch.cuda.synchronize()
            assert ctx.read_counts()["status"] == 0, rep
            assert_same_mesh(verts.cpu().numpy(), quads.cpu().numpy(), ref)
    finally:
        ctx.set_option(OPT_EMIT_BLOCKS_PER_CU, 0)


def test_field_generators_match_cpu(ctx, torch_cuda):
    dims = (70, 51, 33)
    gb = ctx.field_blobs(dims, seed=11, nblobs=9)
    gg = ctx.field_gyroid(dims, period=24.0, seed=12, noise_amp=0.05)
    torch_cuda.cuda.synchronize()
    cb = oracle.field_blobs(dims, 11, 9)
    cg = oracle.field_gyroid(dims, 24.0, 12, 0.05)
    assert np.array_equal(gb.cpu().numpy().view(np.uint32), cb.view(np.uint32))
    assert np.array_equal(gg.cpu().numpy().view(np.uint32), cg.view(np.uint32))
    # several groups of 8 blobs (the last one padded), more blobs than the
    # registers hold (the per-point path), and generator tasks (64 columns of
    # one row in up to 16 planes) cut short in x and in z (23 planes are 16 + 7)
    for dims, nb in (((67, 29, 23), 27), ((130, 17, 9), 32), ((5, 3, 7), 17), ((70, 9, 37), 40)):
        gb = ctx.field_blobs(dims, seed=nb, nblobs=nb, sigma=(2.0, 9.0))
        gg = ctx.field_gyroid(dims, period=7.0, seed=nb, noise_amp=0.3)
        torch_cuda.cuda.synchronize()
        cb = oracle.field_blobs(dims, nb, nb, sigma=(2.0, 9.0))
        cg = oracle.field_gyroid(dims, 7.0, nb, 0.3)
        assert np.array_equal(gb.cpu().numpy().view(np.uint32), cb.view(np.uint32)), (dims, nb)
        assert np.array_equal(gg.cpu().numpy().view(np.uint32), cg.view(np.uint32)), (dims, nb)


def test_host_entry_matches_device(ctx, torch_cuda):
    """mdmc_run (host field in, host mesh out) == device path == oracle."""
    import framework_amd as fa

    f = oracle.field_blobs((57, 61, 49), 21, 7, sigma=(3.0, 9.0))
    v, q = fa.run_host(f, 0.1)
    ref = oracle.run(f, 0.1)
    assert_same_mesh(v, q, ref)
    verts, tris = fa.run_dmc(f, (0, 0, 0), (56, 60, 48), 0.1)
    assert tris.shape == (2 * ref["num_quads"], 3) and tris.dtype == np.uint32
    # triangles (q0,q1,q2),(q2,q3,q0) per quad (dmc.cu:2039-2045), as values
    q = ref["quads"].astype(np.uint32)
    assert np.array_equal(tris[0::2], q[:, [0, 1, 2]]) and np.array_equal(tris[1::2], q[:, [2, 3, 0]])
    assert np.array_equal(verts.view(np.uint32), ref["vertices"].view(np.uint32))


def test_device_entry_with_library_mesh(ctx, torch_cuda):
    """mdmc_run_device (device field in, library-allocated device mesh out,
    on a caller stream) == the context path == the oracle; smoothing and an
    empty surface too."""
    import framework_amd as fa

    f = oracle.field_gyroid((70, 53, 61), period=11.0, seed=4, noise_amp=0.2)
    fd = _dev(torch_cuda, f)
    st = torch_cuda.cuda.Stream()
    v, q, c = fa.run_device(fd, 0.0, stream=st)
    ref = oracle.run(f, 0.0)
    assert c["num_vertices"] == ref["num_vertices"] and c["num_active"] == ref["num_active"]
    assert_same_mesh(v.cpu().numpy(), q.cpu().numpy(), ref)
    vs, qs, _ = fa.run_device(fd, 0.0, num_smooth=2)
    vc, qc, _ = ctx.extract(fd, 0.0, num_smooth=2)
    assert torch_cuda.equal(qs, qc) and torch_cuda.equal(vs.view(torch_cuda.int32), vc.view(torch_cuda.int32))
    ve, qe, ce = fa.run_device(torch_cuda.zeros((9, 8, 7), device="cuda"), 0.5)
    assert ce["num_vertices"] == 0 and ve.shape == (0, 3) and qe.shape == (0, 4)


def test_deterministic(ctx, torch_cuda):
    f = ctx.field_gyroid((90, 90, 90), period=15.0, seed=3, noise_amp=0.1)
    torch_cuda.cuda.synchronize()
    v1, q1, _ = ctx.extract(f, 0.0)
    v2, q2, _ = ctx.extract(f, 0.0)
    assert torch_cuda.equal(q1, q2) and torch_cuda.equal(v1.view(torch_cuda.int32), v2.view(torch_cuda.int32))


def test_capacity_overflow_is_flagged(ctx, torch_cuda):
    torch = torch_cuda
    f = ctx.field_gyroid((60, 60, 60), period=12.0, seed=4, noise_amp=0.05)
    ctx.count(f, 0.0)
    torch.cuda.synchronize()
    full = ctx.read_counts()
    V, Q = full["num_vertices"], full["num_quads"]
    verts = torch.full((V // 2, 3), -7.0, device="cuda")
    quads = torch.full((Q // 2, 4), -7, dtype=torch.int32, device="cuda")
    guard_v = torch.full((1024,), 123.0, device="cuda")
    ctx.emit(f, verts, quads)
    torch.cuda.synchronize()
    st = ctx.read_counts()["status"]
    assert st & 1 and st & 2
    assert bool((guard_v == 123.0).all())


def test_blobs_512_full(ctx, torch_cuda):
    """Config 1 at full size: 512^3, 32 blobs, sigma U[8,32], amp U[-1,1], iso 0.1."""
    f = ctx.field_blobs((512, 512, 512), seed=0, nblobs=32)
    torch_cuda.cuda.synchronize()
    ref = check_against_oracle(ctx, f, 0.1, check_cells=True)
    assert ref["num_active"] > 100000


def _closed_form_counts(torch, f, iso, chunk=64):
    """Active cells (corners not all on one side) and quads (sign-changing
    grid edges whose four-cell ring lies in the grid), counted on the GPU
    from the field alone, plane chunk by plane chunk."""
    Z, Y, X = f.shape
    A = Q = 0
    for z0 in range(0, Z - 1, chunk):
        z1 = min(z0 + chunk, Z - 1)
        b = (f[z0:z1 + 1] < iso).to(torch.int8)
        s = (b[:-1, :-1, :-1] + b[:-1, :-1, 1:] + b[:-1, 1:, :-1] + b[:-1, 1:, 1:] +
             b[1:, :-1, :-1] + b[1:, :-1, 1:] + b[1:, 1:, :-1] + b[1:, 1:, 1:])
        A += int(((s > 0) & (s < 8)).sum())
        # x- and y-edges in point planes z in [z0, z1) with 1 <= z <= Z - 2;
        # z-edges from planes z in [z0, z1)
        lo, hi = max(z0, 1), min(z1 - 1, Z - 2)
        if hi >= lo:
            bb = b[lo - z0:hi - z0 + 1]
            Q += int((bb[:, 1:-1, 1:] != bb[:, 1:-1, :-1]).sum())  # x-edges, y in [1, Y-2]
            Q += int((bb[:, 1:, 1:-1] != bb[:, :-1, 1:-1]).sum())  # y-edges, x in [1, X-2]
        Q += int((b[1:, 1:-1, 1:-1] != b[:-1, 1:-1, 1:-1]).sum())  # z-edges, x, y in [1, dim-2]
        del b, s
    return A, Q


def test_gyroid_1024_surface_dense(ctx, torch_cuda):
    """1024^3 gyroid (period 24) + noise 0.05, iso 0: the surface-dense case.
    Too large for the oracle, so: active-cell and quad counts against closed
    forms from the field, quad indices valid and distinct, vertices in the
    box, a repeat run bitwise equal, and an 80^3 corner window against the
    oracle cell by cell (configuration, variant, vertex count and vertex bits;
    the oracle takes the whole grid's coordinate divisors for the window)."""
    torch = torch_cuda
    n = 1024
    f = ctx.field_gyroid((n, n, n), period=24.0, seed=0, noise_amp=0.05)
    torch.cuda.synchronize()
    verts, quads, cnt = ctx.extract(f, 0.0)
    A, Q = _closed_form_counts(torch, f, 0.0)
    assert cnt["num_active"] == A and cnt["num_quads"] == Q, (cnt, A, Q)
    assert A > 2e8
    V = cnt["num_vertices"]
    assert int(quads.min()) >= 0 and int(quads.max()) < V
    for a, b in ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)):
        assert bool((quads[:, a] != quads[:, b]).all())
    fin = torch.isfinite(verts)
    assert bool(fin.all(dim=1).any())
    vv = torch.where(fin, verts, torch.zeros_like(verts))
    # in the box up to float rounding of i / cells * cells and of the lerp
    assert float(vv.min()) >= -1e-3 and float(vv.max()) <= float(n - 1) + 1e-3
    cells = ctx.cells(cnt["num_active"], host=False)
    # repeat run into fresh buffers: bitwise equal
    v2 = torch.empty_like(verts)
    q2 = torch.empty_like(quads)
    ctx.extract_into(f, 0.0, v2, q2)
    torch.cuda.synchronize()
    assert torch.equal(q2, quads) and torch.equal(v2.view(torch.int32), verts.view(torch.int32))
    del v2, q2
    # corner window vs the oracle
    _window_check(torch, cells, verts, V, n, 80, f[:80, :80, :80].contiguous().cpu().numpy())


def _window_check(torch, cells, verts, V, n, W, win):
    """Cells of an n^3 grid's extraction (tensors of ctx.cells(host=False)
    and the vertex tensor) against the oracle on the grid's W^3 corner
    window (with the grid's coordinate divisors), for cells at least three
    cells from the window's upper faces."""
    dev = verts.device
    ref = oracle.run(win, 0.0, grid_cells=(n - 1, n - 1, n - 1), offset=(0, 0, 0))
    c = n - 1
    ci = cells["cell_index"].to(torch.int64) & 0xFFFFFFFF
    x, y, z = ci % c, (ci // c) % c, ci // (c * c)
    sel = torch.nonzero((x < W - 3) & (y < W - 3) & (z < W - 3)).flatten()
    vb_all = cells["vertex_begin"].to(torch.int64) & 0xFFFFFFFF
    vb_next = torch.cat([vb_all[1:], torch.tensor([V], device=dev, dtype=torch.int64)])
    g_key = (x[sel] + W * (y[sel] + W * z[sel])).cpu().numpy()
    g_cfg = cells["cfg"][sel].cpu().numpy()
    g_l2 = cells["lut2"][sel].cpu().numpy()
    g_vb = vb_all[sel].cpu().numpy()
    g_nv = (vb_next[sel] - vb_all[sel]).cpu().numpy()
    wc = W - 1
    rci = ref["cell_index"].astype(np.int64)
    rx, ry, rz = rci % wc, (rci // wc) % wc, rci // (wc * wc)
    rin = (rx < W - 3) & (ry < W - 3) & (rz < W - 3)
    r_key = (rx + W * (ry + W * rz))[rin]
    r_vb = ref["vertex_begin"].astype(np.int64)
    r_nv = np.append(r_vb[1:], ref["num_vertices"]) - r_vb
    assert len(r_key) > 1000
    assert np.array_equal(np.sort(g_key), np.sort(r_key))
    og, orf = np.argsort(g_key), np.argsort(r_key)
    assert np.array_equal(g_cfg[og], ref["cfg"][rin][orf])
    assert np.array_equal(g_l2[og], ((ref["info"] >> 7) & 1)[rin][orf])
    assert np.array_equal(g_nv[og], r_nv[rin][orf])
    gidx = np.concatenate([np.arange(b, b + k) for b, k in zip(g_vb[og], g_nv[og])])
    g_pos = verts[torch.from_numpy(gidx).to(dev)].cpu().numpy()
    r_pos = np.concatenate([ref["vertices"][b:b + k] for b, k in zip(r_vb[rin][orf], r_nv[rin][orf])])
    # the window ran with the whole grid's coordinate divisors: bit for bit
    assert np.array_equal(g_pos.view(np.uint32), r_pos.view(np.uint32))


def test_hand_derived_winding_on_gpu(ctx, torch_cuda):
    """The hand-derived single-quad rings (tests/test_oracle.py) and the
    outward winding on sphere distance fields, through the HIP path."""
    from tests.test_oracle import WINDING_CASES, _one_edge_field, _signed_volume

    for edge, (p, ring) in WINDING_CASES.items():
        for inside in (-1.0, 1.0):
            verts, quads, cnt = ctx.extract(_dev(torch_cuda, _one_edge_field(p, inside)), 0.0)
            cells = ctx.cells(cnt["num_active"])
            first = dict(zip(cells["cell_index"].tolist(), cells["vertex_begin"].tolist()))
            m = [first[c] for c in ring]
            want = m if inside < 0 else [m[0], m[3], m[2], m[1]]
            assert quads.cpu().numpy().tolist() == [want], (edge, inside)
    r = 21.7
    f = oracle.field_sphere((64, 60, 62), (31.5, 29.0, 30.2), r)
    for sign in (1.0, -1.0):
        verts, quads, _ = ctx.extract(_dev(torch_cuda, sign * f), 0.0)
        vol = _signed_volume(verts.cpu().numpy(), quads.cpu().numpy())
        assert np.sign(vol) == sign and abs(abs(vol) / (4.0 / 3.0 * np.pi * r ** 3) - 1.0) < 0.02, (sign, vol)
