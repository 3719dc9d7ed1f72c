"""Pins for the CPU oracle (CPU only).

Each check is independent of the oracle's own code:
  * classification, 2B/3B correction, vertex/quad counts and vertex_begin
    offsets against a separate numpy formulation (brute force, small grids);
  * quad count against its closed form (sign-changing interior grid edges);
  * every quad references iso vertices of its four ring cells;
  * vertex positions against closed forms (linear fields: all points lie on
    the plane), the FP64 twin, and mirror symmetry;
  * mesh invariants on generic spheres/tori (closed, consistently oriented
    2-manifold, Euler characteristic 2 / 0);
  * hand-built two-cell grids for all 36 2B/3B configurations;
  * the defined behaviours: ties (72 tie edges on the 64^3 sphere, as measured
    on the reference layout), NaN corners, empty surfaces, unset iso vertices.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle

CORNERS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
DIR_OFF = {0: (1, 0, 0), 1: (-1, 0, 0), 2: (0, 1, 0), 3: (0, -1, 0), 4: (0, 0, 1), 5: (0, 0, -1)}
FACE_DIR = [5, 3, 0, 2, 1, 4]


@pytest.fixture(scope="module")
def luts():
    b = oracle.luts_bytes()
    lab = np.frombuffer(b[:6144], dtype=np.uint8).reshape(2, 256, 12)
    nv = np.frombuffer(b[6144:6656], dtype=np.uint8).reshape(2, 256)
    amb = np.frombuffer(b[6656:6912], dtype=np.uint8)
    return lab, nv, amb


def numpy_model(f, iso, luts):
    """Counts and offsets from array operations (no per-cell loops)."""
    lab, nv, amb = luts
    s = (f < iso)
    Z, Y, X = f.shape
    cz, cy, cx = Z - 1, Y - 1, X - 1
    cfg = np.zeros((cz, cy, cx), dtype=np.int64)
    for b, (dx, dy, dz) in enumerate(CORNERS):
        cfg |= s[dz:dz + cz, dy:dy + cy, dx:dx + cx].astype(np.int64) << b
    active = (cfg != 0) & (cfg != 255)
    bit = lambda c, i: (c >> i) & 1
    b6 = bit(cfg, 2) ^ bit(cfg, 6)
    b9 = bit(cfg, 5) ^ bit(cfg, 6)
    b10 = bit(cfg, 7) ^ bit(cfg, 6)
    # 2B/3B: neighbour across the listed face, inside the grid, also 2B/3B
    face = amb[cfg]
    lut2 = np.zeros_like(active)
    is_amb = amb != 0xFF
    zz, yy, xx = np.meshgrid(np.arange(cz), np.arange(cy), np.arange(cx), indexing="ij")
    for f_idx in range(6):
        ox, oy, oz = DIR_OFF[FACE_DIR[f_idx]]
        sel = face == f_idx
        nx, ny, nz = xx + ox, yy + oy, zz + oz
        inside = (nx >= 0) & (ny >= 0) & (nz >= 0) & (nx < cx) & (ny < cy) & (nz < cz)
        ncfg = cfg[np.clip(nz, 0, cz - 1), np.clip(ny, 0, cy - 1), np.clip(nx, 0, cx - 1)]
        lut2 |= sel & inside & is_amb[ncfg]
    nverts = np.where(lut2, nv[1][cfg], nv[0][cfg]) + b6 + b9 + b10
    xin, yin, zin = xx + 1 < cx, yy + 1 < cy, zz + 1 < cz
    nq = (b6 & xin & yin) + (b9 & xin & zin) + (b10 & yin & zin)
    flat = active.reshape(-1)
    nvf = nverts.reshape(-1)[flat]
    vb = np.concatenate([[0], np.cumsum(nvf)[:-1]]) if nvf.size else nvf
    return {
        "cell_index": np.nonzero(flat)[0].astype(np.uint32),
        "cfg": cfg.reshape(-1)[flat].astype(np.uint8),
        "lut2": lut2.reshape(-1)[flat].astype(np.uint8),
        "vertex_begin": vb.astype(np.uint32),
        "num_vertices": int(nvf.sum()),
        "num_quads": int(nq.reshape(-1)[flat].sum()),
    }


def closed_form_quads(f, iso):
    """Sign-changing grid edges whose four surrounding cells all exist."""
    s = f < iso
    qz = (s[1:, 1:-1, 1:-1] != s[:-1, 1:-1, 1:-1]).sum()
    qy = (s[1:-1, 1:, 1:-1] != s[1:-1, :-1, 1:-1]).sum()
    qx = (s[1:-1, 1:-1, 1:] != s[1:-1, 1:-1, :-1]).sum()
    return int(qx + qy + qz)


def random_fields():
    rng = np.random.default_rng(42)
    out = []
    for _ in range(6):
        Z, Y, X = (int(v) for v in rng.integers(3, 14, size=3))
        out.append((rng.standard_normal((Z, Y, X)).astype(np.float32), float(rng.uniform(-0.3, 0.3))))
    out.append((oracle.field_gyroid((23, 19, 17), 6.0, 3, 0.6), 0.0))
    out.append((oracle.field_blobs((30, 26, 22), 5, 6, sigma=(2.0, 6.0)), 0.1))
    return out


@pytest.mark.parametrize("idx", range(8))
def test_counts_and_offsets_match_numpy_model(luts, idx):
    f, iso = random_fields()[idx]
    r = oracle.run(f, iso)
    m = numpy_model(f, iso, luts)
    assert r["num_anomalies"] == 0
    assert np.array_equal(r["cell_index"], m["cell_index"])
    assert np.array_equal(r["cfg"], m["cfg"])
    assert np.array_equal((r["info"] >> 7) & 1, m["lut2"])
    assert np.array_equal(r["vertex_begin"], m["vertex_begin"])
    assert r["num_vertices"] == m["num_vertices"]
    assert r["num_quads"] == m["num_quads"] == closed_form_quads(f, iso)


def test_quads_reference_iso_vertices_of_their_ring(luts):
    _, nv, _ = luts
    f, iso = random_fields()[6]
    r = oracle.run(f, iso)
    Z, Y, X = f.shape
    cx, cy = X - 1, Y - 1
    nviso = np.where((r["info"] >> 7) & 1, nv[1][r["cfg"]], nv[0][r["cfg"]])
    owner = np.full(r["num_vertices"], -1, dtype=np.int64)
    iso_slot = np.zeros(r["num_vertices"], dtype=bool)
    for a in range(r["num_active"]):
        vb = r["vertex_begin"][a]
        owner[vb:vb + nviso[a]] = r["cell_index"][a]
        iso_slot[vb:vb + nviso[a]] = True
    q = r["quads"]
    assert q.min() >= 0 and q.max() < r["num_vertices"]
    assert iso_slot[q].all()
    cells = owner[q]
    ci, cj, ck = cells % cx, (cells // cx) % cy, cells // (cx * cy)
    # the four cells of a quad span exactly a 2x2 block around one edge
    span = [(c.max(axis=1) - c.min(axis=1)) for c in (ci, cj, ck)]
    assert all(((s == 0) | (s == 1)).all() for s in span)
    assert (np.sum([s for s in span], axis=0) == 2).all()


def test_linear_field_vertices_lie_on_the_plane():
    a, b, c, d = 0.37, -0.81, 0.52, -3.1
    Z, Y, X = 21, 19, 23
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    f = (a * x + b * y + c * z + d).astype(np.float64)
    f32 = f.astype(np.float32)
    r = oracle.run(f32, 0.0, fp64=True)
    v = r["vertices"][~np.isnan(r["vertices"]).any(1)]
    resid = a * v[:, 0] + b * v[:, 1] + c * v[:, 2] + d
    assert np.abs(resid).max() < 2e-5  # float32 field and coordinates
    v64 = r["vertices64"][~np.isnan(r["vertices64"]).any(1)]
    resid64 = a * v64[:, 0] + b * v64[:, 1] + c * v64[:, 2] + d
    assert np.abs(resid64).max() < 2e-6  # only the float32 field rounding remains


def test_fp64_twin_error_bound():
    f = oracle.field_sphere((64, 64, 64), (32.0, 32.0, 32.0), 20.0)
    r = oracle.run(f, 0.0, fp64=True)
    ok = ~np.isnan(r["vertices"]).any(1)
    err = np.abs(r["vertices"][ok].astype(np.float64) - r["vertices64"][ok])
    assert err.max() < 1e-5  # voxel units, coordinates < 64
    assert err.max() > 0  # the twin really is a different computation


def test_mirror_symmetry():
    f = oracle.field_sphere((33, 30, 28), (16.0, 14.3, 13.7), 9.4)
    fm = np.ascontiguousarray(f[:, :, ::-1])
    r, rm = oracle.run(f, 0.0), oracle.run(fm, 0.0)
    assert r["num_quads"] == rm["num_quads"]
    a = np.sort(r["vertices"][:, 0])
    b = np.sort(32.0 - rm["vertices"][:, 0])
    assert np.abs(a - b).max() < 1e-4


def mesh_invariants(quads):
    edges = np.concatenate([np.stack([quads[:, i], quads[:, (i + 1) % 4]], 1) for i in range(4)])
    directed = {tuple(e) for e in edges.tolist()}
    assert len(directed) == len(edges), "a directed edge appears twice (orientation)"
    und = np.sort(edges, axis=1)
    _, counts = np.unique(und, axis=0, return_counts=True)
    assert (counts == 2).all(), "not a closed 2-manifold"
    V = len(np.unique(quads))
    return V - len(counts) + len(quads)


@pytest.mark.parametrize("centre,radius", [((20.3, 19.7, 21.1), 13.2), ((15.5, 16.2, 17.9), 9.8)])
def test_sphere_is_closed_oriented_genus0(centre, radius):
    f = oracle.field_sphere((42, 40, 44), centre, radius)
    r = oracle.run(f, 0.0)
    assert r["num_tie_edges"] == 0
    assert mesh_invariants(r["quads"]) == 2


def test_torus_has_euler_characteristic_zero():
    Z, Y, X = 40, 44, 46
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    R, rr = 12.3, 5.1
    q = np.sqrt((x - 22.7) ** 2 + (y - 21.4) ** 2) - R
    f = (np.sqrt(q ** 2 + (z - 19.6) ** 2) - rr).astype(np.float32)
    r = oracle.run(f, 0.0)
    assert mesh_invariants(r["quads"]) == 0


def test_sphere64_reference_counts():
    """Config 0 (64^3 sphere r=20 at (32,32,32), iso 0): 30 exact zeros give 72
    owned edges the reference would call bipolar although both corner bits
    agree; here they are not bipolar."""
    f = oracle.field_sphere((64, 64, 64), (32.0, 32.0, 32.0), 20.0)
    assert int((f == 0).sum()) == 30
    r = oracle.run(f, 0.0)
    assert (r["num_active"], r["num_vertices"], r["num_quads"], r["num_tie_edges"]) == (7472, 14942, 7470, 72)
    assert r["num_anomalies"] == 0 and r["num_unset_vertices"] == 0
    assert r["num_quads"] == closed_form_quads(f, 0.0)


def two_cell_grid(cfg_a, d, other_bits):
    """Cells A and B adjacent in direction d (A -> B); A has configuration
    cfg_a, B's four corners away from the shared face come from other_bits."""
    ox, oy, oz = DIR_OFF[d]
    dims = (2 + abs(ox), 2 + abs(oy), 2 + abs(oz))
    X, Y, Z = dims
    inside = np.zeros((Z, Y, X), dtype=bool)
    ax, ay, az = (max(-ox, 0), max(-oy, 0), max(-oz, 0))
    for b, (dx, dy, dz) in enumerate(CORNERS):
        inside[az + dz, ay + dy, ax + dx] = (cfg_a >> b) & 1
    bx, by, bz = ax + ox, ay + oy, az + oz
    k = 0
    for dx, dy, dz in CORNERS:
        px, py, pz = bx + dx, by + dy, bz + dz
        on_face = (px - ax in (0, 1)) and (py - ay in (0, 1)) and (pz - az in (0, 1))
        if not on_face:
            inside[pz, py, px] = (other_bits >> k) & 1
            k += 1
    f = np.where(inside, -1.0, 1.0).astype(np.float32)
    a_index = ax + (X - 1) * (ay + (Y - 1) * az)
    b_index = bx + (X - 1) * (by + (Y - 1) * bz)
    return f, a_index, b_index


def test_2b3b_pairs_all_36_configs(luts):
    _, _, amb = luts
    cfgs = [c for c in range(256) if amb[c] != 0xFF]
    assert len(cfgs) == 36
    n_set = 0
    for c in cfgs:
        d = FACE_DIR[amb[c]]
        for other in range(16):
            f, ia, ib = two_cell_grid(c, d, other)
            r = oracle.run(f, 0.0)
            idx = {int(ci): n for n, ci in enumerate(r["cell_index"])}
            assert r["cfg"][idx[ia]] == c
            cb = int(r["cfg"][idx[ib]])
            expect = amb[cb] != 0xFF
            assert bool((r["info"][idx[ia]] >> 7) & 1) == expect, (hex(c), other, hex(cb))
            n_set += expect
    assert n_set > 0


def test_nan_corners_are_outside():
    rng = np.random.default_rng(3)
    f = rng.standard_normal((12, 11, 13)).astype(np.float32)
    f[3, 4, 5] = np.nan
    r = oracle.run(f, 0.0)
    assert r["num_anomalies"] == 0
    g = f.copy()
    g[3, 4, 5] = 1.0  # same sign bit as NaN ("not < iso")
    r2 = oracle.run(g, 0.0)
    assert np.array_equal(r["cfg"], r2["cfg"]) and np.array_equal(r["quads"], r2["quads"])


def test_empty_surface():
    r = oracle.run(np.ones((5, 6, 7), dtype=np.float32), 0.0)
    assert r["num_active"] == 0 and r["num_quads"] == 0 and r["vertices"].shape == (0, 3)


def test_unset_iso_vertex_is_nan():
    """Cell (0,0,0) with only corner 0 inside: its group {e0,e3,e4} has every
    owner outside the grid, so the reference leaves it uninitialised."""
    f = np.ones((4, 4, 4), dtype=np.float32)
    f[0, 0, 0] = -1.0
    r = oracle.run(f, 0.0)
    assert r["num_unset_vertices"] == 1
    assert np.isnan(r["vertices"][0]).all() and r["vertices"][0].view(np.uint32)[0] == 0x7FC00000
    assert r["num_quads"] == 0 or not (r["quads"] == 0).any()


def test_threads_do_not_change_results():
    """The per-cell loops run on several threads; the compact numbering and
    the prefix sums stay in cell order, so every output array is the same
    bitwise as with one thread (ties, 2B/3B cells, smoothing included)."""
    rng = np.random.default_rng(4)
    noise = rng.standard_normal((23, 31, 37)).astype(np.float32)
    sphere = oracle.field_sphere((64, 64, 64), (32.0, 32.0, 32.0), 20.0)
    gyro = oracle.field_gyroid((40, 36, 44), period=9.0, seed=2, noise_amp=0.3)
    cases = [(noise, 0.1, 0), (sphere, 0.0, 0), (gyro, 0.0, 2)]
    try:
        for field, iso, ns in cases:
            oracle.set_threads(1)
            one = oracle.run(field, iso, fp64=True, num_smooth=ns)
            oracle.set_threads(8)
            many = oracle.run(field, iso, fp64=True, num_smooth=ns)
            for key, val in one.items():
                if isinstance(val, np.ndarray):
                    assert np.array_equal(val.view(np.uint8), many[key].view(np.uint8)), key
                else:
                    assert val == many[key], key
    finally:
        oracle.set_threads(1)


def test_window_mode_reproduces_the_whole_grid():
    """A window of a grid, extracted with the grid's coordinate divisors and
    its own offset, gives each interior cell (one cell or more from the
    window's lower faces, three or more from its upper faces) the whole grid's
    configuration, variant, vertex count and vertex bits.  Pins the window
    mode by the whole-grid extraction, not by its own formula."""
    full = oracle.field_gyroid((61, 57, 53), 9.0, 21, 0.3)
    Zf, Yf, Xf = full.shape
    gc = (Xf - 1, Yf - 1, Zf - 1)
    box = ((-3.0, 1.0, 2.0), (7.5, 4.0, 9.25))
    ref = oracle.run(full, 0.02, *box)
    o = (9, 5, 13)
    W = (33, 29, 31)
    win = np.ascontiguousarray(full[o[2]:o[2] + W[2], o[1]:o[1] + W[1], o[0]:o[0] + W[0]])
    got = oracle.run(win, 0.02, *box, grid_cells=gc, offset=o)

    def by_cell(r, dims, shift):
        cx, cy = dims[0] - 1, dims[1] - 1
        ci = r["cell_index"].astype(np.int64)
        x, y, z = ci % cx + shift[0], (ci // cx) % cy + shift[1], ci // (cx * cy) + shift[2]
        vb = r["vertex_begin"].astype(np.int64)
        nv = np.append(vb[1:], r["num_vertices"]) - vb
        return {(int(a), int(b), int(c)): (int(r["cfg"][n]), int(r["info"][n]) >> 7, r["vertices"][vb[n]:vb[n] + nv[n]])
                for n, (a, b, c) in enumerate(zip(x, y, z))}

    G = by_cell(ref, (Xf, Yf, Zf), (0, 0, 0))
    Wc = by_cell(got, W, o)
    inner = [c for c in Wc if all(o[d] + 1 <= c[d] <= o[d] + W[d] - 4 for d in range(3))]
    assert len(inner) > 2000
    for c in inner:
        cfg, l2, v = Wc[c]
        rcfg, rl2, rv = G[c]
        assert (cfg, l2) == (rcfg, rl2), c
        assert v.shape == rv.shape and np.array_equal(v.view(np.uint32), rv.view(np.uint32)), c
    # and without a window the mode is the plain extraction
    same = oracle.run(full, 0.02, *box, grid_cells=gc, offset=(0, 0, 0))
    assert np.array_equal(same["quads"], ref["quads"])
    assert np.array_equal(same["vertices"].view(np.uint32), ref["vertices"].view(np.uint32))


# One crossing grid edge per case, derived by hand from the reference's ring
# table (circular_edge_lut, dmc.cu:653-658: edge 6's ring is the cells at
# (x, y) offsets (0,0), (1,0), (1,1), (0,1), edge 9's at (z, x) offsets and
# edge 10's at (y, z) offsets in that order) and its orientation rule
# (use_ccw = v[lower end of the edge] <= iso, dmc.cu:1199-1216; otherwise the
# ring is taken 0, 3, 2, 1, dmc.cu:1634-1644).  On a 3^3-point grid with +1
# everywhere and -1 at one point P, the four cells around P's one interior
# edge are active with one iso vertex each; the owner (cell 0) also holds the
# edge's crossing point, so the cells' first vertices are 0, 2, 3, 4 in cell
# order, and the single quad is (owner, ring 1, ring 2, ring 3) by those
# numbers.  Flipping the field's sign flips the orientation.
WINDING_CASES = {
    # edge: (point P as (x, y, z), ring cells in ccw order as cell indices)
    6: ((1, 1, 0), [0, 1, 3, 2]),
    9: ((1, 0, 1), [0, 4, 5, 1]),
    10: ((0, 1, 1), [0, 2, 6, 4]),
}


def _one_edge_field(p, inside_value):
    f = np.full((3, 3, 3), -inside_value, np.float32)
    f[p[2], p[1], p[0]] = inside_value
    return f


@pytest.mark.parametrize("edge", [6, 9, 10])
def test_hand_derived_ring_and_winding(edge):
    p, ring = WINDING_CASES[edge]
    for inside in (-1.0, 1.0):
        ref = oracle.run(_one_edge_field(p, inside), 0.0)
        assert ref["num_active"] == 4 and ref["num_vertices"] == 5 and ref["num_quads"] == 1
        assert sorted(ref["cell_index"].tolist()) == sorted(ring)
        first = dict(zip(ref["cell_index"].tolist(), ref["vertex_begin"].tolist()))
        m = [first[c] for c in ring]  # the ring's iso vertices in ccw order
        # v at the edge's lower end is the -1/+1 point: ccw when it is <= iso
        want = m if inside < 0 else [m[0], m[3], m[2], m[1]]
        assert ref["quads"].tolist() == [want], (edge, inside, ref["quads"])


def _signed_volume(verts, quads):
    """Volume enclosed by the triangles (q0,q1,q2),(q2,q3,q0) (dmc.cu:2039-2045),
    by the divergence theorem: positive when they face outward (right-hand
    normals away from the enclosed region)."""
    v = verts.astype(np.float64)
    t = np.concatenate([quads[:, [0, 1, 2]], quads[:, [2, 3, 0]]])
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)


def test_winding_faces_increasing_field():
    """The rules above make every quad's right-hand normal point toward
    increasing field values, i.e. out of the region below the isovalue: on a
    signed distance field of a sphere the enclosed volume is positive and
    close to 4/3 pi r^3; on the negated field (the sphere's outside is now
    below the isovalue) it is negative.  A global flip, a misread of the
    use_ccw comparison or of which edge end is tested fails this."""
    r = 11.3
    f = oracle.field_sphere((40, 36, 38), (19.2, 17.6, 18.4), r)
    exact = 4.0 / 3.0 * np.pi * r ** 3
    for sign in (1.0, -1.0):
        ref = oracle.run(sign * f, 0.0)
        vol = _signed_volume(ref["vertices"], ref["quads"])
        assert abs(abs(vol) / exact - 1.0) < 0.02, (sign, vol, exact)
        assert np.sign(vol) == sign, (sign, vol)
