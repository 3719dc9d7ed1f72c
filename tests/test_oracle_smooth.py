"""The oracle's smoothing passes (num_smooth > 0, dmc.cu:1868-1981 and
2157-2169) pinned on the CPU: planar surfaces are fixed points, and an
independent float64 restatement of one smoothing step (its own projection,
signs, split rule with arccos angles and barycentric test) predicts which
edge points move and where."""
from __future__ import annotations

import numpy as np
import pytest

import oracle


def _planes():
    n = 24
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    out = []
    for a, b, c, d in [(0.0, 0.0, 1.0, 11.3), (0.3, -0.2, 1.0, 7.1), (1.0, 0.9, 0.4, 21.7), (-0.5, 1.0, 0.25, 3.3)]:
        out.append((a * x + b * y + c * z - d).astype(np.float32))
    return out


@pytest.mark.parametrize("k", range(4))
def test_planes_are_fixed_points(k):
    f = _planes()[k]
    r0 = oracle.run(f, 0.0)
    r2 = oracle.run(f, 0.0, num_smooth=2)
    assert np.array_equal(r0["quads"], r2["quads"])
    fin = np.isfinite(r0["vertices"])
    assert np.array_equal(fin, np.isfinite(r2["vertices"]))
    assert np.abs(r2["vertices"][fin] - r0["vertices"][fin]).max() < 1e-4


def _ijk(i, n, rng, mn):
    return (np.float32(i) / np.float32(n) * np.float32(rng) + np.float32(mn)).astype(np.float32)


def _quad_owners(r, dims):
    """(cell, owned edge l, edge vertex index, i, j, k) of every quad, in quad order."""
    X, Y, Z = dims
    cx, cy, cz = X - 1, Y - 1, Z - 1
    ci = r["cell_index"].astype(np.int64)
    i, j, k = ci % cx, (ci // cx) % cy, ci // (cx * cy)
    info = r["info"].astype(np.int64)
    vb = r["vertex_begin"].astype(np.int64)
    nv = np.append(vb[1:], r["num_vertices"]) - vb
    niso = nv - ((info & 1) + ((info >> 1) & 1) + ((info >> 2) & 1))
    rows = []
    for l, cond in enumerate([(i + 1 < cx) & (j + 1 < cy), (i + 1 < cx) & (k + 1 < cz), (j + 1 < cy) & (k + 1 < cz)]):
        has = ((info >> l) & 1).astype(bool) & cond
        below = sum(((info >> b) & 1) for b in range(l)) if l else 0
        ev = vb + niso + below
        a = np.nonzero(has)[0]
        rows.append(np.stack([a, np.full_like(a, l), ev[a], i[a], j[a], k[a]], axis=1))
    allr = np.concatenate(rows)
    return allr[np.lexsort((allr[:, 1], allr[:, 0]))]


def _predict(r0, dims):
    """One smoothing step in float64 from the unsmoothed mesh: predicted new
    edge points (nan rows: unchanged) and a mask of quads too close to a
    decision boundary to predict."""
    X, Y, Z = dims
    own = _quad_owners(r0, dims)
    V = r0["vertices"].astype(np.float64)
    q = r0["quads"].astype(np.int64)
    assert len(own) == len(q)
    pred = np.full((len(q), 3), np.nan)
    unsure = np.zeros(len(q), bool)
    axes = {0: (0, 1), 1: (0, 2), 2: (1, 2)}
    for t in range(len(q)):
        a, l, ev, i, j, k = own[t]
        c0, c1 = axes[int(l)]
        P = V[q[t]][:, [c0, c1]]
        cr = []
        for p in range(4):
            m, lf, rt = P[p], P[(p + 3) % 4], P[(p + 1) % 4]
            cr.append((rt[0] - m[0]) * (lf[1] - m[1]) - (lf[0] - m[0]) * (rt[1] - m[1]))
        cr = np.array(cr)
        if np.any(np.abs(cr) < 1e-6):
            unsure[t] = True
            continue
        pos = cr >= 0
        if pos.sum() == 2:
            continue
        if pos.all() or (~pos).all():
            ang = []
            for p in range(4):
                u, w = P[(p + 3) % 4] - P[p], P[(p + 1) % 4] - P[p]
                ang.append(np.arccos(np.clip(u @ w / np.linalg.norm(u) / np.linalg.norm(w), -1, 1)))
            ang = np.array(ang)
            s0 = int(np.argmax(ang))
            if np.sort(ang)[-1] - np.sort(ang)[-2] < 1e-5:
                unsure[t] = True
        else:
            s0 = int(np.nonzero(pos if pos.sum() == 1 else ~pos)[0][0])
        s1 = (s0 + 2) % 4
        coords = [_ijk(i + 1, X - 1, X - 1, 0), _ijk(j + 1, Y - 1, Y - 1, 0), _ijk(k + 1, Z - 1, Z - 1, 0)]
        o = np.array([coords[c0], coords[c1]], np.float64)
        for ta, tb, tc in ((s0, (s0 + 1) % 4, s1), (s1, (s1 + 1) % 4, s0)):
            M = np.array([[P[tb][0] - P[ta][0], P[tc][0] - P[ta][0]], [P[tb][1] - P[ta][1], P[tc][1] - P[ta][1]]])
            try:
                be, ga = np.linalg.solve(M, o - P[ta])
            except np.linalg.LinAlgError:
                unsure[t] = True
                break
            if min(abs(be + 1e-4), abs(ga + 1e-4), abs(be + ga - 1 - 1e-4)) < 1e-4:
                unsure[t] = True
            if be > -1e-4 and ga > -1e-4 and be + ga < 1 + 1e-4:
                pred[t] = (1 - be - ga) * V[q[t][ta]] + be * V[q[t][tb]] + ga * V[q[t][tc]]
                break
    return own, pred, unsure


@pytest.mark.parametrize("case", ["sphere", "gyroid", "blobs"])
def test_one_step_against_a_float64_restatement(case):
    if case == "sphere":
        dims = (30, 33, 28)
        f = oracle.field_sphere(dims, (14.2, 16.1, 13.7), 10.6)
    elif case == "gyroid":
        dims = (26, 22, 24)
        f = oracle.field_gyroid(dims, 9.0, 3, 0.2)
    else:
        dims = (28, 30, 26)
        f = oracle.field_blobs(dims, 5, 7, sigma=(3.0, 8.0))
    iso = 0.0 if case != "blobs" else 0.1
    r0 = oracle.run(f, iso)
    r1 = oracle.run(f, iso, num_smooth=1)
    assert np.array_equal(r0["quads"], r1["quads"])
    own, pred, unsure = _predict(r0, dims)
    ev = own[:, 2]
    want = ~np.isnan(pred[:, 0])  # the restatement moves these (possibly onto themselves)
    sure = ~unsure
    assert sure.sum() > 0.9 * len(sure)
    got = r1["vertices"][ev].astype(np.float64)
    still = sure & ~want
    assert np.array_equal(r1["vertices"][ev][still], r0["vertices"][ev][still])
    ok = sure & want
    assert ok.sum() > 10
    assert np.abs(got[ok] - pred[ok]).max() < 1e-3
    moved = ~np.all(r1["vertices"][ev] == r0["vertices"][ev], axis=1)
    assert moved.sum() > 10
    # a moved edge point stays on its edge's line: the two coordinates across
    # the edge are the edge's trace
    X, Y, Z = dims
    for l, (c0, c1) in {0: (0, 1), 1: (0, 2), 2: (1, 2)}.items():
        sel = moved & (own[:, 1] == l)
        i, j, k = own[sel, 3], own[sel, 4], own[sel, 5]
        tr = [_ijk(i + 1, X - 1, X - 1, 0), _ijk(j + 1, Y - 1, Y - 1, 0), _ijk(k + 1, Z - 1, Z - 1, 0)]
        assert np.abs(got[sel][:, c0] - tr[c0]).max(initial=0) < 1e-3
        assert np.abs(got[sel][:, c1] - tr[c1]).max(initial=0) < 1e-3


def test_iso_vertices_follow_the_moved_edge_points():
    """After smoothing, every iso vertex is again the mean of its group's
    stored edge points (the re-average pass), checked with the owner map."""
    dims = (22, 24, 20)
    f = oracle.field_sphere(dims, (10.3, 11.9, 9.4), 7.7)
    r = oracle.run(f, 0.0, num_smooth=2)
    r0 = oracle.run(f, 0.0)
    own = _quad_owners(r, dims)
    # each quad member is an iso vertex that averages its edges; the edge of a
    # quad is one of them, so a moved edge point moves the four ring vertices
    ev = own[:, 2]
    moved = ~np.all(r["vertices"][ev] == r0["vertices"][ev], axis=1)
    ring = r["quads"][moved].ravel()
    assert moved.any()
    assert not np.all(r["vertices"][ring] == r0["vertices"][ring])
