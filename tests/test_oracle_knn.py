"""Pins of the CPU restatement's query stage against brute force.

KNN (src/tree.hpp:3198-3251): candidate sets S(r), start and stop radii and
the K selected are checked against a vectorised numpy restatement of the
walk for every query; where the K-th candidate distance lies below r* - m
(m bounds how far a particle sits outside its node box) the K nearest of
S(r*) are the K nearest overall and are also checked against an exhaustive
search (distances in the reference's ((dx^2 + dy^2) + dz^2) order).
Densities (src/tree.hpp:3406-3491) are recomputed with the kernels' closed
forms from the restated lists; HOP's densest neighbours (src/tree.hpp:
3711-3760) from brute-force lists, and peaks by pointer chasing."""
import math

import numpy as np
import pytest

import oracle_py as O
from _cases import dense_case

_cache = {}


def _tree(name):
    if name not in _cache:
        snaps = {
            "median": lambda: dense_case(8192, 6, 2, a=3e-4),
            "qs": lambda: dense_case(8192, 5, 7, a=3e-4, nx=96),
        }
        snap = snaps[name]()
        O.set_threads(4)
        t = O.run_pipeline(snap)
        _cache[name] = (snap, t)
    return _cache[name]


def brute_d2(q, pos):
    d = q[:, None, :] - pos[None, :, :]
    return (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]


def brute_knn(q, pos, K, chunk=512):
    idx = np.empty((len(q), K), np.int64)
    d2 = np.empty((len(q), K))
    for s in range(0, len(q), chunk):
        dd = brute_d2(q[s:s + chunk], pos)
        ii = np.arange(pos.shape[0])
        order = np.lexsort((np.broadcast_to(ii, dd.shape), dd), axis=1)[:, :K]
        idx[s:s + chunk] = order
        d2[s:s + chunk] = np.take_along_axis(dd, order, 1)
    return idx, d2


def pinned(ok, d2_bf, margin):
    """queries whose K-th distance lies inside r* - margin and is not tied
    with the (K+1)-th overall"""
    rs = ok["rstop"]
    kth = ok["d2"][:, -1]
    inside = np.sqrt(kth) < rs - margin
    return inside & (ok["ndrop"] == 0)


def _margin(t):
    c, hw = t.root()
    return 2.5 * (2 * hw / 2.0 ** 31) * 4


def numpy_candidates(nodes, q, r):
    """S(r) by a vectorised restatement of the walk (src/tree.hpp:3001-3049,
    3177-3193): a node is visited when every ancestor intersects the sphere
    without being contained and is not a leaf; a visited node that is
    contained, or an intersecting leaf, contributes all its particles."""
    c, hw = nodes["center"], nodes["hw"]
    diag = math.sqrt(3.0)
    m = np.abs(c - q) + (hw * diag)[:, None]
    cont = (m[:, 0] * m[:, 0] + m[:, 1] * m[:, 1]) + m[:, 2] * m[:, 2] < r * r
    dq = c - q
    c2c = (dq[:, 0] * dq[:, 0] + dq[:, 1] * dq[:, 1]) + dq[:, 2] * dq[:, 2]
    tmp = hw * diag + r
    md = np.zeros(len(hw))
    for k in range(3):
        lo, hi = c[:, k] - hw, c[:, k] + hw
        md = md + np.where(q[k] < lo, (q[k] - lo) * (q[k] - lo), np.where(q[k] > hi, (q[k] - hi) * (q[k] - hi), 0.0))
    inter = ~(c2c > tmp * tmp) & ((c2c < r) | (c2c < hw) | (md <= r * r))
    leaf = nodes["num_daughters"] == 0
    take = cont | (inter & leaf)
    desc = ~cont & inter & ~leaf
    vis = np.zeros(len(hw), bool)
    vis[0] = True
    par = nodes["parent"].astype(np.int64)
    for lv in np.unique(nodes["level"])[1:]:
        ids = np.flatnonzero(nodes["level"] == lv)
        vis[ids] = vis[par[ids]] & desc[par[ids]]
    sel = np.flatnonzero(vis & take)
    if len(sel) == 0:
        return np.zeros(0, np.int64)
    return np.concatenate([np.arange(nodes["begin"][i], nodes["end"][i]) for i in sel])


def test_knn_candidate_sets_restated():
    """Candidate sets, stop radii and the K selected, against the numpy
    restatement of the walk for every query (ties at the K-th distance
    excepted)."""
    snap, t = _tree("median")
    K = snap.params["knn"]
    pos = t.particles()["pos"]
    nodes = t.nodes()
    leaves = nodes["leaf_nodes"][np.argsort(nodes["begin"][nodes["leaf_nodes"]])]
    lb = nodes["begin"][leaves]
    rng = np.random.default_rng(3)
    q_idx = rng.choice(t.n, 800, replace=False)
    ok = t.knn(pos[q_idx], K)
    grown = 0
    for r_, i in enumerate(q_idx):
        q = pos[i]
        # start radius: half width of the first node on the leaf-to-root path
        # holding >= K/4 particles (src/tree.hpp:3206-3219)
        nd = int(leaves[np.searchsorted(lb, i, side="right") - 1])
        assert nodes["begin"][nd] <= i < nodes["end"][nd]
        while nodes["end"][nd] - nodes["begin"][nd] < K // 4:
            nd = int(nodes["parent"][nd])
        r0 = math.sqrt(nodes["hw"][nd] * nodes["hw"][nd])
        rs = ok["rstop"][r_]
        k = round(math.log2(rs / r0))
        assert rs == r0 * 2.0 ** k and k >= 0
        S = numpy_candidates(nodes, q, rs)
        assert len(S) == ok["ncand"][r_] and len(S) >= K
        if k > 0:
            grown += 1
            assert len(numpy_candidates(nodes, q, rs / 2)) < K
        d = pos[S] - q
        dd = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        order = np.lexsort((S, dd))
        if ok["ndrop"][r_] == 0:
            assert np.array_equal(np.sort(S[order[:K]]), np.sort(ok["idx"][r_].astype(np.int64)))
            assert np.array_equal(dd[order[:K]], ok["d2"][r_])
    assert grown > 10, grown


def test_knn_matches_brute_force():
    """Queries whose K-th distance lies inside r* - margin: the K nearest of
    S(r*) are the K nearest overall."""
    snap, t = _tree("median")
    K = snap.params["knn"]
    pos = t.particles()["pos"]
    rng = np.random.default_rng(0)
    q_idx = rng.choice(t.n, 600, replace=False)
    q = pos[q_idx]
    ok = t.knn(q, K)
    bi, bd = brute_knn(q, pos, K + 1)
    sel = pinned(ok, bd, _margin(t)) & (bd[:, K - 1] < bd[:, K])
    assert sel.sum() > 30, sel.sum()
    assert np.array_equal(np.sort(ok["idx"][sel], 1), np.sort(bi[sel, :K], 1))
    assert np.array_equal(ok["d2"][sel], bd[sel, :K])
    # every returned distance is the exhaustive distance of that index
    dd = brute_d2(q, pos)
    assert np.array_equal(np.take_along_axis(dd, ok["idx"].astype(np.int64), 1), ok["d2"])
    assert np.all(np.diff(ok["d2"], axis=1) >= 0)
    assert np.all(ok["ncand"] >= K)


def test_point_queries_with_start_radius():
    snap, t = _tree("median")
    K = snap.params["knn"]
    pos = t.particles()["pos"]
    rng = np.random.default_rng(1)
    a = rng.integers(0, t.n, 400)
    b = rng.integers(0, t.n, 400)
    q = 0.5 * (pos[a] + pos[b])
    est = 10.0 ** rng.uniform(-4.5, -2.0, len(q))
    ok = t.knn(q, K, est)
    # the search starts at est and only doubles
    ratio = ok["rstop"] / est
    assert np.all(np.isclose(np.log2(ratio), np.round(np.log2(ratio)), atol=1e-9, rtol=0))
    assert np.all(ratio >= 1)
    nodes = t.nodes()
    for r_ in range(len(q)):
        rs = ok["rstop"][r_]
        S = numpy_candidates(nodes, q[r_], rs)
        assert len(S) == ok["ncand"][r_] and len(S) >= K
        if rs > est[r_]:
            assert len(numpy_candidates(nodes, q[r_], rs / 2)) < K
        if ok["ndrop"][r_] == 0:
            d = pos[S] - q[r_]
            dd = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
            order = np.lexsort((S, dd))
            assert np.array_equal(np.sort(S[order[:K]]), np.sort(ok["idx"][r_].astype(np.int64)))
    bi, bd = brute_knn(q, pos, K + 1)
    sel = pinned(ok, bd, _margin(t)) & (bd[:, K - 1] < bd[:, K])
    assert np.array_equal(np.sort(ok["idx"][sel], 1), np.sort(bi[sel, :K], 1))


def median_density(d2, masses):
    K = len(d2)
    m = 4.0 * d2[K // 2]
    i_star = K // 2
    for j in range(K - 1, K // 2, -1):
        if not d2[j] > m:
            i_star = j
            break
    rho = 0.0
    for j in range(i_star, -1, -1):
        rho = rho + masses[j]
    return rho / ((math.pi * 4.0) / 3.0 * math.pow(d2[i_star], 1.5))


def qs_density(dpos, masses, cell, cell_volume):
    rho = 0.0
    for j in range(len(masses)):
        w = 1.0
        for k in range(3):
            dk = abs(dpos[j][k])
            if dk <= cell[k] / 2.0:
                w = w * (0.75 - (dk * dk) / (cell[k] * cell[k]))
            elif dk < (cell[k] / 2.0) * 3.0:
                b = 1.5 - dk / cell[k]
                w = w * (0.5 * (b * b))
            else:
                w = 0.0
                break
        rho = rho + (masses[j] / cell_volume) * w
    return rho


@pytest.mark.parametrize("which", ["median", "qs"])
def test_density_closed_form(which):
    snap, t = _tree(which)
    K = snap.params["knn"]
    assert t.stage_info()["kernel"] == (0 if which == "median" else 1)
    parts = t.particles()
    pos, mass, rho = parts["pos"], parts["mass"], t.densities()
    rng = np.random.default_rng(2)
    q_idx = np.sort(rng.choice(t.n, 400, replace=False))
    ok = t.knn(pos[q_idx], K)
    nodes = t.nodes()
    d = O.derive(snap.params, snap.n)
    cell = [d.cell_length[k] for k in range(3)]
    checked = 0
    for r in range(len(q_idx)):
        if ok["ndrop"][r]:
            continue
        i = q_idx[r]
        S = numpy_candidates(nodes, pos[i], ok["rstop"][r])
        dv = pos[S] - pos[i]
        dd = (dv[:, 0] * dv[:, 0] + dv[:, 1] * dv[:, 1]) + dv[:, 2] * dv[:, 2]
        sel = S[np.lexsort((S, dd))[:K]]
        dsel = np.sort(dd)[:K]
        if which == "median":
            want = median_density(dsel, mass[sel])
        else:
            want = qs_density(pos[sel] - pos[i], mass[sel], cell, d.cell_volume)
        assert rho[i] == want, (i, rho[i], want)
        checked += 1
    assert checked > 300


def test_hop_against_restated_lists():
    snap, t = _tree("median")
    H = t.stage_info()["hop"]
    thr = t.catalog()["density_threshold"]
    parts = t.particles()
    pos, rho = parts["pos"], t.densities()
    dn, pk = t.hop()
    nodes = t.nodes()
    above = np.flatnonzero(rho >= thr)
    assert len(above) > 100
    # below the threshold a particle is its own densest neighbour
    below = np.flatnonzero(rho < thr)
    assert np.array_equal(dn[below], below)
    rng = np.random.default_rng(4)
    sample = rng.choice(above, min(400, len(above)), replace=False)
    ok = t.knn(pos[sample], H)
    checked = 0
    for r_, i in enumerate(sample):
        if ok["ndrop"][r_]:
            continue
        S = numpy_candidates(nodes, pos[i], ok["rstop"][r_])
        d = pos[S] - pos[i]
        dd = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        nb = S[np.lexsort((S, dd))[:H]]
        best = i
        for j in nb:
            if rho[best] < rho[j] or (rho[best] == rho[j] and best > j):
                best = j
        assert dn[i] == best
        checked += 1
    assert checked > 300
    # peaks: follow the pointers to the fixed point; each step raises (rho, -index)
    for i in above:
        p = i
        for _ in range(len(above) + 1):
            if dn[p] == p:
                break
            assert rho[dn[p]] > rho[p] or (rho[dn[p]] == rho[p] and dn[p] < p)
            p = dn[p]
        assert pk[i] == p
