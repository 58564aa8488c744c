"""Pins of the CPU restatement's tree stage: parameters, ghosts, Morton keys,
key order and the level-major node array, checked against closed forms,
independent numpy constructions and the layout invariants of the reference
(src/tree.hpp:2655-2755, 2849-2886)."""
import math

import numpy as np
import pytest

import oracle_py as O
from _cases import dense_case
from plan_amd import synth


def test_derived_parameters_closed_form():
    p = synth.default_params(262144)
    d = O.derive(p, 262144)
    L = p["box_max"][0] - p["box_min"][0]
    assert d.box_length[0] == L
    assert d.cell_length[0] == L / 128
    assert d.cell_volume == pytest.approx((L / 128) ** 3, rel=1e-15)
    G = p["four_pi_G"] / (4 * math.pi)
    assert d.grav_constant == pytest.approx(G, rel=1e-15)
    M = 0.02 * L * L * math.sqrt(2 * math.pi)
    assert d.mass_total_code_units == pytest.approx(M, rel=1e-15)
    assert d.mass_per_particle[0] == pytest.approx(M / 262144, rel=1e-15)
    assert d.min_trusted_mass_code_unit == pytest.approx(3 * (L / 128) ** 3 / G, rel=1e-14)
    assert d.shear_speed == pytest.approx(1.5 * L, rel=1e-15)
    assert d.max_half_width == L / 2
    assert d.max_ghost_zone_width == 0.025


def test_series_masses_from_the_first_snapshot():
    """A later snapshot of a series keeps the first one's particle masses,
    M_tot * weight / N_first (ReadLisFile sets them at loop_begin only,
    src/tree.hpp:1563-1581), whatever its own count."""
    p = synth.default_params(1000)
    p["type_info"] = [1.0, 3.0]
    first = O.derive(p, 1000)
    later = O.derive(dict(p, series_first_n=1000), 900)
    alone = O.derive(p, 900)
    M = first.mass_total_code_units
    assert later.mass_per_particle[0] == first.mass_per_particle[0] == (0.25 * M) / 1000
    assert later.mass_per_particle[1] == (0.75 * M) / 1000
    assert alone.mass_per_particle[0] == (0.25 * M) / 900


def _ghost_case(seed=0, n=3000, time=0.37):
    rng = np.random.default_rng(seed)
    p = synth.default_params(n, time=time)
    pos = rng.uniform(-0.1, 0.1, (n, 3)).astype(np.float32).astype(np.float64)
    vel = rng.normal(0, 1e-3, (n, 3))
    return p, pos, vel


@pytest.mark.parametrize("time", [0.0, 0.37, 12.5])
def test_ghosts_are_shearing_periodic_images(time):
    p, pos, vel = _ghost_case(time=time)
    n = len(pos)
    dens = np.arange(n, dtype=np.float64)
    gp, gv, gd, gpr, gid = O.make_ghosts(p, pos, vel, dens, np.zeros(n, np.int32), np.arange(n, dtype=np.uint32))
    L = p["box_max"][0] - p["box_min"][0]
    w = L / 2 - 0.025
    lo, hi = -w, w  # box centre is 0
    # expected counts by region membership
    nx = np.sum(pos[:, 0] > hi) + np.sum(pos[:, 0] < lo)
    T = len(gp)
    assert np.array_equal(gid, np.arange(T))
    assert np.array_equal(gp[:n], pos)
    xg = gp[n:n + nx]
    src_x = gd[n:n + nx].astype(int)
    # every x image is shifted by -+Lx and by -+shear*t (mod Ly) in y
    dx = xg[:, 0] - pos[src_x, 0]
    assert np.allclose(np.abs(dx), L, atol=1e-15)
    shift = -np.sign(dx) * 1.5 * L * time
    dy = np.mod(xg[:, 1] - pos[src_x, 1] - shift + L / 2, L) - L / 2
    assert np.allclose(dy, 0, atol=1e-12)
    assert np.all(xg[:, 1] >= p["box_min"][1] - 1e-12) and np.all(xg[:, 1] <= p["box_max"][1] + 1e-12)
    assert np.array_equal(xg[:, 2], pos[src_x, 2])
    # y images: of the x images first, then of the originals; each a +-Ly copy
    yg = gp[n + nx:]
    src_y = gd[n + nx:].astype(int)
    ny_from_x = np.sum(xg[:, 1] < lo) + np.sum(xg[:, 1] > hi)
    ny_from_o = np.sum(pos[:, 1] < lo) + np.sum(pos[:, 1] > hi)
    assert len(yg) == ny_from_x + ny_from_o
    srcpos = np.concatenate([xg[:ny_from_x] * 0, pos[src_y[ny_from_x:]]]) if False else None
    yy = yg[ny_from_x:]
    assert np.allclose(np.abs(yy[:, 1] - pos[src_y[ny_from_x:], 1]), L, atol=1e-15)
    assert np.array_equal(yy[:, 0], pos[src_y[ny_from_x:], 0])
    # every original inside the ghost band has exactly its images
    in_band = (pos[:, 0] > hi) | (pos[:, 0] < lo)
    assert set(src_x.tolist()) == set(np.flatnonzero(in_band).tolist())


def _dilate_py(v):
    r = 0
    for b in range(32):
        if v >> b & 1:
            r |= 1 << (3 * b)
    return r


def _expected_keys(pos, center, hw):
    """Morton keys from the closed form: grid = round-half-even(fl(s*(2^32-2))),
    s = fl((x - lo) * fl(1/(hi - lo))); bit b of axis d -> key bit 3b + d."""
    lo = center - hw
    hi = center + hw
    scale = 1.0 / (hi - lo)
    s = (pos - lo) * scale
    g = np.rint(s * 4294967294.0).astype(np.uint64) & 0xffffffff
    keys = []
    for row in g:
        k = 0
        for dd in range(3):
            k |= _dilate_py(int(row[dd])) << dd
        keys.append(k)
    return keys


def test_double_to_grid_edge_values():
    # closed form at the documented edges: 0, 2^-32, 0.5, nextafter(1, 0)
    vals = np.array([0.0, 2.0 ** -32, 0.5, np.nextafter(1.0, 0.0), 0.25 + 2.0 ** -40])
    g = np.rint(vals * 4294967294.0)
    assert g[0] == 0 and g[1] == 1 and g[2] == 2147483647
    assert g[3] == 4294967294


@pytest.mark.parametrize("seed", [0, 1])
def test_morton_keys_and_order(seed):
    snap = dense_case(3000, 3, seed, a=3e-4)
    t = O.run_pipeline(snap)
    c, hw = t.root()
    w0, w1, w2, pidx = t.keys()
    got = [int(a) | int(b) << 32 | int(cc) << 64 for a, b, cc in zip(w0, w1, w2)]
    # positions of the T input particles in the oracle's build order
    pos = snap.pos.astype(np.float64)
    gp, _, _, _, _ = O.make_ghosts(snap.params, pos, snap.vel.astype(np.float64), snap.density.astype(np.float64),
                                   snap.property_index, snap.id)
    exp = _expected_keys(gp[pidx], c, hw)
    assert got == exp
    # sorted on the 96-bit key
    assert all(got[i] <= got[i + 1] for i in range(len(got) - 1))
    assert sorted(pidx.tolist()) == list(range(len(gp)))


def _tree_invariants(t):
    nd = t.nodes()
    n = t.n
    c, hw = t.root()
    w0, w1, w2, pidx = t.keys()
    keys = np.array([int(a) | int(b) << 32 | int(cc) << 64 for a, b, cc in zip(w0, w1, w2)], dtype=object)
    uniq = np.ones(len(keys), bool)
    uniq[1:] = keys[1:] != keys[:-1]
    keys = keys[uniq]
    assert len(keys) == n
    N = len(nd["hw"])
    lvl = nd["level"].astype(int)
    # level-major, nondecreasing levels; root at level 1
    assert lvl[0] == 1 and np.all(np.diff(lvl) >= 0)
    assert nd["begin"][0] == 0 and nd["end"][0] == n and nd["parent"][0] == 0xffffffff
    # within a level, nodes are in key order and tile the covered particles
    for L in np.unique(lvl):
        idx = np.flatnonzero(lvl == L)
        b = nd["begin"][idx]
        assert np.all(np.diff(b.astype(np.int64)) > 0)
    leaf = nd["num_daughters"] == 0
    # daughters contiguous, ranges partition the parent, digit = orthant
    for v in range(N):
        k = int(nd["num_daughters"][v])
        if k == 0:
            continue
        f = int(nd["first_daughter"][v])
        ch = np.arange(f, f + k)
        assert np.all(nd["parent"][ch] == v)
        assert nd["begin"][ch[0]] == nd["begin"][v] and nd["end"][ch[-1]] == nd["end"][v]
        assert np.all(nd["end"][ch[:-1]] == nd["begin"][ch[1:]])
        assert np.all(lvl[ch] == lvl[v] + 1)
        assert np.all(np.diff(nd["orthant"][ch].astype(int)) > 0)
        h = nd["hw"][ch[0]]
        assert h == 0.5 * nd["hw"][v] if v else h == 0.5 * hw
        for cc in ch:
            o = int(nd["orthant"][cc])
            sign = np.array([1.0 if o >> d & 1 else -1.0 for d in range(3)])
            assert np.array_equal(nd["center"][cc], nd["center"][v] + h * sign)
            shift = int(93 - 3 * (lvl[v] - 1))
            for i in range(nd["begin"][cc], nd["end"][cc]):
                assert (keys[i] >> shift) & 7 == o
    # leaf criterion: at most 8 particles or level 32; internal nodes > 8
    size = nd["end"].astype(np.int64) - nd["begin"]
    internal = ~leaf
    internal[0] = True
    assert np.all(size[internal][1:] > 8)
    assert np.all((size[leaf] <= 8) | (lvl[leaf] == 32))
    # leaves in DFS (begin) order, node2leaf inverse
    ln = nd["leaf_nodes"]
    assert np.all(np.diff(nd["begin"][ln].astype(np.int64)) > 0)
    assert np.array_equal(nd["node2leaf"][ln], np.arange(len(ln)))
    assert np.all(nd["node2leaf"][internal] == -1)
    # particles inside their leaf box, within the reference's tolerance
    P = t.particles()["pos"]
    eps = 2 * hw / 2 ** 31
    for v in ln:
        pts = P[nd["begin"][v]:nd["end"][v]]
        assert np.all(np.abs(pts - nd["center"][v]) <= nd["hw"][v] + eps)
    return nd


@pytest.mark.parametrize("seed", [0, 3])
def test_tree_layout_invariants(seed):
    snap = dense_case(2500, 3, seed, a=2e-4)
    t = O.run_pipeline(snap)
    _tree_invariants(t)


def test_node_count_from_prefixes():
    """Independent construction: level-(L+1) nodes are the distinct L-digit
    prefixes among particles whose level-L node holds more than 8."""
    snap = dense_case(4000, 4, 9, a=2e-4)
    t = O.run_pipeline(snap)
    nd = t.nodes()
    w0, w1, w2, _ = t.keys()
    keys = [int(a) | int(b) << 32 | int(c) << 64 for a, b, c in zip(w0, w1, w2)]
    keys = sorted(set(keys))
    counts = {1: 1}
    active = [keys]
    for L in range(1, 32):
        nxt = []
        c = 0
        for grp in active:
            shift = 93 - 3 * (L - 1)
            runs = {}
            for k in grp:
                runs.setdefault(k >> shift, []).append(k)
            c += len(runs)
            for r in runs.values():
                if len(r) > 8 and L < 31:
                    nxt.append(r)
        counts[L + 1] = c
        active = nxt
        if not active:
            break
    lvl = nd["level"].astype(int)
    for L, c in counts.items():
        assert np.sum(lvl == L) == c, L


def test_sink_particles_merge_duplicates():
    """Particles at identical positions become one sink with summed mass and
    mass-weighted velocity (src/tree.hpp:2591-2651)."""
    rng = np.random.default_rng(5)
    n = 800
    p = synth.default_params(n)
    p["no_ghost"] = 1
    pos = rng.uniform(-0.09, 0.09, (n, 3)).astype(np.float32).astype(np.float64)
    pos[10] = pos[20]
    pos[30] = pos[20]
    vel = rng.normal(0, 1, (n, 3))
    t = O.OracleTree(p, n, pos, vel, np.ones(n), np.zeros(n, np.int32), np.arange(n, dtype=np.uint32))
    assert t.n == n - 2
    assert t.num_sinks() == 1
    P = t.particles()
    s = np.flatnonzero(P["sink"])[0]
    m = O.derive(p, n).mass_per_particle[0]
    assert P["mass"][s] == pytest.approx(3 * m, rel=1e-15)
    np.testing.assert_allclose(P["vel"][s], vel[[10, 20, 30]].mean(0), rtol=1e-12)
