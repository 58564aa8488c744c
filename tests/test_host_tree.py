"""The merge passes' clump tree on the host (plan_host_tree_balls: BuildTree
over the clump centres with the particle tree's root box, and
RecursiveBallSearch, src/tree.hpp:5251-5284, 3275-3293) against the
restatement's tree and ball search on the same points: every query's list of
payloads equal, in the same order.  Point sets cover a single point, leaf-
sized sets, clustered and uniform sets of thousands of points, exact
duplicates (one sink particle, the first point's payload), points on the
root box's faces, and radii from zero to beyond the box.  A run of equal
positions is compared as one particle: which member carries it follows the
sort's tie order, the first point here (the device tree's rule).  Runs on the CPU:
the hook uses no device."""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "plan_amd", "libplan_hip.so")


def product_balls(params, pts4, rc, hw, q4):
    import plan_amd
    from plan_amd import abi

    p = abi.make_params(params)
    pts4 = np.ascontiguousarray(pts4, np.float64)
    q4 = np.ascontiguousarray(q4, np.float64)
    rc = np.ascontiguousarray(rc, np.float64)
    off = np.zeros(len(q4) + 1, np.uint32)
    cap = 1 << 20
    out = np.zeros(cap)
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    rcode = plan_amd.lib().plan_host_tree_balls(C.byref(p), len(pts4), ptr(pts4), ptr(rc), float(hw), ptr(q4), len(q4),
                                                ptr(off), ptr(out), cap)
    assert rcode == 0
    return [out[off[i]:off[i + 1]].copy() for i in range(len(q4))]


def point_sets():
    rng = np.random.default_rng(5)
    sets = {"one": rng.uniform(-0.1, 0.1, (1, 3)), "leaf": rng.uniform(-0.1, 0.1, (8, 3)),
            "nine": rng.uniform(-0.1, 0.1, (9, 3)), "uniform": rng.uniform(-0.1, 0.1, (3000, 3))}
    c = rng.uniform(-0.08, 0.08, (20, 3))
    sets["clustered"] = (c[rng.integers(0, 20, 2500)] + rng.normal(0, 2e-3, (2500, 3))).clip(-0.1, 0.1)
    d = rng.uniform(-0.1, 0.1, (300, 3))
    d[100:140] = d[7]  # one run of 41 equal positions: a sink
    d[200] = d[201]
    sets["duplicates"] = d
    f = rng.uniform(-0.1, 0.1, (400, 3))
    f[:50, 0] = -0.125
    f[50:100, 1] = 0.125
    f[100:120] = [0.0, 0.0, 0.0]
    sets["faces"] = f
    return sets


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
@pytest.mark.parametrize("name", list(point_sets()))
def test_host_clump_tree_matches_the_restatement(name):
    import oracle_py as O
    from plan_amd import synth

    rng = np.random.default_rng(hash(name) % 1000)
    pts = point_sets()[name]
    n = len(pts)
    pts4 = np.concatenate([pts, (1000 + np.arange(n, dtype=np.float64))[:, None]], 1)
    params = synth.default_params(max(n, 8))
    rc, hw = np.zeros(3), 0.125
    radii = np.concatenate([[0.0, 1e-4, 2e-3, 0.01, 0.05, 0.3], rng.uniform(0, 0.02, 40)])
    centres = np.concatenate([pts[rng.integers(0, n, 30)], rng.uniform(-0.12, 0.12, (16, 3))])
    q4 = np.array([[*centres[i % len(centres)], radii[i % len(radii)]] for i in range(len(centres) * 3)])
    got = product_balls(params, pts4, rc, hw, q4)
    want = O.point_tree_balls(params, pts4, rc, hw, q4)
    # a run of equal keys is one particle: here the first point's payload,
    # in the restatement std::sort's pick among them (unspecified order)
    runs = {"duplicates": ([1007.0] + [1000.0 + k for k in range(100, 140)], [1200.0, 1201.0]),
            "faces": ([1000.0 + k for k in range(100, 120)],)}.get(name, ())
    for run in runs:
        for lst in got + want:
            lst[np.isin(lst, run)] = run[0]
    nonempty = 0
    for g, w in zip(got, want):
        assert np.array_equal(g, w)
        nonempty += len(w) > 0
    assert nonempty > len(q4) // 4
    if name == "duplicates":
        # the sink carries the first point's payload; the other 40 are gone
        everything = product_balls(params, pts4, rc, hw, np.array([[0.0, 0.0, 0.0, 1.0]]))[0]
        assert 1007.0 in everything and 1100.0 not in everything and len(everything) == n - 41
