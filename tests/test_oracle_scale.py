"""Pins of the CPU restatement of the density-against-scale pass
(src/analyses.hpp:196-333, src/tree.hpp:3307-3402): without wrapping the
sheared ball count is the brute-force count; radii, levels, the z range and
the count-to-density conversion against closed forms."""
import math

import numpy as np
import pytest

import oracle_py as O
from plan_amd import synth


def scale_case(n=6000, nx=8, time=0.37, seed=4):
    snap = synth.small_case(n, 3, seed=seed, nx=nx)
    snap.params["time"] = time
    return snap


@pytest.fixture(scope="module")
def tree():
    snap = scale_case()
    return snap, O.build_tree(snap)


def test_unwrapped_count_is_brute_force(tree):
    snap, ot = tree
    pos = ot.particles()["pos"]
    rng = np.random.default_rng(0)
    for _ in range(60):
        c = rng.uniform(-0.1, 0.1, 3)
        r = 10 ** rng.uniform(-3, -0.8)
        d = pos - c
        want = int(np.sum(((d[:, 0] * d[:, 0]) + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2] <= r * r))
        assert ot.sheared_count(c, r, [1e9, 1e9, 1e9], 0.0) == want


def test_levels_radii_and_z_range(tree):
    snap, ot = tree
    r = ot.density_vs_scale()
    L = snap.params["box_max"][0] - snap.params["box_min"][0]
    cell = L / snap.params["Nx"][0]
    assert len(r["radii"]) == round(math.log2(L / cell)) + 1
    assert list(r["radii"]) == [cell / 2 * 2 ** i for i in range(len(r["radii"]))]
    z = ot.particles()["pos"][:, 2]
    assert r["z_max"] == z.max()
    assert r["z_min"] - cell <= z.min() < r["z_min"] or r["z_min"] == -cell / 2
    m = O.derive(snap.params, snap.n).mass_per_particle[0]
    counts = r["max_rhop"] * (4.1887902047863910 * r["radii"] ** 3) / m
    np.testing.assert_allclose(counts, np.round(counts), rtol=1e-12)
    assert np.all(r["sums"] > 0) and np.all(counts >= 1)
