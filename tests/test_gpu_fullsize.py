"""The 16.8M-particle benchmark snapshot (256^3 particles, 200 Plummer
clumps, K = 32, H raised to 64): the whole pipeline against the CPU
restatement run on the same snapshot (tree order, densities, HOP pointers and
peaks, the catalog), and through properties that hold at any size: keys
strictly ascending, the level-major node array's range and daughter
invariants, a sample of neighbour searches against brute force (every
particle inside the search radius and closer than the K-th is returned) and
the densities they give, HOP pointers climbing in density to fixed-point
peaks, and the catalog's member lists, masses and Hill radii."""
import os
import time

import numpy as np
import pytest

import fullsize_checks as F

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def run():
    import plan_amd
    from plan_amd import synth

    snap = synth.config_c2()
    print(f"[fullsize] n={snap.n}", flush=True)
    ctx = plan_amd.Context(0)
    ctx.set_particles(snap.params, snap.pos, snap.vel, snap.density, snap.property_index, snap.id)
    T = ctx.make_ghosts()
    info = ctx.build_tree(True)
    cat = ctx.find_planetesimals()
    return snap, ctx, T, info, cat


def test_keys_and_nodes(run):
    snap, ctx, T, info, cat = run
    F.keys_and_nodes(ctx, info)


def test_density_sample_against_brute_force(run):
    snap, ctx, T, info, cat = run
    F.density_sample_against_brute_force(ctx, info, snap.params["knn"])


def test_hop_climbs_to_peaks(run):
    snap, ctx, T, info, cat = run
    F.hop_climbs_to_peaks(ctx, info)


def test_catalog_properties(run):
    snap, ctx, T, info, cat = run
    F.catalog_properties(ctx, snap, cat)


def test_whole_snapshot_against_the_oracle(run):
    """The CPU restatement on all 16,777,216 particles (about 20 s on the
    host's 16 cores): tree order, every density, HOP and the whole catalog
    bit-exact."""
    import oracle_py as O

    snap, ctx, T, info, cat = run
    threads = min(16, os.cpu_count() or 1)
    O.set_threads(threads)
    print(f"[fullsize] oracle on {snap.n} particles, {threads} threads", flush=True)
    t0 = time.time()
    ot = O.run_pipeline(snap)
    print(f"[fullsize] oracle done in {time.time() - t0:.1f} s", flush=True)
    assert ot.n == info.num_particles
    tp = ctx.tree_particles()
    op = ot.particles()
    assert np.array_equal(tp["oid"], op["oid"])
    assert np.array_equal(tp["pos"], op["pos"])
    rho, orho = tp["rho"], op["rho"]
    rel = np.abs(rho - orho) / np.abs(orho)
    exact = float(np.mean(rho == orho))
    print(f"[fullsize] densities: max rel {rel.max():.2e}, bit-exact {exact:.6f}", flush=True)
    assert exact == 1.0, (exact, rel.max())
    dn, pk = ctx.hop_result()
    odn, opk = ot.hop()
    assert np.array_equal(dn, odn) and np.array_equal(pk, opk)
    oc = ot.catalog()
    print(f"[fullsize] catalog: {len(cat['clumps'])} clumps, oracle {len(oc['clumps'])}", flush=True)
    assert [c["peak_index"] for c in cat["clumps"]] == [c["peak_index"] for c in oc["clumps"]]
    assert np.array_equal(cat["member_original_ids"], oc["member_original_ids"])
    for a, b in zip(cat["clumps"], oc["clumps"]):
        assert a["num_particles"] == b["num_particles"]
        for k in ("total_mass", "Hill_radius", "one10th_radius", "two_sigma_mass_radius", "max_distance",
                  "peak_density"):
            assert a[k] == b[k], k
        for k in ("center_of_mass", "vel_com", "J", "J_quarter_Hill"):
            assert tuple(a[k]) == tuple(b[k]), k
    assert cat["Mp_tot"] == oc["Mp_tot"]


def test_series_carries_the_raised_hop(run):
    """A series of this snapshot and the same one after 100,000 particles left
    (16,677,216 < 1.67e7, so its own run keeps H = 32): the first snapshot's
    automatic choice raised H to 64, the reference keeps that in its
    parameters (src/tree.hpp:3583-3586), so the second runs with H = 64 and
    the first snapshot's particle masses; its catalog equals a single
    context given those parameters."""
    import copy

    import plan_amd
    from plan_amd.pipeline import find_clumps_series

    snap, ctx, T, info, cat = run
    assert info.hop_auto == 1
    keep = snap.n - 100_000
    s2 = copy.copy(snap)
    s2.pos, s2.vel, s2.density = snap.pos[:keep], snap.vel[:keep], snap.density[:keep]
    s2.property_index, s2.id = snap.property_index[:keep], snap.id[:keep]
    out = find_clumps_series([snap, s2], contexts=2)
    assert out[0][1]["Mp_tot"] == cat["Mp_tot"]
    p2, cat2 = out[1]
    assert p2["hop"] == 64 and p2["series_first_n"] == snap.n
    one = plan_amd.Context(0)
    try:
        one.set_particles(dict(snap.params, series_first_n=snap.n), s2.pos, s2.vel, s2.density,
                          s2.property_index, s2.id)
        one.make_ghosts()
        assert one.build_tree().hop_auto == 0
        one.set_num_hop(64)
        want = one.find_planetesimals()
    finally:
        one.close()
    print(f"[fullsize] series: second snapshot {keep} particles, {len(cat2['clumps'])} clumps", flush=True)
    assert [c["peak_index"] for c in cat2["clumps"]] == [c["peak_index"] for c in want["clumps"]]
    assert np.array_equal(cat2["member_original_ids"], want["member_original_ids"])
    assert cat2["Mp_tot"] == want["Mp_tot"]
