"""Size-independent properties of a full snapshot's pipeline run, shared by
the 16.8M and 134M tests: keys strictly ascending, the level-major node
array's range and daughter invariants, sampled neighbour searches against
brute force and the densities they give, HOP pointers climbing in density to
fixed-point peaks, and the catalog's member lists, masses, centres and Hill
radii."""
import math

import numpy as np
import pytest

def keys_and_nodes(ctx, info):
    w0, w1, w2 = ctx.keys()
    k = (w2.astype(np.uint64) << np.uint64(32)) | w1.astype(np.uint64)
    # strictly ascending 96-bit keys (sinks merged): high 64 bits, then the low word
    hi_up = k[1:] > k[:-1]
    eq = k[1:] == k[:-1]
    assert np.all(hi_up | (eq & (w0[1:] > w0[:-1])))
    nd = ctx.nodes()
    lvl, beg, end, fd, ndt = nd["level"], nd["begin"], nd["end"], nd["first_daughter"], nd["num_daughters"]
    assert beg[0] == 0 and end[0] == info.num_particles and lvl[0] == 1
    assert np.all(np.diff(lvl.astype(np.int64)) >= 0)  # level-major
    inner = np.flatnonzero(ndt > 0)
    first, last = fd[inner], fd[inner] + ndt[inner] - 1
    assert np.array_equal(beg[first], beg[inner]) and np.array_equal(end[last], end[inner])
    assert np.all(lvl[first] == lvl[inner] + 1)
    leaves = ndt == 0
    assert np.all((end[leaves] - beg[leaves] <= 8) | (lvl[leaves] == 32))
    assert int(np.sum(end[leaves] - beg[leaves])) == info.num_particles


def density_sample_against_brute_force(ctx, info, K, nsample=64, tag="fullsize"):
    """The reference's neighbours are the K nearest of a candidate set that
    holds every particle within the final search radius r*
    (src/tree.hpp:3198-3251).  So, for a sample of particles: the returned
    distances are those of the returned particles, every particle closer
    than both r* and the K-th returned distance (brute force) is returned,
    and the median-kernel density follows from them
    (src/tree.hpp:3473-3491)."""
    tp = ctx.tree_particles()
    pos, rho, mass = tp["pos"], tp["rho"], tp["mass"]
    order = np.argsort(pos[:, 0], kind="stable")
    xs = pos[order, 0]
    rng = np.random.default_rng(0)
    sample = rng.choice(info.num_particles, nsample, replace=False).astype(np.uint32)
    got = ctx.knn(K, particle_idx=sample)
    inside = 0
    for j, i in enumerate(sample):
        srt = np.argsort(got["d2"][j], kind="stable")
        idx, gd = got["idx"][j][srt], got["d2"][j][srt]
        d = pos[idx] - pos[i]
        assert np.array_equal((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2], gd), i
        r = min(math.sqrt(gd[-1]), got["rstop"][j]) * (1 + 1e-9)
        cand = order[np.searchsorted(xs, pos[i, 0] - r):np.searchsorted(xs, pos[i, 0] + r, side="right")]
        d = pos[cand] - pos[i]
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        close = cand[(d2 < gd[-1]) & (d2 <= got["rstop"][j] ** 2)]
        assert np.all(np.isin(close, idx)), i
        inside += len(close)
        m_med = 4 * gd[K // 2]
        last = int(np.flatnonzero(gd <= m_med)[-1])
        msum = 0.0
        for q in idx[:last + 1][::-1]:
            msum += mass[q]
        ref = msum / (4.1887902047863910 * math.pow(gd[last], 1.5))
        assert rho[i] == pytest.approx(ref, rel=1e-12)
    print(f"[{tag}] {len(sample)} sampled queries, {inside} neighbours confirmed by brute force", flush=True)
    assert inside > 0


def hop_climbs_to_peaks(ctx, info):
    dn, pk = ctx.hop_result()
    rho = ctx.tree_particles()["rho"]
    n = info.num_particles
    above = pk < n
    assert np.all(rho[dn[above]] >= rho[above])
    peaks = np.unique(pk[above])
    assert np.all(dn[peaks] == peaks) and np.all(pk[peaks] == peaks)


def catalog_properties(ctx, snap, cat, min_clumps=10):
    tp = ctx.tree_particles()
    G = snap.params["four_pi_G"] / (4 * math.pi)
    assert len(cat["clumps"]) > min_clumps
    members = cat["members"]
    assert len(np.unique(members)) == len(members)  # disjoint member lists
    assert np.array_equal(cat["member_original_ids"], tp["oid"][members])
    tot = 0.0
    for c in cat["clumps"]:
        m = members[c["member_offset"]:c["member_offset"] + c["num_members"]]
        assert c["total_mass"] == pytest.approx(float(np.sum(tp["mass"][m])), rel=1e-10)
        assert c["Hill_radius"] == pytest.approx((G * c["total_mass"] / 3.0) ** (1 / 3), rel=1e-12)
        com = np.sum(tp["pos"][m] * tp["mass"][m][:, None], axis=0) / np.sum(tp["mass"][m])
        np.testing.assert_allclose(c["center_of_mass"], com, rtol=0, atol=1e-12)
        d = np.linalg.norm(tp["pos"][m] - np.array(c["center_of_mass"]), axis=1)
        assert np.all(np.diff(d) >= -1e-15)  # members in distance order
        assert d[-1] <= c["Hill_radius"] * (1 + 1e-12)
        tot += c["total_mass"]
    assert cat["Mp_tot"] == pytest.approx(tot, rel=1e-12)
