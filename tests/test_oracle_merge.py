"""Hand-worked pins of the oracle's multi-clump loops: the two greedy merge
loops (src/tree.hpp:3863-4152), real-peak re-keying (4156-4182), envelope
ripping (4229-4249), ghost-clump removal (4258-4268) and the erasure of the
clump keyed by tree index 0 that the merge loops' zero-filled delete lists
cause (3874 -> 4000-4003, 4019 -> 4146-4149).

Every case is built from blobs of particles on the x axis whose radii,
masses and densities are chosen so that the outcome can be worked out by
hand, with wide margins; the saddle-point density is supplied by the test
(its KNN estimate is pinned in test_oracle_knn.py), so each case decides the
scheduling logic alone.  The expected catalogs are stated as clump keys and
member sets, never recomputed by the oracle's own formulas.
"""
import numpy as np
import pytest

import oracle_py as O
from plan_amd import synth

N_REAL = 10 ** 6          # particle mass from the real count: m = M_tot / N
SHELLS = (1 / 8, 1 / 6, 1 / 4, 1.0)   # blob radii in units of s (6 axis points each)


def params(**kw):
    p = synth.default_params(N_REAL, nx=128)
    p.update(no_ghost=1)
    p.update(kw)
    return p


def hill(params_, n_members):
    """R_H = (G M / 3 / Omega^2)^(1/3), Omega = 1 (src/tree.hpp:4715-4717)."""
    return (synth.grav_constant(params_) * n_members * synth.particle_mass(params_, N_REAL) / 3.0) ** (1 / 3)


class World:
    """Blobs: a centre particle plus 6 points on the axes at each radius of
    SHELLS x s (25 particles, centre of mass at the centre by symmetry,
    mean d^2 = 0.265 s^2, so IsPositionDispersion2Large(0.55) is false and
    one10th = s for equal member densities).  Velocities follow the shear
    flow v_y = q Omega x, plus a blob velocity."""

    def __init__(self, p):
        self.p = p
        self.pos, self.vel, self.rho = [], [], []
        self.blobs = []
        self.radius = []      # one10th radius of each blob (its outer shell)

    def blob(self, cx, s, peak_rho=2000.0, v=(0.0, 0.0, 0.0), cy=0.0, cz=0.0, shells=SHELLS):
        first = len(self.pos)
        pts = [(0.0, 0.0, 0.0)]
        # a blob-specific 0.1% stretch keeps points of different blobs apart
        s = s * (1.0 + 1e-3 * len(self.blobs))
        for f in shells:
            for k in range(3):
                for sg in (-1.0, 1.0):
                    o = [0.0, 0.0, 0.0]
                    o[k] = sg * f * s
                    pts.append(tuple(o))
        for i, o in enumerate(pts):
            x = (cx + o[0], cy + o[1], cz + o[2])
            self.pos.append(x)
            self.vel.append((v[0], 1.5 * x[0] + v[1], v[2]))
            self.rho.append(peak_rho if i == 0 else 1000.0)
        ids = list(range(first, len(self.pos)))
        self.blobs.append(ids)
        self.radius.append(s * max(shells))
        return len(self.blobs) - 1

    def points(self, xyz, rho=1000.0):
        first = len(self.pos)
        for x in xyz:
            self.pos.append(tuple(x))
            self.vel.append((0.0, 1.5 * x[0], 0.0))
            self.rho.append(rho)
        self.blobs.append(list(range(first, len(self.pos))))
        self.radius.append(None)
        return len(self.blobs) - 1

    def build(self):
        pos = np.array(self.pos, np.float64)
        vel = np.array(self.vel, np.float64)
        n = len(pos)
        t = O.OracleTree(self.p, N_REAL, pos, vel, np.full(n, 1e12), np.zeros(n, np.int32),
                         np.arange(n, dtype=np.uint32))
        oid = t.particles()["oid"]
        assert t.n == n, "constructed particles must not form sinks"
        self.tidx = np.empty(n, np.int64)
        self.tidx[oid] = np.arange(n)
        rho = np.empty(n)
        rho[self.tidx] = self.rho
        t.set_densities(rho=rho)
        self.tree = t
        return t

    def key(self, b):
        """tree index of blob b's centre (its peak)"""
        return int(self.tidx[self.blobs[b][0]])

    def members(self, b):
        return [int(self.tidx[i]) for i in self.blobs[b]]

    def run(self, clumps, stages, saddle=None, thr=100.0, **kw):
        """clumps: {key: [blobs]} or a list of blobs, each a clump keyed by
        its centre"""
        if not isinstance(clumps, dict):
            clumps = {self.key(b): [b] for b in clumps}
        spec = {k: [m for b in bs for m in self.members(b)] for k, bs in clumps.items()}
        return self.tree.clump_stages(spec, stages, thr, saddle=saddle, **kw)


def member_sets(cat):
    out = {}
    for c in cat["clumps"]:
        o, n = c["member_offset"], c["num_members"]
        out[c["peak_index"]] = set(int(x) for x in cat["members"][o:o + n])
    return out


NO_SADDLE = lambda pt, est: 0.0   # every saddle is deep: only the contains-centre rule merges


def test_eaten_clump_cannot_eat_in_the_same_pass():
    """Keys A < B < D along x.  A contains B's centre and B contains D's,
    A and D neither contain each other nor share a dense saddle.  In key
    order A eats B; B, eaten, cannot eat D in that pass; the merged A+B does
    not reach D in the next one.  A transitive closure (union-find over the
    pair predicate) would merge all three."""
    w = World(params())
    a = w.blob(-0.006, 0.004)
    b = w.blob(-0.003, 0.002)
    d = w.blob(0.0005, 0.004)
    w.build()
    assert w.key(a) < w.key(b) < w.key(d)
    cat = w.run([a, b, d], ["merge_a"], saddle=NO_SADDLE)
    got = member_sets(cat)
    assert got == {w.key(a): set(w.members(a)) | set(w.members(b)), w.key(d): set(w.members(d))}


def test_contains_centre_merges_without_boundness():
    """cd < max(r1, r2) merges even a pair flying apart (the boundness test
    only guards the overlapping branch, src/tree.hpp:3912-3921)."""
    w = World(params())
    a = w.blob(-0.004, 0.003)
    b = w.blob(-0.002, 0.001, v=(5.0, 0.0, 0.0))
    w.build()
    cat = w.run([a, b], ["merge_a"], saddle=NO_SADDLE)
    assert member_sets(cat) == {w.key(a): set(w.members(a)) | set(w.members(b))}


@pytest.mark.parametrize("factor, merged", [(1 - 1e-12, False), (1.0, True), (1 + 1e-12, True)])
def test_overlap_branch_saddle_threshold_and_point(factor, merged):
    """max(r1, r2) <= cd < r1 + r2 with a bound pair: merged iff the saddle
    density is not below 2.5 rho_thr.  The saddle point lies r1 / (r1 + r2)
    of the way from the predator's centre to the prey's, and its KNN starts
    at min(r1, r2) (src/tree.hpp:4880-4892)."""
    thr = 100.0
    w = World(params())
    a = w.blob(-0.004, 0.003)
    b = w.blob(0.0005, 0.002)
    w.build()
    calls = []

    def saddle(pt, est):
        calls.append((pt.copy(), est))
        return 2.5 * thr * factor

    cat = w.run([a, b], ["merge_a"], saddle=saddle, thr=thr)
    ka, kb = w.key(a), w.key(b)
    want = {ka: set(w.members(a)) | set(w.members(b))} if merged else {ka: set(w.members(a)), kb: set(w.members(b))}
    assert member_sets(cat) == want
    # A asks first (key order); without a merge B asks about A in its turn
    assert len(calls) == (1 if merged else 2)
    pt, est = calls[0]
    ca, cb, ra, rb = -0.004, 0.0005, w.radius[a], w.radius[b]
    assert abs(est - min(ra, rb)) < 1e-15
    np.testing.assert_allclose(pt, [ca + (cb - ca) * ra / (ra + rb), 0, 0], atol=1e-15)
    if not merged:
        pt2, est2 = calls[1]
        np.testing.assert_allclose(pt2, [cb + (ca - cb) * rb / (ra + rb), 0, 0], atol=1e-15)


def test_overlap_branch_needs_a_bound_pair():
    """The same overlapping pair flying apart is never merged, and its saddle
    density is not even asked for (src/tree.hpp:3916)."""
    w = World(params())
    a = w.blob(-0.004, 0.003)
    b = w.blob(0.0005, 0.002, v=(0.0, 0.0, 3.0))
    w.build()
    calls = []
    cat = w.run([a, b], ["merge_a"], saddle=lambda pt, est: calls.append(1) or 1e30)
    assert len(cat["clumps"]) == 2 and calls == []


def test_search_radius_adds_the_largest_radius():
    """A small clump first in key order finds a large one whose radius
    covers its centre only through the ball radius r1 + max r
    (src/tree.hpp:3891): it eats the larger clump and keeps its key; the
    prey's denser peak becomes the clump's peak and the predator's own peak
    goes to the subpeaks, so re-keying (which looks only at the subpeaks)
    leaves the key on the less dense particle."""
    w = World(params())
    a = w.blob(-0.0035, 0.0005, peak_rho=1500.0)
    b = w.blob(0.0, 0.004, peak_rho=3000.0)
    w.build()
    ka, kb = w.key(a), w.key(b)
    assert ka < kb
    assert 0.0035 > 2 * 0.0005     # beyond r1 + r1: only max r brings B into A's ball
    cat = w.run([a, b], ["merge_a", "rekey"], saddle=NO_SADDLE)
    assert member_sets(cat) == {ka: set(w.members(a)) | set(w.members(b))}
    assert cat["clumps"][0]["subpeaks"] == [ka]


@pytest.mark.parametrize("rho_c, new_key", [(2500.0, "c"), (2000.0, "a")])
def test_rekey_takes_a_denser_subpeak_strictly(rho_c, new_key):
    """A (peak 2000) eats B (peak 3000) in pass 1: peak -> B, subpeaks [A].
    The merged clump eats C in pass 2: its peak B is denser, so C goes to the
    subpeaks [A, C].  Re-keying compares the key A with its subpeaks only and
    moves the key to C when C is strictly denser than A, appending the old
    key: subpeaks [A, C, A]; with equal densities the key stays."""
    w = World(params())
    a = w.blob(-0.006, 0.004, peak_rho=2000.0)
    b = w.blob(-0.003, 0.002, peak_rho=3000.0)
    c = w.blob(-0.0005, 0.001, peak_rho=rho_c)
    w.build()
    ka, kb, kc = w.key(a), w.key(b), w.key(c)
    assert ka < kb < kc
    cat = w.run([a, b, c], ["merge_a", "rekey"], saddle=NO_SADDLE)
    allm = set(w.members(a)) | set(w.members(b)) | set(w.members(c))
    if new_key == "c":
        assert member_sets(cat) == {kc: allm}
        assert cat["clumps"][0]["subpeaks"] == [ka, kc, ka]
    else:
        assert member_sets(cat) == {ka: allm}
        assert cat["clumps"][0]["subpeaks"] == [ka, kc]


def test_hill_loop_recomputes_the_radius_after_each_merge():
    """Merge loop B with r = 0.75 R_H: A eats B (pass 1), the merged mass
    raises A's Hill radius by 2^(1/3) at once (src/tree.hpp:4123-4124), so in
    pass 2 its larger radius covers D's centre.  B, eaten in pass 1, could
    not eat D although it covered D's centre."""
    p = params()
    h = hill(p, 25)
    w = World(p)
    s = 0.3 * h      # compact enough that the merged clumps pass IsPositionDispersion2Large
    a = w.blob(0.0, s)
    b = w.blob(0.29 * h, s)
    d = w.blob(1.0 * h, s)
    w.build()
    assert w.key(a) < w.key(b) < w.key(d)
    cat = w.run([a, b, d], ["merge_b"], saddle=NO_SADDLE, hill_frac=0.75)
    allm = set(w.members(a)) | set(w.members(b)) | set(w.members(d))
    assert member_sets(cat) == {w.key(a): allm}
    np.testing.assert_allclose(cat["clumps"][0]["Hill_radius"], hill(p, 75), rtol=1e-12)


def test_key_zero_is_erased_after_a_merge_loop():
    """The delete lists are zero-filled to the clump count and erased key by
    key, so a clump keyed by tree index 0 that survives a merge loop is
    erased with the deleted ones; clumps merged into it go with it."""
    for stages in (["merge_a"], ["merge_b"]):
        w = World(params())
        e = w.blob(-0.08, 0.002)
        f = w.blob(0.02, 0.002)
        g = w.blob(0.05, 0.002)
        w.build()
        low = min(w.members(e))
        assert low == 0
        cat = w.run({0: [e], w.key(f): [f], w.key(g): [g]}, stages, saddle=NO_SADDLE, hill_frac=0.75)
        assert member_sets(cat) == {w.key(f): set(w.members(f)), w.key(g): set(w.members(g))}
        assert w.tree.key0_erased() == 1
    # key 0 eats a neighbour: both are gone
    w = World(params())
    e = w.blob(-0.08, 0.003)
    f = w.blob(-0.078, 0.001)
    g = w.blob(0.05, 0.002)
    w.build()
    cat = w.run({0: [e], w.key(f): [f], w.key(g): [g]}, ["merge_a"], saddle=NO_SADDLE)
    assert member_sets(cat) == {w.key(g): set(w.members(g))}
    # no clump keyed 0: nothing extra is erased
    cat = w.run([f, g], ["merge_a"], saddle=NO_SADDLE)
    assert len(cat["clumps"]) == 2 and w.tree.key0_erased() == 0


def test_envelope_ripping_cuts_at_the_hill_radius_until_it_holds():
    """A bound core (25 members within s << R_H) with 6 members at d1 and 6
    more at d2: R_H(37 m) < d2, so the first cut keeps d < R_H; then
    R_H(31 m) < d1 and the second cut leaves the core, whose R_H covers it
    (src/tree.hpp:4229-4249).  Members at rest in the shear frame stay bound
    through each unbinding."""
    p = params()
    w = World(p)
    h25, h31, h37 = hill(p, 25), hill(p, 31), hill(p, 37)
    d1, d2 = 1.05 * h31, 1.1 * h37
    assert d1 < h37 < d2 and h31 < d1
    core = w.blob(0.0, 0.05 * h25)
    env = w.points([(sg * d1, 0, 0) for sg in (-1, 1)] + [(0, sg * d1, 0) for sg in (-1, 1)] +
                   [(0, 0, sg * d1) for sg in (-1, 1)] +
                   [(sg * d2 * 0.6, sg * d2 * 0.8, 0) for sg in (-1, 1)] +
                   [(sg * d2 * 0.6, -sg * d2 * 0.8, 0) for sg in (-1, 1)] +
                   [(0, 0, sg * d2) for sg in (-1, 1)])
    w.build()
    key = w.key(core)
    cat = w.run({key: [core, env]}, ["unbind"], saddle=NO_SADDLE)
    assert member_sets(cat) == {key: set(w.members(core)) | set(w.members(env))}
    cat = w.run({key: [core, env]}, ["unbind", "rip"], saddle=NO_SADDLE)
    assert member_sets(cat) == {key: set(w.members(core))}
    np.testing.assert_allclose(cat["clumps"][0]["Hill_radius"], h25, rtol=1e-12)
    # an envelope inside the Hill sphere is kept
    w = World(p)
    core = w.blob(0.0, 0.05 * h25)
    env = w.points([(sg * 0.5 * h31, 0, 0) for sg in (-1, 1)] + [(0, sg * 0.5 * h31, 0) for sg in (-1, 1)] +
                   [(0, 0, sg * 0.5 * h31) for sg in (-1, 1)])
    w.build()
    key = w.key(core)
    cat = w.run({key: [core, env]}, ["unbind", "rip"], saddle=NO_SADDLE)
    assert member_sets(cat) == {key: set(w.members(core)) | set(w.members(env))}


def _exact_one_member_centre(m, x):
    """kinematics of a one-member clump: (0 + x m) / m (src/tree.hpp:4431-4442)"""
    return (np.float64(0.0) + np.float64(x) * np.float64(m)) / np.float64(m)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_ghost_clump_removal_keeps_the_box_faces(axis):
    """Clumps whose centre lies on box_max or box_min are kept (inclusive
    comparisons, src/tree.hpp:4258-4268); one ulp beyond either face is
    removed."""
    p = params()
    m = synth.particle_mass(p, N_REAL)
    lo, hi = p["box_min"][axis], p["box_max"][axis]
    faces = [hi, lo, np.nextafter(hi, 1.0), np.nextafter(lo, -1.0)]
    for x in faces:
        assert _exact_one_member_centre(m, x) == x, "choose N_REAL so that the centre is exact"
    w = World(p)
    blobs = []
    for j, x in enumerate(faces):
        q = [0.01 * (j + 1), -0.02 + 0.01 * j, 0.03 - 0.01 * j]
        q[axis] = x
        blobs.append(w.points([q]))
    w.build()
    keys = [int(w.tidx[w.blobs[b][0]]) for b in blobs]
    cat = w.run({k: [b] for k, b in zip(keys, blobs)}, ["ghost"], saddle=NO_SADDLE)
    assert sorted(member_sets(cat)) == sorted(keys[:2])
