"""Closed-form pins of the oracle's clump operations (the reference's
Planetesimal / PlanetesimalList members, src/tree.hpp:4431-4909).

Each case is built so that a plausible slip in the restatement (a dropped
term, a flipped sign, the wrong member, the wrong mass in a denominator)
changes the outcome: the expected values come from the geometry of the case,
not from re-running the formula."""
import math

import numpy as np
import pytest

import oracle_py as O
from plan_amd import synth

P = synth.default_params(4096)


def consts():
    d = O.derive(P, 4096)
    return d.grav_constant, d.Omega, d.q


def test_constants_are_the_reference_defaults():
    G, Om, q = consts()
    assert Om == 1.0 and q == 1.5            # src/global.hpp:749-757
    assert G == pytest.approx(P["four_pi_G"] / (4 * math.pi), rel=1e-15)


# ------------------------------------------------------------ SortParticles

def line_clump(masses, xs):
    pos = np.zeros((len(xs), 3))
    pos[:, 0] = xs
    return pos


def test_sort_particles_order_com_and_radii():
    """Ten members on the x axis with centre of mass exactly 0 and distances
    1..9 and 360; the median-distance member (the 6th nearest, x = 6) sets the
    tenth-density threshold (src/tree.hpp:4465), not the peak."""
    xs = [1, 2, 3, 4, 5, 6, 7, 8, 9, -360.0]
    m = [1.0] * 9 + [0.125]                  # 0.125 * 360 = 45 = 1 + ... + 9
    rho = [100, 90, 80, 70, 60, 50, 4, 6, 3, 2.0]
    vel = np.zeros((10, 3))
    vel[:, 1] = np.arange(10) * 0.25
    r = O.clump_op(P, "sort", line_clump(m, xs), vel, m, rho)
    assert list(r["order"]) == list(range(10))
    assert r["M"] == 9.125
    assert np.all(r["X"] == 0.0)
    assert r["V"][1] == pytest.approx((0.25 * sum(range(9)) + 0.125 * 0.25 * 9) / 9.125, rel=1e-15)
    # tenth = 50 / 10 = 5: inner = first member (by distance) below 5 -> x = 7;
    # outer = last member above 5 -> x = 8; outer/inner < 2.5 -> min(7.5, two_sigma)
    assert r["inner"] == 7.0 and r["outer"] == 8.0
    # outermost 5 % of the mass (0.456): 0.125 (x = 360) + 1 (x = 9) crosses it at x = 9
    assert r["two_sigma"] == 9.0
    assert r["one10th"] == 7.5
    # the peak's density (100) would give inner 7, outer 6 and one10th 6.5


def test_radii_wide_envelope_uses_outer_or_two_sigma():
    """outer / inner > 2.5 switches one10th to min(outer, two_sigma)."""
    xs = [1, 2, 3, 4, 5, 6, 7, 8, 9, -360.0]
    m = [1.0] * 9 + [0.125]
    rho = [100, 3, 80, 70, 60, 50, 40, 30, 6, 2.0]   # tenth 5: inner = x2, outer = x9
    r = O.clump_op(P, "sort", line_clump(m, xs), np.zeros((10, 3)), m, rho)
    assert r["inner"] == 2.0 and r["outer"] == 9.0
    assert r["one10th"] == min(9.0, r["two_sigma"]) == 9.0


def test_dispersion_rule():
    """IsPositionDispersion2Large: mean d^2 > (0.55 d_max)^2."""
    # eight members near distance 1 and two at distance 2, COM 0: mean d^2 =
    # 1.725 > (0.55 * 2)^2 = 1.21 -> too large
    xs = [1, -1, 1, -1, 1, -1, 1, -1, 2, -2.0]
    pos = np.zeros((10, 3))
    pos[:, 0] = xs
    pos[:8, 1] = [0.5, 0.5, -0.5, -0.5, 0.25, 0.25, -0.25, -0.25]
    d2 = (pos ** 2).sum(1)
    r = O.clump_op(P, "dispersion", pos, np.zeros((10, 3)), np.ones(10))
    assert d2.mean() > (0.55 * np.sqrt(d2.max())) ** 2
    assert r["dispersion"] is True
    # a compact core with one far member: mean d^2 well below the bound
    pos2 = np.zeros((10, 3))
    pos2[:9, 0] = np.linspace(-0.1, 0.1, 9)
    pos2[9, 0] = 3.0
    pos2[:, 0] -= pos2[:, 0].mean()
    r2 = O.clump_op(P, "dispersion", pos2, np.zeros((10, 3)), np.ones(10))
    assert r2["dispersion"] is False


# ---------------------------------------------------- RemoveUnboundParticles

def two_body(vB, mB=1e-3, d=None):
    G, Om, q = consts()
    if d is None:
        d = (4 * G * 1.0 / (q * q * Om * Om)) ** (1 / 3)
    pos = np.array([[0.0, 0, 0], [d, 0, 0]])
    vel = np.zeros((2, 3))
    vel[1] = vB
    return pos, vel, np.array([1.0, mB]), d


def test_unbind_shear_frame_sign():
    """Member B at distance d on the x axis, d^3 = 4 G M / (q Omega)^2.  At rest
    in the shear flow (v_y = q Omega d) it is bound (e = -G/d); at rest in the
    box frame, or moving against the shear, it leaves (e = (q Omega d)^2/2M - G/d
    = G/d > 0).  A flipped shear sign swaps the outcomes."""
    G, Om, q = consts()
    _, _, _, d = two_body(0)
    for vy, kept in ((q * Om * d, 2), (0.0, 1), (2 * q * Om * d, 1)):
        pos, vel, m, _ = two_body([0.0, vy, 0.0])
        r = O.clump_op(P, "unbind", pos, vel, m)
        assert len(r["order"]) == kept, vy


def test_unbind_kinetic_term_uses_total_mass_before_removal():
    """e = 0.5 / M_total |dv|^2 - G / d with M_total including the tested member
    (src/tree.hpp:4656-4659): with m_B = 0.5 and |dv|^2 = 2.5 G / d the member is
    bound (0.5 / 1.5 * 2.5 = 0.83 < 1); with the remaining mass (1.0) it
    would not be (1.25 > 1)."""
    G, Om, q = consts()
    d = 0.01
    v = math.sqrt(2.5 * G / d)
    pos = np.array([[0.0, 0, 0], [d, 0, 0]])
    vel = np.zeros((2, 3))
    vel[1] = [0.0, q * Om * d, v]             # shear-compensated, plus v along z
    r = O.clump_op(P, "unbind", pos, vel, np.array([1.0, 0.5]))
    assert len(r["order"]) == 2
    vel[1, 2] = math.sqrt(3.5 * G / d)       # 0.5 / 1.5 * 3.5 = 1.17 > 1: leaves
    r = O.clump_op(P, "unbind", pos, vel, np.array([1.0, 0.5]))
    assert len(r["order"]) == 1


def test_unbind_never_tests_the_innermost_member():
    """Two members flying apart: only the outer one is tested; the innermost
    stays even though it is as unbound (the loop stops at i = 1)."""
    G, Om, q = consts()
    d = 0.01
    pos = np.array([[0.0, 0, 0], [d, 0, 0]])
    vel = np.array([[0.0, 0, -50.0], [0.0, q * Om * d, 50.0]])
    r = O.clump_op(P, "unbind", pos, vel, np.array([1.0, 0.9]))
    assert list(r["order"]) == [0]
    assert r["M"] == 1.0 and np.all(r["X"] == 0.0)


def test_unbind_updates_the_state_after_each_removal():
    """A fast outer member C is removed first; the next member B is then tested
    against A alone.  Against A + C (no state update) B would leave: C's
    momentum shifts the centre-of-mass velocity by far more than B's escape
    speed."""
    G, Om, q = consts()
    pos = np.array([[0.0, 0, 0], [0.01, 0, 0], [0.0, 0.0, 0.05]])
    vel = np.zeros((3, 3))
    vel[1, 1] = q * Om * 0.01                 # B co-moving with A in the shear frame
    vel[2, 0] = 50.0                          # C: far above escape speed
    m = np.array([1.0, 1e-3, 0.5])
    r = O.clump_op(P, "unbind", pos, vel, m)
    assert sorted(r["order"]) == [0, 1]
    assert r["M"] == pytest.approx(1.001, rel=1e-15)


# ------------------------------------------------------------- Hill radius

def test_hill_radius_closed_form():
    G, Om, q = consts()
    m = np.array([0.75, 0.5, 0.25])
    pos = np.array([[0.0, 0, 0], [1e-3, 0, 0], [0, 2e-3, 0]])
    r = O.clump_op(P, "hill", pos, np.zeros((3, 3)), m)
    assert r["hill"] == pytest.approx((G * 1.5 / (3 * Om * Om)) ** (1 / 3), rel=1e-15)


# -------------------------------------------------------- angular momentum

def test_angular_momentum_of_rotating_rings():
    """Two coplanar rings spinning at omega about the centre, velocities given
    in the shear frame.  Each member contributes m (omega + Omega) R^2 to J_z
    (the Omega (dx^2 + dy^2) term is the frame rotation); J is normalised by
    M R_H^2 Omega, and J_quarter keeps members within R_H / 4."""
    G, Om, q = consts()
    n, M = 16, 2.0
    RH = (G * M / (3 * Om * Om)) ** (1 / 3)
    R1, R2, w = RH / 8, RH / 2, 3.0
    ang = np.arange(n) * 2 * np.pi / n
    pos, vel = [], []
    for R in (R1, R2):
        x, y = R * np.cos(ang), R * np.sin(ang)
        pos.append(np.stack([x, y, np.zeros(n)], 1))
        vel.append(np.stack([-w * y, w * x + q * Om * x, np.zeros(n)], 1))
    pos, vel = np.concatenate(pos), np.concatenate(vel)
    m = np.full(2 * n, M / (2 * n))
    r = O.clump_op(P, "J", pos, vel, m)
    assert r["hill"] == pytest.approx(RH, rel=1e-14)
    jz = (w + Om) * (R1 * R1 + R2 * R2) / 2 / (RH * RH * Om)
    jq = (w + Om) * (R1 * R1) / 2 / (RH * RH * Om)
    assert r["J"][2] == pytest.approx(jz, rel=1e-12)
    assert r["Jq"][2] == pytest.approx(jq, rel=1e-12)
    assert abs(r["J"][0]) < 1e-12 * jz and abs(r["J"][1]) < 1e-12 * jz


# --------------------------------------------------------------- rebinding

def test_rebind_skips_flagged_and_low_density_candidates():
    G, Om, q = consts()
    st = [1.0, 0, 0, 0, 0, 0, 0]
    pos = np.array([[0.01, 0, 0]] * 3)
    vel = np.array([[0.0, q * Om * 0.01, 0]] * 3)  # all three bound to the clump
    acc, _ = O.rebind_candidates(P, st, pos, vel, [1e-3] * 3, [10.0, 1.0, 10.0], [1, 0, 0], thr=5.0)
    assert list(acc) == [False, False, True]   # flagged, below threshold, joins


def test_rebind_centre_update_uses_new_mass_on_both_sides():
    """The reference updates X' = (X M' + m x) / M' with M' = M + m
    (src/tree.hpp:4697-4698), i.e. X' = X + m x / M', instead of
    (X M + m x) / M'.  From X = (1, 0, 0), M = 1, a member of mass 0.5 at
    x = 1.02 gives X'_x = 1 + 0.5 * 1.02 / 1.5 = 1.34."""
    G, Om, q = consts()
    st = [1.0, 1.0, 0.0, 0.0, 0.0, 0.25, 0.0]
    pos = np.array([[1.02, 0, 0]])
    vel = np.array([[0.0, 0.25 - q * Om * (1.0 - 1.02), 0]])   # co-moving in the shear frame
    acc, s2 = O.rebind_candidates(P, st, pos, vel, [0.5], [10.0], [0], thr=5.0)
    assert acc[0]
    assert s2[0] == 1.5
    assert s2[1] == pytest.approx(1.0 + 0.5 * 1.02 / 1.5, rel=1e-15)
    assert s2[1] != pytest.approx((1.0 + 0.5 * 1.02) / 1.5, rel=1e-3)
    assert s2[5] == pytest.approx(0.25 + 0.5 * vel[0, 1] / 1.5, rel=1e-15)


def test_rebind_later_candidates_see_the_updated_state():
    """Candidates are tested in order against the state after each join: a
    second member that is only bound to the heavier clump joins after the first
    one, and the order of the list decides it."""
    G, Om, q = consts()
    d = 0.01
    m1, m2 = 1.0, 1e-3
    # e2 = 0.5/(M+m2) v^2 - G/d2 with v^2 between the two thresholds
    d2 = 0.02
    v_lo = math.sqrt(2 * G / d2 * (1.0 + m2))      # bound iff M >= 1 + ...
    v = 1.2 * v_lo                                # needs M > 1.2 (1 + m2) - m2
    st = [1.0, 0, 0, 0, 0, 0, 0]
    pos = np.array([[d, 0, 0], [0, 0, d2]])
    vel = np.array([[0.0, q * Om * d, 0], [0.0, 0.0, v]])
    acc, s2 = O.rebind_candidates(P, st, pos, vel, [m1, m2], [10.0, 10.0], [0, 0], thr=5.0)
    assert list(acc) == [True, True]
    acc, _ = O.rebind_candidates(P, st, pos[::-1], vel[::-1], [m2, m1], [10.0, 10.0], [0, 0], thr=5.0)
    assert list(acc) == [False, True]


# ------------------------------------------------------------------ merging

def test_merge_two_clumps_closed_form():
    """MergeAnotherPlanetesimal: mass-weighted centre and velocity, the denser
    peak wins (ties keep the predator's), the other becomes a subpeak, members
    re-sorted by distance to the new centre."""
    pos = np.array([[0.0, 0, 0], [0.25, 0, 0], [4.0, 0, 0], [4.5, 0, 0]])
    vel = np.array([[1.0, 0, 0], [1.0, 0, 0], [0, 2.0, 0], [0, 2.0, 0]])
    m = np.array([1.0, 1.0, 2.0, 2.0])
    rho = np.array([5.0, 1.0, 7.0, 1.0])
    r = O.merge_two(P, 2, pos, vel, m, rho, peak_a=0, peak_b=2)
    assert r["M"] == 6.0
    assert r["X"][0] == pytest.approx((2 * 0.125 + 4 * 4.25) / 6, rel=1e-15)
    assert r["V"][0] == pytest.approx(2 / 6, rel=1e-15) and r["V"][1] == pytest.approx(8 / 6, rel=1e-15)
    assert r["peak"] == 2 and r["subpeak"] == 0
    X = r["X"][0]
    assert list(r["order"]) == list(np.argsort(np.abs(pos[:, 0] - X), kind="stable"))
    rho[2] = 5.0                                  # tie: the predator keeps its peak
    r = O.merge_two(P, 2, pos, vel, m, rho, peak_a=0, peak_b=2)
    assert r["peak"] == 0 and r["subpeak"] == 2


def test_gravitationally_bound_pair():
    """IsGravitationallyBound: 0.5/(Ma+Mb) |Va - s dx - Vb|^2 - G/d < 0, with
    the shear s = (0, q Omega, 0) and dx = Xa_x - Xb_x (src/tree.hpp:4867-4874)."""
    G, Om, q = consts()
    d = 0.02
    a = [1.0, d, 0, 0, 0, 0, 0]
    b = [0.5, 0, 0, 0, 0, 0, 0]
    b_comoving = list(b)
    b_comoving[5] = -q * Om * d                   # Vb = Va - s dx: no relative motion
    assert O.clumps_bound(P, a, b_comoving)
    b_rest = list(b)                              # |dv| = q Omega d
    ek = 0.5 / 1.5 * (q * Om * d) ** 2
    assert O.clumps_bound(P, a, b_rest) == (ek - G / d < 0)
    fast = list(b)
    fast[6] = math.sqrt(2 * 1.5 * G / d) * 1.01   # just above escape along z
    fast[5] = -q * Om * d
    assert not O.clumps_bound(P, a, fast)
    fast[6] = math.sqrt(2 * 1.5 * G / d) * 0.99
    assert O.clumps_bound(P, a, fast)


# ---------------------------------------------------------------- filters

def test_small_mass_low_peak_and_grid_density_veto():
    """RemoveSmallMassAndLowPeak (src/tree.hpp:3524-3546): a clump goes when
    its mass is below the trusted mass, its peak's density below the peak
    threshold, or when none of its first min(n, K) members has a grid
    density (ath) at or above the peak threshold (src/tree.hpp:3533-3540)."""
    K, mthr, pthr = 4, 2.5, 30.0
    #            kept     light    low peak  ath veto  ath beyond K  ath exactly thr
    sizes = [5, 2, 5, 5, 6, 5]
    off = np.concatenate([[0], np.cumsum(sizes)])
    n = off[-1]
    mass = np.ones(n)
    mass[off[1]:off[2]] = 1.0                    # 2 members -> M = 2 < 2.5
    rho = np.full(n, 40.0)
    rho[off[2]] = 20.0                           # peak density below 30
    ath = np.full(n, 1.0)
    ath[off[0] + 2] = 31.0                       # kept: a member within the first K
    ath[off[1]] = 99.0
    ath[off[2]] = 99.0
    ath[off[4] + 4] = 99.0                       # only the 5th member: beyond K = 4
    ath[off[5] + 3] = 30.0                       # >= threshold counts
    keys = off[:-1].copy()                       # the first member is the peak
    keep = O.filter_clumps(off, keys, mass, rho, ath, K, mthr, pthr)
    assert list(keep) == [True, False, False, False, False, True]


def test_ranking_and_cap():
    """Ranking by (mass, peak index) and the num_peaks cap drop the lightest;
    equal masses fall by the smaller peak index first."""
    keys = [40, 10, 30, 20, 50]
    mass = [3.0, 1.0, 2.0, 2.0, 5.0]
    keep, tot, mx = O.rank_clumps(keys, mass, 3)
    # order (1,10) (2,20) (2,30) (3,40) (5,50): drop peaks 10 and 20
    assert list(keep) == [True, False, True, False, True]
    assert tot == 10.0 and mx == 5.0
    keep, tot, mx = O.rank_clumps(keys, mass, 0)
    assert keep.all() and tot == 13.0 and mx == 5.0
