"""Pins of the FP64 pairwise-boundness oracle (an extension of the path;
no reference counterpart): two- and three-body closed forms, the shear term,
an independent numpy brute force of the whole iteration, and the Plummer
potential energy."""
import math

import numpy as np
import pytest

import oracle_py as O
from plan_amd import synth

G = 0.7


def test_two_body_closed_form():
    # two unit masses at distance 2, at rest in the shear frame
    pos = np.array([[0.0, 0, 0], [2.0, 0, 0]])
    vel = np.array([[0.0, 0.0 * 1.5, 0], [0.0, 1.5 * 2.0 * 0.5 * 0 + 1.5 * 1.0 * 2.0 / 2.0 * 0, 0]])
    # choose velocities so that dv = 0 exactly: v_y = V_y - qO (X_x - x)
    X = 1.0
    vel[:, 1] = -1.5 * (X - pos[:, 0])
    vel[:, 1] -= vel[:, 1].mean()
    bound, o = O.boundness([0, 2], pos, vel, np.ones(2), G)
    assert bound.all()
    assert o["E_pot"][0] == pytest.approx(-G * 1 * 1 / 2.0, rel=1e-15)
    assert o["E_kin"][0] == pytest.approx(0.0, abs=1e-15)
    assert o["iterations"][0] == 1


def test_two_body_unbound_pair_leaves():
    # relative speed above escape: both particles have e > 0 and leave
    pos = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    m = np.ones(2)
    v_esc = math.sqrt(2 * G * 2 / 1.0)
    vel = np.array([[0.0, 0, -v_esc], [0.0, 0, v_esc]])
    bound, o = O.boundness([0, 2], pos, vel, m, G)
    # e_i = (v_esc)^2 / 2 - G = 2G - G > 0 for both
    assert not bound.any()
    assert o["n_bound"][0] == 0


def test_three_body_one_leaves():
    pos = np.array([[0.0, 0, 0], [0.1, 0, 0], [0.0, 0.1, 0]])
    m = np.array([1.0, 1.0, 1e-3])
    vel = np.zeros((3, 3))
    vel[2, 2] = 50.0            # the light one is far above escape speed
    bound, o = O.boundness([0, 3], pos, vel, m, G, Omega=0.0)
    assert list(bound) == [True, True, False]
    assert o["iterations"][0] == 2
    # final pair: E_pot = -G m1 m2 / d, E_kin = 0 (both at rest about their COM)
    assert o["E_pot"][0] == pytest.approx(-G / 0.1, rel=1e-14)
    assert o["E_kin"][0] == pytest.approx(0.0, abs=1e-15)


def numpy_boundness(pos, vel, m, G, qO, max_iter=64):
    B = np.arange(len(m))
    for it in range(1, max_iter + 1):
        if len(B) == 0:
            return B, 0.0, 0.0, it
        p, v, mm = pos[B], vel[B], m[B]
        M = mm.sum()
        X = (mm[:, None] * p).sum(0) / M
        V = (mm[:, None] * v).sum(0) / M
        d = p[:, None, :] - p[None, :, :]
        r = np.sqrt((d * d).sum(-1))
        with np.errstate(divide="ignore"):
            inv = np.where(r > 0, 1.0 / r, 0.0)
        phi = -G * (inv * mm[None, :]).sum(1)
        dv = v - V
        dv[:, 1] += qO * (X[0] - p[:, 0])
        kin = 0.5 * (dv * dv).sum(1)
        ek = (mm * kin).sum()
        ep = 0.5 * (mm * phi).sum()
        keep = ~(kin + phi > 0)
        if keep.all():
            return B, ek, ep, it
        B = B[keep]
    return B, ek, ep, max_iter


def test_against_numpy_brute_force():
    batch = synth.boundness_batch(n_clumps=12, n_min=16, n_max=600, seed=5)
    bound, o = O.boundness(batch.offsets, batch.pos, batch.vel, batch.mass, batch.G, batch.Omega, batch.q)
    for c in range(batch.num_clumps):
        a, b = int(batch.offsets[c]), int(batch.offsets[c + 1])
        B, ek, ep, it = numpy_boundness(batch.pos[a:b].astype(np.float64), batch.vel[a:b].astype(np.float64),
                                        batch.mass[a:b], batch.G, batch.q * batch.Omega)
        assert np.array_equal(np.flatnonzero(bound[a:b]), B)
        assert o["iterations"][c] == it
        assert o["E_pot"][c] == pytest.approx(ep, rel=1e-12)
        assert o["E_kin"][c] == pytest.approx(ek, rel=1e-12)
    # the interlopers (hot shell) mostly leave, the Plummer members mostly stay
    n_int = batch.n_interlopers
    assert bound.sum() < batch.offsets[-1] - 0.5 * n_int.sum()


def test_plummer_potential_energy():
    rng = np.random.default_rng(0)
    n, a, M = 4000, 1.0, 1.0
    x, v, _ = synth.plummer_members(rng, n, a, M, G)
    _, o = O.boundness([0, n], x, v, np.full(n, M / n), G, Omega=0.0, max_iter=1)
    # -3 pi G M^2 / (32 a) for the untruncated sphere (tails cut at 10a change
    # it by < 1%; sampling noise ~ 1/sqrt(n))
    assert o["E_pot"][0] == pytest.approx(-3 * math.pi * G * M * M / (32 * a), rel=0.04)
    # the sampled velocities are the equilibrium ones: 2K/|W| ~ 1
    assert 2 * o["E_kin"][0] / abs(o["E_pot"][0]) == pytest.approx(1.0, abs=0.1)


def test_virial_rescale_puts_energy_near_zero():
    batch = synth.boundness_batch(n_clumps=4, n_min=200, n_max=400, seed=9, virial_targets={1: 1.0})
    a, b = int(batch.offsets[1]), int(batch.offsets[2])
    ek, ep = synth.pairwise_energies_fp64(batch.pos[a:b], batch.vel[a:b], batch.mass[a:b], batch.G)
    assert abs(ek / abs(ep) - 1.0) < 2e-3
