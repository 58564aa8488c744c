"""Pins of the CPU restatement of the catalog files (src/tree.hpp:4935-5222):
the offset statistics against closed forms (no rotation when J is along z,
members along J land on the z axis after the rotation, offsets below 1e-32
count as 1e-32), and the text layout parsed back against the values."""
import math

import numpy as np
import pytest

import oracle_py as O


def _clump(peak, mass, com, J, off, n, **kw):
    return dict(peak_index=peak, num_members=n, num_particles=n, member_offset=off, total_mass=mass,
                center_of_mass=tuple(com), vel_com=(0.5, -0.25, 0.125), Hill_radius=0.01 * mass, J=tuple(J),
                J_quarter_Hill=(0.0, 0.0, 0.0), one10th_radius=0.0, two_sigma_mass_radius=0.0, max_distance=0.0,
                peak_density=0.0, **kw)


def test_offsets_without_rotation():
    rng = np.random.default_rng(0)
    com = (0.01, -0.02, 0.03)
    pos = np.array(com) + rng.normal(0, 1e-3, (101, 3))
    geo, med = O.peak_offsets(com, (0.0, 0.0, 2.5), pos)
    off = np.abs(pos - np.array(com))
    assert med == list(np.median(off, axis=0))
    np.testing.assert_allclose(geo, np.exp(np.log(off).mean(axis=0)), rtol=1e-13)
    geo2, med2 = O.peak_offsets(com, (0.0, 0.0, 2.5), pos[:100])  # even count: mean of the middle two
    assert med2 == list(np.median(off[:100], axis=0))


def test_members_along_J_end_on_z():
    J = (1.0, 2.0, -0.5)
    u = np.array(J) / np.linalg.norm(J)
    com = (0.0, 0.0, 0.0)
    pos = np.outer(np.linspace(1e-3, 2e-3, 11), u)
    geo, med = O.peak_offsets(com, J, pos)
    np.testing.assert_allclose(med[2], np.median(np.linspace(1e-3, 2e-3, 11)), rtol=1e-14)
    assert med[0] < 1e-17 and med[1] < 1e-17
    # the transposed rotation would leave the members off the z axis
    assert med[2] > 1e-3


def test_tiny_offsets_count_as_1e_minus_32():
    com = (0.1, 0.2, 0.3)
    pos = np.array([com, (0.1 + 1e-3, 0.2 + 1e-3, 0.3 + 1e-3)])
    geo, _ = O.peak_offsets(com, (0.0, 0.0, 1.0), pos)
    for d in range(3):
        want = 10.0 ** ((-32 + math.log10(abs(pos[1][d] - com[d]))) / 2)
        assert geo[d] == pytest.approx(want, rel=1e-13)


def _catalog():
    c0 = _clump(7, 2.0, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 0, 3)
    c1 = _clump(3, 1.0, (0.5, 0.5, 0.5), (1.0, 0.0, 0.0), 3, 2)
    return dict(clumps=[c1, c0], members=np.array([0, 1, 2, 3, 4], np.uint32),
                member_original_ids=np.array([10, 11, 12, 13, 14], np.uint32), density_threshold=1.0,
                peak_density_threshold=3.0, clump_mass_threshold=0.0, Mp_max=2.0, Mp_tot=3.0)


def test_peaks_text_layout():
    rng = np.random.default_rng(1)
    pos = rng.normal(0, 1e-2, (5, 3))
    cat = _catalog()
    cat["clumps"][0]["total_mass"] = 2.0  # equal masses: ranked by peak index
    txt = O.peaks_text(cat, lambda t: [pos[t]])
    lines = txt.splitlines(keepends=True)
    assert lines[0] == O.PEAKS_HEADER and len(lines[0]) == 24 + 17 * 24 + 1
    assert [int(l[:12]) for l in lines[1:]] == [3, 7]
    for l, c in zip(lines[1:], [cat["clumps"][0], cat["clumps"][1]]):
        assert len(l) == 24 + 17 * 24 + 1
        vals = [float(l[24 + 24 * k:48 + 24 * k]) for k in range(17)]
        assert int(l[12:24]) == c["num_members"]
        np.testing.assert_allclose(vals[:11], [c["total_mass"], c["Hill_radius"], *c["center_of_mass"],
                                                *c["vel_com"], *c["J"]], rtol=1e-15)


def test_clump_particles_text_round_trips():
    rng = np.random.default_rng(2)
    tp = dict(pos=rng.normal(0, 1, (5, 3)), vel=rng.normal(0, 1, (5, 3)), mass=rng.uniform(1, 2, 5))
    cat = _catalog()
    txt = O.clump_particles_text(cat, 1, tp)
    rows = [l.split() for l in txt.splitlines()]
    assert [int(r[0]) for r in rows] == [10, 11, 12]
    for r, t in zip(rows, [0, 1, 2]):
        assert [float(x) for x in r[1:]] == [*tp["pos"][t], *tp["vel"][t], tp["mass"][t]]
    assert len(O.clump_particles_text(cat, 1, tp, rate=2).splitlines()) == 2


def test_basic_results_text():
    cat = _catalog()
    txt = O.basic_results_text(cat, 12.5, 4.0, 0.05, True)
    head, row = txt.splitlines()
    assert head.split() == ["#time", "N_peak", "M_p,max,code", "M_p,tot,code", "M_p,max,frac", "M_p,tot,frac",
                            "M_p,max,Ceres", "M_p,tot,Ceres"]
    v = [float(x) for x in row.split()]
    mp = 0.5 * 1.3e24
    np.testing.assert_allclose(v, [12.5, 2, 2.0, 3.0, 0.5, 0.75, 2.0 * mp / 9.4e20, 3.0 * mp / 9.4e20], rtol=1e-6)
    assert len(row) == 8 * 15
