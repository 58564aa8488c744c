"""Several snapshots on concurrent device contexts (plan_amd.pipeline).

Run independently, each catalog (and each catalog file written from the
worker) equals the one a single context gives.  Run as a series, the
snapshots get the parameters the reference's loop hands them (the first
snapshot's box, species weights and particle masses; H carried), and every
catalog equals the CPU restatement's loop over the same snapshots."""
import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same_catalog(cat, want):
    assert [c["peak_index"] for c in cat["clumps"]] == [c["peak_index"] for c in want["clumps"]]
    assert np.array_equal(cat["member_original_ids"], want["member_original_ids"])
    assert [c["total_mass"] for c in cat["clumps"]] == [c["total_mass"] for c in want["clumps"]]
    assert cat["Mp_tot"] == want["Mp_tot"]


def test_independent_snapshots_match_single_context(tmp_path):
    import plan_amd
    from _cases import CASES
    from plan_amd import synth
    from plan_amd.pipeline import find_clumps_series

    snaps = [CASES["n32k"](), CASES["n8k"](), CASES["c1"]()]
    lis = synth.write_lis(snaps[2], str(tmp_path / "c1"), num_files=2, seed=3)
    items = [snaps[0], snaps[1], lis]

    def write(ctx, params, cat):
        ctx.write_peaks(cat, tmp_path / f"peaks_{len(cat['members'])}_{cat['Mp_tot']!r}.txt")

    print(f"[pipeline] sizes {[s.n for s in snaps]}", flush=True)
    out = find_clumps_series(items[:2], contexts=2, write=write, series=False) + \
        find_clumps_series(items[2:], params=snaps[2].params, contexts=2, write=write, series=False)
    for s, (p, cat) in zip(snaps, out):
        ctx = plan_amd.Context(0)
        ctx.set_particles(s.params, s.pos, s.vel, s.density, s.property_index, s.id)
        ctx.make_ghosts()
        ctx.build_tree()
        want = ctx.find_planetesimals()
        assert len(want["clumps"]) > 0
        _same_catalog(cat, want)
        ctx.write_peaks(want, tmp_path / "want.txt")
        got = tmp_path / f"peaks_{len(cat['members'])}_{cat['Mp_tot']!r}.txt"
        assert got.read_text() == (tmp_path / "want.txt").read_text()


def _lose_particles(snap, keep):
    """The snapshot after some particles left the run: the first `keep`
    particles, ids unchanged."""
    s = copy.deepcopy(snap)
    s.pos, s.vel, s.density = s.pos[:keep].copy(), s.vel[:keep].copy(), s.density[:keep].copy()
    s.property_index, s.id = s.property_index[:keep].copy(), s.id[:keep].copy()
    s.params = dict(s.params, time=s.params.get("time", 0.0) + 0.25)
    return s


def test_series_matches_the_oracle_loop():
    """Three snapshots of 32768, 8192 and 31000 particles (two species in the
    first): the later two take the first one's masses (M_tot / 32768 per
    unit weight), so their densities, thresholds and catalogs differ from
    runs of their own; each catalog equals the restatement's loop."""
    import oracle_py as O
    from _cases import CASES
    from plan_amd.pipeline import find_clumps_series

    s0 = CASES["species_time"]()
    s1 = CASES["n8k"]()
    s2 = _lose_particles(s0, 31000)
    snaps = [s0, s1, s2]
    out = find_clumps_series(snaps, contexts=2)
    trees = O.run_series(snaps)
    assert [p["series_first_n"] for p, _ in out[1:]] == [s0.n, s0.n]
    for (p, cat), t in zip(out, trees):
        _same_catalog(cat, t.catalog())
    # the later snapshots' masses come from the first's count
    alone = O.run_pipeline(s2)
    assert alone.catalog()["Mp_tot"] != trees[2].catalog()["Mp_tot"]
    print(f"[series] clumps {[len(c['clumps']) for _, c in out]}", flush=True)


def test_lis_series_keeps_the_first_header_box(tmp_path):
    """.lis snapshots of a series: the second file's header box (wider here)
    is ignored, as ReadLisFile ignores it after loop_begin; catalogs equal
    the restatement's loop over the same particles with the first box."""
    import oracle_py as O
    from _cases import CASES
    from plan_amd import synth
    from plan_amd.pipeline import find_clumps_series

    s0 = CASES["c1"]()
    s1 = _lose_particles(s0, s0.n - 5000)
    s1.params = dict(s1.params, box_min=[-0.1005] * 3, box_max=[0.1005] * 3)
    paths = [synth.write_lis(s0, str(tmp_path / "a"), num_files=2, seed=3),
             synth.write_lis(s1, str(tmp_path / "b"), num_files=1, seed=4)]
    out = find_clumps_series(paths, params=s0.params, contexts=2)
    p0, p1 = out[0][0], out[1][0]
    assert p1["box_min"] == p0["box_min"] and p1["box_max"] == p0["box_max"]
    assert p1["series_first_n"] == s0.n
    # the restatement's loop on the same particles and the header values
    # the first file carries (float32 box, time)
    r0 = dict(s0.params, box_min=p0["box_min"], box_max=p0["box_max"], time=p0["time"])
    r1 = dict(s1.params, time=p1["time"])
    a = copy.copy(s0)
    a.params = r0
    b = copy.copy(s1)
    b.params = r1
    trees = O.run_series([a, b])
    for (p, cat), t in zip(out, trees):
        _same_catalog(cat, t.catalog())
