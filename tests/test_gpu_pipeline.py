"""Several snapshots on concurrent device contexts (plan_amd.pipeline): each
catalog, and each catalog file written from the worker, equals the one a
single context gives."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_series_matches_single_context(tmp_path):
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
    out = find_clumps_series(items[:2], contexts=2, write=write) + \
        find_clumps_series(items[2:], params=snaps[2].params, contexts=2, write=write)
    for s, (p, cat) in zip(snaps, out):
        ctx = plan_amd.Context(0)
        ctx.set_particles(s.params, s.pos, s.vel, s.density, s.property_index, s.id)
        ctx.make_ghosts()
        ctx.build_tree()
        want = ctx.find_planetesimals()
        assert len(want["clumps"]) > 0
        assert [c["peak_index"] for c in cat["clumps"]] == [c["peak_index"] for c in want["clumps"]]
        assert np.array_equal(cat["member_original_ids"], want["member_original_ids"])
        assert cat["Mp_tot"] == want["Mp_tot"]
        ctx.write_peaks(want, tmp_path / "want.txt")
        got = tmp_path / f"peaks_{len(cat['members'])}_{cat['Mp_tot']!r}.txt"
        assert got.read_text() == (tmp_path / "want.txt").read_text()
