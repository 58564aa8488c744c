"""ReadLisFile on the device (plan_read_lis, plan_set_particles_lis) against
the CPU restatement: the particle set bit-exact, the header's fields, the
same catalog as the arrays loaded directly, and the error paths."""
import numpy as np
import pytest

from test_oracle_lis import _hand_file

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def snap():
    from plan_amd import synth
    return synth.small_case(20000, 4, seed=3)


def _same_set(got, want):
    for k in ("pos", "vel", "density", "property_index", "id"):
        assert np.array_equal(got[k], want[k]), k


@pytest.mark.parametrize("num_files,high_bits", [(1, False), (3, True), (8, False)])
def test_read_lis_matches_oracle(tmp_path, snap, num_files, high_bits):
    import oracle_py as O
    import plan_amd
    from plan_amd import synth

    paths = synth.write_lis(snap, str(tmp_path / "s"), num_files=num_files, seed=11, high_bits=high_bits)
    print(f"[lis] files={num_files} high_bits={high_bits} n={snap.n}", flush=True)
    want = O.read_lis(paths)
    ctx = plan_amd.Context(0)
    keys = {k: v for k, v in snap.params.items() if k not in ("box_min", "box_max", "time", "type_info")}
    keys.update(box_min=[0.0] * 3, box_max=[1.0] * 3, type_info=[1.0])
    params, dt, n = ctx.read_lis(keys, paths)
    assert n == snap.n
    assert params["box_min"] == list(want["box_min"]) and params["box_max"] == list(want["box_max"])
    assert params["time"] == want["time"] and dt == want["dt"] and params["type_info"] == list(want["type_info"])
    _same_set(ctx.particle_set(), want)


def test_hand_made_file_small_sort(tmp_path):
    """Five records (the one-workgroup sort), key cuts and header fields."""
    import oracle_py as O
    import plan_amd
    from plan_amd import synth

    p = str(tmp_path / "hand.lis")
    _hand_file(p)
    want = O.read_lis(p)
    ctx = plan_amd.Context(0)
    params, dt, n = ctx.read_lis(synth.default_params(5, nx=2), p)
    assert n == 5 and params["type_info"] == [3.0, 1.0]
    _same_set(ctx.particle_set(), want)


def test_records_in_memory(snap):
    import oracle_py as O
    import plan_amd
    from plan_amd import synth

    rec, _ = synth.lis_records(snap, num_cpus=6, seed=5, high_bits=True)
    ctx = plan_amd.Context(0)
    assert ctx.set_particles_lis(snap.params, rec) == snap.n
    got = ctx.particle_set()
    assert np.array_equal(got["pos"], snap.pos.astype(np.float64))
    assert np.array_equal(got["vel"], snap.vel.astype(np.float64))
    assert np.array_equal(got["density"], snap.density.astype(np.float64))
    assert np.array_equal(got["id"], np.arange(snap.n, dtype=np.uint32))


def test_catalog_same_as_arrays(tmp_path):
    """Ghosts -> BuildTree -> FindPlanetesimals after plan_read_lis gives the
    catalog of the same snapshot loaded as arrays."""
    import plan_amd
    from _cases import CASES
    from plan_amd import synth

    snap = CASES["n32k"]()
    paths = synth.write_lis(snap, str(tmp_path / "e"), num_files=2, seed=1)
    a = plan_amd.Context(0)
    ps = plan_amd.ParticleSet.ReadLisFile(paths, snap.params, a)
    ds = plan_amd.DataSet(ps, a)
    ps.MakeGhostParticles(a)
    ds.tree.BuildTree(ps.params, ps)
    got = ds.tree.FindPlanetesimals(ds, -1)
    want = plan_amd.find_clumps(snap)
    assert len(want["clumps"]) > 0
    assert [c["peak_index"] for c in got["clumps"]] == [c["peak_index"] for c in want["clumps"]]
    assert np.array_equal(got["member_original_ids"], want["member_original_ids"])
    assert [c["total_mass"] for c in got["clumps"]] == [c["total_mass"] for c in want["clumps"]]


def test_errors(tmp_path, snap):
    import plan_amd
    from plan_amd import synth

    ctx = plan_amd.Context(0)
    with pytest.raises(plan_amd.PlanError, match="Failed to open"):
        ctx.read_lis(snap.params, str(tmp_path / "missing.lis"))
    short = str(tmp_path / "short.lis")
    _hand_file(short, n_declared=6)
    with pytest.raises(plan_amd.PlanError, match="fewer records"):
        ctx.read_lis(snap.params, short)
    with pytest.raises(plan_amd.PlanError, match="no particle set"):
        ctx.make_ghosts()
    rec, _ = synth.lis_records(snap, seed=2)
    rec["property_index"][17] = 1  # one species only
    with pytest.raises(plan_amd.PlanError, match="property_index"):
        ctx.set_particles_lis(snap.params, rec)
    with pytest.raises(plan_amd.PlanError, match="no particle set"):
        ctx.make_ghosts()
