"""The density-against-scale pass on the device tree (plan_density_vs_scale)
against the CPU restatement: radii, densities and the per-level sums of
every search's count equal, and the reference's max_rhop_vs_scale lines."""
import numpy as np
import pytest

from test_oracle_scale import scale_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,nx,time", [(6000, 8, 0.37), (20000, 16, 0.0), (20000, 16, 2.9), (3000, 4, 1.0)])
def test_density_vs_scale(tmp_path, n, nx, time):
    import oracle_py as O
    import plan_amd

    snap = scale_case(n=n, nx=nx, time=time)
    print(f"[scale] n={n} nx={nx} time={time}", flush=True)
    want = O.build_tree(snap).density_vs_scale()
    ctx = plan_amd.Context(0)
    ctx.set_particles(snap.params, snap.pos, snap.vel, snap.density, snap.property_index, snap.id)
    ctx.make_ghosts()
    ctx.build_tree()
    path = tmp_path / "max_rhop_vs_scale.txt"
    got = ctx.density_vs_scale(path, header=True)
    assert np.array_equal(got["radii"], want["radii"])
    assert np.array_equal(got["sums"], want["sums"])
    assert np.array_equal(got["max_rhop"], want["max_rhop"])
    text = "%15s" % "#time" + "".join("%15.6e" % r for r in want["radii"]) + "\n"
    text += "%15.6e" % time + "".join("%15.6e" % v for v in want["max_rhop"]) + "\n"
    assert path.read_text() == text
