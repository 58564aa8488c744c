"""The 134,217,728-particle configuration (512^3 particles, 85% uniform and
15% in 2,000 Plummer clumps, a in [2e-4, 3e-3], K = 64, H raised to 64):
209.7M tree particles held on one GPU.  Too large for the CPU restatement in
a test, so it is checked through the size-independent properties (the same
mix at 2^22 particles is compared with the restatement in test_gpu_parity,
case c5_4m).  Its merge passes and rebinding run ball searches whose node
frontiers exceed a million entries, the buffers' regrowth path."""
import time

import numpy as np
import pytest

import fullsize_checks as F

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def run():
    import plan_amd
    from plan_amd import synth

    t0 = time.time()
    snap = synth.config_c5()
    print(f"[c5] n={snap.n} generated in {time.time() - t0:.1f} s", flush=True)
    ctx = plan_amd.Context(0)
    ctx.set_particles(snap.params, snap.pos, snap.vel, snap.density, snap.property_index, snap.id)
    T = ctx.make_ghosts()
    info = ctx.build_tree(True)
    print(f"[c5] tree: {info.num_particles} particles, {info.num_nodes} nodes", flush=True)
    t0 = time.time()
    cat = ctx.find_planetesimals()
    print(f"[c5] FindPlanetesimals {time.time() - t0:.2f} s: {len(cat['clumps'])} clumps, "
          f"device memory {ctx.device_memory() / 2**30:.1f} GiB, stats {ctx.clump_stats()}", flush=True)
    return snap, ctx, T, info, cat


def test_keys_and_nodes(run):
    snap, ctx, T, info, cat = run
    F.keys_and_nodes(ctx, info)


def test_density_sample_against_brute_force(run):
    snap, ctx, T, info, cat = run
    assert snap.params["knn"] == 64
    F.density_sample_against_brute_force(ctx, info, 64, nsample=32, tag="c5")


def test_hop_climbs_to_peaks(run):
    snap, ctx, T, info, cat = run
    F.hop_climbs_to_peaks(ctx, info)


def test_catalog_properties(run):
    snap, ctx, T, info, cat = run
    F.catalog_properties(ctx, snap, cat, min_clumps=100)
