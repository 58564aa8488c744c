"""Seeded snapshot cases shared by the parity tests."""
from plan_amd import synth


def dense_case(n, n_clumps, seed, a=4e-4, frac=0.3, knn=32, hop=None, nx=None, decades=1.0):
    import math
    nx = nx or max(2, int(math.isqrt(n // 4)))
    return synth.make_snapshot(n, n_clumps, frac, a_range=(a * 0.5, a * 2), mass_decades=decades, seed=seed,
                               nx=nx, knn=knn, hop=hop)


def duplicates_case():
    """A clumpy snapshot with 100 particles at one position (one sink, a run
    of equal keys longer than the sort's run fix-up takes) and 40 particles a
    few float32 steps apart in x (equal high 64 key bits, distinct low word)."""
    import numpy as np
    snap = dense_case(8192, 6, 8, a=3e-4)
    pos = snap.pos
    pos[100:200] = pos[100]
    base = pos[300].copy()
    for k in range(40):
        pos[300 + k] = base
        pos[300 + k, 0] = np.nextafter(base[0], np.float32(1), dtype=np.float32) if k else base[0]
        for _ in range(k - 1):
            pos[300 + k, 0] = np.nextafter(pos[300 + k, 0], np.float32(1), dtype=np.float32)
    return snap


def species_time_case():
    """Two species with mass weights 1 : 3 (per-type particle masses,
    src/tree.hpp:1571-1582) at time 0.37, so the ghost images in y move with
    the shear (src/tree.hpp:1660-1675)."""
    import numpy as np
    snap = dense_case(32768, 10, 11, a=4e-4)
    snap.params["type_info"] = [1.0, 3.0]
    snap.params["time"] = 0.37
    snap.property_index = np.random.default_rng(11).integers(0, 2, snap.n).astype(np.int32)
    return snap


def options_case():
    """The parameter-file options of the path set away from their defaults:
    --No_Ghost (src/tree.hpp:1627-1629), num_peaks = 4 (the ranking cap,
    src/tree.hpp:4301-4316), clump_diffuse_threshold = 0.45 and
    Hill_fraction_for_merge = 0.5 (src/global.cpp:199-219)."""
    snap = dense_case(32768, 12, 12, a=4e-4)
    snap.params.update(no_ghost=1, num_peaks=4, clump_diffuse_threshold=0.45, Hill_fraction_for_merge=0.5)
    return snap


def c5_small_case():
    """The 134M configuration's mix (85% uniform, 15% in 2,000 Plummer clumps,
    a in [2e-4, 3e-3], K = 64) at 2^22 particles, H = 64 as the full size
    takes it (src/tree.hpp:3585-3586)."""
    snap = synth.config_c5(seed=0, n=1 << 22)
    snap.params.update(hop=64, fixed_hop=1)
    return snap


def lattice_case(knn=32):
    """51^3 particles on the cubic lattice k 2^-8, |k| <= 25: every pairwise
    difference is exact, so squared distances tie exactly and most K-th
    neighbours are tied (the index tie-break decides)."""
    import numpy as np
    k = np.arange(-25, 26, dtype=np.float64) * 2.0 ** -8
    g = np.stack(np.meshgrid(k, k, k, indexing="ij"), -1).reshape(-1, 3)
    n = g.shape[0]
    params = synth.default_params(n, nx=64, knn=knn)
    vel = np.zeros((n, 3), np.float32)
    vel[:, 1] = (1.5 * g[:, 0]).astype(np.float32)
    dens = np.full(n, params["Z"] * np.sqrt(2 * np.pi) / 0.2, np.float32)
    return synth.Snapshot(pos=g.astype(np.float32), vel=vel, density=dens, property_index=np.zeros(n, np.int32),
                          id=np.arange(n, dtype=np.uint32), params=params)


def key0_case():
    """A clump keyed by tree index 0 that survives merge loop A, so the
    reference's zero-filled delete list erases it (src/tree.hpp:3874 ->
    4000-4003).  Without ghosts the particle at box_min in all three
    coordinates is the first in Morton order; 60 duplicates there make it a
    heavy sink, 16 particles about 1e-5 away along the three axes and one at
    4e-5 on the diagonal have it as their densest neighbour (its K/2-th
    neighbour is nearer than theirs), and the group is compact enough to
    pass the dispersion filters."""
    import numpy as np
    snap = dense_case(32768, 10, 13, a=4e-4)
    snap.params.update(no_ghost=1)
    rng = np.random.default_rng(13)
    c = np.float64(np.float32(-0.1))
    eps = 1e-5
    pts = [np.array([c, c, c])] * 60
    for ax, cnt in ((0, 6), (1, 5), (2, 5)):
        for _ in range(cnt):
            o = np.abs(rng.normal(0, 0.03, 3)) * eps
            o[ax] = eps * (1 + 0.05 * rng.uniform())
            pts.append(c + o)
    pts.append(c + 4 * eps / np.sqrt(3) * np.ones(3))
    pts = np.array(pts).astype(np.float32)
    k = len(pts)
    snap.pos[:k] = pts
    snap.vel[:k] = 0
    snap.vel[:k, 1] = 1.5 * pts[:, 0]
    snap.density[:k] = 1e6
    return snap


CASES = {
    # tree small enough for the one-workgroup sorts of BuildTree and HOP
    "n1k": lambda: dense_case(900, 2, 9, a=2e-4),
    "n4k": lambda: dense_case(4096, 4, 1, a=2e-4),
    "n8k": lambda: dense_case(8192, 6, 2, a=3e-4),
    "n32k": lambda: dense_case(32768, 10, 3, a=4e-4),
    "n65k": lambda: dense_case(65536, 16, 4, a=5e-4),
    "n32k_k16": lambda: dense_case(32768, 10, 5, a=4e-4, knn=16),
    "n32k_k64_h64": lambda: dense_case(32768, 10, 6, a=4e-4, knn=64, hop=64),
    "n16k_qs": lambda: dense_case(16384, 6, 7, a=3e-4, nx=128),
    # neighbour counts other than 8, 16, 32, 64 (the kernels keep the kq
    # nearest of the next of those): K = 48 with a fixed H = 24, and K = 3
    # (start node of one leaf, kq / 4 = 0) with H = 5
    "n32k_k48_h24": lambda: dense_case(32768, 10, 14, a=4e-4, knn=48, hop=24),
    "n8k_k3_h5": lambda: dense_case(8192, 6, 15, a=3e-4, knn=3, hop=5),
    # the benchmark's mix (200 Plummer clumps in 10% of the particles) at 2^20
    "c2_1m": lambda: synth.config_c2(seed=3, n=1 << 20),
    # the reference's CPU-runnable configuration: 262,144 particles, 16
    # Plummer spheres in 10%, K = 32, seed 0
    "c1": lambda: synth.config_c1(seed=0),
    "dup": duplicates_case,
    "species_time": species_time_case,
    "options": options_case,
    "key0": key0_case,
    "c5_4m": c5_small_case,
    "lattice": lattice_case,
    "lattice_k20": lambda: lattice_case(knn=20),
}
