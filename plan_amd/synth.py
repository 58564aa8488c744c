"""Deterministic synthetic particle snapshots for the clump-finding path.

The snapshots mimic what the reference reads from an Athena ``.lis`` file
(src/tree.hpp:1310-1336, 1384-1417): float32 positions and velocities, a
float32 grid density per particle, a property index, and ids.  The
generator places a uniform background plus Plummer spheres in the shearing
box [-0.1, 0.1]^3:

* positions: Plummer radii by the inverse cumulative mass
  r = a / sqrt(u^(-2/3) - 1), truncated at 10 a; x and y are wrapped
  periodically, clump centres keep 10 a_max away from the z faces (z is not
  periodic in the reference, src/tree.hpp:1626-1754).
* velocities: Gaussian background dispersion plus, for clump members, an
  isotropic speed drawn from the Plummer distribution function by rejection
  (Aarseth, Henon & Wielen 1974, A&A 37, 183), v = q v_esc(r) with
  g(q) = q^2 (1 - q^2)^(7/2); then the shear flow v_y += q Omega x
  (src/global.hpp:749-757).
* the ``.lis`` density field (ath_density): the mean particle density plus
  the member's own Plummer density, so clump cores pass the reference's
  3 x threshold peak veto (src/tree.hpp:3533-3540).

``four_pi_G`` is chosen so the reference's density threshold
2 Omega^2 / (pi G) (src/tree.hpp:3633) equals ``thr_over_mean`` times the
mean particle density Z sqrt(2 pi) / L_z of the box.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

BOX = np.float32(0.1)


@dataclass
class Snapshot:
    pos: np.ndarray          # (N, 3) float32
    vel: np.ndarray          # (N, 3) float32
    density: np.ndarray      # (N,) float32  (.lis "density" -> ath_density)
    property_index: np.ndarray  # (N,) int32
    id: np.ndarray           # (N,) uint32  (index after the (cpu_id, id_in_run) sort)
    params: dict = field(default_factory=dict)
    clump_sizes: np.ndarray | None = None
    clump_centres: np.ndarray | None = None
    clump_scale: np.ndarray | None = None

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])


def default_params(n: int, nx: int = 128, knn: int = 32, hop: int | None = None,
                   thr_over_mean: float = 10.0, time: float = 0.0) -> dict:
    """Parameter-file keys and header values for a snapshot of n particles."""
    bmin = float(np.float32(-0.1))
    bmax = float(np.float32(0.1))
    L = bmax - bmin
    Z = 0.02
    mean = Z * math.sqrt(2 * math.pi) / L
    four_pi_G = 4 * math.pi * 2.0 / (math.pi * thr_over_mean * mean)
    p = dict(box_min=[bmin] * 3, box_max=[bmax] * 3, time=time, Z=Z, Nx=[nx, nx, nx],
             four_pi_G=four_pi_G, knn=knn, hop=32 if hop is None else hop,
             fixed_hop=0 if hop is None else 1, num_peaks=0,
             clump_diffuse_threshold=-1.0, Hill_fraction_for_merge=-1.0,
             no_ghost=0, type_info=[1.0], n=n)
    return p


def grav_constant(params: dict) -> float:
    return params["four_pi_G"] / (4.0 * math.pi)


def particle_mass(params: dict, n: int) -> float:
    L0 = params["box_max"][0] - params["box_min"][0]
    L1 = params["box_max"][1] - params["box_min"][1]
    return params["Z"] * L0 * L1 * math.sqrt(2 * math.pi) / n


def _plummer_radii(rng, n, a, rmax_over_a=10.0):
    umax = (1.0 + 1.0 / rmax_over_a ** 2) ** -1.5
    u = rng.uniform(1e-12, umax, n)
    return a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)


def _iso(rng, n):
    z = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2 * math.pi, n)
    s = np.sqrt(1.0 - z * z)
    return np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1)


def _plummer_q(rng, n):
    out = np.empty(n)
    filled = 0
    while filled < n:
        m = max(2 * (n - filled), 1024)
        x = rng.uniform(0.0, 1.0, m)
        y = rng.uniform(0.0, 0.1, m)
        ok = x[y < x * x * (1.0 - x * x) ** 3.5]
        take = min(ok.size, n - filled)
        out[filled:filled + take] = ok[:take]
        filled += take
    return out


def plummer_members(rng, n, a, M, G):
    """Positions (relative) and velocities of an n-member Plummer sphere."""
    r = _plummer_radii(rng, n, a)
    x = _iso(rng, n) * r[:, None]
    vesc = np.sqrt(2.0 * G * M / np.sqrt(r * r + a * a))
    v = _iso(rng, n) * (_plummer_q(rng, n) * vesc)[:, None]
    rho = 3.0 * M / (4.0 * math.pi * a ** 3) * (1.0 + (r / a) ** 2) ** -2.5
    return x, v, rho


def _wrap_xy(x, lo, L):
    return lo + np.mod(x - lo, L)


def _to_box_f32(x64):
    x = x64.astype(np.float32)
    lo, hi = np.float32(-0.1), np.float32(0.1)
    below_hi = np.nextafter(hi, np.float32(0.0))
    x = np.where(x >= hi, below_hi, x)
    x = np.where(x < lo, lo, x)
    return x.astype(np.float32)


def make_snapshot(n: int, n_clumps: int, clump_frac: float, a_range=(0.002, 0.002),
                  mass_decades: float = 0.0, seed: int = 0, sigma: float = 1e-3,
                  nx: int = 128, knn: int = 32, hop: int | None = None,
                  thr_over_mean: float = 10.0) -> Snapshot:
    """Uniform background + Plummer spheres; see the module docstring."""
    rng = np.random.default_rng(seed)
    params = default_params(n, nx=nx, knn=knn, hop=hop, thr_over_mean=thr_over_mean)
    G = grav_constant(params)
    m_par = particle_mass(params, n)
    n_cl = int(round(n * clump_frac)) if n_clumps > 0 else 0
    n_bg = n - n_cl
    lo, L = -0.1, 0.2
    pos = np.empty((n, 3))
    vel = rng.normal(0.0, sigma, (n, 3))
    dens = np.full(n, (params["Z"] * math.sqrt(2 * math.pi) / L), dtype=np.float64)
    pos[:n_bg] = rng.uniform(lo, lo + L, (n_bg, 3))
    sizes = np.zeros(0, dtype=np.int64)
    centres = np.zeros((0, 3))
    scale = np.zeros(0)
    if n_cl > 0:
        w = 10.0 ** rng.uniform(0.0, mass_decades, n_clumps) if mass_decades > 0 else np.ones(n_clumps)
        sizes = np.floor(w / w.sum() * n_cl).astype(np.int64)
        sizes[: n_cl - sizes.sum()] += 1
        a0, a1 = a_range
        scale = np.exp(rng.uniform(math.log(a0), math.log(a1), n_clumps)) if a1 > a0 else np.full(n_clumps, a0)
        zpad = 10.0 * a1
        centres = np.stack([rng.uniform(lo, lo + L, n_clumps), rng.uniform(lo, lo + L, n_clumps),
                            rng.uniform(lo + zpad, lo + L - zpad, n_clumps)], axis=1)
        off = n_bg
        for c in range(n_clumps):
            k = int(sizes[c])
            if k == 0:
                continue
            x, v, rho = plummer_members(rng, k, float(scale[c]), k * m_par, G)
            pos[off:off + k] = centres[c] + x
            vel[off:off + k] += v
            dens[off:off + k] += rho
            off += k
    pos[:, 0] = _wrap_xy(pos[:, 0], lo, L)
    pos[:, 1] = _wrap_xy(pos[:, 1], lo, L)
    pos32 = _to_box_f32(pos)
    vel[:, 1] += 1.5 * pos32[:, 0].astype(np.float64)
    snap = Snapshot(pos=pos32, vel=vel.astype(np.float32), density=dens.astype(np.float32),
                    property_index=np.zeros(n, dtype=np.int32), id=np.arange(n, dtype=np.uint32),
                    params=params, clump_sizes=sizes, clump_centres=centres, clump_scale=scale)
    return snap


# Named workloads (sizes and mixes as the benchmark configurations state them).
def config_c1(seed: int = 0) -> Snapshot:
    """262,144 particles, 90% uniform, 10% in 16 Plummer spheres a=0.002, K=32."""
    return make_snapshot(262144, 16, 0.10, a_range=(0.002, 0.002), seed=seed, knn=32)


def config_c2(seed: int = 0, n: int = 16777216) -> Snapshot:
    """256^3 particles, 200 Plummer clumps, masses log-uniform over 2 decades,
    a log-uniform in [5e-4, 4e-3], K=32."""
    return make_snapshot(n, 200, 0.10, a_range=(5e-4, 4e-3), mass_decades=2.0, seed=seed, knn=32)


def config_c5(seed: int = 0, n: int = 134217728) -> Snapshot:
    """512^3 particles, 15% in 2000 clumps, a in [2e-4, 3e-3], K=64."""
    return make_snapshot(n, 2000, 0.15, a_range=(2e-4, 3e-3), mass_decades=2.0, seed=seed, knn=64)


def small_case(n: int, n_clumps: int, seed: int = 0, nx: int | None = None, knn: int = 32,
               a: float = 0.004, hop: int | None = None) -> Snapshot:
    """Small snapshot for parity tests.  nx defaults to the largest grid for
    which the reference still picks the median spherical kernel
    (N >= 4 Nx Ny, src/tree.hpp:3583)."""
    if nx is None:
        nx = max(2, int(math.isqrt(n // 4)))
    return make_snapshot(n, n_clumps, 0.3, a_range=(a, a), seed=seed, nx=nx, knn=knn, hop=hop)


def to_plan_params_dict(snap: Snapshot) -> dict:
    return dict(snap.params)


# ---------------------------------------------------------------- .lis files

# One particle record of an Athena .lis file (src/tree.hpp:1384-1417): 44
# bytes, little-endian, no padding.
LIS_RECORD = np.dtype([("pos", "<f4", (3,)), ("vel", "<f4", (3,)), ("density", "<f4"),
                       ("property_index", "<i4"), ("id_in_run", "<i8"), ("cpu_id", "<i4")])


def lis_header(snap: Snapshot, n: int, dt: float = 1e-3) -> bytes:
    """float coor_lim[12] (grid lo/hi x3, then domain lo/hi x3), int
    num_types, float type_info[num_types], float time, float dt, long N."""
    p = snap.params
    lim = []
    for k in range(3):
        lim += [p["box_min"][k], p["box_max"][k]]
    ti = p.get("type_info", [1.0])
    return (np.array(lim + lim, "<f4").tobytes() + np.array([len(ti)], "<i4").tobytes()
            + np.array(ti, "<f4").tobytes() + np.array([p.get("time", 0.0), dt], "<f4").tobytes()
            + np.array([n], "<i8").tobytes())


def lis_records(snap: Snapshot, num_cpus: int = 4, seed: int = 0, high_bits: bool = False):
    """Records whose (cpu_id, id_in_run) order is the snapshot's order:
    particle i belongs to process i * num_cpus // n, with ids in the run
    increasing by 1-3 inside a process.  With high_bits the bits the reader
    drops (above 16 of cpu_id, above 32 of id_in_run) are filled at random.
    Returns (records in random order, process of each record)."""
    rng = np.random.default_rng(seed)
    n = snap.n
    cpu = (np.arange(n, dtype=np.int64) * num_cpus) // n
    step = rng.integers(1, 4, n)
    idr = np.empty(n, np.int64)
    for c in range(num_cpus):
        m = cpu == c
        idr[m] = np.cumsum(step[m])
    rec = np.zeros(n, LIS_RECORD)
    rec["pos"] = snap.pos
    rec["vel"] = snap.vel
    rec["density"] = snap.density
    rec["property_index"] = snap.property_index
    rec["id_in_run"] = idr + (rng.integers(0, 1 << 30, n) << 32 if high_bits else 0)
    rec["cpu_id"] = cpu + (rng.integers(0, 1 << 14, n) << 16 if high_bits else 0)
    order = rng.permutation(n)
    return rec[order], cpu[order]


def write_lis(snap: Snapshot, base: str, num_files: int = 1, seed: int = 0, dt: float = 1e-3,
              high_bits: bool = False) -> list:
    """The snapshot as one combined .lis file (num_files = 1, four processes'
    records mixed) or as one file per process; returns the paths in process
    order."""
    rec, cpu = lis_records(snap, num_cpus=4 if num_files == 1 else num_files, seed=seed, high_bits=high_bits)
    if num_files == 1:
        path = f"{base}.lis"
        with open(path, "wb") as f:
            f.write(lis_header(snap, len(rec), dt))
            f.write(rec.tobytes())
        return [path]
    paths = []
    for c in range(num_files):
        part = rec[cpu == c]
        path = f"{base}.{c:04d}.lis"
        with open(path, "wb") as f:
            f.write(lis_header(snap, len(part), dt))
            f.write(part.tobytes())
        paths.append(path)
    return paths


# ---------------------------------------------------------------- boundness batches

@dataclass
class ClumpBatch:
    """Clumps in CSR form for the pairwise-boundness entry points."""
    offsets: np.ndarray     # num_clumps + 1
    pos: np.ndarray         # members x 3 (float32 values)
    vel: np.ndarray         # members x 3 (float32 values)
    mass: np.ndarray        # members
    n_interlopers: np.ndarray
    G: float
    Omega: float = 1.0
    q: float = 1.5

    @property
    def num_clumps(self):
        return len(self.offsets) - 1


def power_law_sizes(rng, n, n_min, n_max, slope):
    """n sizes with dN/dlnN proportional to N^slope on [n_min, n_max]."""
    a, b = math.log(n_min), math.log(n_max)
    u = rng.uniform(0.0, 1.0, n)
    ea, eb = math.exp(slope * a), math.exp(slope * b)
    x = np.log(ea + u * (eb - ea)) / slope
    return np.clip(np.rint(np.exp(x)).astype(np.int64), n_min, n_max)


def boundness_batch(n_clumps: int = 5000, n_min: int = 32, n_max: int = 2_000_000, slope: float = -0.6,
                    interloper_max: float = 0.3, hill_frac=(0.1, 0.3), seed: int = 0,
                    virial_targets: dict | None = None) -> ClumpBatch:
    """Plummer clumps with power-law member counts and 0..interloper_max
    extra unbound interlopers from a hot Gaussian shell (radius 2a, width
    a/2, velocity dispersion 3x the local escape speed).  Particle mass and G
    follow the 16.8M-snapshot parameters (Omega = 1, q = 1.5); each clump's
    scale radius is a log-uniform fraction hill_frac of its Hill radius
    (G M / 3 Omega^2)^(1/3), so the sphere is bound against the shear.
    Positions and velocities are float32 values.  virial_targets {clump: ratio} rescales the velocities of
    those clumps about their mean so that E_kin / |E_pot| (pairwise, FP64)
    equals the ratio (1 puts the total energy at zero)."""
    rng = np.random.default_rng(seed)
    params = default_params(16777216)
    G = grav_constant(params)
    m_par = particle_mass(params, 16777216)
    sizes = power_law_sizes(rng, n_clumps, n_min, n_max, slope)
    frac = rng.uniform(0.0, interloper_max, n_clumps)
    n_int = np.rint(sizes * frac).astype(np.int64)
    tot = sizes + n_int
    offsets = np.zeros(n_clumps + 1, np.int64)
    offsets[1:] = np.cumsum(tot)
    nm = int(offsets[-1])
    pos = np.empty((nm, 3), np.float32)
    vel = np.empty((nm, 3), np.float32)
    hill = (G * sizes * m_par / 3.0) ** (1.0 / 3.0)
    scale = hill * np.exp(rng.uniform(math.log(hill_frac[0]), math.log(hill_frac[1]), n_clumps))
    centre = rng.uniform(-0.1, 0.1, (n_clumps, 3))
    for c in range(n_clumps):
        k, ki, a = int(sizes[c]), int(n_int[c]), float(scale[c])
        M = k * m_par
        x, v, _ = plummer_members(rng, k, a, M, G)
        r = np.abs(rng.normal(2.0 * a, 0.5 * a, ki))
        xi = _iso(rng, ki) * r[:, None]
        vesc = np.sqrt(2.0 * G * M / np.sqrt(r * r + a * a))
        vi = rng.normal(0.0, 1.0, (ki, 3)) * (3.0 * vesc)[:, None]
        o = int(offsets[c])
        pos[o:o + k] = (centre[c] + x).astype(np.float32)
        pos[o + k:o + k + ki] = (centre[c] + xi).astype(np.float32)
        vel[o:o + k] = v.astype(np.float32)
        vel[o + k:o + k + ki] = vi.astype(np.float32)
    mass = np.full(nm, m_par)
    batch = ClumpBatch(offsets=offsets.astype(np.uint32), pos=pos, vel=vel, mass=mass, n_interlopers=n_int, G=G)
    for c, ratio in (virial_targets or {}).items():
        rescale_to_virial(batch, c, ratio)
    return batch


def pairwise_energies_fp64(pos, vel, mass, G, Omega=1.0, q=1.5):
    """E_kin (shear-corrected, about the centre of mass) and the pairwise
    E_pot of one member set, binary64 brute force in numpy."""
    p = np.asarray(pos, np.float64)
    v = np.asarray(vel, np.float64)
    m = np.asarray(mass, np.float64)
    M = m.sum()
    X = (m[:, None] * p).sum(0) / M
    V = (m[:, None] * v).sum(0) / M
    dv = v - V
    dv[:, 1] += q * Omega * (X[0] - p[:, 0])
    ek = 0.5 * (m * (dv * dv).sum(1)).sum()
    ep = 0.0
    for s in range(0, len(m), 2048):
        d = p[s:s + 2048, None, :] - p[None, :, :]
        r = np.sqrt((d * d).sum(-1))
        with np.errstate(divide="ignore"):
            inv = np.where(r > 0, 1.0 / r, 0.0)
        ep += -0.5 * G * (m[s:s + 2048] * (inv * m[None, :]).sum(1)).sum()
    return ek, ep


def rescale_to_virial(batch: ClumpBatch, c: int, ratio: float):
    """Scale clump c's velocity dispersion about its mass-weighted mean so
    that E_kin / |E_pot| = ratio for the whole member set."""
    o0, o1 = int(batch.offsets[c]), int(batch.offsets[c + 1])
    m = batch.mass[o0:o1]
    # the shear term does not scale with the dispersion: a few fixed-point steps
    for _ in range(6):
        ek, ep = pairwise_energies_fp64(batch.pos[o0:o1], batch.vel[o0:o1], m, batch.G, batch.Omega, batch.q)
        if ek <= 0 or abs(ek / abs(ep) - ratio) < 1e-5:
            return
        lam = math.sqrt(ratio * abs(ep) / ek)
        v = batch.vel[o0:o1].astype(np.float64)
        V = (m[:, None] * v).sum(0) / m.sum()
        batch.vel[o0:o1] = (V + lam * (v - V)).astype(np.float32)
