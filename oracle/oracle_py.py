"""ctypes binding of the CPU restatement (liboracle).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the parity checker / CPU baseline.  The
product under plan_amd/ never imports it.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
import weakref

import numpy as np

from plan_amd.abi import PlanCatalog, PlanDerived, catalog_to_py, make_params

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libplan_oracle.so")
_lib = None

P = np.ctypeslib.ndpointer
_d = P(dtype=np.float64, flags="C_CONTIGUOUS")
_u32 = P(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32 = P(dtype=np.int32, flags="C_CONTIGUOUS")
_u8 = P(dtype=np.uint8, flags="C_CONTIGUOUS")
_u16 = P(dtype=np.uint16, flags="C_CONTIGUOUS")
# the saddle-point density callback of orc_clump_stages: (point[3], est, user) -> density
SADDLE_FN = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double), C.c_double, C.c_void_p)

# stage bits of orc_clump_stages (steps 4-10 of FindPlanetesimals, in this order)
STAGES = dict(merge_a=1, merge_b=2, rekey=4, unbind=8, rebind=16, filter6=32, rip=64, filter8=128, ghost=256,
              J=512, rank=1024)


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        vp = C.c_void_p
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_derive.argtypes = [vp, C.c_uint32, vp]
        L.orc_make_ghosts.restype = C.c_uint32
        L.orc_make_ghosts.argtypes = [vp, C.c_uint32, _d, _d, _d, _i32, _u32, vp, vp, vp, vp, vp]
        L.orc_tree_build.restype = vp
        L.orc_tree_build.argtypes = [vp, C.c_uint32, C.c_uint32, _d, _d, _d, _i32, _u32]
        L.orc_tree_free.argtypes = [vp]
        for f in ("orc_tree_num_particles", "orc_tree_num_nodes", "orc_tree_num_leaves", "orc_tree_num_sinks"):
            getattr(L, f).restype = C.c_uint32
            getattr(L, f).argtypes = [vp]
        L.orc_tree_root.argtypes = [vp, _d, _d]
        L.orc_tree_keys.argtypes = [vp, _u32, _u32, _u32, _u32]
        L.orc_tree_particles.argtypes = [vp, _d, _d, _d, _d, _d, _u32, _u8]
        L.orc_tree_nodes.argtypes = [vp, _d, _d, _u32, _u32, _u32, _u32, _u8, _u8, _u16, _i32, _i32]
        L.orc_knn.restype = C.c_uint32
        L.orc_knn.argtypes = [vp, _d, C.c_uint32, C.c_uint32, vp, _u32, _d, _d, _d, _u32, _u32]
        L.orc_knn_candidates.restype = C.c_uint32
        L.orc_knn_candidates.argtypes = [vp, _d, C.c_uint32, C.c_double, _u32, _d, C.c_uint32]
        L.orc_ball.restype = C.c_uint32
        L.orc_ball.argtypes = [vp, _d, C.c_double, _u32, C.c_uint32]
        L.orc_point_tree_balls.restype = C.c_uint32
        L.orc_point_tree_balls.argtypes = [vp, C.c_uint32, _d, _d, C.c_double, _d, C.c_uint32, _u32, _d, C.c_uint32]
        L.orc_find_planetesimals.argtypes = [vp]
        L.orc_density_vs_scale.restype = C.c_uint32
        L.orc_density_vs_scale.argtypes = [vp, C.c_uint32, _d, _d, P(dtype=np.uint64, flags="C_CONTIGUOUS"), _d]
        L.orc_sheared_count.restype = C.c_uint32
        L.orc_sheared_count.argtypes = [vp, _d, C.c_double, _d, C.c_double]
        L.orc_tree_members.restype = C.c_uint32
        L.orc_tree_members.argtypes = [vp, C.c_uint32, _d, C.c_uint32]
        L.orc_densities.argtypes = [vp, _d]
        L.orc_hop.argtypes = [vp, _u32, _u32]
        L.orc_stage_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                     C.POINTER(C.c_uint32)]
        L.orc_catalog.restype = C.POINTER(PlanCatalog)
        L.orc_catalog.argtypes = [vp]
        L.orc_catalog_free.argtypes = [C.POINTER(PlanCatalog)]
        L.orc_boundness.argtypes = [C.c_uint32, _u32, _d, _d, _d, C.c_double, C.c_double, C.c_double, C.c_uint32,
                                    _u8, _d]
        L.orc_unbind_once.restype = C.c_uint32
        L.orc_unbind_once.argtypes = [vp, C.c_uint32, _d, _d, _d, _u32, _d]
        L.orc_clump_op.restype = C.c_uint32
        L.orc_clump_op.argtypes = [vp, C.c_int, C.c_uint32, _d, _d, _d, _d, _d, _u32, _d]
        L.orc_rebind_candidates.restype = C.c_uint32
        L.orc_rebind_candidates.argtypes = [vp, _d, C.c_uint32, _d, _d, _d, _d, _u8, C.c_double, _u8]
        L.orc_merge_two.restype = C.c_uint32
        L.orc_merge_two.argtypes = [vp, C.c_uint32, C.c_uint32, _d, _d, _d, _d, C.c_uint32, C.c_uint32, _u32, _d]
        L.orc_clumps_bound.restype = C.c_int
        L.orc_clumps_bound.argtypes = [vp, _d, _d]
        L.orc_filter_clumps.argtypes = [C.c_uint32, _u32, _u32, _d, _d, _d, C.c_uint, C.c_double, C.c_double, _u8]
        L.orc_rank_clumps.argtypes = [C.c_uint32, _u32, _d, C.c_uint32, _u8, _d]
        L.orc_tree_set_densities.argtypes = [vp, vp, vp]
        L.orc_clump_stages.restype = C.c_uint32
        L.orc_clump_stages.argtypes = [vp, C.c_uint32, _u32, _u32, _u32, C.c_uint, _d, SADDLE_FN, vp]
        L.orc_key0_erased.restype = C.c_uint32
        L.orc_key0_erased.argtypes = [vp]
        L.orc_clump_subpeaks.restype = C.c_uint32
        L.orc_clump_subpeaks.argtypes = [vp, C.c_uint32, _u32, C.c_uint32]
        _lib = L
    return _lib


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def derive(params: dict, n: int) -> PlanDerived:
    p = make_params(params)
    d = PlanDerived()
    lib().orc_derive(C.byref(p), n, C.byref(d))
    return d


def point_tree_balls(params: dict, pts4, root_center, half_width, queries4):
    """BuildClumpTree's tree over pts4 (x, y, z, payload) with the given root
    box and one ball search per query (x, y, z, r): per query the payloads
    found, in traversal order."""
    p = make_params(params)
    pts4 = np.ascontiguousarray(pts4, dtype=np.float64).reshape(-1, 4)
    q4 = np.ascontiguousarray(queries4, dtype=np.float64).reshape(-1, 4)
    rc = np.ascontiguousarray(root_center, dtype=np.float64)
    off = np.zeros(len(q4) + 1, np.uint32)
    cap = 1 << 20
    out = np.zeros(cap)
    n = lib().orc_point_tree_balls(C.byref(p), len(pts4), pts4.ravel(), rc, float(half_width), q4.ravel(), len(q4),
                                   off, out, cap)
    assert n <= cap
    return [out[off[i]:off[i + 1]].copy() for i in range(len(q4))]


def make_ghosts(params: dict, pos, vel, dens, prop, ids):
    """pos/vel: (N,3) float64; returns arrays for the T particles."""
    p = make_params(params)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    dens = np.ascontiguousarray(dens, dtype=np.float64)
    prop = np.ascontiguousarray(prop, dtype=np.int32)
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    n = pos.shape[0]
    L = lib()
    T = L.orc_make_ghosts(C.byref(p), n, pos.ravel(), vel.ravel(), dens, prop, ids, None, None, None, None, None)
    op = np.empty((T, 3)); ov = np.empty((T, 3)); od = np.empty(T)
    opr = np.empty(T, dtype=np.int32); oid = np.empty(T, dtype=np.uint32)
    L.orc_make_ghosts(C.byref(p), n, pos.ravel(), vel.ravel(), dens, prop, ids,
                      op.ctypes.data, ov.ctypes.data, od.ctypes.data, opr.ctypes.data, oid.ctypes.data)
    return op, ov, od, opr, oid


class OracleTree:
    """A built oracle tree over T particles (N real)."""

    def __init__(self, params: dict, n_real: int, pos, vel, dens, prop, ids):
        self.params = params
        self._p = make_params(params)
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        self.h = lib().orc_tree_build(C.byref(self._p), n_real, pos.shape[0], pos.ravel(), vel.ravel(),
                                      np.ascontiguousarray(dens, dtype=np.float64),
                                      np.ascontiguousarray(prop, dtype=np.int32),
                                      np.ascontiguousarray(ids, dtype=np.uint32))
        # freed when the object dies or, at the latest, before the interpreter
        # tears the module down
        weakref.finalize(self, lib().orc_tree_free, self.h)

    @property
    def n(self):
        return lib().orc_tree_num_particles(self.h)

    def root(self):
        c = np.zeros(3); hw = np.zeros(1)
        lib().orc_tree_root(self.h, c, hw)
        return c, float(hw[0])

    def keys(self):
        n = self.n
        w = [np.empty(n, dtype=np.uint32) for _ in range(4)]
        lib().orc_tree_keys(self.h, *w)
        return w  # w0, w1, w2, particle index

    def particles(self):
        n = self.n
        pos = np.empty((n, 3)); vel = np.empty((n, 3)); m = np.empty(n); ath = np.empty(n); rho = np.empty(n)
        oid = np.empty(n, dtype=np.uint32); sink = np.empty(n, dtype=np.uint8)
        lib().orc_tree_particles(self.h, pos.ravel(), vel.ravel(), m, ath, rho, oid, sink)
        return dict(pos=pos, vel=vel, mass=m, ath=ath, rho=rho, oid=oid, sink=sink)

    def nodes(self):
        L = lib()
        nn = L.orc_tree_num_nodes(self.h)
        nl = L.orc_tree_num_leaves(self.h)
        out = dict(center=np.empty((nn, 3)), hw=np.empty(nn), begin=np.empty(nn, np.uint32),
                   end=np.empty(nn, np.uint32), parent=np.empty(nn, np.uint32),
                   first_daughter=np.empty(nn, np.uint32), num_daughters=np.empty(nn, np.uint8),
                   level=np.empty(nn, np.uint8), orthant=np.empty(nn, np.uint16),
                   node2leaf=np.empty(nn, np.int32), leaf_nodes=np.empty(max(nl, 1), np.int32))
        L.orc_tree_nodes(self.h, out["center"].ravel(), out["hw"], out["begin"], out["end"], out["parent"],
                         out["first_daughter"], out["num_daughters"], out["level"], out["orthant"],
                         out["node2leaf"], out["leaf_nodes"])
        out["leaf_nodes"] = out["leaf_nodes"][:nl]
        return out

    def num_sinks(self):
        return lib().orc_tree_num_sinks(self.h)

    def knn(self, q, K, est=None):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
        nq = q.shape[0]
        idx = np.empty((nq, K), np.uint32); d2 = np.empty((nq, K)); rk = np.empty(nq); rs = np.empty(nq)
        nc = np.empty(nq, np.uint32); nd = np.empty(nq, np.uint32)
        e = None
        if est is not None:
            est = np.ascontiguousarray(est, dtype=np.float64)
            e = est.ctypes.data
        bad = lib().orc_knn(self.h, q.ravel(), nq, K, e, idx.ravel(), d2.ravel(), rk, rs, nc, nd)
        if bad:
            raise RuntimeError(f"oracle knn: start node missing for query {bad - 1}")
        return dict(idx=idx, d2=d2, rk=rk, rstop=rs, ncand=nc, ndrop=nd)

    def knn_candidates(self, q, K, est=0.0, cap=1 << 16):
        """S(r*) of one query: (tree indices, squared distances)."""
        idx = np.empty(cap, np.uint32); d2 = np.empty(cap)
        n = lib().orc_knn_candidates(self.h, np.ascontiguousarray(q, np.float64), int(K), float(est), idx, d2, cap)
        assert n <= cap
        return idx[:n].copy(), d2[:n].copy()

    def ball(self, c, r, cap=1 << 22):
        out = np.empty(cap, np.uint32)
        n = lib().orc_ball(self.h, np.ascontiguousarray(c, dtype=np.float64), float(r), out, cap)
        return out[:min(n, cap)]

    def find_planetesimals(self):
        lib().orc_find_planetesimals(self.h)

    def densities(self):
        r = np.empty(self.n)
        lib().orc_densities(self.h, r)
        return r

    def hop(self):
        n = self.n
        dn = np.empty(n, np.uint32); pk = np.empty(n, np.uint32)
        lib().orc_hop(self.h, dn, pk)
        return dn, pk

    def stage_info(self):
        k = C.c_int(); h = C.c_uint32(); a = C.c_uint32(); b = C.c_uint32()
        lib().orc_stage_info(self.h, C.byref(k), C.byref(h), C.byref(a), C.byref(b))
        return dict(kernel=k.value, hop=h.value, naive_clumps=a.value, after_merge=b.value)

    def density_vs_scale(self):
        """BasicAnalysesWithTree's density-vs-scale pass (src/analyses.hpp:196-333)."""
        L = lib()
        cap = 64
        radii, rhop, sums, zr = np.empty(cap), np.empty(cap), np.empty(cap, np.uint64), np.empty(2)
        n = L.orc_density_vs_scale(self.h, cap, radii, rhop, sums, zr)
        return dict(radii=radii[:n].copy(), max_rhop=rhop[:n].copy(), sums=sums[:n].copy(), z_min=zr[0], z_max=zr[1])

    def sheared_count(self, c, r, maxd, shear):
        return lib().orc_sheared_count(self.h, np.ascontiguousarray(c, np.float64), float(r),
                                       np.ascontiguousarray(maxd, np.float64), float(shear))

    def members(self, i: int):
        """Positions of the particles merged into tree particle i, in sorted order."""
        n = lib().orc_tree_members(self.h, int(i), np.empty(3), 0)
        out = np.empty((n, 3))
        lib().orc_tree_members(self.h, int(i), out.ravel(), n)
        return out

    def key0_erased(self) -> int:
        """Clumps keyed by tree index 0 that the merge loops' zero-filled
        delete lists erased (src/tree.hpp:3874/4000, 4019/4146)."""
        return int(lib().orc_key0_erased(self.h))

    def set_densities(self, rho=None, ath=None):
        """Replace new_density / ath_density of the tree particles (tree order)."""
        r = None if rho is None else np.ascontiguousarray(rho, np.float64)
        a = None if ath is None else np.ascontiguousarray(ath, np.float64)
        lib().orc_tree_set_densities(self.h, None if r is None else r.ctypes.data,
                                     None if a is None else a.ctypes.data)

    def clump_stages(self, clumps, stages, density_threshold, hill_frac=0.75, mass_threshold=0.0, diffuse=0.35,
                     saddle=None):
        """Steps 4-10 of FindPlanetesimals on constructed naive clumps.
        clumps: {key (tree index): member tree indices}; stages: names of
        STAGES; saddle(point, est) -> density replaces the KNN density at
        saddle points.  Returns the catalog with each clump's subpeaks."""
        keys = sorted(clumps)
        off = np.zeros(len(keys) + 1, np.uint32)
        mem = []
        for i, k in enumerate(keys):
            mem.extend(int(m) for m in clumps[k])
            off[i + 1] = len(mem)
        bits = 0
        for s in stages:
            bits |= STAGES[s]
        cb = SADDLE_FN(0) if saddle is None else SADDLE_FN(
            lambda pt, est, user: float(saddle(np.array([pt[0], pt[1], pt[2]]), est)))
        settings = np.array([density_threshold, hill_frac, mass_threshold, diffuse], np.float64)
        L = lib()
        L.orc_clump_stages(self.h, len(keys), np.array(keys, np.uint32), off, np.array(mem or [0], np.uint32),
                           bits, settings, cb, None)
        cat = self.catalog()
        for i, c in enumerate(cat["clumps"]):
            buf = np.empty(64, np.uint32)
            n = L.orc_clump_subpeaks(self.h, i, buf, 64)
            c["subpeaks"] = [int(x) for x in buf[:min(n, 64)]]
        return cat

    def catalog(self):
        L = lib()
        c = L.orc_catalog(self.h)
        try:
            return catalog_to_py(c.contents)
        finally:
            L.orc_catalog_free(c)


def build_tree(snap, params=None):
    """Ghosts -> tree on a synth.Snapshot (no FindPlanetesimals)."""
    params = params or snap.params
    gp, gv, gd, gpr, gid = make_ghosts(params, snap.pos.astype(np.float64), snap.vel.astype(np.float64),
                                       snap.density.astype(np.float64), snap.property_index, snap.id)
    return OracleTree(params, snap.n, gp, gv, gd, gpr, gid)


def run_pipeline(snap, params=None, threads: int | None = None):
    """Ghosts -> tree -> FindPlanetesimals on a synth.Snapshot; returns the tree."""
    params = params or snap.params
    if threads:
        set_threads(threads)
    pos = snap.pos.astype(np.float64)
    vel = snap.vel.astype(np.float64)
    dens = snap.density.astype(np.float64)
    gp, gv, gd, gpr, gid = make_ghosts(params, pos, vel, dens, snap.property_index, snap.id)
    t = OracleTree(params, snap.n, gp, gv, gd, gpr, gid)
    t.find_planetesimals()
    return t


def series_params(first: dict, first_n: int, params: dict, hop: int) -> dict:
    """Parameters of a later snapshot of a series run in one process: box and
    species weights of the first snapshot, particle masses from its count
    (ReadLisFile derives them at loop_begin only, src/tree.hpp:1563-1581),
    and H as the previous FindPlanetesimals left it in the parameters
    (src/tree.hpp:3583-3586)."""
    p = dict(params)
    p.update(box_min=list(first["box_min"]), box_max=list(first["box_max"]),
             type_info=list(first.get("type_info", [1.0])), series_first_n=int(first_n), hop=int(hop))
    return p


def run_series(snaps, params=None, threads: int | None = None):
    """The reference's loop over snapshots in one process (src/main.cpp:56-79):
    each snapshot through ghosts, BuildTree and FindPlanetesimals, with the
    parameters the earlier snapshots left behind.  Returns the trees."""
    trees = []
    first = first_n = hop = None
    for s in snaps:
        p = params or s.params
        if first is None:
            first, first_n = p, s.n
        else:
            p = series_params(first, first_n, p, hop)
        t = run_pipeline(s, p, threads)
        hop = t.stage_info()["hop"]
        trees.append(t)
    return trees


def boundness(offsets, pos, vel, mass, G, Omega=1.0, q=1.5, max_iter=64):
    """FP64 brute-force pairwise boundness of clumps in CSR form; returns
    (bound flags, per-clump dict arrays)."""
    offsets = np.ascontiguousarray(offsets, np.uint32)
    nc = len(offsets) - 1
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    vel = np.ascontiguousarray(vel, np.float64).reshape(-1, 3)
    mass = np.ascontiguousarray(mass, np.float64)
    bound = np.zeros(len(mass), np.uint8)
    out = np.zeros((nc, 8))
    lib().orc_boundness(nc, offsets, pos.ravel(), vel.ravel(), mass, G, Omega, q, max_iter, bound, out.ravel())
    return bound.astype(bool), dict(n_bound=out[:, 0].astype(np.int64), iterations=out[:, 1].astype(np.int64),
                                    E_kin=out[:, 2], E_pot=out[:, 3], mass=out[:, 4], com=out[:, 5:8])


# ---- single clump operations (closed-form pins of the clump logic) --------

_OPS = {"sort": 0, "unbind": 1, "hill": 2, "J": 3, "dispersion": 4}


def _clump_out(n, order, out):
    return dict(order=order[:n].copy(), M=out[0], X=out[1:4].copy(), V=out[4:7].copy(), inner=out[7],
                outer=out[8], one10th=out[9], two_sigma=out[10], hill=out[11], J=out[12:15].copy(),
                Jq=out[15:18].copy(), dispersion=bool(out[18]))


def clump_op(params: dict, op: str, pos, vel, mass, rho=None, ath=None) -> dict:
    """SortParticles and then RemoveUnboundParticles / CalculateHillRadius /
    CalculateAngularMomentum / IsPositionDispersion2Large on one clump given
    as arrays (members in order)."""
    p = make_params(params)
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    n = pos.shape[0]
    vel = np.ascontiguousarray(vel, np.float64).reshape(n, 3)
    mass = np.ascontiguousarray(mass, np.float64)
    rho = np.ones(n) if rho is None else np.ascontiguousarray(rho, np.float64)
    ath = np.zeros(n) if ath is None else np.ascontiguousarray(ath, np.float64)
    order = np.zeros(max(n, 1), np.uint32)
    out = np.zeros(24)
    k = lib().orc_clump_op(C.byref(p), _OPS[op], n, pos.ravel(), vel.ravel(), mass, rho, ath, order, out)
    return _clump_out(k, order, out)


def rebind_candidates(params: dict, state, pos, vel, mass, rho, skip, thr):
    """The candidate loop of SearchBoundParticlesWithinHillRadius for a
    clump state (M, X, V) and candidates in ball order."""
    p = make_params(params)
    st = np.ascontiguousarray(state, np.float64).copy()
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    n = pos.shape[0]
    acc = np.zeros(max(n, 1), np.uint8)
    lib().orc_rebind_candidates(C.byref(p), st, n, pos.ravel(),
                                np.ascontiguousarray(vel, np.float64).ravel(),
                                np.ascontiguousarray(mass, np.float64), np.ascontiguousarray(rho, np.float64),
                                np.ascontiguousarray(skip, np.uint8), float(thr), acc)
    return acc[:n].astype(bool), st


def merge_two(params: dict, na: int, pos, vel, mass, rho, peak_a: int, peak_b: int) -> dict:
    """MergeAnotherPlanetesimal of members [na, n) into members [0, na)."""
    p = make_params(params)
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    n = pos.shape[0]
    order = np.zeros(n, np.uint32)
    out = np.zeros(24)
    k = lib().orc_merge_two(C.byref(p), na, n - na, pos.ravel(), np.ascontiguousarray(vel, np.float64).ravel(),
                            np.ascontiguousarray(mass, np.float64), np.ascontiguousarray(rho, np.float64),
                            peak_a, peak_b, order, out)
    r = _clump_out(k, order, out)
    r["peak"] = int(out[19])
    r["subpeak"] = int(out[20])
    return r


def clumps_bound(params: dict, sa, sb) -> bool:
    """IsGravitationallyBound of two clump states (M, X[3], V[3])."""
    p = make_params(params)
    return bool(lib().orc_clumps_bound(C.byref(p), np.ascontiguousarray(sa, np.float64),
                                       np.ascontiguousarray(sb, np.float64)))


def filter_clumps(offsets, keys, mass, rho, ath, K, mass_thr, peak_thr):
    """RemoveSmallMassAndLowPeak on clumps given as member ranges."""
    off = np.ascontiguousarray(offsets, np.uint32)
    nc = len(off) - 1
    keep = np.zeros(nc, np.uint8)
    lib().orc_filter_clumps(nc, off, np.ascontiguousarray(keys, np.uint32), np.ascontiguousarray(mass, np.float64),
                            np.ascontiguousarray(rho, np.float64), np.ascontiguousarray(ath, np.float64), int(K),
                            float(mass_thr), float(peak_thr), keep)
    return keep.astype(bool)


def rank_clumps(keys, mass, num_peaks):
    """Ranking and cap: kept flags, Mp_tot, Mp_max."""
    keys = np.ascontiguousarray(keys, np.uint32)
    keep = np.zeros(len(keys), np.uint8)
    out = np.zeros(2)
    lib().orc_rank_clumps(len(keys), keys, np.ascontiguousarray(mass, np.float64), int(num_peaks), keep, out)
    return keep.astype(bool), out[0], out[1]


# ---------------------------------------------------------------- .lis ingest

def read_lis(paths):
    """ParticleSet<3>::ReadLisFile restated with numpy (src/tree.hpp:1540-1583).

    Header of the first file (ReadSingleLisFile src/tree.hpp:1439-1466,
    ReadMultipleLisFile 1306-1337): float coor_lim[12], int num_types, float
    type_info[num_types], float time, float dt, long N.  The other files'
    counts sit at byte (14 + num_types) * 4 + 4 (1344-1350).  Records of 44
    bytes follow every header (1384-1417, 1488-1512): float pos[3], float
    vel[3], float density, int property_index, long id_in_run (kept as its
    low 32 bits), int cpu_id (kept as its low 16 bits); files are read in the
    order given.  Particles are then sorted by (cpu_id, id_in_run) (1552-1557,
    a stable sort here: the reference leaves equal keys in unspecified order)
    and id = index (1558-1560); the box is the header's domain limits
    (1565-1566).  Returns a dict of float64 pos/vel/density, int32
    property_index, uint32 id, and the header values."""
    if isinstance(paths, (str, os.PathLike)):
        paths = [paths]
    bodies = []
    head = None
    for k, p in enumerate(paths):
        raw = np.fromfile(p, dtype=np.uint8)
        if k == 0:
            cl = raw[:48].view("<f4").astype(np.float64)
            nt = int(raw[48:52].view("<i4")[0])
            ti = raw[52:52 + 4 * nt].view("<f4").astype(np.float64)
            time, dt = raw[52 + 4 * nt:60 + 4 * nt].view("<f4").astype(np.float64)
            head = dict(coor_lim=cl, num_types=nt, type_info=ti, time=float(time), dt=float(dt))
        nt = head["num_types"]
        off = (14 + nt) * 4 + 4
        n = int(raw[off:off + 8].view("<i8")[0])
        body = raw[off + 8:off + 8 + 44 * n]
        if body.size != 44 * n:
            raise ValueError(f"{p}: fewer records than the header states")
        bodies.append(body.reshape(n, 44))
    rec = np.concatenate(bodies)
    f32 = lambda a, b: np.ascontiguousarray(rec[:, a:b]).view("<f4").astype(np.float64)
    pos, vel, dens = f32(0, 12), f32(12, 24), f32(24, 28)[:, 0]
    prop = np.ascontiguousarray(rec[:, 28:32]).view("<i4")[:, 0]
    idr = np.ascontiguousarray(rec[:, 32:36]).view("<u4")[:, 0]
    cpu = np.ascontiguousarray(rec[:, 40:42]).view("<u2")[:, 0]
    order = np.lexsort((idr, cpu))
    cl = head["coor_lim"]
    return dict(pos=pos[order], vel=vel[order], density=dens[order], property_index=prop[order].astype(np.int32),
                id=np.arange(len(order), dtype=np.uint32), box_min=cl[[6, 8, 10]], box_max=cl[[7, 9, 11]], **head)


# ---------------------------------------------------------------- catalog files

def _c_div(a, b):
    """IEEE division (0/0 = nan, x/0 = inf), as C does it."""
    with np.errstate(all="ignore"):
        return float(np.float64(a) / np.float64(b))


def _c_acos(x):
    return math.nan if math.isnan(x) or abs(x) > 1 else math.acos(x)


def peak_offsets(com, J, pos):
    """geo_mean_offset and median_offset (src/tree.hpp:5027-5104): offsets of
    the member positions pos (sinks expanded) from com, rotated by
    R = Rz(phi) Ry(theta) with theta = acos(J_z / |J|), phi = atan2(J_y, J_x)
    (offset'_d = sum_a offset_a R[a][d]), then per axis the geometric mean of
    |offset'| as 10^(sum log10 / n) with values below 1e-32 counted as -32,
    and the median (mean of the two middle values for even n)."""
    n2 = 0.0
    for k in range(3):
        n2 += J[k] * J[k]
    theta = _c_acos(_c_div(J[2], math.sqrt(n2)))
    phi = math.atan2(J[1], J[0])
    rz = [[math.cos(phi), -math.sin(phi), 0.0], [math.sin(phi), math.cos(phi), 0.0], [0.0, 0.0, 1.0]]
    ry = [[math.cos(theta), 0.0, math.sin(theta)], [0.0, 1.0, 0.0], [-math.sin(theta), 0.0, math.cos(theta)]]
    rot = [[0.0] * 3 for _ in range(3)]
    for a in range(3):
        for b in range(3):
            acc = 0.0
            for d in range(3):
                acc += rz[a][d] * ry[d][b]
            rot[a][b] = acc
    npar = len(pos)
    off = [[0.0] * npar for _ in range(3)]
    for i in range(npar):
        t = [float(pos[i][k]) - com[k] for k in range(3)]
        for d in range(3):
            acc = 0.0
            for a in range(3):
                acc += t[a] * rot[a][d]
            off[d][i] = acc
    geo, med = [0.0] * 3, [0.0] * 3
    half = npar // 2
    for d in range(3):
        acc = 0.0
        for i in range(npar):
            v = abs(off[d][i])
            acc = acc - 32 if v < 1e-32 else acc + math.log10(v)
        geo[d] = math.pow(10.0, acc / npar)
        srt = sorted(abs(v) for v in off[d])
        med[d] = srt[half]
        if npar % 2 == 0:
            med[d] = (med[d] + srt[half - 1]) / 2.0
    return geo, med


PEAKS_HEADER = ("%12s%12s" % ("#peak_id", "Npar") + "".join("%24s" % h for h in (
    "total_mass", "Hill_radius", "center_of_mass[x]", "center_of_mass[y]", "center_of_mass[z]", "vel_COM[x]",
    "vel_COM[y]", "vel_COM[z]", "J[x]", "J[y]", "J[z]/(M R_H^2 Omega)", "geo_mean_offset[x]", "geo_mean_offset[y]",
    "geo_mean_offset[z]", "median_offset[x]", "median_offset[y]", "median_offset[z]")) + "\n")


def peaks_text(cat: dict, members_of) -> str:
    """OutputPlanetesimalsInfo's peaks table (src/tree.hpp:5004-5150) for a
    catalog dict: clumps in peaks_and_masses order (ascending mass, then peak
    index, src/tree.hpp:4291-4299); setw(12) integers, then setprecision(16)
    setw(24) values in the default float format (= printf %24.16g).
    members_of(tree_index) gives the positions a tree particle stands for."""
    order = sorted(range(len(cat["clumps"])), key=lambda i: (cat["clumps"][i]["total_mass"],
                                                              cat["clumps"][i]["peak_index"]))
    out = [PEAKS_HEADER]
    for i in order:
        c = cat["clumps"][i]
        mem = cat["members"][c["member_offset"]:c["member_offset"] + c["num_members"]]
        pos = [p for t in mem for p in members_of(int(t))]
        geo, med = peak_offsets(c["center_of_mass"], c["J"], pos)
        vals = [c["total_mass"], c["Hill_radius"], *c["center_of_mass"], *c["vel_com"], *c["J"], *geo, *med]
        out.append("%12u%12u" % (c["peak_index"], len(pos)) + "".join("%24.16g" % v for v in vals) + "\n")
    return "".join(out)


def clump_particles_text(cat: dict, i: int, tree_particles: dict, rate: int = 1) -> str:
    """OutputSinglePlanetesimal (src/tree.hpp:5191-5222) for clump i: every
    rate-th member in distance order, setw(16) original id, then scientific
    setprecision(16) setw(24) position, velocity, mass (= %24.16e)."""
    c = cat["clumps"][i]
    lines = []
    rate = max(rate, 1)
    for j in range(0, c["num_members"], rate):
        t = int(cat["members"][c["member_offset"] + j])
        oid = int(cat["member_original_ids"][c["member_offset"] + j])
        vals = [*tree_particles["pos"][t], *tree_particles["vel"][t], tree_particles["mass"][t]]
        lines.append("%16u" % oid + "".join("%24.16e" % float(v) for v in vals) + "\n")
    return "".join(lines)


def basic_results_text(cat: dict, time: float, mass_total: float, G_tilde: float, header: bool) -> str:
    """WriteBasicResults (src/tree.hpp:4935-4953): width 15, scientific with
    the default precision 6 (= %15.6e); largest and total clump mass summed
    in peak-index order; mass_physical = (G_tilde / 0.1) 1.3e24 and the Ceres
    mass 9.4e20 (src/global.cpp:40-42, src/global.hpp:803-809)."""
    out = ""
    if header:
        out = "".join("%15s" % h for h in ("#time", "N_peak", "M_p,max,code", "M_p,tot,code", "M_p,max,frac",
                                            "M_p,tot,frac", "M_p,max,Ceres", "M_p,tot,Ceres")) + "\n"
    mmax = mtot = 0.0
    for c in cat["clumps"]:
        mmax = max(mmax, c["total_mass"])
        mtot += c["total_mass"]
    mp = (G_tilde / 0.1) * 1.3e24
    vals = [mmax, mtot, mmax / mass_total, mtot / mass_total, mmax * mp / 9.4e20, mtot * mp / 9.4e20]
    return out + "%15.6e%15u" % (time, len(cat["clumps"])) + "".join("%15.6e" % v for v in vals) + "\n"
