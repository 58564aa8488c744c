"""plan_amd — MI355X (gfx950) implementation of PLAN's clump-finding path.

The compute runs in libplan_hip.so (hand-written HIP kernels behind the C ABI
declared in include/plan_hip.h).  This module is a thin host mirror of the
reference's C++ interface for that path:

* ``ParticleSet``  — ParticleSet<3> after ReadLisFile (src/tree.hpp:1200-1621)
  with ``MakeGhostParticles`` (src/tree.hpp:1626-1754);
* ``BHtree``       — ``BuildTree`` (src/tree.hpp:2759) and
  ``FindPlanetesimals`` (src/tree.hpp:3576);
* ``DataSet``      — the aggregate the reference passes around (src/tree.hpp:5298-5320).

There is no CPU fallback: every call goes to the device library, and loading
fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .abi import (PlanBoundnessClump, PlanBoundnessIn, PlanBoundnessStats, PlanCatalog, PlanParams, catalog_to_py,
                  make_params, py_to_catalog)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libplan_hip.so")
_lib = None

P = np.ctypeslib.ndpointer
_d = P(dtype=np.float64, flags="C_CONTIGUOUS")
_f = P(dtype=np.float32, flags="C_CONTIGUOUS")
_u32 = P(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32 = P(dtype=np.int32, flags="C_CONTIGUOUS")
_u8 = P(dtype=np.uint8, flags="C_CONTIGUOUS")
_u16 = P(dtype=np.uint16, flags="C_CONTIGUOUS")
vp = C.c_void_p


class TreeInfo(C.Structure):
    _fields_ = [("num_particles", C.c_uint32), ("num_nodes", C.c_uint32), ("num_leaf_nodes", C.c_uint32),
                ("num_sink_particles", C.c_uint32), ("depth", C.c_uint32), ("hop_auto", C.c_uint32),
                ("root_center", C.c_double * 3), ("half_width", C.c_double), ("level_count", C.c_int32 * 32)]


EXPORTS = [
    "plan_ctx_create", "plan_ctx_destroy", "plan_last_error", "plan_set_particles", "plan_set_particles_f32",
    "plan_make_ghost_particles", "plan_build_tree", "plan_knn_density", "plan_hop", "plan_find_planetesimals",
    "plan_catalog_free", "plan_get_tree_particles", "plan_get_keys", "plan_get_nodes", "plan_get_hop",
    "plan_knn_query", "plan_density_at", "plan_ball_query", "plan_stage_times", "plan_last_density_kernel_ms",
    "plan_keep_input", "plan_restore_input", "plan_device_memory", "plan_clump_stats",
    "plan_last_density_radius_ms", "plan_boundness_load", "plan_boundness_load_catalog", "plan_boundness_run",
    "plan_boundness_result", "plan_set_knn_list_limit", "plan_read_lis", "plan_set_particles_lis",
    "plan_get_particle_set", "plan_write_peaks", "plan_write_clump_particles", "plan_write_basic_results",
    "plan_density_vs_scale", "plan_knn_query_noalloc", "plan_set_num_hop", "plan_host_tree_balls",
]


class PlanError(RuntimeError):
    pass


def lib():
    """Load libplan_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PlanError(f"{LIB_PATH} not built: run __graft_entry__.build() or python plan_amd/build.py")
    L = C.CDLL(LIB_PATH)
    L.plan_ctx_create.argtypes = [C.POINTER(vp), C.c_int]
    L.plan_ctx_destroy.argtypes = [vp]
    L.plan_last_error.restype = C.c_char_p
    L.plan_last_error.argtypes = [vp]
    L.plan_set_particles.argtypes = [vp, C.POINTER(PlanParams), C.c_uint32, _d, _d, _d, _i32, _u32]
    L.plan_set_particles_f32.argtypes = [vp, C.POINTER(PlanParams), C.c_uint32, _f, _f, _f, _i32, _u32]
    L.plan_make_ghost_particles.argtypes = [vp, C.POINTER(C.c_uint32)]
    L.plan_build_tree.argtypes = [vp, C.c_int, C.POINTER(TreeInfo)]
    L.plan_knn_density.argtypes = [vp, C.c_uint32, C.c_int]
    L.plan_hop.argtypes = [vp, C.c_uint32, C.c_double, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.plan_find_planetesimals.argtypes = [vp, C.POINTER(C.POINTER(PlanCatalog))]
    L.plan_catalog_free.argtypes = [C.POINTER(PlanCatalog)]
    L.plan_get_tree_particles.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
    L.plan_get_keys.argtypes = [vp, vp, vp, vp]
    L.plan_get_nodes.argtypes = [vp] + [vp] * 11
    L.plan_get_hop.argtypes = [vp, vp, vp]
    L.plan_knn_query.argtypes = [vp, C.c_uint32, vp, vp, vp, C.c_uint32, vp, vp, vp]
    L.plan_knn_query_noalloc.argtypes = [vp, C.c_uint32, vp, vp, vp, C.c_uint32, vp, vp, vp]
    L.plan_density_at.argtypes = [vp, C.c_uint32, _d, _d, C.c_uint32, C.c_int, _d]
    L.plan_ball_query.argtypes = [vp, _d, C.c_double, vp, C.c_uint32, C.POINTER(C.c_uint32)]
    L.plan_stage_times.argtypes = [vp, _d, C.c_int]
    L.plan_last_density_kernel_ms.argtypes = [vp, C.POINTER(C.c_double)]
    L.plan_keep_input.argtypes = [vp]
    L.plan_restore_input.argtypes = [vp]
    L.plan_device_memory.argtypes = [vp, C.POINTER(C.c_uint64)]
    L.plan_clump_stats.argtypes = [vp, _d, C.c_int]
    L.plan_last_density_radius_ms.argtypes = [vp, C.POINTER(C.c_double)]
    L.plan_boundness_load.argtypes = [vp, C.POINTER(PlanBoundnessIn)]
    L.plan_boundness_load_catalog.argtypes = [vp, C.POINTER(PlanCatalog)]
    L.plan_boundness_run.argtypes = [vp, C.c_int, C.c_uint32, C.POINTER(PlanBoundnessStats)]
    L.plan_boundness_result.argtypes = [vp, vp, C.POINTER(PlanBoundnessClump)]
    L.plan_set_knn_list_limit.argtypes = [vp, C.c_uint32]
    L.plan_host_tree_balls.argtypes = [vp, C.c_uint32, vp, vp, C.c_double, vp, C.c_uint32, vp, vp, C.c_uint32]
    L.plan_set_num_hop.argtypes = [vp, C.c_uint32]
    L.plan_read_lis.argtypes = [vp, C.POINTER(PlanParams), C.POINTER(C.c_char_p), C.c_uint32,
                                C.POINTER(C.c_double), C.POINTER(C.c_uint32)]
    L.plan_set_particles_lis.argtypes = [vp, C.POINTER(PlanParams), C.c_uint32, vp]
    L.plan_get_particle_set.argtypes = [vp, C.c_uint32, vp, vp, vp, vp, vp, C.POINTER(C.c_uint32)]
    L.plan_density_vs_scale.argtypes = [vp, C.c_uint32, vp, vp, vp, C.POINTER(C.c_uint32), C.c_char_p, C.c_int]
    L.plan_write_peaks.argtypes = [vp, C.POINTER(PlanCatalog), C.c_char_p]
    L.plan_write_clump_particles.argtypes = [vp, C.POINTER(PlanCatalog), C.c_char_p, C.c_uint32]
    L.plan_write_basic_results.argtypes = [vp, C.POINTER(PlanCatalog), C.c_char_p, C.c_int]
    for f in EXPORTS:
        if f in ("plan_ctx_destroy", "plan_catalog_free", "plan_last_error"):
            continue
        getattr(L, f).restype = C.c_int
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data


class Context:
    """One device context (one HIP stream, one snapshot's device memory)."""

    def __init__(self, device: int = 0):
        self._h = vp()
        rc = lib().plan_ctx_create(C.byref(self._h), int(device))
        if rc != 0:
            raise PlanError(f"plan_ctx_create failed with code {rc} (no usable HIP device?)")
        self.info = None
        self.n = 0

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().plan_ctx_destroy(self._h)
            self._h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            msg = lib().plan_last_error(self._h).decode(errors="replace")
            raise PlanError(f"{what} failed (code {rc}): {msg}")

    # --- loading and building -------------------------------------------------
    def set_particles(self, params: dict, pos, vel, density, prop=None, ids=None):
        n = int(np.asarray(pos).shape[0])
        prop = np.zeros(n, np.int32) if prop is None else np.ascontiguousarray(prop, np.int32)
        ids = np.arange(n, dtype=np.uint32) if ids is None else np.ascontiguousarray(ids, np.uint32)
        self._p = make_params(params)
        if np.asarray(pos).dtype == np.float32:
            self._chk(lib().plan_set_particles_f32(self._h, C.byref(self._p), n,
                                                   np.ascontiguousarray(pos, np.float32).ravel(),
                                                   np.ascontiguousarray(vel, np.float32).ravel(),
                                                   np.ascontiguousarray(density, np.float32), prop, ids),
                      "plan_set_particles_f32")
        else:
            self._chk(lib().plan_set_particles(self._h, C.byref(self._p), n,
                                               np.ascontiguousarray(pos, np.float64).ravel(),
                                               np.ascontiguousarray(vel, np.float64).ravel(),
                                               np.ascontiguousarray(density, np.float64), prop, ids),
                      "plan_set_particles")
        return n

    def read_lis(self, params: dict, paths) -> tuple:
        """ReadLisFile: one combined .lis file or one per process (in process
        order).  params holds the parameter-file keys; returns (params with the
        header's box, time and species weights, dt, N)."""
        if isinstance(paths, (str, os.PathLike)):
            paths = [paths]
        enc = [os.fsencode(p) for p in paths]
        arr = (C.c_char_p * len(enc))(*enc)
        self._p = make_params(params)
        dt = C.c_double()
        n = C.c_uint32()
        self._chk(lib().plan_read_lis(self._h, C.byref(self._p), arr, len(enc), C.byref(dt), C.byref(n)),
                  "plan_read_lis")
        out = dict(params)
        p = self._p
        out["box_min"] = [p.box_min[k] for k in range(3)]
        out["box_max"] = [p.box_max[k] for k in range(3)]
        out["time"] = p.time
        out["type_info"] = [p.type_info[k] for k in range(p.num_types)]
        out["n"] = n.value
        return out, dt.value, n.value

    def set_particles_lis(self, params: dict, records) -> int:
        """Particles from .lis records already in memory (44 bytes each, file order)."""
        buf = np.ascontiguousarray(np.frombuffer(records, dtype=np.uint8) if isinstance(records, (bytes, bytearray))
                                   else records.view(np.uint8).ravel())
        if buf.size % 44:
            raise PlanError("records must be a whole number of 44-byte .lis records")
        n = buf.size // 44
        self._p = make_params(params)
        self._chk(lib().plan_set_particles_lis(self._h, C.byref(self._p), n, buf.ctypes.data),
                  "plan_set_particles_lis")
        return n

    def particle_set(self, cap: int | None = None) -> dict:
        """The loaded particle set (with ghosts after make_ghosts)."""
        cnt = C.c_uint32()
        self._chk(lib().plan_get_particle_set(self._h, 0, None, None, None, None, None, C.byref(cnt)),
                  "plan_get_particle_set")
        n = cnt.value if cap is None else min(cap, cnt.value)
        out = dict(pos=np.empty((n, 3)), vel=np.empty((n, 3)), density=np.empty(n), property_index=np.empty(n, np.int32),
                   id=np.empty(n, np.uint32))
        self._chk(lib().plan_get_particle_set(self._h, n, *[out[k].ctypes.data for k in
                                                            ("pos", "vel", "density", "property_index", "id")],
                                              C.byref(cnt)), "plan_get_particle_set")
        return out

    def make_ghosts(self) -> int:
        t = C.c_uint32()
        self._chk(lib().plan_make_ghost_particles(self._h, C.byref(t)), "plan_make_ghost_particles")
        return t.value

    def build_tree(self, check_pos: bool = True) -> TreeInfo:
        info = TreeInfo()
        self._chk(lib().plan_build_tree(self._h, 1 if check_pos else 0, C.byref(info)), "plan_build_tree")
        self.info = info
        self.n = info.num_particles
        return info

    def knn_density(self, K: int, kernel: int = -1):
        self._chk(lib().plan_knn_density(self._h, int(K), int(kernel)), "plan_knn_density")

    def hop(self, H: int, threshold: float):
        a = C.c_uint32(); g = C.c_uint32()
        self._chk(lib().plan_hop(self._h, int(H), float(threshold), C.byref(a), C.byref(g)), "plan_hop")
        return a.value, g.value

    def find_planetesimals(self, keep_for_boundness: bool = False) -> dict:
        """FindPlanetesimals; with keep_for_boundness the catalog's clumps are
        also staged for the pairwise-boundness entry points."""
        out = C.POINTER(PlanCatalog)()
        self._chk(lib().plan_find_planetesimals(self._h, C.byref(out)), "plan_find_planetesimals")
        try:
            if keep_for_boundness:
                self._chk(lib().plan_boundness_load_catalog(self._h, out), "plan_boundness_load_catalog")
                self._bnd_nc = out.contents.num_clumps
                self._bnd_nm = out.contents.num_members_total
            return catalog_to_py(out.contents)
        finally:
            lib().plan_catalog_free(out)

    # --- analyses on the tree ------------------------------------------------------

    def density_vs_scale(self, path=None, header: bool = False) -> dict:
        """BasicAnalysesWithTree's density against scale on the built tree:
        radii, max_rhop and the per-level sums of the counts; with path the
        reference's max_rhop_vs_scale lines are appended there."""
        cap = 64
        radii, rhop, sums = np.empty(cap), np.empty(cap), np.empty(cap, np.uint64)
        n = C.c_uint32()
        self._chk(lib().plan_density_vs_scale(self._h, cap, radii.ctypes.data, rhop.ctypes.data, sums.ctypes.data,
                                              C.byref(n), None if path is None else os.fsencode(path),
                                              1 if header else 0), "plan_density_vs_scale")
        k = n.value
        return dict(radii=radii[:k].copy(), max_rhop=rhop[:k].copy(), sums=sums[:k].copy())

    # --- catalog files -----------------------------------------------------------

    def write_peaks(self, cat: dict, path) -> None:
        """The peaks table of OutputPlanetesimalsInfo for a catalog of this
        context's current tree."""
        c, keep = py_to_catalog(cat)
        self._chk(lib().plan_write_peaks(self._h, C.byref(c), os.fsencode(path)), "plan_write_peaks")

    def write_clump_particles(self, cat: dict, directory, sampling_rate: int = 1) -> None:
        """OutputSinglePlanetesimal for every clump: directory/<peak_index>.txt."""
        c, keep = py_to_catalog(cat)
        self._chk(lib().plan_write_clump_particles(self._h, C.byref(c), os.fsencode(directory), int(sampling_rate)),
                  "plan_write_clump_particles")

    def write_basic_results(self, cat: dict, path, header: bool = False) -> None:
        """WriteBasicResults: one summary line appended to path."""
        c, keep = py_to_catalog(cat)
        self._chk(lib().plan_write_basic_results(self._h, C.byref(c), os.fsencode(path), 1 if header else 0),
                  "plan_write_basic_results")

    # --- pairwise boundness (extension) -----------------------------------------
    FP64, FP32 = 0, 1

    def boundness_load(self, batch):
        """Stage a synth.ClumpBatch (or any object with offsets, pos, vel,
        mass, G, Omega, q) in device memory."""
        self._bnd_keep = (np.ascontiguousarray(batch.offsets, np.uint32),
                          np.ascontiguousarray(batch.pos, np.float64).reshape(-1, 3),
                          np.ascontiguousarray(batch.vel, np.float64).reshape(-1, 3),
                          np.ascontiguousarray(batch.mass, np.float64))
        off, pos, vel, m = self._bnd_keep
        b = PlanBoundnessIn(len(off) - 1, off.ctypes.data, pos.ctypes.data, vel.ctypes.data, m.ctypes.data,
                            float(batch.G), float(batch.Omega), float(batch.q))
        self._chk(lib().plan_boundness_load(self._h, C.byref(b)), "plan_boundness_load")
        self._bnd_nc = len(off) - 1
        self._bnd_nm = int(off[-1])
        self._bnd_keep = None

    def boundness_run(self, form: int, max_iter: int = 64) -> dict:
        st = PlanBoundnessStats()
        self._chk(lib().plan_boundness_run(self._h, int(form), int(max_iter), C.byref(st)), "plan_boundness_run")
        return dict(ms=st.ms, pair_evaluations=st.pair_evaluations, refit_clumps=st.refit_clumps,
                    graph_replays=st.graph_replays, tile_shape=st.tile_shape, tune_ms=list(st.tune_ms))

    def boundness_result(self, num_members: int | None = None, num_clumps: int | None = None):
        nm = self._bnd_nm if num_members is None else num_members
        nc = self._bnd_nc if num_clumps is None else num_clumps
        bound = np.zeros(nm, np.uint8)
        arr = (PlanBoundnessClump * max(nc, 1))()
        self._chk(lib().plan_boundness_result(self._h, bound.ctypes.data, arr), "plan_boundness_result")
        out = dict(n_in=np.array([arr[i].n_in for i in range(nc)]), n_bound=np.array([arr[i].n_bound for i in range(nc)]),
                   iterations=np.array([arr[i].iterations for i in range(nc)]),
                   mass=np.array([arr[i].mass for i in range(nc)]),
                   com=np.array([list(arr[i].center_of_mass) for i in range(nc)]).reshape(-1, 3),
                   E_kin=np.array([arr[i].E_kin for i in range(nc)]), E_pot=np.array([arr[i].E_pot for i in range(nc)]))
        return bound.astype(bool), out

    # --- inspection -------------------------------------------------------------
    def tree_particles(self) -> dict:
        n = self.n
        o = dict(pos=np.empty((n, 3)), vel=np.empty((n, 3)), mass=np.empty(n), ath=np.empty(n), rho=np.empty(n),
                 oid=np.empty(n, np.uint32), sinkcnt=np.empty(n, np.uint32))
        self._chk(lib().plan_get_tree_particles(self._h, _ptr(o["pos"]), _ptr(o["vel"]), _ptr(o["mass"]),
                                                _ptr(o["ath"]), _ptr(o["rho"]), _ptr(o["oid"]), _ptr(o["sinkcnt"])),
                  "plan_get_tree_particles")
        return o

    def keys(self):
        n = self.n
        w = [np.empty(n, np.uint32) for _ in range(3)]
        self._chk(lib().plan_get_keys(self._h, *[_ptr(x) for x in w]), "plan_get_keys")
        return w

    def nodes(self) -> dict:
        nn = self.info.num_nodes
        nl = self.info.num_leaf_nodes
        o = dict(center=np.empty((nn, 3)), hw=np.empty(nn), begin=np.empty(nn, np.uint32),
                 end=np.empty(nn, np.uint32), parent=np.empty(nn, np.uint32),
                 first_daughter=np.empty(nn, np.uint32), num_daughters=np.empty(nn, np.uint8),
                 level=np.empty(nn, np.uint8), orthant=np.empty(nn, np.uint16),
                 node2leaf=np.empty(nn, np.int32), leaf_nodes=np.empty(max(nl, 1), np.int32))
        self._chk(lib().plan_get_nodes(self._h, *[_ptr(o[k]) for k in ("center", "hw", "begin", "end", "parent",
                                                                       "first_daughter", "num_daughters", "level",
                                                                       "orthant", "node2leaf", "leaf_nodes")]),
                  "plan_get_nodes")
        o["leaf_nodes"] = o["leaf_nodes"][:nl]
        return o

    def hop_result(self):
        n = self.n
        dn = np.empty(n, np.uint32); pk = np.empty(n, np.uint32)
        self._chk(lib().plan_get_hop(self._h, _ptr(dn), _ptr(pk)), "plan_get_hop")
        return dn, pk

    def knn_device(self, K: int, nq: int, particle_idx: int = 0, pts: int = 0, est: int = 0, idx: int = 0,
                   d2: int = 0, rstop: int = 0):
        """plan_knn_query_noalloc: the same queries on DEVICE buffers given by
        address (0 = none), without any device allocation; nq is limited to
        PLAN_NOALLOC_MAX_QUERIES."""
        a = lambda x: vp(int(x)) if x else None
        self._chk(lib().plan_knn_query_noalloc(self._h, int(nq), a(particle_idx), a(pts), a(est), int(K), a(idx),
                                               a(d2), a(rstop)), "plan_knn_query_noalloc")

    def knn(self, K: int, particle_idx=None, pts=None, est=None):
        if particle_idx is not None:
            q = np.ascontiguousarray(particle_idx, np.uint32)
            nq = q.size
            e = None if est is None else np.ascontiguousarray(est, np.float64)
            pts_ = None
        else:
            pts_ = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
            nq = pts_.shape[0]
            q = None
            e = np.ascontiguousarray(est, np.float64)
        idx = np.empty((nq, K), np.uint32); d2 = np.empty((nq, K)); rs = np.empty(nq)
        self._chk(lib().plan_knn_query(self._h, nq, _ptr(q), _ptr(pts_), _ptr(e), int(K), _ptr(idx), _ptr(d2),
                                       _ptr(rs)), "plan_knn_query")
        return dict(idx=idx, d2=d2, rstop=rs)

    def density_at(self, pts, est, K: int, kernel: int = 0):
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        est = np.ascontiguousarray(est, np.float64)
        rho = np.empty(pts.shape[0])
        self._chk(lib().plan_density_at(self._h, pts.shape[0], pts.ravel(), est, int(K), int(kernel), rho),
                  "plan_density_at")
        return rho

    def ball(self, center, radius, cap=1 << 22):
        out = np.empty(cap, np.uint32)
        cnt = C.c_uint32()
        self._chk(lib().plan_ball_query(self._h, np.ascontiguousarray(center, np.float64), float(radius),
                                        _ptr(out), cap, C.byref(cnt)), "plan_ball_query")
        return out[:min(cnt.value, cap)]

    def stage_times(self) -> dict:
        t = np.zeros(6)
        self._chk(lib().plan_stage_times(self._h, t, 6), "plan_stage_times")
        return dict(zip(("ghosts", "build", "density", "hop", "clumps", "find_total"), t.tolist()))

    def clump_stats(self) -> dict:
        v = np.zeros(56)
        self._chk(lib().plan_clump_stats(self._h, v, 56), "plan_clump_stats")
        names = ["num_above", "hop_groups", "naive", "after_merge", "merge_passes", "pair_candidates",
                 "saddle_queries", "rebind_searches", "d2h_ms", "rebind_reruns", "ms_sort", "ms_mergeA", "ms_mergeB",
                 "ms_rekey", "ms_unbind", "ms_rebind", "ms_filters", "ms_rip", "ms_J", "ms_rank",
                 "rebind_iters", "rebind_max_iters", "rebind_max_n", "rebind_max_ms",
                 "ms_merge_trees", "ms_merge_saddle", "ms_merge_apply",
                 "ms_rebind_supersets", "ms_rebind_parallel", "ms_rebind_validate",
                 "density_overflow_waves", "hop_overflow_waves",
                 "rebind_max_exact", "rebind_max_exact_ms", "rebind_max_sort_ms",
                 "rebind_work_ms", "rebind_exact_total_ms", "rebind_max_slot",
                 "rebind_last_end_ms", "rebind_last_slot", "rebind_last_ms",
                 "density_resumed_queries", "hop_resumed_queries", "key0_erased",
                 "density_pow_checks", "density_pow_fixes", "ms_catalog", "ms_clump_logic",
                 "rebind_max_scan_ms", "rebind_max_energy_ms", "rebind_max_live", "rebind_max_evals",
                 "rebind_scan_total_ms", "rebind_energy_total_ms", "rebind_sort_total_ms",
                 "rebind_setup_total_ms"]
        return {k: float(x) for k, x in zip(names, v)}

    def keep_input(self):
        self._chk(lib().plan_keep_input(self._h), "plan_keep_input")

    def restore_input(self):
        self._chk(lib().plan_restore_input(self._h), "plan_restore_input")

    def device_memory(self) -> int:
        v = C.c_uint64()
        self._chk(lib().plan_device_memory(self._h, C.byref(v)), "plan_device_memory")
        return v.value

    def set_num_hop(self, num_hop: int):
        """H the next find_planetesimals starts from (a series carries the
        previous snapshot's, as the reference's parameters do)."""
        self._chk(lib().plan_set_num_hop(self._h, int(num_hop)), "plan_set_num_hop")

    def set_knn_list_limit(self, max_entries: int):
        """Test hook: cap the per-wave KNN include list (0 = full capacity)
        so that small inputs take the overflow paths."""
        self._chk(lib().plan_set_knn_list_limit(self._h, int(max_entries)), "plan_set_knn_list_limit")

    def density_kernel_ms(self) -> float:
        v = C.c_double()
        self._chk(lib().plan_last_density_kernel_ms(self._h, C.byref(v)), "plan_last_density_kernel_ms")
        return v.value

    def density_radius_ms(self) -> float:
        v = C.c_double()
        self._chk(lib().plan_last_density_radius_ms(self._h, C.byref(v)), "plan_last_density_radius_ms")
        return v.value


# ---------------------------------------------------------------- reference-shaped facade

class ParticleSet:
    """ParticleSet<3> after ReadLisFile (particles sorted by (cpu_id, id_in_run), id = index)."""

    def __init__(self, pos, vel, density, property_index=None, ids=None, params: dict | None = None):
        self.pos = pos
        self.vel = vel
        self.density = density
        self.property_index = property_index
        self.ids = ids
        self.params = params
        self.num_particles = 0 if pos is None else int(np.asarray(pos).shape[0])
        self.num_total_particles = self.num_particles
        self._ctx = None

    @classmethod
    def from_snapshot(cls, snap):
        return cls(snap.pos, snap.vel, snap.density, snap.property_index, snap.id, snap.params)

    @classmethod
    def ReadLisFile(cls, paths, params: dict, ctx: Context):
        """ReadLisFile (src/tree.hpp:1540-1583) straight into ctx's device
        memory: one combined file or one per process, in process order.  The
        header's box, time and species weights update params; the particle
        arrays stay on the device."""
        params, dt, n = ctx.read_lis(params, paths)
        ps = cls(None, None, None, params=params)
        ps.num_particles = ps.num_total_particles = n
        ps.time, ps.dt = params["time"], dt
        ps._ctx = ctx
        return ps

    def _bind(self, ctx: Context):
        if self._ctx is not ctx:
            if self.pos is None:
                raise PlanError("this particle set was read into another context; read it again")
            ctx.set_particles(self.params, self.pos, self.vel, self.density, self.property_index, self.ids)
            self._ctx = ctx

    def MakeGhostParticles(self, ctx: Context):
        """Ghost images in the shearing-periodic ghost zone (src/tree.hpp:1626)."""
        self._bind(ctx)
        self.num_total_particles = ctx.make_ghosts()
        return self.num_total_particles


class BHtree:
    """Device octree with the reference's BuildTree / FindPlanetesimals entry points."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or Context(0)
        self.info = None

    def BuildTree(self, paras: dict, particle_set: ParticleSet, quiet: bool = False, check_pos: bool = True):
        particle_set._bind(self.ctx)
        self.info = self.ctx.build_tree(check_pos)
        particle_set._ctx = None  # consumed, like the reference (src/tree.hpp:2820)
        return self.info

    def FindPlanetesimals(self, ds: "DataSet", loop_count: int = -1):
        cat = self.ctx.find_planetesimals()
        ds.planetesimal_list = cat
        return cat


class DataSet:
    """particle_set + tree + planetesimal_list (src/tree.hpp:5298-5320)."""

    def __init__(self, particle_set: ParticleSet, ctx: Context | None = None):
        self.particle_set = particle_set
        self.tree = BHtree(ctx)
        self.planetesimal_list = None


def find_clumps(snap, device: int = 0, ctx: Context | None = None) -> dict:
    """Ghosts -> BuildTree -> FindPlanetesimals for a synth.Snapshot."""
    ctx = ctx or Context(device)
    ps = ParticleSet.from_snapshot(snap)
    ds = DataSet(ps, ctx)
    ps.MakeGhostParticles(ctx)
    ds.tree.BuildTree(snap.params, ps)
    return ds.tree.FindPlanetesimals(ds, -1)
