"""plan_knn_query_noalloc (device buffers, fixed scratch array, no device
allocation) against plan_knn_query on the same tree: identical index lists,
squared distances and stop radii, for particle and point queries, on two
calls in a row with the second smaller, and sizes beyond the limit refused."""
import numpy as np
import pytest

from _cases import CASES

pytestmark = pytest.mark.gpu

MAX_Q = 16384


@pytest.fixture(scope="module")
def tree():
    import plan_amd

    snap = CASES["n65k"]()
    print(f"[no_alloc] n65k n={snap.n}", flush=True)
    ctx = plan_amd.Context(0)
    ctx.set_particles(snap.params, snap.pos, snap.vel, snap.density, snap.property_index, snap.id)
    ctx.make_ghosts()
    ctx.build_tree(True)
    return ctx


class _Hip:
    """The caller's side: device buffers from the HIP runtime the library
    itself uses (libamdhip64), allocated and freed by the test."""

    def __init__(self):
        import ctypes as C

        self.C = C
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.C.c_void_p()
        assert self.rt.hipMalloc(self.C.byref(p), max(a.nbytes, 16)) == 0
        self.live.append(p.value)
        assert self.rt.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1) == 0
        return p.value

    def empty(self, nbytes):
        p = self.C.c_void_p()
        assert self.rt.hipMalloc(self.C.byref(p), max(nbytes, 16)) == 0
        self.live.append(p.value)
        return p.value

    def get(self, ptr, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.rt.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def _device_run(ctx, K, pidx=None, pts=None, est=None):
    hip = _Hip()
    try:
        nq = len(pidx) if pidx is not None else len(pts)
        q = hip.put(np.asarray(pidx, np.uint32)) if pidx is not None else 0
        p = hip.put(np.asarray(pts, np.float64)) if pts is not None else 0
        e = hip.put(np.asarray(est, np.float64)) if est is not None else 0
        idx, d2, rs = hip.empty(nq * K * 4), hip.empty(nq * K * 8), hip.empty(nq * 8)
        ctx.knn_device(K, nq, q, p, e, idx, d2, rs)
        return hip.get(idx, (nq, K), np.uint32), hip.get(d2, (nq, K), np.float64), hip.get(rs, nq, np.float64)
    finally:
        hip.free()


@pytest.mark.parametrize("K", [32, 64])
def test_particle_queries_match_plan_knn_query(tree, K):
    ctx = tree
    n = ctx.info.num_particles
    rng = np.random.default_rng(K)
    for nq in (MAX_Q, 1000):  # two calls in a row, the second smaller
        q = np.sort(rng.choice(n, nq, replace=False)).astype(np.uint32)
        print(f"[no_alloc] particle queries K={K} nq={nq}", flush=True)
        want = ctx.knn(K, particle_idx=q)
        got = _device_run(ctx, K, pidx=q)
        assert np.array_equal(got[0], want["idx"])
        assert np.array_equal(got[1], want["d2"])
        assert np.array_equal(got[2], want["rstop"])


def test_point_queries_match_plan_knn_query(tree):
    ctx = tree
    K = 32
    parts = ctx.tree_particles()
    rng = np.random.default_rng(5)
    for nq in (4096, 300):
        src = rng.choice(parts["pos"].shape[0], nq, replace=False)
        pts = parts["pos"][src] + rng.normal(0, 2e-4, (nq, 3))
        est = np.full(nq, 1e-3)
        print(f"[no_alloc] point queries K={K} nq={nq}", flush=True)
        want = ctx.knn(K, pts=pts, est=est)
        got = _device_run(ctx, K, pts=pts, est=est)
        assert np.array_equal(got[0], want["idx"])
        assert np.array_equal(got[1], want["d2"])
        assert np.array_equal(got[2], want["rstop"])


def test_limits_are_refused(tree):
    import plan_amd

    ctx = tree
    with pytest.raises(plan_amd.PlanError):
        ctx.knn_device(32, MAX_Q + 1, particle_idx=1)   # refused before any device access
    with pytest.raises(plan_amd.PlanError):
        _device_run(ctx, 32, pidx=np.array([0, ctx.info.num_particles], np.uint32))
