// query.hip — tree queries on gfx950: approximate-KNN with the reference's
// candidate-set semantics, density kernels, densest-neighbour search, and
// ball search.
//
// Semantics restated from the reference: KNN_Search_OpenMP
// (src/tree.hpp:3198-3251), RecursiveKNN_OpenMP (3177-3193),
// SphereContainNode / SphereNodeIntersect (3001-3049, including the squared
// vs. linear comparison of 3030), MedianSphericalDensityKernel (3473-3491),
// QseudoQuadraticSplinesKernel (3427-3434), the densest-neighbour rule
// (3711-3760) and RecursiveBallSearch (3275-3293).
//
// Mapping: one wavefront serves 64 queries (consecutive tree particles, so
// spatially close).  The wave walks the tree once per radius in lock step:
// the node index is wave-uniform (scalar loads), each lane evaluates the
// reference's predicates for its own query and radius, and the wave descends
// into a node if any lane needs to.  A lane that includes or rejects a node
// covers the node's particle range [begin, end) and stays idle until the walk
// leaves it.  Leaf particles are read once per wave (uniform address) and
// every including lane offers them to its own bounded max-heap in LDS.
// Heap order is (d^2, tree index), so exact distance ties at the K-th
// neighbour resolve to the smaller index.
#include "ctx.h"

namespace plan {

struct TreeView {
    const double4 *pm;
    const double *rho;
    const NodeGeo *ngeo;
    const NodeTop *ntop;
    const uint32_t *ninfo;
    const uint32_t *skip;
    const uint32_t *leaf_of;
    uint32_t n;
    uint32_t nn;
    double to_diag;
    double cell_len[3], half_cell[3], cell_sq[3], three_half_cell[3];
    double cell_volume;
    /* bound on how far a particle can sit outside its node's box (Morton
     * quantisation of the key grid), with a wide allowance for rounding */
    double margin_abs;
};

static TreeView view_of(const DevTree &t, const plan_derived &d)
{
    TreeView v;
    v.pm = t.pm;
    v.rho = t.rho;
    v.ngeo = t.ngeo;
    v.ntop = t.ntop;
    v.ninfo = t.ninfo;
    v.skip = t.skip;
    v.leaf_of = t.leaf_of;
    v.n = t.n;
    v.nn = t.num_nodes;
    v.to_diag = std::sqrt(3.0);
    for (int k = 0; k < 3; ++k) {
        v.cell_len[k] = d.cell_length[k];
        v.half_cell[k] = d.cell_length[k] / 2.;
        v.cell_sq[k] = d.cell_length[k] * d.cell_length[k];
        v.three_half_cell[k] = v.half_cell[k] * 3.;
    }
    v.cell_volume = d.cell_volume;
    v.margin_abs = 2.5 * t.epsilon;
    return v;
}

/* ((0 + dx^2) + dy^2) + dz^2 with dx = a - b (SmallVec::Norm2 order) */
__device__ __forceinline__ double dist2(double ax, double ay, double az, double bx, double by, double bz)
{
    double dx = __dsub_rn(ax, bx), dy = __dsub_rn(ay, by), dz = __dsub_rn(az, bz);
    return __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
}

__device__ __forceinline__ bool sphere_contain(const NodeGeo &g, double qx, double qy, double qz, double r,
                                               double to_diag)
{
    double t = __dmul_rn(g.hw, to_diag);
    double mx = __dadd_rn(fabs(__dsub_rn(g.cx, qx)), t);
    double my = __dadd_rn(fabs(__dsub_rn(g.cy, qy)), t);
    double mz = __dadd_rn(fabs(__dsub_rn(g.cz, qz)), t);
    double s = __dadd_rn(__dadd_rn(__dmul_rn(mx, mx), __dmul_rn(my, my)), __dmul_rn(mz, mz));
    return s < __dmul_rn(r, r);
}

__device__ __forceinline__ bool sphere_intersect(const NodeGeo &g, double qx, double qy, double qz, double r,
                                                 double to_diag)
{
    double c2c = dist2(g.cx, g.cy, g.cz, qx, qy, qz);
    double t = __dadd_rn(__dmul_rn(g.hw, to_diag), r);
    if (c2c > __dmul_rn(t, t)) return false;
    if (c2c < r || c2c < g.hw) return true;
    double md = 0;
    double lo, hi;
    lo = __dsub_rn(g.cx, g.hw);
    hi = __dadd_rn(g.cx, g.hw);
    if (qx < lo) md = __dadd_rn(md, __dmul_rn(__dsub_rn(qx, lo), __dsub_rn(qx, lo)));
    else if (qx > hi) md = __dadd_rn(md, __dmul_rn(__dsub_rn(qx, hi), __dsub_rn(qx, hi)));
    lo = __dsub_rn(g.cy, g.hw);
    hi = __dadd_rn(g.cy, g.hw);
    if (qy < lo) md = __dadd_rn(md, __dmul_rn(__dsub_rn(qy, lo), __dsub_rn(qy, lo)));
    else if (qy > hi) md = __dadd_rn(md, __dmul_rn(__dsub_rn(qy, hi), __dsub_rn(qy, hi)));
    lo = __dsub_rn(g.cz, g.hw);
    hi = __dadd_rn(g.cz, g.hw);
    if (qz < lo) md = __dadd_rn(md, __dmul_rn(__dsub_rn(qz, lo), __dsub_rn(qz, lo)));
    else if (qz > hi) md = __dadd_rn(md, __dmul_rn(__dsub_rn(qz, hi), __dsub_rn(qz, hi)));
    return md <= __dmul_rn(r, r);
}

/* Geometric node tests for exact distance queries (a particle lies within
 * m of its node's box): the whole box within r - m of q (farthest corner),
 * or some point of the box within r + m of q. */
__device__ __forceinline__ bool geo_contain(const NodeGeo &g, double qx, double qy, double qz, double r, double m)
{
    const double rr = r - m;
    if (!(rr > 0)) return false;
    const double mx = fabs(g.cx - qx) + g.hw, my = fabs(g.cy - qy) + g.hw, mz = fabs(g.cz - qz) + g.hw;
    return mx * mx + my * my + mz * mz < rr * rr;
}

__device__ __forceinline__ bool geo_intersect(const NodeGeo &g, double qx, double qy, double qz, double r, double m)
{
    const double ex = fmax(fabs(g.cx - qx) - g.hw, 0.0), ey = fmax(fabs(g.cy - qy) - g.hw, 0.0),
                 ez = fmax(fabs(g.cz - qz) - g.hw, 0.0);
    const double rr = r + m;
    return ex * ex + ey * ey + ez * ez <= rr * rr;
}

/* x^1.5 correctly rounded in all but astronomically rare cases:
 * x * sqrt(x) carried in double-double.  s = RN(sqrt x); e = x - s^2 exactly
 * (one fused multiply-add); sqrt(x) = s + e / (2 s) to first order;
 * x s = p_hi + p_lo exactly; the result rounds p_hi + (p_lo + x e / (2 s))
 * once. */
__device__ __forceinline__ double pow15(double x)
{
    if (!(x > 0.0)) return 0.0;
    double s = __dsqrt_rn(x);
    double e = __fma_rn(-s, s, x);
    double ph = __dmul_rn(x, s);
    double pl = __fma_rn(x, s, -ph);
    double c = __ddiv_rn(__dmul_rn(x, e), __dadd_rn(s, s));
    return __dadd_rn(ph, __dadd_rn(pl, c));
}

/* per-lane bounded max-heap on (d2, idx) in LDS, lane-interleaved */
template <int K> struct LaneHeap {
    double (*d)[64];
    uint32_t (*ix)[64];
    int lane;
    int size;
    double top_d;
    uint32_t top_i;

    __device__ __forceinline__ static bool less(double ad, uint32_t ai, double bd, uint32_t bi)
    {
        return ad < bd || (ad == bd && ai < bi);
    }
    __device__ __forceinline__ void insert(double v, uint32_t id)
    {
        if (size < K) {
            int j = size++;
            while (j > 0) {
                int p = (j - 1) >> 1;
                double pd = d[p][lane];
                uint32_t pi = ix[p][lane];
                if (less(pd, pi, v, id)) {
                    d[j][lane] = pd;
                    ix[j][lane] = pi;
                    j = p;
                } else
                    break;
            }
            d[j][lane] = v;
            ix[j][lane] = id;
            top_d = d[0][lane];
            top_i = ix[0][lane];
        } else if (less(v, id, top_d, top_i)) {
            sift_down_from_root(v, id, K);
            top_d = d[0][lane];
            top_i = ix[0][lane];
        }
    }
    __device__ __forceinline__ void sift_down_from_root(double v, uint32_t id, int lim) { sift_down(0, v, id, lim); }
    __device__ __forceinline__ void sift_down(int j, double v, uint32_t id, int lim)
    {
        while (true) {
            int l = 2 * j + 1;
            if (l >= lim) break;
            int c = l;
            double cd = d[l][lane];
            uint32_t ci = ix[l][lane];
            if (l + 1 < lim) {
                double rd = d[l + 1][lane];
                uint32_t ri = ix[l + 1][lane];
                if (less(cd, ci, rd, ri)) {
                    c = l + 1;
                    cd = rd;
                    ci = ri;
                }
            }
            if (less(v, id, cd, ci)) {
                d[j][lane] = cd;
                ix[j][lane] = ci;
                j = c;
            } else
                break;
        }
        d[j][lane] = v;
        ix[j][lane] = id;
    }
    /* unordered append while filling, then one bottom-up heapify */
    __device__ __forceinline__ void push_unordered(double v, uint32_t id)
    {
        d[size][lane] = v;
        ix[size][lane] = id;
        if (++size == K) {
            for (int j = K / 2 - 1; j >= 0; --j) sift_down(j, d[j][lane], ix[j][lane], K);
            top_d = d[0][lane];
            top_i = ix[0][lane];
        }
    }
    __device__ __forceinline__ void offer(double v, uint32_t id)
    {
        if (size < K) push_unordered(v, id);
        else if (less(v, id, top_d, top_i)) {
            sift_down(0, v, id, K);
            top_d = d[0][lane];
            top_i = ix[0][lane];
        }
    }
    /* in-place heap sort to ascending order */
    __device__ void sort_ascending()
    {
        for (int end = size - 1; end > 0; --end) {
            double vd = d[end][lane];
            uint32_t vi = ix[end][lane];
            d[end][lane] = d[0][lane];
            ix[end][lane] = ix[0][lane];
            sift_down_from_root(vd, vi, end);
        }
    }
};

/* Up to 8 consecutive nodes (a sibling group: siblings are contiguous in the
 * level-major array) staged in LDS by one vector load and then read by the
 * whole wave; pre-order walks visit siblings back to back, so one fetch
 * serves a group instead of one dependent scalar load per node. */
struct NodeCache {
    NodeGeo g[8];
    NodeTop t[8];
    uint32_t info[8];
    uint32_t skip[8];
};

__device__ __forceinline__ void cache_fetch(const TreeView &tv, NodeCache &c, uint32_t v, uint32_t &lo, uint32_t &hi)
{
    const int lane = threadIdx.x;
    __syncthreads();
    if (lane < 8 && v + lane < tv.nn) {
        c.g[lane] = tv.ngeo[v + lane];
        c.t[lane] = tv.ntop[v + lane];
        c.info[lane] = tv.ninfo[v + lane];
        c.skip[lane] = tv.skip[v + lane];
    }
    __syncthreads();
    lo = v;
    hi = min(v + 8u, tv.nn);
}

/* Lock-step pre-order walk for the 64 lanes of a wave at per-lane radius r.
 * COUNT: adds |S_q(r)| per lane.  GATHER: offers every candidate of S_q(r)
 * to the lane's heap; candidate particles of a node are staged in LDS 64 at
 * a time with coalesced loads. */
template <int K, bool GATHER>
__device__ __forceinline__ void walk(const TreeView &tv, NodeCache &nc, double4 *stage, double qx, double qy,
                                     double qz, double r, bool active, uint32_t &count, LaneHeap<K> &heap,
                                     uint32_t *contained = nullptr)
{
    const int lane = threadIdx.x;
    uint32_t cov_end = active ? 0u : 0xffffffffu;
    uint32_t nd = 0, lo = 1, hi = 0;
    uint32_t steps = 0;
    while (nd != PLAN_NONE && steps++ <= tv.nn) { /* pre-order visits each node at most once */
        nd = uniform_u32(nd);
        if (nd < lo || nd >= hi) cache_fetch(tv, nc, nd, lo, hi);
        const uint32_t s = nd - lo;
        const NodeGeo g = nc.g[s];
        const NodeTop t = nc.t[s];
        const bool leaf = (nc.info[s] & 0xffu) == 0u;
        int act = 0; /* 1 include, 2 descend, 3 reject */
        bool whole = false;
        if (t.begin >= cov_end) {
            if (sphere_contain(g, qx, qy, qz, r, tv.to_diag)) {
                act = 1;
                whole = true;
            } else if (sphere_intersect(g, qx, qy, qz, r, tv.to_diag)) act = leaf ? 1 : 2;
            else act = 3;
            if (act != 2) cov_end = t.end;
        }
        if (!GATHER && contained && whole) *contained += t.end - t.begin;
        if (GATHER) {
            if (__ballot(act == 1)) {
                for (uint32_t b = t.begin; b < t.end; b += 64) {
                    const uint32_t cnt = min(64u, t.end - b);
                    __syncthreads();
                    if ((uint32_t)lane < cnt) stage[lane] = tv.pm[b + lane];
                    __syncthreads();
                    if (act == 1)
                        for (uint32_t j = 0; j < cnt; ++j) {
                            const double4 x = stage[j];
                            heap.insert(dist2(qx, qy, qz, x.x, x.y, x.z), b + j);
                        }
                }
            }
        } else {
            if (act == 1) count += t.end - t.begin;
        }
        nd = __ballot(act == 2) ? t.first_daughter : nc.skip[s];
    }
}

/* query position and the reference's start radius: hw of the first node on
 * the leaf-to-root path holding >= K/4 particles (src/tree.hpp:3206-3219),
 * or the caller's estimate */
template <int K>
__device__ __forceinline__ void query_setup(const TreeView &tv, uint32_t q, const uint32_t *qidx, const double4 *qpts,
                                            const double *est, double &qx, double &qy, double &qz, double &r,
                                            uint32_t &self, uint32_t q0 = 0)
{
    if (qpts) {
        double4 p = qpts[q];
        qx = p.x;
        qy = p.y;
        qz = p.z;
        self = 0;
    } else {
        self = qidx ? qidx[q] : q0 + q;
        double4 p = tv.pm[self];
        qx = p.x;
        qy = p.y;
        qz = p.z;
    }
    if (est) {
        r = est[q];
    } else {
        uint32_t node = tv.leaf_of[self];
        while (tv.ntop[node].end - tv.ntop[node].begin < (uint32_t)(K / 4)) node = tv.ntop[node].parent;
        double hw = tv.ngeo[node].hw;
        r = __dsqrt_rn(__dmul_rn(hw, hw));
    }
}

/* Pass 1: radius schedule r, 2r, 4r ... until |S(r)| >= K (node sizes only,
 * no heap; K <= n guarantees an end, 2200 doublings exceed any finite start). */
template <int K>
__global__ __launch_bounds__(64) void k_knn_radius(TreeView tv, uint32_t nq, const uint32_t *__restrict__ qidx,
                                                   const double4 *__restrict__ qpts, const double *__restrict__ est,
                                                   const double *__restrict__ rt, double *__restrict__ r_out,
                                                   uint32_t *__restrict__ c_out)
{
    __shared__ NodeCache nc;
    const int lane = threadIdx.x;
    const uint32_t q = blockIdx.x * 64 + lane;
    const bool valid = q < nq;
    double qx = 0, qy = 0, qz = 0, r = 0;
    uint32_t self = 0;
    if (valid) query_setup<K>(tv, q, qidx, qpts, est, qx, qy, qz, r, self);
    LaneHeap<K> none;
    /* a point query resolved by k_point_tight: S(est) holds >= K particles,
     * so the schedule stops at est */
    const bool tight = valid && rt && rt[q] > 0;
    bool done = !valid || tight;
    int rounds = 0;
    uint32_t inside = tight ? (uint32_t)K : 0u;
    while (__ballot(!done) && rounds++ < 2200) {
        uint32_t cnt = 0, cin = 0;
        walk<K, false>(tv, nc, nullptr, qx, qy, qz, r, !done, cnt, none, &cin);
        if (!done) {
            if (cnt >= (uint32_t)K) {
                done = true;
                inside = cin;
            } else {
                r = __dmul_rn(r, 2.0);
            }
        }
    }
    if (valid) {
        r_out[q] = r;
        c_out[q] = inside;
    }
}

/* Exact count of particles with d^2 <= r^2 from q, by a lock-step walk with
 * the geometric node tests (whole nodes counted by size, boundary leaves
 * particle by particle). */
__device__ __forceinline__ uint32_t walk_geo_count(const TreeView &tv, NodeCache &nc, double qx, double qy, double qz,
                                                   double r, bool active)
{
    uint32_t cov_end = active ? 0u : 0xffffffffu;
    uint32_t nd = 0, lo = 1, hi = 0, steps = 0, count = 0;
    const double r2 = r * r, m = tv.margin_abs;
    while (nd != PLAN_NONE && steps++ <= tv.nn) {
        nd = uniform_u32(nd);
        if (nd < lo || nd >= hi) cache_fetch(tv, nc, nd, lo, hi);
        const uint32_t s = nd - lo;
        const NodeGeo g = nc.g[s];
        const NodeTop t = nc.t[s];
        const bool leaf = (nc.info[s] & 0xffu) == 0u;
        int act = 0; /* 1 whole, 4 boundary leaf, 2 descend, 3 reject */
        if (t.begin >= cov_end) {
            if (geo_contain(g, qx, qy, qz, r, m)) act = 1;
            else if (geo_intersect(g, qx, qy, qz, r, m)) act = leaf ? 4 : 2;
            else act = 3;
            if (act != 2) cov_end = t.end;
        }
        if (act == 1) count += t.end - t.begin;
        if (act == 4)
            for (uint32_t p = t.begin; p < t.end; ++p) {
                const double4 x = tv.pm[p];
                if (dist2(qx, qy, qz, x.x, x.y, x.z) <= r2) ++count;
            }
        nd = __ballot(act == 2) ? t.first_daughter : nc.skip[s];
    }
    return count;
}

/* Point queries (saddle points, user points) with a start radius est: when
 * some radius r_t <= est - margin already holds >= K particles within
 * distance r_t, every particle within r_t belongs to the reference's
 * candidate set S(est) (src/tree.hpp:3160-3230), so the K smallest
 * (d^2, index) of S(est) are the K nearest overall and can be gathered from
 * the ball of radius r_t instead of all of S(est).  r_t starts at the half
 * width of the deepest node holding q with >= K/4 particles and doubles.
 * rt_out[q] = r_t, or 0 when no such radius exists (the query then takes
 * the candidate-set path). */
template <int K>
__global__ __launch_bounds__(64) void k_point_tight(TreeView tv, uint32_t nq, const double4 *__restrict__ qpts,
                                                    const double *__restrict__ est, double *__restrict__ rt_out)
{
    __shared__ NodeCache nc;
    const int lane = threadIdx.x;
    const uint32_t q = blockIdx.x * 64 + lane;
    const bool valid = q < nq;
    double qx = 0, qy = 0, qz = 0, e = 0, r = 0, lim = 0;
    if (valid) {
        const double4 p = qpts[q];
        qx = p.x;
        qy = p.y;
        qz = p.z;
        e = est[q];
        lim = e - (tv.margin_abs + 1e-12 * e);
        uint32_t node = 0;
        for (int lvl = 0; lvl < 32; ++lvl) {
            const uint32_t nd = tv.ninfo[node] & 0xffu;
            if (nd == 0) break;
            const uint32_t fd = tv.ntop[node].first_daughter;
            uint32_t next = PLAN_NONE;
            for (uint32_t c = 0; c < nd; ++c) {
                const NodeGeo g = tv.ngeo[fd + c];
                if (fabs(qx - g.cx) <= g.hw && fabs(qy - g.cy) <= g.hw && fabs(qz - g.cz) <= g.hw) {
                    next = fd + c;
                    break;
                }
            }
            if (next == PLAN_NONE) break;
            const NodeTop t = tv.ntop[next];
            if (t.end - t.begin < (uint32_t)(K / 4 > 0 ? K / 4 : 1)) break;
            node = next;
        }
        r = tv.ngeo[node].hw;
    }
    bool done = !valid || !(r > 0) || !(r <= lim);
    bool ok = false;
    int rounds = 0;
    while (__ballot(!done) && rounds++ < 64) {
        const uint32_t cnt = walk_geo_count(tv, nc, qx, qy, qz, r, !done);
        if (!done) {
            if (cnt >= (uint32_t)K) {
                done = true;
                ok = true;
            } else {
                r = r * 2.0;
                if (!(r <= lim)) done = true;
            }
        }
    }
    if (valid) rt_out[q] = ok ? r : 0.0;
}

/* One entry of a wave's include list: a node whose particles some lanes take
 * whole (contained node or intersecting leaf), with the mask of those lanes. */
struct ListEnt {
    uint32_t begin, end;
    uint64_t mask;
};
static constexpr int LCAP = 512;

struct WalkState {
    uint32_t nd, cov_end, lo, hi;
};

/* Resumable lock-step walk that records the include list.  Returns true when
 * the walk is complete, false when the list is full; the node that did not
 * fit is re-evaluated on resume (lane state is only updated after it is
 * recorded). */
__device__ __forceinline__ bool walk_fill(const TreeView &tv, NodeCache &nc, WalkState &w, double qx, double qy,
                                          double qz, double r, bool geo, ListEnt *lst, uint32_t &nent)
{
    const int lane = threadIdx.x;
    uint32_t guard = 0;
    while (w.nd != PLAN_NONE && guard++ <= tv.nn) {
        const uint32_t nd = uniform_u32(w.nd);
        if (nd < w.lo || nd >= w.hi) cache_fetch(tv, nc, nd, w.lo, w.hi);
        const uint32_t s = nd - w.lo;
        const NodeGeo g = nc.g[s];
        const NodeTop t = nc.t[s];
        const bool leaf = (nc.info[s] & 0xffu) == 0u;
        int act = 0;
        if (t.begin >= w.cov_end) {
            if (geo) {
                if (geo_contain(g, qx, qy, qz, r, tv.margin_abs)) act = 1;
                else if (geo_intersect(g, qx, qy, qz, r, tv.margin_abs)) act = leaf ? 1 : 2;
                else act = 3;
            } else {
                if (sphere_contain(g, qx, qy, qz, r, tv.to_diag)) act = 1;
                else if (sphere_intersect(g, qx, qy, qz, r, tv.to_diag)) act = leaf ? 1 : 2;
                else act = 3;
            }
        }
        const uint64_t inc = __ballot(act == 1);
        if (inc) {
            if (nent == LCAP) return false;
            if (lane == 0) lst[nent] = ListEnt{t.begin, t.end, inc};
            ++nent;
        }
        if (act == 1 || act == 3) w.cov_end = t.end;
        w.nd = __ballot(act == 2) ? t.first_daughter : nc.skip[s];
    }
    return true;
}

/* Stream the particles of the recorded entries 64 at a time (coalesced,
 * next chunk prefetched into registers while the current one is used) and
 * call f(lane-included?, d2, tree index) for each staged particle. */
template <class F>
__device__ __forceinline__ void stream_list(const TreeView &tv, const ListEnt *lst, uint32_t nent, uint32_t *lpre,
                                            double4 *stage, uint32_t *stage_p, uint64_t *stage_m, double qx,
                                            double qy, double qz, F f)
{
    const int lane = threadIdx.x;
    __syncthreads();
    {
        /* prefix of entry sizes, two entries per lane */
        uint32_t e0 = 2 * lane, e1 = 2 * lane + 1;
        uint32_t s0 = e0 < nent ? lst[e0].end - lst[e0].begin : 0u;
        uint32_t s1 = e1 < nent ? lst[e1].end - lst[e1].begin : 0u;
        uint32_t sum = s0 + s1, inc = sum;
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t tt = __shfl_up(inc, o, 64);
            if (lane >= o) inc += tt;
        }
        uint32_t ex = inc - sum;
        if (e0 < nent) lpre[e0] = ex;
        if (e1 < nent) lpre[e1] = ex + s0;
        if (lane == 63) lpre[nent] = inc;
    }
    __syncthreads();
    const uint32_t total = lpre[nent];
    auto locate = [&](uint32_t fpos, uint32_t &pidx, uint64_t &m) {
        uint32_t lo = 0, hi = nent; /* lpre[lo] <= fpos < lpre[hi] */
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (lpre[mid] <= fpos) lo = mid;
            else hi = mid;
        }
        pidx = lst[lo].begin + (fpos - lpre[lo]);
        m = lst[lo].mask;
    };
    double4 nx = make_double4(0, 0, 0, 0);
    uint32_t np = 0;
    uint64_t nm = 0;
    if ((uint32_t)lane < total) {
        locate(lane, np, nm);
        nx = tv.pm[np];
    }
    for (uint32_t b = 0; b < total; b += 64) {
        const uint32_t cnt = min(64u, total - b);
        __syncthreads();
        stage[lane] = nx;
        stage_p[lane] = np;
        stage_m[lane] = nm;
        __syncthreads();
        const uint32_t fn = b + 64 + lane;
        if (fn < total) {
            locate(fn, np, nm);
            nx = tv.pm[np];
        }
        for (uint32_t j = 0; j < cnt; ++j) {
            const bool mine = (stage_m[j] >> lane) & 1ull;
            if (mine) {
                const double4 x = stage[j];
                f(dist2(qx, qy, qz, x.x, x.y, x.z), stage_p[j]);
            }
        }
    }
}

/* The K smallest (d^2, index) seen so far, ascending, in registers.  Every
 * array index is a compile-time constant (fully unrolled), so the list never
 * touches LDS or scratch; an insertion is one branch-free pass of compares
 * and selects, identical for all lanes that take it. */
template <int K> struct RegList {
    double d[K];
    uint32_t i[K];

    __device__ __forceinline__ static bool less(double ad, uint32_t ai, double bd, uint32_t bi)
    {
        return ad < bd || (ad == bd && ai < bi);
    }
    __device__ __forceinline__ void init()
    {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            d[j] = __longlong_as_double(0x7ff0000000000000ll);
            i[j] = 0xffffffffu;
        }
    }
    __device__ __forceinline__ bool accepts(double x, uint32_t xi) const { return less(x, xi, d[K - 1], i[K - 1]); }
    /* insert (x, xi): one ascending pass keeps the smaller of (x, entry j)
     * at j and carries the larger on; the largest of the K + 1 drops out */
    __device__ __forceinline__ void insert(double x, uint32_t xi)
    {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const bool lt = less(x, xi, d[j], i[j]);
            const double lo = lt ? x : d[j];
            const double hi = lt ? d[j] : x;
            const uint32_t ilo = lt ? xi : i[j];
            const uint32_t ihi = lt ? i[j] : xi;
            d[j] = lo;
            i[j] = ilo;
            x = hi;
            xi = ihi;
        }
    }
};

/* Stream this lane's entries of the recorded include list into the list. */
template <int K>
__device__ __forceinline__ void stream_entries(const TreeView &tv, const ListEnt *lst, uint32_t nent, double qx,
                                               double qy, double qz, double T, RegList<K> &best)
{
    const int lane = threadIdx.x;
    uint32_t e = 0, p = 0, pe = 0;
    while (p >= pe && e < nent) {
        const ListEnt le = lst[e++];
        if ((le.mask >> lane) & 1ull) {
            p = le.begin;
            pe = le.end;
        }
    }
    /* the list is only touched outside divergent code: one insertion per
     * iteration for the whole wave, with a +inf key on lanes that have
     * nothing to add, so its registers are updated in place */
    while (__ballot(p < pe)) {
        double d2 = __longlong_as_double(0x7ff0000000000000ll);
        uint32_t pi = 0xffffffffu;
        if (p < pe) {
            const double4 x = tv.pm[p];
            const double dd = dist2(qx, qy, qz, x.x, x.y, x.z);
            if (dd <= T && best.accepts(dd, p)) {
                d2 = dd;
                pi = p;
            }
            if (++p >= pe) {
                while (p >= pe && e < nent) {
                    const ListEnt le = lst[e++];
                    if ((le.mask >> lane) & 1ull) {
                        p = le.begin;
                        pe = le.end;
                    }
                }
            }
        }
        if (__ballot(pi != 0xffffffffu)) best.insert(d2, pi);
    }
}

/* Pass 2: the K smallest (d^2, index) of S(r*) per query, written as
 * lists in [j][slot] layout (coalesced across lanes).  The walk is lock-step
 * over the wave and records the nodes some lane takes whole; each lane then
 * streams only its own entries' particles.  RESUME = false: one walk whose
 * list must fit LCAP entries (a wave that overflows sets its flags and is
 * redone with RESUME = true, which alternates walking and streaming).
 * Slots: q (or qslot[q] when given). */
template <int K, bool RESUME>
__global__ __launch_bounds__(64) void k_knn_gather(TreeView tv, uint32_t nq, const uint32_t *__restrict__ qidx,
                                                   const double4 *__restrict__ qpts, const double *__restrict__ r_in,
                                                   const uint32_t *__restrict__ c_in, const double *__restrict__ rt_in,
                                                   double *__restrict__ ld2, uint32_t *__restrict__ lidx, uint32_t q0,
                                                   uint32_t nslots, const uint32_t *__restrict__ qslot,
                                                   uint32_t *__restrict__ ovf)
{
    __shared__ NodeCache nc;
    __shared__ ListEnt lst[LCAP];
    const int lane = threadIdx.x;
    const uint32_t q = blockIdx.x * 64 + lane;
    const bool valid = q < nq;
    double qx = 0, qy = 0, qz = 0, r = 0;
    uint32_t self = 0;
    const uint32_t slot = valid ? (qslot ? qslot[q] : q) : 0u;
    if (valid) query_setup<K>(tv, q, qidx, qpts, r_in, qx, qy, qz, r, self, q0);
    /* Candidates in nodes the sphere contains lie within r* (a contained
     * node's box is inside the sphere with a margin far above the Morton
     * quantisation); when there are >= K of them the K nearest are within
     * r*, and only d^2 <= (r* + 1e-9)^2 can enter the list. */
    double T = __longlong_as_double(0x7ff0000000000000ll);
    if (valid && c_in[q] >= (uint32_t)K) {
        const double rr = __dadd_rn(r, 1e-9);
        T = __dmul_rn(rr, rr);
    }
    /* a point query with a tight radius (k_point_tight): >= K particles lie
     * within r_t, and the K nearest of S(est) are the K nearest within r_t */
    const bool geo = valid && rt_in && rt_in[q] > 0;
    if (geo) {
        r = rt_in[q];
        T = r * r;
    }
    WalkState w{0u, valid ? 0u : 0xffffffffu, 1u, 0u};
    RegList<K> best;
    if (!RESUME) {
        uint32_t nent = 0;
        const bool fin = walk_fill(tv, nc, w, qx, qy, qz, r, geo, lst, nent);
        if (!fin) {
            if (valid) ovf[q] = 1u;
            return;
        }
        __syncthreads();
        best.init();
        stream_entries<K>(tv, lst, nent, qx, qy, qz, T, best);
    } else {
        best.init();
        bool fin = false;
        while (!fin) {
            uint32_t nent = 0;
            fin = walk_fill(tv, nc, w, qx, qy, qz, r, geo, lst, nent);
            __syncthreads();
            stream_entries<K>(tv, lst, nent, qx, qy, qz, T, best);
            __syncthreads();
        }
    }
    if (!valid) return;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        ld2[(size_t)j * nslots + slot] = best.d[j];
        lidx[(size_t)j * nslots + slot] = best.i[j];
    }
}

__global__ void k_ovf_compact(const uint32_t *__restrict__ ovf, const uint32_t *__restrict__ pos, uint32_t n,
                              uint32_t *__restrict__ list)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && ovf[i]) list[pos[i]] = i;
}

/* the overflow redo: queries named by slot, with the batch's query arrays */
__global__ void k_ovf_gather_args(const uint32_t *__restrict__ slots, uint32_t m, const uint32_t *__restrict__ qidx,
                                  const double4 *__restrict__ qpts, const double *__restrict__ r_in,
                                  const uint32_t *__restrict__ c_in, const double *__restrict__ rt_in, uint32_t q0,
                                  uint32_t *__restrict__ qi2, double4 *__restrict__ qp2, double *__restrict__ r2,
                                  uint32_t *__restrict__ c2, double *__restrict__ rt2)
{
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const uint32_t q = slots[k];
    if (qpts) qp2[k] = qpts[q];
    else qi2[k] = qidx ? qidx[q] : q0 + q;
    r2[k] = r_in[q];
    rt2[k] = rt_in ? rt_in[q] : 0.0;
    c2[k] = c_in[q];
}

/* Uses of the lists, one thread per query: 0 density, 1 densest neighbour,
 * 2 row-major copy of the lists. */
template <int K>
__global__ __launch_bounds__(256) void k_knn_finish(TreeView tv, uint32_t nq, const uint32_t *__restrict__ qidx,
                                                    const double4 *__restrict__ qpts,
                                                    const double *__restrict__ ld2, const uint32_t *__restrict__ lidx,
                                                    int mode, int kernel, const double *__restrict__ rho_in,
                                                    double *__restrict__ rho_out, uint32_t *__restrict__ densest,
                                                    uint32_t *__restrict__ idx_out, double *__restrict__ d2_out,
                                                    uint32_t q0)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const uint32_t self = qpts ? 0u : (qidx ? qidx[q] : q0 + q);
    if (mode == 1) {
        double bv = rho_in[self];
        uint32_t bi = self;
        for (int j = 0; j < K; ++j) {
            const uint32_t id = lidx[(size_t)j * nq + q];
            const double rj = rho_in[id];
            if (bv < rj) {
                bi = id;
                bv = rj;
            } else if (bv == rj && bi > id) {
                bi = id;
            }
        }
        densest[q] = bi;
        return;
    }
    if (mode == 2 || idx_out) {
        for (int j = 0; j < K; ++j) {
            if (idx_out) idx_out[(size_t)q * K + j] = lidx[(size_t)j * nq + q];
            if (d2_out) d2_out[(size_t)q * K + j] = ld2[(size_t)j * nq + q];
        }
        if (mode == 2) return;
    }
    double rho = 0;
    if (kernel == 0) {
        /* MedianSphericalDensityKernel (ascending list): m = 4 d2[K/2];
         * i* = last index with d2 <= m; masses summed from i* down to 0 */
        const double m = __dmul_rn(4.0, ld2[(size_t)(K / 2) * nq + q]);
        int is = K / 2;
        for (int j = K - 1; j > K / 2; --j)
            if (!(ld2[(size_t)j * nq + q] > m)) {
                is = j;
                break;
            }
        const double tm = ld2[(size_t)is * nq + q];
        for (int j = is; j >= 0; --j) rho = __dadd_rn(rho, tv.pm[lidx[(size_t)j * nq + q]].w);
        const double four_pi_over_three = (3.14159265358979323846 * 4.) / 3.;
        rho = __ddiv_rn(rho, __dmul_rn(four_pi_over_three, pow15(tm)));
    } else {
        /* QseudoQuadraticSplinesKernel; point queries use particle 0 as "self" */
        const double4 sp = tv.pm[self];
        for (int j = 0; j < K; ++j) {
            const double4 x = tv.pm[lidx[(size_t)j * nq + q]];
            double dd[3] = {fabs(__dsub_rn(x.x, sp.x)), fabs(__dsub_rn(x.y, sp.y)), fabs(__dsub_rn(x.z, sp.z))};
            double w = 1;
            for (int k = 0; k < 3; ++k) {
                if (dd[k] <= tv.half_cell[k]) {
                    w = __dmul_rn(w, __dsub_rn(0.75, __ddiv_rn(__dmul_rn(dd[k], dd[k]), tv.cell_sq[k])));
                } else if (dd[k] < tv.three_half_cell[k]) {
                    double b = __dsub_rn(1.5, __ddiv_rn(dd[k], tv.cell_len[k]));
                    w = __dmul_rn(w, __dmul_rn(0.5, __dmul_rn(b, b)));
                } else {
                    w = 0.;
                    break;
                }
            }
            rho = __dadd_rn(rho, __dmul_rn(__ddiv_rn(x.w, tv.cell_volume), w));
        }
    }
    rho_out[qpts || qidx ? q : self] = rho;
}

template <int K>
static void launch_knn(const DevTree &t, const TreeView &tv, uint32_t nq, const uint32_t *qidx, const double4 *qpts,
                       const double *est, int mode, int kernel, const double *rho_in, KnnOut out, hipStream_t s)
{
    /* the neighbour lists of one batch live in HBM: 12 K bytes per query */
    const uint32_t batch = (uint32_t)std::min<uint64_t>(nq, (uint64_t)(1u << 24));
    Arena &a = const_cast<DevTree &>(t).arena;
    double *rbuf = a.get_t<double>("knn_r", (size_t)nq);
    uint32_t *cbuf = a.get_t<uint32_t>("knn_c", (size_t)nq);
    double *ld2 = a.get_t<double>("knn_ld2", (size_t)batch * K);
    uint32_t *lidx = a.get_t<uint32_t>("knn_lidx", (size_t)batch * K);
    double *rt = nullptr;
    if (out.ev) HIPCHK(hipEventRecord(out.ev[0], s));
    if (qpts) {
        rt = a.get_t<double>("knn_rt", (size_t)nq);
        hipLaunchKernelGGL(k_point_tight<K>, dim3(ceil_div(nq, 64)), dim3(64), 0, s, tv, nq, qpts, est, rt);
        LAUNCHCHK();
    }
    hipLaunchKernelGGL(k_knn_radius<K>, dim3(ceil_div(nq, 64)), dim3(64), 0, s, tv, nq, qidx, qpts, est, rt, rbuf,
                       cbuf);
    LAUNCHCHK();
    if (out.ev) HIPCHK(hipEventRecord(out.ev[1], s));
    for (uint32_t b0 = 0; b0 < nq; b0 += batch) {
        const uint32_t nb = std::min(batch, nq - b0);
        const uint32_t *qi = qidx ? qidx + b0 : nullptr;
        const double4 *qp = qpts ? qpts + b0 : nullptr;
        const uint32_t q0 = (qidx || qpts) ? 0u : b0; /* all-particle pass: particles b0.. */
        const size_t o = (qidx || qpts) ? b0 : 0;     /* outputs indexed by query, else by particle */
        uint32_t *ovf = a.get_t<uint32_t>("knn_ovf", (size_t)nb + 1);
        uint32_t *opos = a.get_t<uint32_t>("knn_opos", (size_t)nb + 1);
        uint32_t *otot = a.get_t<uint32_t>("knn_otot", 4);
        HIPCHK(hipMemsetAsync(ovf, 0, (size_t)nb * 4, s));
        hipLaunchKernelGGL((k_knn_gather<K, false>), dim3(ceil_div(nb, 64)), dim3(64), 0, s, tv, nb, qi, qp,
                           rbuf + b0, cbuf + b0, rt ? rt + b0 : nullptr, ld2, lidx, q0, nb, nullptr, ovf);
        LAUNCHCHK();
        scan_exclusive_u32(a, ovf, opos, nb, s, otot);
        uint32_t novf = 0;
        HIPCHK(hipMemcpyAsync(&novf, otot, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (novf) {
            uint32_t *slots = a.get_t<uint32_t>("knn_oslot", novf);
            uint32_t *qi2 = a.get_t<uint32_t>("knn_oqi", novf);
            double4 *qp2 = a.get_t<double4>("knn_oqp", novf);
            double *r2 = a.get_t<double>("knn_or", novf);
            uint32_t *c2 = a.get_t<uint32_t>("knn_oc", novf);
            double *rt2 = a.get_t<double>("knn_ort", novf);
            hipLaunchKernelGGL(k_ovf_compact, dim3(ceil_div(nb, 256)), dim3(256), 0, s, ovf, opos, nb, slots);
            LAUNCHCHK();
            hipLaunchKernelGGL(k_ovf_gather_args, dim3(ceil_div(novf, 256)), dim3(256), 0, s, slots, novf, qi, qp,
                               rbuf + b0, cbuf + b0, rt ? rt + b0 : nullptr, q0, qi2, qp2, r2, c2, rt2);
            LAUNCHCHK();
            hipLaunchKernelGGL((k_knn_gather<K, true>), dim3(ceil_div(novf, 64)), dim3(64), 0, s, tv, novf,
                               qp ? nullptr : qi2, qp ? qp2 : nullptr, r2, c2, rt ? rt2 : nullptr, ld2, lidx, 0u,
                               nb, slots, ovf);
            LAUNCHCHK();
        }
        if (out.ev && b0 == 0) HIPCHK(hipEventRecord(out.ev[2], s));
        hipLaunchKernelGGL(k_knn_finish<K>, dim3(ceil_div(nb, 256)), dim3(256), 0, s, tv, nb, qi, qp, ld2, lidx,
                           mode, kernel, rho_in, out.rho ? out.rho + o : nullptr,
                           out.densest ? out.densest + b0 : nullptr, out.idx ? out.idx + (size_t)b0 * K : nullptr,
                           out.d2 ? out.d2 + (size_t)b0 * K : nullptr, q0);
        LAUNCHCHK();
    }
    if (out.rstop) HIPCHK(hipMemcpyAsync(out.rstop, rbuf, (size_t)nq * 8, hipMemcpyDeviceToDevice, s));
}

void dev_knn(const DevTree &t, uint32_t K, int mode, int kernel, uint32_t nq, const uint32_t *qidx,
             const double4 *qpts, const double *est, const double *rho_in, const plan_derived &d, KnnOut out,
             hipStream_t s)
{
    if (nq == 0) return;
    if (K > t.n) throw std::runtime_error("KNN: k exceeds the number of tree particles");
    if (qpts && !est) throw std::runtime_error("KNN: point queries need a start radius");
    TreeView tv = view_of(t, d);
    switch (K) {
    case 8: launch_knn<8>(t, tv, nq, qidx, qpts, est, mode, kernel, rho_in, out, s); break;
    case 16: launch_knn<16>(t, tv, nq, qidx, qpts, est, mode, kernel, rho_in, out, s); break;
    case 32: launch_knn<32>(t, tv, nq, qidx, qpts, est, mode, kernel, rho_in, out, s); break;
    case 64: launch_knn<64>(t, tv, nq, qidx, qpts, est, mode, kernel, rho_in, out, s); break;
    default: throw std::runtime_error("KNN: supported k are 8, 16, 32, 64");
    }
}

/* ------------------------------------------------------------ ball search */

/* one wave per query; FILL=false counts, FILL=true writes indices in the
 * reference's traversal order (ascending tree index) */
template <bool FILL>
__global__ __launch_bounds__(64) void k_ball(TreeView tv, uint32_t nq, const double4 *__restrict__ cr,
                                             uint32_t *__restrict__ counts, const uint32_t *__restrict__ offsets,
                                             uint32_t *__restrict__ out, const double *__restrict__ rho,
                                             double rho_min, int filter)
{
    const uint32_t q = blockIdx.x;
    if (q >= nq) return;
    const int lane = threadIdx.x;
    const double4 c = cr[q];
    const double r = c.w;
    const double r2 = __dmul_rn(r, r);
    uint32_t cnt = 0;
    uint32_t *o = FILL ? out + offsets[q] : nullptr;
    const uint64_t lt = lanemask_lt();
    uint32_t nd = 0;
    uint32_t steps = 0;
    while (nd != PLAN_NONE && steps++ <= tv.nn) {
        nd = uniform_u32(nd);
        const NodeGeo g = tv.ngeo[nd];
        const NodeTop t = tv.ntop[nd];
        const bool leaf = (tv.ninfo[nd] & 0xffu) == 0u;
        int act;
        if (sphere_contain(g, c.x, c.y, c.z, r, tv.to_diag)) act = 1;
        else if (sphere_intersect(g, c.x, c.y, c.z, r, tv.to_diag)) act = leaf ? 4 : 2;
        else act = 3;
        if (act == 1 || act == 4) {
            for (uint32_t b = t.begin; b < t.end; b += 64) {
                uint32_t p = b + lane;
                bool take = p < t.end;
                if (take && act == 4) {
                    double4 x = tv.pm[p];
                    take = dist2(c.x, c.y, c.z, x.x, x.y, x.z) <= r2;
                }
                if (take && filter) take = rho[p] >= rho_min;
                uint64_t m = __ballot(take);
                if (FILL && take) o[cnt + __popcll(m & lt)] = p;
                cnt += __popcll(m);
            }
        }
        if (act == 2) nd = t.first_daughter;
        else nd = tv.skip[nd];
    }
    if (!FILL && lane == 0) counts[q] = cnt;
}

void dev_ball(const DevTree &t, uint32_t nq, const double4 *centre_r, uint32_t *counts, const uint32_t *offsets,
              uint32_t *out, const double *rho, double rho_min, int filter, hipStream_t s)
{
    if (nq == 0) return;
    plan_derived d{};
    d.cell_length[0] = d.cell_length[1] = d.cell_length[2] = 1;
    d.cell_volume = 1;
    TreeView tv = view_of(t, d);
    if (out) {
        hipLaunchKernelGGL(k_ball<true>, dim3(nq), dim3(64), 0, s, tv, nq, centre_r, counts, offsets, out, rho,
                           rho_min, filter);
    } else {
        hipLaunchKernelGGL(k_ball<false>, dim3(nq), dim3(64), 0, s, tv, nq, centre_r, counts, offsets, out, rho,
                           rho_min, filter);
    }
    LAUNCHCHK();
}

}  // namespace plan
