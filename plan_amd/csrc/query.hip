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
    __device__ __forceinline__ void sift_down_from_root(double v, uint32_t id, int lim)
    {
        int j = 0;
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

/* Lock-step walk of the tree for the 64 lanes of a wave.  COUNT mode adds
 * |S_q(r)| per lane; gather mode offers every candidate to the lane heap. */
template <int K, bool GATHER>
__device__ __forceinline__ void walk(const TreeView &tv, double qx, double qy, double qz, double r, bool active,
                                     uint32_t &count, LaneHeap<K> &heap)
{
    uint32_t cov_end = active ? 0u : 0xffffffffu;
    uint32_t nd = 0;
    uint32_t steps = 0;
    while (nd != PLAN_NONE && steps++ <= tv.nn) { /* pre-order visits each node at most once */
        nd = uniform_u32(nd);
        const NodeGeo g = tv.ngeo[nd];
        const NodeTop t = tv.ntop[nd];
        const bool leaf = (tv.ninfo[nd] & 0xffu) == 0u;
        int act = 0; /* 1 include, 2 descend, 3 reject */
        if (t.begin >= cov_end) {
            if (sphere_contain(g, qx, qy, qz, r, tv.to_diag)) act = 1;
            else if (sphere_intersect(g, qx, qy, qz, r, tv.to_diag)) act = leaf ? 1 : 2;
            else act = 3;
            if (act != 2) cov_end = t.end;
        }
        if (GATHER) {
            if (__ballot(act == 1)) {
                for (uint32_t p = t.begin; p < t.end; ++p) {
                    const double4 x = tv.pm[uniform_u32(p)];
                    if (act == 1) heap.insert(dist2(qx, qy, qz, x.x, x.y, x.z), p);
                }
            }
        } else {
            if (act == 1) count += t.end - t.begin;
        }
        if (__ballot(act == 2)) nd = t.first_daughter;
        else nd = tv.skip[nd];
    }
}

template <int K>
__global__ __launch_bounds__(64) void k_knn(TreeView tv, uint32_t nq, const uint32_t *__restrict__ qidx,
                                            const double4 *__restrict__ qpts, const double *__restrict__ est,
                                            int mode, int kernel, const double *__restrict__ rho_in,
                                            double *__restrict__ rho_out, uint32_t *__restrict__ densest,
                                            uint32_t *__restrict__ idx_out, double *__restrict__ d2_out,
                                            double *__restrict__ rstop_out)
{
    __shared__ double hd[K][64];
    __shared__ uint32_t hix[K][64];
    const int lane = threadIdx.x;
    const uint32_t q = blockIdx.x * 64 + lane;
    const bool valid = q < nq;
    double qx = 0, qy = 0, qz = 0, r = 0;
    uint32_t self = 0;
    if (valid) {
        if (qpts) {
            double4 p = qpts[q];
            qx = p.x;
            qy = p.y;
            qz = p.z;
        } else {
            self = qidx ? qidx[q] : q;
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
    LaneHeap<K> heap;
    heap.d = hd;
    heap.ix = hix;
    heap.lane = lane;
    heap.size = 0;
    heap.top_d = 0;
    heap.top_i = 0;
    /* radius schedule: r, 2r, 4r ... until |S(r)| >= K (K <= n guarantees an end;
     * 2200 doublings exceed any finite binary64 start radius) */
    bool done = !valid;
    int rounds = 0;
    while (__ballot(!done) && rounds++ < 2200) {
        uint32_t cnt = 0;
        walk<K, false>(tv, qx, qy, qz, r, !done, cnt, heap);
        if (!done) {
            if (cnt >= (uint32_t)K) done = true;
            else r = __dmul_rn(r, 2.0);
        }
    }
    uint32_t dummy = 0;
    walk<K, true>(tv, qx, qy, qz, r, valid, dummy, heap);
    if (!valid) return;
    if (rstop_out) rstop_out[q] = r;
    if (mode == 1) {
        double best = rho_in[self];
        uint32_t bi = self;
        for (int j = 0; j < K; ++j) {
            uint32_t id = hix[j][lane];
            double rj = rho_in[id];
            if (best < rj) {
                bi = id;
                best = rj;
            } else if (best == rj && bi > id) {
                bi = id;
            }
        }
        densest[q] = bi;
        return;
    }
    heap.sort_ascending();
    if (mode == 2 || idx_out) {
        for (int j = 0; j < K; ++j) {
            if (idx_out) idx_out[(size_t)q * K + j] = hix[j][lane];
            if (d2_out) d2_out[(size_t)q * K + j] = hd[j][lane];
        }
        if (mode == 2) return;
    }
    double rho = 0;
    if (kernel == 0) {
        /* MedianSphericalDensityKernel */
        double tm = __dmul_rn(4.0, hd[K / 2][lane]);
        int i = K - 1;
        for (; i >= 0; --i)
            if (!(hd[i][lane] > tm)) break;
        tm = hd[i][lane];
        for (; i >= 0; --i) rho = __dadd_rn(rho, tv.pm[hix[i][lane]].w);
        const double four_pi_over_three = (3.14159265358979323846 * 4.) / 3.;
        rho = __ddiv_rn(rho, __dmul_rn(four_pi_over_three, pow15(tm)));
    } else {
        /* QseudoQuadraticSplinesKernel; point queries use particle 0 as "self" */
        const double4 sp = tv.pm[qpts ? 0u : self];
        for (int j = 0; j < K; ++j) {
            const double4 x = tv.pm[hix[j][lane]];
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
static void launch_knn(const TreeView &tv, uint32_t nq, const uint32_t *qidx, const double4 *qpts, const double *est,
                       int mode, int kernel, const double *rho_in, KnnOut out, hipStream_t s)
{
    hipLaunchKernelGGL(k_knn<K>, dim3(ceil_div(nq, 64)), dim3(64), 0, s, tv, nq, qidx, qpts, est, mode, kernel,
                       rho_in, out.rho, out.densest, out.idx, out.d2, out.rstop);
    LAUNCHCHK();
}

void dev_knn(const DevTree &t, uint32_t K, int mode, int kernel, uint32_t nq, const uint32_t *qidx,
             const double4 *qpts, const double *est, const double *rho_in, const plan_derived &d, KnnOut out,
             hipStream_t s)
{
    if (nq == 0) return;
    if (K > t.n) throw std::runtime_error("KNN: k exceeds the number of tree particles");
    if (!qpts && est) {
        /* particle queries with explicit start radius are allowed */
    }
    if (qpts && !est) throw std::runtime_error("KNN: point queries need a start radius");
    TreeView tv = view_of(t, d);
    switch (K) {
    case 8: launch_knn<8>(tv, nq, qidx, qpts, est, mode, kernel, rho_in, out, s); break;
    case 16: launch_knn<16>(tv, nq, qidx, qpts, est, mode, kernel, rho_in, out, s); break;
    case 32: launch_knn<32>(tv, nq, qidx, qpts, est, mode, kernel, rho_in, out, s); break;
    case 64: launch_knn<64>(tv, nq, qidx, qpts, est, mode, kernel, rho_in, out, s); break;
    case 128: launch_knn<128>(tv, nq, qidx, qpts, est, mode, kernel, rho_in, out, s); break;
    default: throw std::runtime_error("KNN: supported k are 8, 16, 32, 64, 128");
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
