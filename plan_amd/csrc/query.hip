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

#include <cmath>

namespace plan {

TreeView tree_view(const DevTree &t, const plan_derived &d)
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
    v.list_limit = t.list_limit;
    v.mass_uniform = t.mass_uniform;
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

/* The reference's node decision for one (node, query) pair, evaluated by
 * all lanes of a wave: take the node's particles (0 = the sphere contains
 * the node, 1 = intersecting leaf), 2 = descend, 3 = reject.  Same binary64 operations as
 * sphere_contain and sphere_intersect.  The bounding-sphere rejection, the
 * containment test and the squared-vs-linear comparisons decide nearly every
 * pair; the box distance is computed only when some lane of the wave needs
 * it (a wave-uniform branch).  An axis where the query lies inside the box
 * adds +0 to the box distance, which leaves the sum unchanged. */
__device__ __forceinline__ int classify_sphere(const NodeGeo &g, bool leaf, double qx, double qy, double qz, double r,
                                               double to_diag, bool active = true)
{
    const double t = __dmul_rn(g.hw, to_diag);
    const double dx = __dsub_rn(g.cx, qx), dy = __dsub_rn(g.cy, qy), dz = __dsub_rn(g.cz, qz);
    const double r2 = __dmul_rn(r, r);
    const double mx = __dadd_rn(fabs(dx), t), my = __dadd_rn(fabs(dy), t), mz = __dadd_rn(fabs(dz), t);
    const double sc = __dadd_rn(__dadd_rn(__dmul_rn(mx, mx), __dmul_rn(my, my)), __dmul_rn(mz, mz));
    const bool contain = sc < r2;
    const double c2c = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
    const double tt = __dadd_rn(t, r);
    const bool far = c2c > __dmul_rn(tt, tt);
    const bool quirk = c2c < r || c2c < g.hw;
    bool inter = !far && quirk;
    const bool need_box = active && !far && !quirk && !contain;
    if (__ballot(need_box)) {
        const double lx = __dsub_rn(g.cx, g.hw), hx = __dadd_rn(g.cx, g.hw);
        const double ly = __dsub_rn(g.cy, g.hw), hy = __dadd_rn(g.cy, g.hw);
        const double lz = __dsub_rn(g.cz, g.hw), hz = __dadd_rn(g.cz, g.hw);
        const double ex = qx < lx ? __dsub_rn(qx, lx) : (qx > hx ? __dsub_rn(qx, hx) : 0.0);
        const double ey = qy < ly ? __dsub_rn(qy, ly) : (qy > hy ? __dsub_rn(qy, hy) : 0.0);
        const double ez = qz < lz ? __dsub_rn(qz, lz) : (qz > hz ? __dsub_rn(qz, hz) : 0.0);
        const double md = __dadd_rn(__dadd_rn(__dmul_rn(ex, ex), __dmul_rn(ey, ey)), __dmul_rn(ez, ez));
        inter = inter || (!far && md <= r2);
    }
    return contain ? 0 : (inter ? (leaf ? 1 : 2) : 3);
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
 * x s = p_hi + p_lo exactly; the result r rounds p_hi + (p_lo + x e / (2 s))
 * once.  rem = (p_hi - r) + (p_lo + x e / (2 s)) is the exact x^1.5 minus r
 * to far below 0.001 ulp.  The same operations on the host and the device
 * (no contraction), so both give the same r bit for bit. */
__host__ __device__ inline double pow15_rn(double x, double &rem)
{
#pragma clang fp contract(off)
    rem = 0.0;
    if (!(x > 0.0)) return 0.0;
#if defined(__HIP_DEVICE_COMPILE__)
    const double s = __dsqrt_rn(x);
    const double e = __fma_rn(-s, s, x);
    const double ph = x * s;
    const double pl = __fma_rn(x, s, -ph);
#else
    const double s = std::sqrt(x);
    const double e = std::fma(-s, s, x);
    const double ph = x * s;
    const double pl = std::fma(x, s, -ph);
#endif
    const double xe = x * e;
    const double c = xe / (s + s);
    const double t = pl + c;
    const double r = ph + t;
    const double hd = ph - r;
    rem = hd + t;
    return r;
}

/* pow15_rn with near = true when the C library's pow may round x^1.5 the
 * other way: the reference calls it (src/tree.hpp:3489), and it is not
 * correctly rounded.  Its stated worst case (exp within 0.511 ulp, log within
 * 1.5 * 2^-68 relative, so about 0.011 + 7e-5 |ln x| ulp before the final
 * rounding for y = 1.5) bounds how far from a rounding midpoint the exact
 * value can lie and still be rounded the other way; the band below is 0.012 +
 * 5e-5 (|e| + 1) ulp with e the binary exponent of x (|ln x| <= 0.7 (|e| + 1)).
 * Results that are powers of two (the spacing below halves) or tiny are
 * always flagged.  Flagged densities are checked on the host against that
 * same libm call. */
__device__ __forceinline__ double pow15_near(double x, bool &near)
{
    double rem;
    const double r = pow15_rn(x, rem);
    near = false;
    if (!(x > 0.0)) return r;
    const uint64_t rb = (uint64_t)__double_as_longlong(r);
    const uint64_t ex = rb >> 52;
    if ((rb & 0xfffffffffffffull) == 0ull || ex <= 53ull) {
        near = true;
    } else {
        const int ee = (int)(((uint64_t)__double_as_longlong(x) >> 52) & 0x7ffull) - 1023;
        const double band = 0.012 + 5e-5 * (double)(abs(ee) + 1);
        const double ulp = __longlong_as_double((long long)((ex - 52ull) << 52));
        near = 0.5 - fabs(rem) / ulp < band;
    }
    return r;
}

/* The nodes of one sibling group (siblings are contiguous in the
 * level-major array), staged in LDS by one vector load and then read by the
 * whole wave. */
struct NodeCache {
    NodeGeo g[8];
    NodeTop t[8];
    uint32_t info[8];
};

/* query position and the reference's start radius: hw of the first node on
 * the leaf-to-root path holding >= kq/4 particles (src/tree.hpp:3206-3219),
 * or the caller's estimate */
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
        while (tv.ntop[node].end - tv.ntop[node].begin < tv.kq / 4u) node = tv.ntop[node].parent;
        double hw = tv.ngeo[node].hw;
        r = __dsqrt_rn(__dmul_rn(hw, hw));
    }
}

/* One entry of a wave's include list: a node whose particles some lanes take
 * whole (contained node or intersecting leaf), with the mask of those lanes
 * (12 bytes; the node's particle range is read from the node array when the
 * entry is expanded). */
struct ListEnt {
    uint32_t node, mlo, mhi;
    __device__ __forceinline__ uint64_t mask() const { return ((uint64_t)mhi << 32) | mlo; }
};
/* include-list capacity (entries) of one wave: larger K walks larger spheres.
 * The list is most of a wave's LDS, so the capacity trades occupancy against
 * the waves that overflow it and are redone in smaller pieces.  K = 64 (the
 * 134M-particle density pass, HOP): 640 entries (10.2 KB, 16 waves per CU);
 * with 384, 512, 576, 640 and 768 entries that density launch of 209.9M
 * queries took 1093, 789, 765, 711 and 847 ms. */
template <int K> struct ListCap {
    static constexpr int value = K >= 64 ? 640 : 320;
};
/* the densest-neighbour pass (H = 64 over the particles above threshold)
 * keeps 512 entries at K = 64: at 16.8M particles its launch takes 4.68 ms
 * with 512 and 4.97 ms with 640 */
template <int K> struct HopCap {
    static constexpr int value = K >= 64 ? 512 : ListCap<K>::value;
};

/* LDS of the group-wise walk: a stack of sibling groups still to visit
 * (first node, count, mask of the queries that descended into their
 * parent), the rank -> query table of the group being evaluated and the
 * per-child result masks.  The stack depth is sized so that a K <= 32
 * select wave fits 6 KB of LDS (26 waves per CU instead of 22 at 7 KB); a
 * walk that needs more is treated as a list overflow. */
constexpr int WALK_STACK = 108;
/* a step tests each lane's own query against the children when at least
 * this many queries descended into the group */
constexpr uint32_t PER_LANE_MIN = 32;
/* A depth-first walk holds at most 7 pending sibling groups per level of
 * its current path (32 levels) plus the root: 256 groups never overflow. */
constexpr int WALK_STACK_FULL = 256;
static_assert(WALK_STACK_FULL >= 1 + 7 * 31, "a full-depth walk fits the stack");
template <int STACK> struct DenseWalkLdsT {
    uint32_t sfd[STACK];
    uint64_t smask[STACK];
    uint8_t snd[STACK];
    uint8_t rk[64];
    uint64_t inc[8];
    uint64_t desc[8];
    uint32_t scnt[64]; /* list walks: particles taken per query (packed steps) */
};
using DenseWalkLds = DenseWalkLdsT<WALK_STACK>;

/* Workgroups are dispatched round-robin over the 8 XCDs (block b runs on
 * XCD b % 8), each with its own L2.  Neighbouring waves of Morton-ordered
 * queries share nodes and candidate particles, so block b is given the
 * logical wave that makes each XCD walk one contiguous range of waves. */
constexpr uint32_t NUM_XCD = 8;
__device__ __forceinline__ uint32_t xcd_local_wave(uint32_t b, uint32_t nb)
{
    const uint32_t x = b % NUM_XCD, j = b / NUM_XCD;
    const uint32_t q = nb / NUM_XCD, rem = nb % NUM_XCD;
    return x * q + min(x, rem) + j;
}

__device__ __forceinline__ uint32_t mbcnt64(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

enum class WalkEnd { done, list_full, stack_full };

/* What a group-wise walk records: the include list with the reference's
 * sphere tests (LIST) or with the exact-ball tests of point queries with a
 * tight radius (LIST_GEO); per query the sizes of the nodes taken and of
 * those the sphere contains (COUNT); per query the exact number of particles
 * within r (GEOCOUNT: whole nodes inside the ball by size, boundary leaves
 * particle by particle). */
enum WalkMode { WALK_LIST, WALK_LIST_GEO, WALK_COUNT, WALK_GEOCOUNT };

/* The geometric decision for one (node, query) pair of the exact ball:
 * 0 = the node lies inside, 1 = boundary leaf, 2 = descend, 3 = outside. */
__device__ __forceinline__ int classify_geo(const NodeGeo &g, bool leaf, double qx, double qy, double qz, double r,
                                            double m)
{
    if (geo_contain(g, qx, qy, qz, r, m)) return 0;
    if (geo_intersect(g, qx, qy, qz, r, m)) return leaf ? 1 : 2;
    return 3;
}

/* The walk of RecursiveKNN_OpenMP (src/tree.hpp:3177-3193) for the queries
 * of the groups on the LDS stack (lane = query, radius r per lane), one
 * sibling group per step: the (child, query) pairs of a group whose parent
 * some queries descended into are packed densely over the lanes and
 * classified, then the children some query takes go to the include list
 * (with the mask of those queries) and the children some query descends
 * into push their own groups.  The set of (node, query) decisions is exactly
 * the recursive walk's.  The counting modes keep no list: COUNT adds per
 * lane the sizes of the nodes it takes (sc) and of those its sphere contains
 * (cc); GEOCOUNT adds to qcnt[query] the particles within r.  When the list
 * is full the group being decided goes back on the stack, so a later call
 * continues the walk (sp is the stack height). */
template <int CAP, int STACK, int MODE>
__device__ WalkEnd walk_dense_run(const TreeView &tv, NodeCache &nc, DenseWalkLdsT<STACK> &D, uint64_t *cont,
                                  ListEnt *lst, uint32_t &nent, double qx, double qy, double qz, double r,
                                  uint32_t &sp, uint32_t &sc, uint32_t &cc, uint32_t *qcnt = nullptr)
{
    constexpr bool LIST = MODE == WALK_LIST || MODE == WALK_LIST_GEO;
    constexpr bool GEO = MODE == WALK_LIST_GEO || MODE == WALK_GEOCOUNT;
    const int lane = threadIdx.x;
    while (sp > 0) {
        __syncthreads();
        --sp;
        const uint32_t fd = uniform_u32(D.sfd[sp]);
        const uint32_t nd = uniform_u32(D.snd[sp]);
        const uint64_t Mv = D.smask[sp];
        const uint64_t M = ((uint64_t)uniform_u32((uint32_t)(Mv >> 32)) << 32) | uniform_u32((uint32_t)Mv);
        if ((uint32_t)lane < nd) {
            nc.g[lane] = tv.ngeo[fd + lane];
            nc.t[lane] = tv.ntop[fd + lane];
            nc.info[lane] = tv.ninfo[fd + lane];
        }
        const uint32_t cnt = (uint32_t)__popcll(M);
        /* Many queries on one group (the upper levels, about 60% of all pairs at
         * 16.8M particles): each lane tests its own query against the
         * children in turn and ballots give the masks, with no lane
         * exchanges or LDS atomics.  Few queries: the (child, query) pairs
         * are packed densely over the lanes. */
        constexpr bool CAN_LANE = MODE == WALK_LIST || MODE == WALK_COUNT;
        const bool per_lane = CAN_LANE && cnt >= PER_LANE_MIN;
        if (!per_lane) {
            if ((M >> lane) & 1ull) D.rk[mbcnt64(M)] = (uint8_t)lane;
            if (lane < 8) {
                D.inc[lane] = 0;
                D.desc[lane] = 0;
                if (MODE == WALK_COUNT) cont[lane] = 0;
            }
        }
        __syncthreads();
        uint64_t I = 0, Dm = 0;
        if (per_lane) {
            const bool mine = (M >> lane) & 1ull;
            for (uint32_t c = 0; c < nd; ++c) {
                const NodeGeo g = nc.g[c];
                /* a child whose centre lies farther than t + r from the
                 * query along one axis fails the bounding-sphere test
                 * (c2c >= dx^2 > (t + r)^2, with a margin far above
                 * rounding); when that holds for every query of the group
                 * the child is rejected for all of them at once */
                const double lim = __dmul_rn(__dadd_rn(__dmul_rn(g.hw, tv.to_diag), r), 1.0 + 1e-12);
                const bool near = fabs(__dsub_rn(g.cx, qx)) <= lim && fabs(__dsub_rn(g.cy, qy)) <= lim &&
                                  fabs(__dsub_rn(g.cz, qz)) <= lim;
                if (!__ballot(mine && near)) continue;
                const bool leaf = (nc.info[c] & 0xffu) == 0u;
                const int act = classify_sphere(g, leaf, qx, qy, qz, r, tv.to_diag, mine);
                const uint64_t tk = __ballot(mine && act <= 1);
                const uint64_t dn = __ballot(mine && act == 2);
                if constexpr (MODE == WALK_COUNT || MODE == WALK_LIST) {
                    const uint32_t sz = nc.t[c].end - nc.t[c].begin;
                    if (mine && act <= 1) sc += sz;
                    if (MODE == WALK_COUNT && mine && act == 0) cc += sz;
                }
                if ((uint32_t)lane == c) {
                    I = tk;
                    Dm = dn;
                }
            }
        } else {
            const uint32_t total = nd * cnt;
            /* hardware reciprocal (1 ulp): (p + 0.5) rc stays within 2^-11 of
             * (p + 0.5) / cnt for p < 512, far inside the 0.5 / cnt margin */
            const float rc = __builtin_amdgcn_rcpf((float)cnt);
            for (uint32_t base = 0; base < total; base += 64) {
                const uint32_t p = base + (uint32_t)lane;
                const bool pv = p < total;
                /* p / cnt, exact for p < 512 (half a step of margin) */
                uint32_t c = (uint32_t)(((float)p + 0.5f) * rc);
                c = pv ? c : 0u;
                const uint32_t k = pv ? p - c * cnt : 0u;
                const int j = (int)D.rk[k];
                const double xj = __shfl(qx, j, 64), yj = __shfl(qy, j, 64), zj = __shfl(qz, j, 64),
                             rj = __shfl(r, j, 64);
                const NodeGeo g = nc.g[c];
                const bool leaf = (nc.info[c] & 0xffu) == 0u;
                int act;
                if constexpr (GEO) act = classify_geo(g, leaf, xj, yj, zj, rj, tv.margin_abs);
                else act = classify_sphere(g, leaf, xj, yj, zj, rj, tv.to_diag, pv);
                if constexpr (MODE == WALK_GEOCOUNT) {
                    if (pv && act == 0) atomicAdd(&qcnt[j], nc.t[c].end - nc.t[c].begin);
                    if (pv && act == 1) {
                        const NodeTop t = nc.t[c];
                        const double r2 = rj * rj;
                        uint32_t in = 0;
                        for (uint32_t q = t.begin; q < t.end; ++q) {
                            const double4 x = tv.pm[q];
                            if (dist2(xj, yj, zj, x.x, x.y, x.z) <= r2) ++in;
                        }
                        atomicAdd(&qcnt[j], in);
                    }
                } else {
                    if (pv && act <= 1) atomicOr((unsigned long long *)&D.inc[c], 1ull << j);
                    if (MODE == WALK_LIST && pv && act <= 1) atomicAdd(&D.scnt[j], nc.t[c].end - nc.t[c].begin);
                    if (MODE == WALK_COUNT && pv && act == 0) atomicOr((unsigned long long *)&cont[c], 1ull << j);
                }
                if (pv && act == 2) atomicOr((unsigned long long *)&D.desc[c], 1ull << j);
            }
            __syncthreads();
            if constexpr (MODE == WALK_COUNT) {
                for (uint32_t c = 0; c < nd; ++c) {
                    const uint32_t sz = nc.t[c].end - nc.t[c].begin;
                    if ((D.inc[c] >> lane) & 1ull) sc += sz;
                    if ((cont[c] >> lane) & 1ull) cc += sz;
                }
            }
            if ((uint32_t)lane < nd) {
                I = D.inc[lane];
                Dm = D.desc[lane];
            }
        }
        NodeTop t{0u, 0u, 0u, 0u};
        uint32_t cnd = 0;
        if ((uint32_t)lane < nd) {
            t = nc.t[lane];
            cnd = nc.info[lane] & 0xffu;
        }
        if constexpr (LIST) {
            const uint64_t mi = __ballot(I != 0ull);
            const uint32_t ni = (uint32_t)__popcll(mi);
            if (nent + ni > min((uint32_t)CAP, tv.list_limit)) {
                ++sp; /* the group is decided again by the next call */
                return WalkEnd::list_full;
            }
            if (I != 0ull) lst[nent + mbcnt64(mi)] = ListEnt{fd + (uint32_t)lane, (uint32_t)I, (uint32_t)(I >> 32)};
            nent += ni;
        }
        const uint64_t md = __ballot(Dm != 0ull);
        const uint32_t nD = (uint32_t)__popcll(md);
        if (sp + nD > (uint32_t)STACK) return WalkEnd::stack_full;
        __syncthreads();
        if (Dm != 0ull) {
            /* reversed, so the first child's group is popped first */
            const uint32_t slot = sp + nD - 1u - mbcnt64(md);
            D.sfd[slot] = t.first_daughter;
            D.snd[slot] = (uint8_t)cnd;
            D.smask[slot] = Dm;
        }
        sp += nD;
    }
    return WalkEnd::done;
}

/* start a walk at the root for the queries of mask M0 */
template <int STACK> __device__ __forceinline__ uint32_t walk_dense_start(DenseWalkLdsT<STACK> &D, uint64_t M0)
{
    __syncthreads();
    D.scnt[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        D.sfd[0] = 0;
        D.snd[0] = 1;
        D.smask[0] = M0;
    }
    return M0 ? 1u : 0u;
}

/* the whole walk into one list; false when the list or the stack is full.
 * cnt: per lane, the particles of the entries recorded for its query (the
 * size of S(r)) */
template <int CAP>
__device__ bool walk_dense(const TreeView &tv, NodeCache &nc, DenseWalkLds &D, ListEnt *lst, uint32_t &nent, double qx,
                           double qy, double qz, double r, uint64_t M0, uint32_t &cnt)
{
    uint32_t sp = walk_dense_start(D, M0), sc = 0, cc = 0;
    const bool ok =
        walk_dense_run<CAP, WALK_STACK, WALK_LIST>(tv, nc, D, nullptr, lst, nent, qx, qy, qz, r, sp, sc, cc) ==
        WalkEnd::done;
    __syncthreads();
    cnt = sc + D.scnt[threadIdx.x];
    return ok;
}

/* LDS of the radius-schedule and tight-radius kernels */
struct CountLds {
    NodeCache nc;
    DenseWalkLdsT<WALK_STACK_FULL> dw;
    uint64_t cont[8];
    uint32_t qcnt[64];
};

/* Pass 1: radius schedule r, 2r, 4r ... until |S(r)| >= kq (node sizes only,
 * no heap; kq <= n guarantees an end, 2200 doublings exceed any finite start). */
__global__ __launch_bounds__(64) void k_knn_radius(TreeView tv, uint32_t nq, const uint32_t *__restrict__ qidx,
                                                   const double4 *__restrict__ qpts, const double *__restrict__ est,
                                                   const double *__restrict__ rt, double *__restrict__ r_out,
                                                   uint32_t *__restrict__ c_out, uint32_t qpw)
{
    __shared__ CountLds L;
    const int lane = threadIdx.x;
    const uint32_t q = blockIdx.x * qpw + lane;
    const bool valid = (uint32_t)lane < qpw && q < nq;
    double qx = 0, qy = 0, qz = 0, r = 0;
    uint32_t self = 0;
    if (valid) query_setup(tv, q, qidx, qpts, est, qx, qy, qz, r, self);
    /* a point query resolved by k_point_tight: S(est) holds >= kq particles,
     * so the schedule stops at est */
    const bool tight = valid && rt && rt[q] > 0;
    bool done = !valid || tight;
    int rounds = 0;
    uint32_t inside = tight ? tv.kq : 0u;
    while (__ballot(!done) && rounds++ < 2200) {
        uint32_t cnt = 0, cin = 0, nent = 0;
        uint32_t sp = walk_dense_start(L.dw, __ballot(!done));
        walk_dense_run<0, WALK_STACK_FULL, WALK_COUNT>(tv, L.nc, L.dw, L.cont, nullptr, nent, qx, qy, qz, r, sp, cnt,
                                                      cin);
        if (!done) {
            if (cnt >= tv.kq) {
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

/* Point queries (saddle points, user points) with a start radius est: when
 * some radius r_t <= est - margin already holds >= kq particles within
 * distance r_t, every particle within r_t belongs to the reference's
 * candidate set S(est) (src/tree.hpp:3160-3230), so the kq smallest
 * (d^2, index) of S(est) are the kq nearest overall and can be gathered from
 * the ball of radius r_t instead of all of S(est).  r_t starts at the half
 * width of the deepest node holding q with >= max(kq/4, 1) particles and doubles.
 * rt_out[q] = r_t, or 0 when no such radius exists (the query then takes
 * the candidate-set path).  The counts are exact (boundary leaves particle
 * by particle) and come from the group-wise walk with the ball tests. */
__global__ __launch_bounds__(64) void k_point_tight(TreeView tv, uint32_t nq, const double4 *__restrict__ qpts,
                                                    const double *__restrict__ est, double *__restrict__ rt_out,
                                                    uint32_t qpw)
{
    __shared__ CountLds L;
    const int lane = threadIdx.x;
    const uint32_t q = blockIdx.x * qpw + lane;
    const bool valid = (uint32_t)lane < qpw && q < nq;
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
            if (t.end - t.begin < max(tv.kq / 4u, 1u)) break;
            node = next;
        }
        r = tv.ngeo[node].hw;
    }
    bool done = !valid || !(r > 0) || !(r <= lim);
    bool ok = false;
    int rounds = 0;
    while (__ballot(!done) && rounds++ < 64) {
        L.qcnt[lane] = 0;
        uint32_t sc = 0, cc = 0, nent = 0;
        uint32_t sp = walk_dense_start(L.dw, __ballot(!done));
        walk_dense_run<0, WALK_STACK_FULL, WALK_GEOCOUNT>(tv, L.nc, L.dw, nullptr, nullptr, nent, qx, qy, qz, r, sp,
                                                         sc, cc, L.qcnt);
        __syncthreads();
        const uint32_t cnt = L.qcnt[lane];
        if (!done) {
            if (cnt >= tv.kq) {
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

__global__ void k_ovf_compact(const uint32_t *__restrict__ ovf, const uint32_t *__restrict__ pos, uint32_t n,
                              uint32_t *__restrict__ list)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && ovf[i]) list[pos[i]] = i;
}

/* waves of f queries -> the waves of one query that cover them */
__global__ void k_wave_expand(const uint32_t *__restrict__ waves, uint32_t n, uint32_t f, uint32_t *__restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n * f) out[i] = waves[i / f] * f + i % f;
}

/* ---------------------------------------------- wave-per-query selection */

__device__ __forceinline__ bool key_less(double a, uint32_t ai, double b, uint32_t bi)
{
    return a < b || (a == b && ai < bi);
}

/* value of a wave-uniform lane j */
__device__ __forceinline__ double lane_f64(double v, int j)
{
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)b, j);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(b >> 32), j);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ uint32_t lane_u32(uint32_t v, int j) { return __builtin_amdgcn_readlane(v, j); }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

/* inclusive sum over the 64 lanes with DPP: shifts by 1, 2, 4, 8 within
 * rows of 16 lanes (lanes shifted in from outside the row add 0), then the
 * last lane of row 0 added to row 1 and of row 2 to row 3 (row_bcast:15),
 * and the last lane of row 1 to rows 2 and 3 (row_bcast:31) */
__device__ __forceinline__ uint32_t wave_incl_sum_dpp(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);
    return v;
}

/* lane l reads lane l ^ J: DPP for J = 1, 2 (quad permutes) and 8 (row
 * rotate by 8), swizzle for 4 and 16, bpermute for 32 */
template <int J> __device__ __forceinline__ uint32_t xor_lane(uint32_t v)
{
    if constexpr (J == 1) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
    else if constexpr (J == 2) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);
    else if constexpr (J == 4) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x101F);
    else if constexpr (J == 8) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, false);
    else if constexpr (J == 16) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x401F);
    else return (uint32_t)__shfl_xor((int)v, J, 64);
}
/* lanes l whose compare-exchange keeps the smaller key: for a bitonic stage
 * (k, j) those with ((l & k) == 0) == ((l & j) == 0); for a half-cleaner
 * (k = 0) those with (l & j) == 0 */
constexpr uint64_t keep_low_mask(int k, int j)
{
    uint64_t m = 0;
    for (int l = 0; l < 64; ++l) {
        const bool asc = k == 0 || (l & k) == 0;
        if (asc == ((l & j) == 0)) m |= 1ull << l;
    }
    return m;
}
/* The selection sorts 32-bit keys: the top 26 bits of d^2 below the sign
 * (exponent and 15 mantissa bits; d^2 >= 0, so its bits order like its
 * value) with the key's source lane in the low 6 bits.  These keys are
 * distinct, a compare-exchange is one min or max, and only the source lane
 * travels; the (d^2, index) pairs follow by one permute at the end.  Keys
 * whose top bits agree may come out in the wrong exact order (only among
 * themselves: the truncation is monotone), which exact_fixup repairs. */
__device__ __forceinline__ uint32_t trunc_key(double d, int lane)
{
    const uint32_t hi = (uint32_t)((uint64_t)__double_as_longlong(d) >> 32);
    return ((hi << 1) & ~63u) | (uint32_t)lane;
}

/* one compare-exchange of the 32-bit keys with lane l ^ J (bitonic stage
 * (K2, J), or a half-cleaner for K2 = 0) */
template <int K2, int J> __device__ __forceinline__ void cex_key(uint32_t &k)
{
    const uint32_t pk = xor_lane<J>(k);
    const bool low = __builtin_amdgcn_inverse_ballot_w64(keep_low_mask(K2, J));
    /* keys are distinct: keep-low lanes take the partner's key when it is
     * smaller, the others when it is larger */
    k = ((pk < k) == low) ? pk : k;
}

/* Odd-even transposition passes on the exact (d^2, index) keys until no
 * pair is out of order.  Runs only when two neighbours share their truncated
 * key, and then moves only keys within such runs (the runs are already in
 * place relative to each other), so one or two passes usually end it. */
__device__ __forceinline__ void exact_fixup(double &d, uint32_t &i)
{
    const int lane = threadIdx.x & 63;
    for (;;) {
        bool moved = false;
#pragma unroll
        for (int ph = 0; ph < 2; ++ph) {
            const int partner = ph == 0 ? (lane ^ 1) : ((lane & 1) ? lane + 1 : lane - 1);
            const bool ok = partner >= 0 && partner < 64;
            const int src = ok ? partner : lane;
            const double pd = __shfl(d, src, 64);
            const uint32_t pi = (uint32_t)__shfl((int)i, src, 64);
            const bool p_less = key_less(pd, pi, d, i);
            const bool m_less = key_less(d, i, pd, pi);
            /* the lower lane keeps the smaller key */
            const bool sw = ok && (lane < partner ? p_less : m_less);
            d = sw ? pd : d;
            i = sw ? pi : i;
            moved = moved || sw;
        }
        if (!__ballot(moved)) break;
    }
}

/* after a key sort: fetch the (d^2, index) pairs in key order, repair runs
 * of equal truncated keys */
__device__ __forceinline__ void keys_to_pairs(uint32_t k, double &d, uint32_t &i)
{
    const int lane = threadIdx.x & 63;
    const int src = (int)(k & 63u);
    d = __shfl(d, src, 64);
    i = (uint32_t)__shfl((int)i, src, 64);
    /* (+inf, NONE) padding shares one key and needs no order */
    const uint32_t t = k >> 6;
    const uint32_t tn = (uint32_t)__shfl_down((int)t, 1, 64);
    if (__ballot(lane < 63 && t == tn && d < __longlong_as_double(0x7ff0000000000000ll))) exact_fixup(d, i);
}

/* Sort of one (d^2, index) key per lane over the 64 lanes, ascending in lane
 * order (bitonic network on the truncated keys, then the exact repair). */
__device__ __forceinline__ void sort64(double &d, uint32_t &i)
{
    uint32_t k = trunc_key(d, threadIdx.x & 63);
    cex_key<2, 1>(k);
    cex_key<4, 2>(k); cex_key<4, 1>(k);
    cex_key<8, 4>(k); cex_key<8, 2>(k); cex_key<8, 1>(k);
    cex_key<16, 8>(k); cex_key<16, 4>(k); cex_key<16, 2>(k); cex_key<16, 1>(k);
    cex_key<32, 16>(k); cex_key<32, 8>(k); cex_key<32, 4>(k); cex_key<32, 2>(k); cex_key<32, 1>(k);
    cex_key<64, 32>(k); cex_key<64, 16>(k); cex_key<64, 8>(k); cex_key<64, 4>(k); cex_key<64, 2>(k); cex_key<64, 1>(k);
    keys_to_pairs(k, d, i);
}

/* (bd, bi) and (cd, ci) sorted ascending over the lanes: (bd, bi) becomes
 * the 64 smallest keys of both, sorted.  The lane-wise minimum against the
 * reversed second list (exact keys) keeps the 64 smallest and is bitonic;
 * six half-cleaners on the truncated keys sort it up to runs of equal
 * truncated keys (a bitonic sequence stays bitonic for every threshold
 * between distinct truncated keys), which the exact repair orders. */
__device__ __forceinline__ void merge64(double &bd, uint32_t &bi, double cd, uint32_t ci)
{
    const int lane = threadIdx.x & 63;
    const double rd = __shfl(cd, 63 - lane, 64);
    const uint32_t ri = __shfl(ci, 63 - lane, 64);
    const bool t = (rd < bd) | ((rd == bd) & (ri < bi));
    bd = t ? rd : bd;
    bi = t ? ri : bi;
    uint32_t k = trunc_key(bd, lane);
    cex_key<0, 32>(k); cex_key<0, 16>(k); cex_key<0, 8>(k); cex_key<0, 4>(k); cex_key<0, 2>(k); cex_key<0, 1>(k);
    keys_to_pairs(k, bd, bi);
}

/* Selection scratch of the query being selected: its list entries (sent),
 * the marks of first candidate positions (start) and up to 64 staged
 * candidates (sd, si) waiting to be sorted. */
template <int CAP> struct SelectScratch {
    uint16_t sent[CAP];
    uint8_t start[64];
    double sd[64];
    uint32_t si[64];
};
/* LDS of one select wave: node cache, the include list of the walk and the
 * selection scratch (sharing its space with the walk's stack) */
template <int CAP> struct SelectLds {
    NodeCache nc;
    ListEnt lst[CAP];
    union {
        SelectScratch<CAP> ss;
        DenseWalkLds dw;
    };
};
static_assert(sizeof(SelectLds<ListCap<32>::value>) <= 6144, "a K <= 32 select wave should fit 6 KB of LDS");
/* the few waves that walk in several fills keep the walk's stack beside the
 * selection's arrays, so the walk resumes after each selection */
template <int CAP> struct ResumeLds {
    NodeCache nc;
    ListEnt lst[CAP];
    SelectScratch<CAP> ss;
    DenseWalkLdsT<WALK_STACK_FULL> dw;
    uint64_t cont[8];
};

/* 64 candidates (one per lane, (+inf, NONE) where none) into the running
 * best-64 (bd, bi) */
__device__ __forceinline__ void feed64(double d, uint32_t id, double &bd, uint32_t &bi, bool &first)
{
    sort64(d, id);
    if (first) {
        bd = d;
        bi = id;
        first = false;
    } else {
        merge64(bd, bi, d, id);
    }
}

/* Feed the candidates of query bit j of the recorded list to the running
 * best-64 (bd, bi).  One ballot pass over the list finds the query's entries
 * (their list positions to sent).  They are expanded 64 entries at a time
 * (nearly every query has fewer): lane r reads entry r's particle range from
 * the node array, a prefix sum of the sizes gives each entry's first
 * candidate position, and the candidates are laid out densely over the
 * lanes, 64 per round (a ballot over the marked first positions gives the
 * entry of each position; the entry's first particle comes from its lane).
 * The first 64 candidates within T are sorted at once; later ones that can
 * still enter the kq best (below the current kq-th key) are staged in LDS
 * and sorted and merged 64 at a time. */
template <int K, class LDS>
__device__ __forceinline__ void select_list(const TreeView &tv, LDS &L, uint32_t nent, int j, double qx,
                                            double qy, double qz, double T, double &bd, uint32_t &bi, bool &first)
{
    const int lane = threadIdx.x & 63;
    const double inf = __longlong_as_double(0x7ff0000000000000ll);
    auto &S = L.ss;
    nent = __builtin_amdgcn_readfirstlane(nent);
    uint32_t M = 0;
    __syncthreads();
    for (uint32_t e0 = 0; e0 < nent; e0 += 64) {
        const uint32_t e = e0 + (uint32_t)lane;
        bool has = false;
        if (e < nent) has = (L.lst[e].mask() >> j) & 1ull;
        const uint64_t bm = __ballot(has);
        if (has) S.sent[M + mbcnt64(bm)] = (uint16_t)e;
        M += (uint32_t)__popcll(bm);
    }
    __syncthreads();
    uint32_t ns = 0; /* staged candidates */
    for (uint32_t g0 = 0; g0 < M; g0 += 64) {
        /* lane r < Mg holds entry g0 + r: its first particle b0, size, first
         * candidate position p0 and base (particle index = base + position,
         * modulo 2^32); node ranges are never empty */
        const uint32_t Mg = min(M - g0, 64u);
        uint32_t b0 = 0, sz = 0;
        if ((uint32_t)lane < Mg) {
            const NodeTop t = tv.ntop[L.lst[S.sent[g0 + lane]].node];
            b0 = t.begin;
            sz = t.end - t.begin;
        }
        const uint32_t incl = wave_incl_sum_dpp(sz);
        const uint32_t C = lane_u32(incl, 63);
        const uint32_t p0 = incl - sz;
        const uint32_t base = b0 - p0;
        const bool starts = (uint32_t)lane < Mg && sz > 0u;
        for (uint32_t k0 = 0; k0 < C; k0 += 64) {
            __syncthreads();
            S.start[lane] = 0;
            __syncthreads();
            if (starts && p0 >= k0 && p0 - k0 < 64u) S.start[p0 - k0] = 1;
            __syncthreads();
            const uint64_t sm = __ballot(S.start[lane] != 0);
            const uint32_t before = (uint32_t)__popcll(__ballot(starts && p0 < k0));
            /* the entry holding position k0 + lane: the entries starting at
             * or below it, less one */
            const uint32_t rk = before + mbcnt64(sm) + (uint32_t)((sm >> lane) & 1ull) - 1u;
            const uint32_t pb = (uint32_t)__shfl((int)base, (int)(rk & 63u), 64);
            const uint32_t c = k0 + (uint32_t)lane;
            double d = inf;
            uint32_t id = PLAN_NONE;
            if (c < C) {
                const uint32_t p = pb + c;
                /* always below n: keeps a wrong rank from reading outside
                 * the particle arrays */
                if (p < tv.n) {
                    const double4 x = tv.pm[p];
                    const double dd = dist2(qx, qy, qz, x.x, x.y, x.z);
                    if (dd <= T) {
                        d = dd;
                        id = p;
                    }
                }
            }
            if (first) {
                feed64(d, id, bd, bi, first);
                continue;
            }
            /* only keys below the current kq-th can enter the kq best */
            const double kd = lane_f64(bd, (int)tv.kq - 1);
            const uint32_t ki = lane_u32(bi, (int)tv.kq - 1);
            const bool v = id != PLAN_NONE && key_less(d, id, kd, ki);
            const uint64_t vm = __ballot(v);
            const uint32_t nv = (uint32_t)__popcll(vm);
            if (nv == 0) continue;
            const uint32_t off = ns + mbcnt64(vm);
            __syncthreads();
            if (v && off < 64u) {
                S.sd[off] = d;
                S.si[off] = id;
            }
            if (ns + nv < 64u) {
                ns += nv;
                continue;
            }
            __syncthreads();
            feed64(S.sd[lane], S.si[lane], bd, bi, first);
            __syncthreads();
            if (v && off >= 64u) {
                S.sd[off - 64u] = d;
                S.si[off - 64u] = id;
            }
            ns = ns + nv - 64u;
        }
    }
    if (ns > 0) {
        __syncthreads();
        const bool h = (uint32_t)lane < ns;
        feed64(h ? S.sd[lane] : inf, h ? S.si[lane] : PLAN_NONE, bd, bi, first);
    }
    __syncthreads();
}

/* What a query's sorted neighbour list (lane l holds the l-th nearest, l < kq)
 * is used for: the densest neighbour (HOP), the density (median spherical or
 * QS kernel) and/or the lists themselves.  Sums run in the reference's
 * order (src/tree.hpp:3320-3420).  The median kernel's last step, the same
 * for every query, is left to median_density: its x (d^2 of the i*-th
 * neighbour) and mass sum go to lane `slot` (tm, num), so one pass does it
 * for the wave's queries a lane each. */
template <int K>
__device__ __forceinline__ void finish_query(const TreeView &tv, double bd, uint32_t bi, uint32_t q, uint32_t self,
                                             bool byq, int mode, int kernel, const double *__restrict__ rho_in,
                                             double *__restrict__ rho_out, uint32_t *__restrict__ densest,
                                             uint32_t *__restrict__ idx_out, double *__restrict__ d2_out, int slot,
                                             double &tm_out, double &num_out)
{
    const int lane = threadIdx.x & 63;
    const int kq = (int)tv.kq;
    const bool mine = lane < kq;
    if (mode == 1) {
        /* maximum of (rho, -index) over self and the neighbours */
        double v = -__longlong_as_double(0x7ff0000000000000ll);
        uint32_t id = PLAN_NONE;
        if (mine && bi != PLAN_NONE) {
            v = rho_in[bi];
            id = bi;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double pv = __shfl_xor(v, o, 64);
            const uint32_t pid = __shfl_xor(id, o, 64);
            if (pv > v || (pv == v && pid < id)) {
                v = pv;
                id = pid;
            }
        }
        if (lane == 0) {
            const double sv = rho_in[self];
            uint32_t best = self;
            if (sv < v || (sv == v && id < self)) best = id;
            densest[q] = best;
        }
        return;
    }
    if (mode == 2 || idx_out) {
        if (mine) {
            if (idx_out) idx_out[(size_t)q * kq + lane] = bi;
            if (d2_out) d2_out[(size_t)q * kq + lane] = bd;
        }
        if (mode == 2) return;
    }
    double rho = 0;
    if (kernel == 0) {
        /* MedianSphericalDensityKernel: m = 4 d2[kq/2]; i* = last index with
         * d2 <= m; masses summed from i* down to 0 */
        const double m = __dmul_rn(4.0, lane_f64(bd, kq / 2));
        const uint64_t in = __ballot(lane > kq / 2 && lane < kq && !(bd > m));
        const int is = in ? 63 - __clzll((long long)in) : kq / 2;
        const double tm = lane_f64(bd, is);
        if (tv.msum && tv.mass_uniform) {
            /* every mass is m0: the sum of i* + 1 of them is the table's */
            rho = tv.msum[is + 1];
        } else if (const double w = (lane <= is) ? tv.pm[bi].w : 0.0; tv.msum && !__ballot(lane <= is && w != tv.m0)) {
            /* i* + 1 equal masses: their sum in that order is the table's */
            rho = tv.msum[is + 1];
        } else {
            /* masses from i* down to 0; the lanes above i* add +0 to the
             * leading zero, which changes nothing (K >= kq) */
#pragma unroll
            for (int j = K - 1; j >= 0; --j) rho = __dadd_rn(rho, lane_f64(w, j));
        }
        if (lane == slot) {
            tm_out = tm;
            num_out = rho;
        }
        return;
    } else {
        /* QseudoQuadraticSplinesKernel; point queries use particle 0 as "self" */
        double term = 0;
        if (mine) {
            const double4 sp = tv.pm[self];
            const double4 x = tv.pm[bi];
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
            term = __dmul_rn(__ddiv_rn(x.w, tv.cell_volume), w);
        }
        /* lanes from kq on add +0 after the sum, which changes nothing */
        for (int j = 0; j < K; ++j) rho = __dadd_rn(rho, lane_f64(term, j));
    }
    if (lane == 0) rho_out[byq ? q : self] = rho;
}

/* The median kernel's density of this lane's query, num / (4 pi / 3 x^1.5)
 * (src/tree.hpp:3489), with x^1.5 correctly rounded; densities whose x^1.5
 * the C library's pow may round the other way are listed for the host. */
__device__ __forceinline__ void median_density(const TreeView &tv, double tm, double num, uint32_t q, uint32_t self,
                                               bool byq, double *__restrict__ rho_out)
{
    const double four_pi_over_three = (3.14159265358979323846 * 4.) / 3.;
    bool near;
    const double rho = __ddiv_rn(num, __dmul_rn(four_pi_over_three, pow15_near(tm, near)));
    if (near && tv.pw_count) {
        const uint32_t slot = atomicAdd(tv.pw_count, 1u);
        if (slot < tv.pw_cap) {
            tv.pw_idx[slot] = byq ? q : self;
            tv.pw_x[slot] = tm;
            tv.pw_num[slot] = num;
        }
    }
    rho_out[byq ? q : self] = rho;
}

/* The kq nearest of S(r*) per query, and their use, in one kernel (K, the
 * list length the kernel is built for, is kq rounded up to 8, 16, 32 or 64).
 * A wave walks the tree once in lock step for its 64 queries (lane =
 * query), recording the nodes some lane takes whole; then it selects for one
 * query at a time, all 64 lanes on that query's candidates.
 * RESUME = false: a wave whose list overflows its capacity flags itself
 * (ovf[wave]) and stops.  RESUME = true: the flagged waves (waves[block])
 * walk in several list fills; each query's K best so far are kept in HBM
 * (sd2/sidx, K per query) between fills. */
template <int K, bool RESUME, bool FUSE, int CAP = ListCap<K>::value>
__global__ __launch_bounds__(64) void k_knn_select(TreeView tv, uint32_t nq, const uint32_t *__restrict__ qidx,
                                                   const double4 *__restrict__ qpts, double *__restrict__ r_io,
                                                   uint32_t *__restrict__ c_io, const double *__restrict__ rt_in,
                                                   uint32_t q0, const uint32_t *__restrict__ waves,
                                                   uint32_t *__restrict__ ovf, double *__restrict__ sd2,
                                                   uint32_t *__restrict__ sidx, int mode, int kernel,
                                                   const double *__restrict__ rho_in, double *__restrict__ rho_out,
                                                   uint32_t *__restrict__ densest, uint32_t *__restrict__ idx_out,
                                                   double *__restrict__ d2_out, uint32_t qpw)
{
    using Lds = std::conditional_t<RESUME, ResumeLds<CAP>, SelectLds<CAP>>;
    __shared__ Lds L;
    const int lane = threadIdx.x;
    /* RESUME: waves[] lists the waves to redo.  Otherwise waves[] (when
     * given) lists 64-query waves to split into 64 / qpw waves each. */
    const uint32_t lb = xcd_local_wave(blockIdx.x, gridDim.x);
    uint32_t wv = lb;
    if (RESUME) {
        wv = waves[lb];
    } else if (waves) {
        const uint32_t sub = 64u / qpw;
        wv = waves[lb / sub] * sub + lb % sub;
    }
    const uint32_t q = wv * qpw + lane;
    const bool valid = (uint32_t)lane < qpw && q < nq;
    double qx = 0, qy = 0, qz = 0, r = 0;
    uint32_t self = 0;
    if (valid) query_setup(tv, q, qidx, qpts, FUSE ? nullptr : r_io, qx, qy, qz, r, self, q0);
    const double inf = __longlong_as_double(0x7ff0000000000000ll);
    const uint32_t nvalid = (uint32_t)__popcll(__ballot(valid));
    uint32_t cin = 0;
    uint32_t nent = 0;
    bool filled = false;
    if constexpr (FUSE && !RESUME) {
        /* the radius schedule r0, 2 r0, ... until |S(r)| >= kq
         * (src/tree.hpp:3206-3223): the first round walks every query at r0;
         * each later round walks only the queries still short of kq
         * candidates, at their doubled radius, after taking their bits out of
         * the previous round's entries (and dropping entries left with no
         * query), so the list holds each query's entries of its last round.
         * When such a round fills the list, the wave starts the list again
         * and walks every query at its current radius in each round. */
        const uint64_t V = __ballot(valid);
        uint64_t U = V;
        bool whole = false;
        nent = 0;
        for (int round = 0; round < 2200; ++round) {
            if (whole) nent = 0;
            const uint32_t e0 = nent;
            const uint64_t W = whole ? V : U;
            uint32_t sc = 0;
            if (!walk_dense<CAP>(tv, L.nc, L.dw, L.lst, nent, qx, qy, qz, r, W, sc)) {
                if (!whole && round > 0) {
                    whole = true;
                    continue;
                }
                if (lane == 0) ovf[wv] = 1u;
                return;
            }
            __syncthreads();
            const bool need = ((W >> lane) & 1ull) && sc < tv.kq;
            const uint64_t N = __ballot(need);
            if (!N) break;
            if (!whole) {
                /* in place: an entry only moves down, past those already read */
                uint32_t w = e0;
                for (uint32_t x0 = e0; x0 < nent; x0 += 64) {
                    const uint32_t e = x0 + (uint32_t)lane;
                    ListEnt le{0u, 0u, 0u};
                    if (e < nent) {
                        le = L.lst[e];
                        le.mlo &= ~(uint32_t)N;
                        le.mhi &= ~(uint32_t)(N >> 32);
                    }
                    const bool kp = (le.mlo | le.mhi) != 0u;
                    const uint64_t keep = __ballot(kp);
                    __syncthreads();
                    if (kp) L.lst[w + mbcnt64(keep)] = le;
                    w += (uint32_t)__popcll(keep);
                }
                nent = w;
                __syncthreads();
            }
            if (need) r = __dmul_rn(r, 2.0);
            U = N;
        }
        filled = true;
        if (valid) {
            r_io[q] = r;
            c_io[q] = 0u;
        }
    } else if constexpr (FUSE) {
        /* waves that walk in several fills: the radius schedule r0, 2 r0, ...
         * until |S(r)| >= kq (src/tree.hpp:3206-3223) by counting walks, which
         * also sum the particles of the nodes each sphere contains */
        const uint64_t V = __ballot(valid);
        for (int round = 0; round < 2200; ++round) {
            uint32_t sc = 0, cc = 0;
            uint32_t sp = walk_dense_start(L.dw, V);
            walk_dense_run<CAP, WALK_STACK_FULL, WALK_COUNT>(tv, L.nc, L.dw, L.cont, L.lst, nent, qx, qy, qz, r, sp,
                                                             sc, cc);
            const bool need = valid && sc < tv.kq;
            if (!need) cin = cc;
            if (!__ballot(need)) break;
            __syncthreads();
            if (need) r = __dmul_rn(r, 2.0);
        }
        if (valid) {
            r_io[q] = r;
            c_io[q] = cin;
        }
    } else {
        if (valid) cin = c_io[q];
    }
    /* Candidates in nodes the sphere contains lie within r* (a contained
     * node's box is inside the sphere with a margin far above the Morton
     * quantisation); when there are >= kq of them the kq nearest are within
     * r*, and only d^2 <= (r* + 1e-9)^2 can enter the list.  A point query
     * with a tight radius (k_point_tight) has >= kq particles within r_t and
     * takes the kq nearest within r_t. */
    double T = inf;
    if (valid && cin >= tv.kq) {
        const double rr = __dadd_rn(r, 1e-9);
        T = __dmul_rn(rr, rr);
    }
    /* point queries with a tight radius take the exact-ball tests; they run
     * one query per wave, so the choice is the wave's */
    const bool geo = valid && rt_in && rt_in[q] > 0;
    if (geo) {
        r = rt_in[q];
        T = r * r;
    }
    const bool wgeo = __ballot(geo) != 0ull;
    if constexpr (!RESUME) {
        if (!filled) {
            nent = 0;
            uint32_t sp = walk_dense_start(L.dw, __ballot(valid)), sc = 0, cc = 0;
            const WalkEnd e =
                wgeo ? walk_dense_run<CAP, WALK_STACK, WALK_LIST_GEO>(tv, L.nc, L.dw, nullptr, L.lst, nent, qx, qy, qz,
                                                                       r, sp, sc, cc)
                     : walk_dense_run<CAP, WALK_STACK, WALK_LIST>(tv, L.nc, L.dw, nullptr, L.lst, nent, qx, qy, qz, r,
                                                                  sp, sc, cc);
            if (e != WalkEnd::done) {
                if (lane == 0) ovf[wv] = 1u;
                return;
            }
        }
        __syncthreads();
        double d_tm = 0, d_num = 0;
        for (uint32_t j = 0; j < nvalid; ++j) {
            double bd = inf;
            uint32_t bi = PLAN_NONE;
            bool first = true;
            select_list<K>(tv, L, nent, (int)j, lane_f64(qx, j), lane_f64(qy, j), lane_f64(qz, j), lane_f64(T, j), bd,
                        bi, first);
            finish_query<K>(tv, bd, bi, wv * qpw + j, lane_u32(self, j), qpts || qidx, mode, kernel, rho_in,
                            rho_out, densest, idx_out, d2_out, (int)j, d_tm, d_num);
        }
        if (mode == 0 && kernel == 0 && valid) median_density(tv, d_tm, d_num, q, self, qpts || qidx, rho_out);
    } else {
        double *bd2 = sd2 + (size_t)blockIdx.x * qpw * K;
        uint32_t *bix = sidx + (size_t)blockIdx.x * qpw * K;
        if (lane < K)
            for (uint32_t j = 0; j < nvalid; ++j) {
                bd2[j * K + lane] = inf;
                bix[j * K + lane] = PLAN_NONE;
            }
        bool fin = false;
        uint32_t sp = walk_dense_start(L.dw, __ballot(valid));
        while (!fin) {
            /* the group walk resumes where the full list stopped it */
            nent = 0;
            uint32_t sc = 0, cc = 0;
            const WalkEnd e =
                wgeo ? walk_dense_run<CAP, WALK_STACK_FULL, WALK_LIST_GEO>(tv, L.nc, L.dw, nullptr, L.lst, nent, qx, qy,
                                                                            qz, r, sp, sc, cc)
                     : walk_dense_run<CAP, WALK_STACK_FULL, WALK_LIST>(tv, L.nc, L.dw, nullptr, L.lst, nent, qx, qy,
                                                                       qz, r, sp, sc, cc);
            fin = e != WalkEnd::list_full;
            __syncthreads();
            for (uint32_t j = 0; j < nvalid; ++j) {
                double bd = lane < K ? bd2[j * K + lane] : inf;
                uint32_t bi = lane < K ? bix[j * K + lane] : PLAN_NONE;
                bool first = false;
                select_list<K>(tv, L, nent, (int)j, lane_f64(qx, j), lane_f64(qy, j), lane_f64(qz, j), lane_f64(T, j),
                            bd, bi, first);
                if (lane < K) {
                    bd2[j * K + lane] = bd;
                    bix[j * K + lane] = bi;
                }
            }
            __syncthreads();
        }
        double d_tm = 0, d_num = 0;
        for (uint32_t j = 0; j < nvalid; ++j) {
            const double bd = lane < K ? bd2[j * K + lane] : inf;
            const uint32_t bi = lane < K ? bix[j * K + lane] : PLAN_NONE;
            finish_query<K>(tv, bd, bi, wv * qpw + j, lane_u32(self, j), qpts || qidx, mode, kernel, rho_in,
                            rho_out, densest, idx_out, d2_out, (int)j, d_tm, d_num);
        }
        if (mode == 0 && kernel == 0 && valid) median_density(tv, d_tm, d_num, q, self, qpts || qidx, rho_out);
    }
}

template <int K, bool FUSE>
static void launch_select(const TreeView &tv, Arena &a, uint32_t nb, const uint32_t *qi, const double4 *qp,
                          double *rbuf, uint32_t *cbuf, const double *rt, uint32_t q0, int mode, int kernel,
                          const double *rho_in, double *rho_o, uint32_t *dn_o, uint32_t *idx_o, double *d2_o,
                          KnnOut &out, uint32_t qpw, hipStream_t s)
{
    const uint32_t nw = ceil_div(nb, qpw);
    uint32_t *ovf = a.get_t<uint32_t>("knn_ovf", (size_t)nw + 1);
    uint32_t *opos = a.get_t<uint32_t>("knn_opos", (size_t)nw + 1);
    uint32_t *otot = a.get_t<uint32_t>("knn_otot", 4);
    HIPCHK(hipMemsetAsync(ovf, 0, (size_t)nw * 4, s));
    if (out.ev) HIPCHK(hipEventRecord(out.ev[1], s));
    if (mode == 1)
        hipLaunchKernelGGL((k_knn_select<K, false, FUSE, HopCap<K>::value>), dim3(nw), dim3(64), 0, s, tv, nb, qi, qp,
                           rbuf, cbuf, rt, q0, nullptr, ovf, nullptr, nullptr, mode, kernel, rho_in, rho_o, dn_o, idx_o,
                           d2_o, qpw);
    else
        hipLaunchKernelGGL((k_knn_select<K, false, FUSE>), dim3(nw), dim3(64), 0, s, tv, nb, qi, qp, rbuf, cbuf, rt,
                           q0, nullptr, ovf, nullptr, nullptr, mode, kernel, rho_in, rho_o, dn_o, idx_o, d2_o, qpw);
    LAUNCHCHK();
    if (out.ev) HIPCHK(hipEventRecord(out.ev[2], s));
    scan_exclusive_u32(a, ovf, opos, nw, s, otot);
    uint32_t novf = 0;
    HIPCHK(hipMemcpyAsync(&novf, otot, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(plan::stream_wait(s));
    if (novf) {
        uint32_t *waves = a.get_t<uint32_t>("knn_owaves", novf);
        hipLaunchKernelGGL(k_ovf_compact, dim3(ceil_div(nw, 256)), dim3(256), 0, s, ovf, opos, nw, waves);
        LAUNCHCHK();
        uint32_t rq = qpw, nr = novf;
        if (qpw == 64) {
            /* a wave whose include list overflowed covers 64 neighbouring
             * queries whose spheres span many nodes: redo it as 8 waves of
             * 8 queries (a smaller union of spheres; at 16.8M particles no
             * density wave and 56 densest-neighbour queries of 1.69M still
             * overflow, against 144 and 336 with 4 waves of 16); only the
             * waves that still overflow take the resumable path */
            const uint32_t sub = 8, nw2 = nw * sub;
            rq = 64 / sub;
            uint32_t *ovf2 = a.get_t<uint32_t>("knn_ovf2", (size_t)nw2 + 1);
            uint32_t *opos2 = a.get_t<uint32_t>("knn_opos2", (size_t)nw2 + 1);
            HIPCHK(hipMemsetAsync(ovf2, 0, (size_t)nw2 * 4, s));
            hipLaunchKernelGGL((k_knn_select<K, false, FUSE>), dim3(novf * sub), dim3(64), 0, s, tv, nb, qi, qp, rbuf,
                               cbuf, rt, q0, waves, ovf2, nullptr, nullptr, mode, kernel, rho_in, rho_o, dn_o, idx_o,
                               d2_o, rq);
            LAUNCHCHK();
            scan_exclusive_u32(a, ovf2, opos2, nw2, s, otot);
            HIPCHK(hipMemcpyAsync(&nr, otot, 4, hipMemcpyDeviceToHost, s));
            HIPCHK(plan::stream_wait(s));
            if (nr) {
                waves = a.get_t<uint32_t>("knn_owaves2", nr);
                hipLaunchKernelGGL(k_ovf_compact, dim3(ceil_div(nw2, 256)), dim3(256), 0, s, ovf2, opos2, nw2, waves);
                LAUNCHCHK();
            }
        }
        if (nr && rq > 1) {
            /* a wave that still overflows after the split covers queries deep
             * in a dense core: each of its queries walks in a wave of its own,
             * so the few left run side by side instead of in one long wave */
            uint32_t *w1 = a.get_t<uint32_t>("knn_owaves1", (size_t)nr * rq);
            hipLaunchKernelGGL(k_wave_expand, dim3(ceil_div(nr * rq, 256)), dim3(256), 0, s, waves, nr, rq, w1);
            LAUNCHCHK();
            waves = w1;
            nr *= rq;
            rq = 1;
        }
        if (nr) {
            double *sd2 = a.get_t<double>("knn_osd2", (size_t)nr * rq * K);
            uint32_t *sidx = a.get_t<uint32_t>("knn_osidx", (size_t)nr * rq * K);
            hipLaunchKernelGGL((k_knn_select<K, true, FUSE>), dim3(nr), dim3(64), 0, s, tv, nb, qi, qp, rbuf, cbuf, rt,
                               q0, waves, ovf, sd2, sidx, mode, kernel, rho_in, rho_o, dn_o, idx_o, d2_o, rq);
            LAUNCHCHK();
        }
        if (out.nres) *out.nres += nr;
    }
    if (out.novf) *out.novf += novf;
}

/* Particle queries with the tree's start radius run the radius schedule inside
 * the select kernel; point queries (and particle queries with a caller's
 * start radius) run k_point_tight / k_knn_radius first.  Particle queries
 * are consecutive in Morton order, so 64 of them share one wave's walk;
 * point queries (saddle points between clumps, callers' points) are
 * scattered, so each gets a wave of its own. */
/* densities whose x^1.5 the C library's pow rounds the other way: the
 * median kernel's quotient again with that pow (slot: entry of the
 * re-check list) */
__global__ void k_pow_fix(const uint32_t *__restrict__ slot, const double *__restrict__ p, uint32_t n,
                          const uint32_t *__restrict__ pw_idx, const double *__restrict__ pw_num,
                          double *__restrict__ rho)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t e = slot[i];
    const double four_pi_over_three = (3.14159265358979323846 * 4.) / 3.;
    rho[pw_idx[e]] = __ddiv_rn(pw_num[e], __dmul_rn(four_pi_over_three, p[i]));
}

/* msum[k] = ((0 + m) + m) + ... with k terms (k <= 64): the median kernel's
 * mass sum when its i* + 1 neighbours all have mass m, in its order */
__global__ void k_mass_sums(double m, double *__restrict__ msum)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    msum[0] = s;
    for (int k = 1; k <= 64; ++k) {
        s = __dadd_rn(s, m);
        msum[k] = s;
    }
}

template <int K>
static void launch_knn(const DevTree &t, const TreeView &tv, uint32_t nq, const uint32_t *qidx, const double4 *qpts,
                       const double *est, int mode, int kernel, const double *rho_in, KnnOut out, hipStream_t s,
                       Arena *scratch)
{
    Arena &a = scratch ? *scratch : const_cast<DevTree &>(t).arena;
    double *rbuf = a.get_t<double>("knn_r", (size_t)nq);
    uint32_t *cbuf = a.get_t<uint32_t>("knn_c", (size_t)nq);
    const bool fuse = !qpts && !est;
    const uint32_t q0 = 0u;
    const uint32_t qpw = qpts ? 1u : 64u;
    if (out.ev) HIPCHK(hipEventRecord(out.ev[0], s));
    if (fuse) {
        launch_select<K, true>(tv, a, nq, qidx, nullptr, rbuf, cbuf, nullptr, q0, mode, kernel, rho_in, out.rho,
                               out.densest, out.idx, out.d2, out, qpw, s);
    } else {
        double *rt = nullptr;
        if (qpts) {
            rt = a.get_t<double>("knn_rt", (size_t)nq);
            hipLaunchKernelGGL(k_point_tight, dim3(ceil_div(nq, qpw)), dim3(64), 0, s, tv, nq, qpts, est, rt, qpw);
            LAUNCHCHK();
        }
        hipLaunchKernelGGL(k_knn_radius, dim3(ceil_div(nq, qpw)), dim3(64), 0, s, tv, nq, qidx, qpts, est, rt,
                           rbuf, cbuf, qpw);
        LAUNCHCHK();
        launch_select<K, false>(tv, a, nq, qidx, qpts, rbuf, cbuf, rt, q0, mode, kernel, rho_in, out.rho,
                                out.densest, out.idx, out.d2, out, qpw, s);
    }
    if (out.rstop) HIPCHK(hipMemcpyAsync(out.rstop, rbuf, (size_t)nq * 8, hipMemcpyDeviceToDevice, s));
}

void dev_knn(const DevTree &t, uint32_t K, int mode, int kernel, uint32_t nq, const uint32_t *qidx,
             const double4 *qpts, const double *est, const double *rho_in, const plan_derived &d, KnnOut out,
             hipStream_t s, Arena *scratch)
{
    if (nq == 0) return;
    if (!knn_k_supported(K)) throw std::runtime_error("KNN: k must be in 1..64");
    if (K > t.n) throw std::runtime_error("KNN: k exceeds the number of tree particles");
    if (qpts && !est) throw std::runtime_error("KNN: point queries need a start radius");
    TreeView tv = tree_view(t, d);
    tv.kq = K;
    if (mode == 0 && kernel == 0) {
        /* sums of equal masses for the median kernel (species 0's mass) */
        Arena &a = scratch ? *scratch : const_cast<DevTree &>(t).arena;
        double *ms = a.get_t<double>("knn_msum", 65);
        hipLaunchKernelGGL(k_mass_sums, dim3(1), dim3(64), 0, s, d.mass_per_particle[0], ms);
        LAUNCHCHK();
        tv.msum = ms;
        tv.m0 = d.mass_per_particle[0];
        if (out.rho) {
            /* densities whose x^1.5 lies near a rounding midpoint, recomputed
             * below with the C library's pow (each query finishes once) */
            tv.pw_count = a.get_t<uint32_t>("knn_pw_count", 1);
            tv.pw_cap = nq;
            tv.pw_idx = a.get_t<uint32_t>("knn_pw_idx", nq);
            tv.pw_x = a.get_t<double>("knn_pw_x", nq);
            tv.pw_num = a.get_t<double>("knn_pw_num", nq);
            HIPCHK(hipMemsetAsync(tv.pw_count, 0, 4, s));
        }
    }
    if (K <= 8) launch_knn<8>(t, tv, nq, qidx, qpts, est, mode, kernel, rho_in, out, s, scratch);
    else if (K <= 16) launch_knn<16>(t, tv, nq, qidx, qpts, est, mode, kernel, rho_in, out, s, scratch);
    else if (K <= 32) launch_knn<32>(t, tv, nq, qidx, qpts, est, mode, kernel, rho_in, out, s, scratch);
    else launch_knn<64>(t, tv, nq, qidx, qpts, est, mode, kernel, rho_in, out, s, scratch);
    if (tv.pw_count) {
        /* the reference's pow(x, 1.5) (src/tree.hpp:3489) from the C library
         * for the listed x; the few that differ from the correctly rounded
         * value the kernel used are sent back and their densities redone */
        uint32_t cnt = 0;
        HIPCHK(hipMemcpyAsync(&cnt, tv.pw_count, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(plan::stream_wait(s));
        if (cnt > tv.pw_cap) throw std::runtime_error("KNN: density re-check list overflow");
        if (cnt) {
            const double *x = (const double *)const_cast<DevTree &>(t).host_stage.get((size_t)cnt * 8);
            HIPCHK(hipMemcpyAsync((void *)x, tv.pw_x, (size_t)cnt * 8, hipMemcpyDeviceToHost, s));
            HIPCHK(plan::stream_wait(s));
            std::vector<uint32_t> bad_slot;
            std::vector<double> bad_p;
#pragma omp parallel
            {
                std::vector<uint32_t> ls;
                std::vector<double> lp;
#pragma omp for schedule(static) nowait
                for (long i = 0; i < (long)cnt; ++i) {
                    double rem;
                    const double g = std::pow(x[i], 1.5);
                    if (g != pow15_rn(x[i], rem)) {
                        ls.push_back((uint32_t)i);
                        lp.push_back(g);
                    }
                }
#pragma omp critical
                {
                    bad_slot.insert(bad_slot.end(), ls.begin(), ls.end());
                    bad_p.insert(bad_p.end(), lp.begin(), lp.end());
                }
            }
            const uint32_t nb = (uint32_t)bad_slot.size();
            if (nb) {
                Arena &a = scratch ? *scratch : const_cast<DevTree &>(t).arena;
                uint32_t *ds = a.get_t<uint32_t>("knn_pw_bad", nb);
                double *dp = a.get_t<double>("knn_pw_p", nb);
                HIPCHK(hipMemcpyAsync(ds, bad_slot.data(), (size_t)nb * 4, hipMemcpyHostToDevice, s));
                HIPCHK(hipMemcpyAsync(dp, bad_p.data(), (size_t)nb * 8, hipMemcpyHostToDevice, s));
                hipLaunchKernelGGL(k_pow_fix, dim3(ceil_div(nb, 256)), dim3(256), 0, s, ds, dp, nb, tv.pw_idx,
                                   tv.pw_num, out.rho);
                LAUNCHCHK();
                HIPCHK(plan::stream_wait(s));
            }
            if (out.npow) *out.npow += nb;
        }
        if (out.nchk) *out.nchk += cnt;
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

/* ---------------------------------------------- level-parallel ball search
 *
 * The reference's BallSearch (src/tree.hpp:3275-3304) visits a node when every
 * ancestor intersected the sphere without containing it; a visited node that
 * the sphere contains contributes all its particles, a visited intersecting
 * leaf those within r.  The visited set does not depend on the visiting order,
 * so it is expanded level by level for all queries at once (one thread per
 * (query, node)), the contributing ranges are sorted by (query, first
 * particle) — the reference's traversal order — and the candidates of all
 * ranges are tested and compacted in parallel.  A large ball thus spreads over
 * the whole GPU instead of one wave. */

struct BallTask {
    uint32_t q, begin, end, leaf;
};

__global__ void k_bfs_seed(uint32_t nq, uint2 *__restrict__ f, uint32_t *__restrict__ counts)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nq) f[q] = make_uint2(q, 0u);
    if (q == 0) counts[0] = nq;
}

__global__ void k_bfs_expand(TreeView tv, const double4 *__restrict__ cr, const uint2 *__restrict__ f,
                             const uint32_t *__restrict__ fcount, uint2 *__restrict__ g, uint32_t *__restrict__ gcount,
                             uint32_t gcap, BallTask *__restrict__ tasks, uint32_t *__restrict__ tcount, uint32_t tcap,
                             uint32_t *__restrict__ overflow)
{
    /* after an overflow the frontier holds fewer entries than its count says
     * (and possibly none valid): stop, the search is redone with capacities
     * from the counts */
    if (*overflow) return;
    const uint32_t n = min(*fcount, gcap);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint2 it = f[k];
        const double4 c = cr[it.x];
        const NodeGeo geo = tv.ngeo[it.y];
        const NodeTop top = tv.ntop[it.y];
        const uint32_t nd = tv.ninfo[it.y] & 0xffu;
        int act;
        if (sphere_contain(geo, c.x, c.y, c.z, c.w, tv.to_diag)) act = 1;
        else if (sphere_intersect(geo, c.x, c.y, c.z, c.w, tv.to_diag)) act = nd == 0 ? 4 : 2;
        else act = 3;
        if (act == 1 || act == 4) {
            const uint32_t slot = atomicAdd(tcount, 1u);
            if (slot < tcap) tasks[slot] = BallTask{it.x, top.begin, top.end, act == 4 ? 1u : 0u};
            else *overflow = 1u;
        } else if (act == 2) {
            const uint32_t slot = atomicAdd(gcount, nd);
            if (slot + nd <= gcap)
                for (uint32_t d = 0; d < nd; ++d) g[slot + d] = make_uint2(it.x, top.first_daughter + d);
            else *overflow = 1u;
        }
    }
}

__global__ void k_bfs_task_keys(const BallTask *__restrict__ tasks, uint32_t n, uint4 *__restrict__ keys)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) keys[k] = make_uint4(tasks[k].begin, tasks[k].q, 0u, k);
}

/* One-workgroup bitonic sort of up to SMALL_SORT_N records by the 64-bit key
 * (y:x), for the short task lists of ball searches, where the eight passes
 * of the radix sort cost mostly launch latency.  The keys (query, begin) of
 * a task list are distinct, so the order is the radix sort's.  p is a power
 * of two >= n. */
constexpr uint32_t SMALL_SORT_N = 4096, SMALL_SORT_T = 1024;
__global__ __launch_bounds__(SMALL_SORT_T) void k_small_sort64(uint4 *__restrict__ keys, uint32_t n, uint32_t p)
{
    __shared__ uint64_t sk[SMALL_SORT_N];
    __shared__ uint32_t sv[SMALL_SORT_N];
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < p; i += SMALL_SORT_T) {
        if (i < n) {
            const uint4 v = keys[i];
            sk[i] = ((uint64_t)v.y << 32) | v.x;
            sv[i] = v.w;
        } else {
            sk[i] = ~0ull;
            sv[i] = PLAN_NONE;
        }
    }
    __syncthreads();
    for (uint32_t k = 2; k <= p; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = t; i < p; i += SMALL_SORT_T) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const uint64_t a = sk[i], b = sk[l];
                    if ((a > b) == ((i & k) == 0)) {
                        sk[i] = b;
                        sk[l] = a;
                        const uint32_t x = sv[i];
                        sv[i] = sv[l];
                        sv[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    for (uint32_t i = t; i < n; i += SMALL_SORT_T)
        keys[i] = make_uint4((uint32_t)sk[i], (uint32_t)(sk[i] >> 32), 0u, sv[i]);
}

__global__ void k_bfs_task_sizes(const BallTask *__restrict__ tasks, const uint4 *__restrict__ keys, uint32_t n,
                                 BallTask *__restrict__ sorted, uint32_t *__restrict__ sizes)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const BallTask t = tasks[keys[k].w];
    sorted[k] = t;
    sizes[k] = t.end - t.begin;
}

/* one thread per candidate: its task by binary search over the size prefix */
__global__ void k_bfs_candidates(TreeView tv, const double4 *__restrict__ cr, const BallTask *__restrict__ sorted,
                                 const uint32_t *__restrict__ pre, uint32_t ntask, uint32_t ncand,
                                 const double *__restrict__ rho, double rho_min, int filter,
                                 uint32_t *__restrict__ flag, uint32_t *__restrict__ pidx)
{
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < ncand; k += gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = ntask;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (pre[mid] <= k) lo = mid;
            else hi = mid;
        }
        const BallTask t = sorted[lo];
        const uint32_t p = t.begin + (k - pre[lo]);
        bool take = true;
        if (t.leaf) {
            const double4 c = cr[t.q];
            const double4 x = tv.pm[p];
            take = dist2(c.x, c.y, c.z, x.x, x.y, x.z) <= __dmul_rn(c.w, c.w);
        }
        if (take && filter) take = rho[p] >= rho_min;
        flag[k] = take ? 1u : 0u;
        pidx[k] = p;
    }
}

__global__ void k_bfs_compact(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pos,
                              const uint32_t *__restrict__ pidx, uint32_t ncand, uint32_t *__restrict__ out)
{
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < ncand; k += gridDim.x * blockDim.x)
        if (flag[k]) out[pos[k]] = pidx[k];
}

/* per task: (query, output position of its first candidate) */
__global__ void k_bfs_task_out(const BallTask *__restrict__ sorted, const uint32_t *__restrict__ pre,
                               const uint32_t *__restrict__ opos, uint32_t ntask, uint2 *__restrict__ tq)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ntask) tq[k] = make_uint2(sorted[k].q, opos[pre[k]]);
}

void dev_ball_lists(const DevTree &t, Arena &a, const std::vector<double4> &queries, const double *rho,
                    double rho_min, int filter, std::vector<std::vector<uint32_t>> &out, hipStream_t s)
{
    const uint32_t nq = (uint32_t)queries.size();
    out.assign(nq, {});
    if (nq == 0 || t.n == 0) return;
    plan_derived d{};
    d.cell_length[0] = d.cell_length[1] = d.cell_length[2] = 1;
    d.cell_volume = 1;
    const TreeView tv = tree_view(t, d);
    double4 *cr = a.get_t<double4>("bl_q", nq);
    HIPCHK(hipMemcpyAsync(cr, queries.data(), nq * sizeof(double4), hipMemcpyHostToDevice, s));
    const int levels = t.depth + 1;
    uint32_t gcap = std::max<uint32_t>(1u << 20, 8 * nq), tcap = std::max<uint32_t>(1u << 20, 8 * nq);
    if (t.list_limit != 0xffffffffu) /* test hook: start small, exercise the regrowth */
        gcap = tcap = std::max<uint32_t>(64u, t.list_limit * 8u);
    for (int attempt = 0; attempt < 8; ++attempt) {
        uint2 *f0 = a.get_t<uint2>("bl_f0", gcap);
        uint2 *f1 = a.get_t<uint2>("bl_f1", gcap);
        BallTask *tasks = a.get_t<BallTask>("bl_tasks", tcap);
        uint32_t *cnt = a.get_t<uint32_t>("bl_cnt", 40);
        HIPCHK(hipMemsetAsync(cnt, 0, 40 * 4, s));
        hipLaunchKernelGGL(k_bfs_seed, dim3(ceil_div(nq, 256)), dim3(256), 0, s, nq, f0, cnt);
        LAUNCHCHK();
        /* cnt[0..33]: frontier sizes per level, cnt[36] tasks, cnt[37] overflow */
        for (int L = 0; L < levels && L < 34; ++L) {
            uint2 *f = (L & 1) ? f1 : f0;
            uint2 *g = (L & 1) ? f0 : f1;
            hipLaunchKernelGGL(k_bfs_expand, dim3(1024), dim3(256), 0, s, tv, cr, f, cnt + L, g, cnt + L + 1, gcap,
                               tasks, cnt + 36, tcap, cnt + 37);
            LAUNCHCHK();
        }
        uint32_t h[40];
        HIPCHK(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
        HIPCHK(plan::stream_wait(s));
        if (h[37]) {
            /* the counts up to the overflowing level are exact requests;
             * grow at least 4x so a later level that was not reached fits
             * on the next attempt more likely */
            uint32_t need = 0;
            for (int L = 0; L < 34; ++L) need = std::max(need, h[L]);
            const uint64_t g2 = std::max<uint64_t>((uint64_t)gcap * 4, (uint64_t)need + need / 4);
            const uint64_t t2 = std::max<uint64_t>((uint64_t)tcap * 4, (uint64_t)h[36] + h[36] / 4);
            if (g2 > 0xffffffffull / 2 || t2 > 0xffffffffull / 2)
                throw std::runtime_error("ball search: node frontier exceeds the work capacity");
            gcap = (uint32_t)g2;
            tcap = (uint32_t)t2;
            continue;
        }
        const uint32_t nt = h[36];
        if (nt == 0) return;
        uint4 *k0 = a.get_t<uint4>("bl_k0", nt), *k1 = a.get_t<uint4>("bl_k1", nt);
        hipLaunchKernelGGL(k_bfs_task_keys, dim3(ceil_div(nt, 256)), dim3(256), 0, s, tasks, nt, k0);
        LAUNCHCHK();
        if (nt <= SMALL_SORT_N) {
            uint32_t p = 2;
            while (p < nt) p <<= 1;
            hipLaunchKernelGGL(k_small_sort64, dim3(1), dim3(SMALL_SORT_T), 0, s, k0, nt, p);
            LAUNCHCHK();
        } else {
            radix_sort_96(a, &k0, &k1, nt, 2, s);
        }
        BallTask *sorted = a.get_t<BallTask>("bl_sorted", nt);
        uint32_t *sizes = a.get_t<uint32_t>("bl_sizes", nt + 1);
        uint32_t *pre = a.get_t<uint32_t>("bl_pre", nt + 1);
        hipLaunchKernelGGL(k_bfs_task_sizes, dim3(ceil_div(nt, 256)), dim3(256), 0, s, tasks, k0, nt, sorted, sizes);
        LAUNCHCHK();
        scan_exclusive_u32(a, sizes, pre, nt, s, pre + nt);
        uint32_t ncand = 0;
        HIPCHK(hipMemcpyAsync(&ncand, pre + nt, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(plan::stream_wait(s));
        uint32_t *flag = a.get_t<uint32_t>("bl_flag", (size_t)ncand + 1);
        uint32_t *pidx = a.get_t<uint32_t>("bl_pidx", (size_t)ncand + 1);
        uint32_t *opos = a.get_t<uint32_t>("bl_opos", (size_t)ncand + 1);
        const unsigned gb = std::min<unsigned>(4096, ceil_div(std::max<uint32_t>(ncand, 1), 256));
        hipLaunchKernelGGL(k_bfs_candidates, dim3(gb), dim3(256), 0, s, tv, cr, sorted, pre, nt, ncand, rho, rho_min,
                           filter, flag, pidx);
        LAUNCHCHK();
        scan_exclusive_u32(a, flag, opos, ncand, s, opos + ncand);
        uint32_t total = 0;
        HIPCHK(hipMemcpyAsync(&total, opos + ncand, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(plan::stream_wait(s));
        uint32_t *res = a.get_t<uint32_t>("bl_out", (size_t)total + 1);
        hipLaunchKernelGGL(k_bfs_compact, dim3(gb), dim3(256), 0, s, flag, opos, pidx, ncand, res);
        LAUNCHCHK();
        uint2 *tq = a.get_t<uint2>("bl_tq", nt);
        hipLaunchKernelGGL(k_bfs_task_out, dim3(ceil_div(nt, 256)), dim3(256), 0, s, sorted, pre, opos, nt, tq);
        LAUNCHCHK();
        std::vector<uint2> htq(nt);
        std::vector<uint32_t> hres(total);
        HIPCHK(hipMemcpyAsync(htq.data(), tq, (size_t)nt * sizeof(uint2), hipMemcpyDeviceToHost, s));
        if (total) HIPCHK(hipMemcpyAsync(hres.data(), res, (size_t)total * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(plan::stream_wait(s));
        /* tasks are grouped by query: a query's results run from its first
         * task's output position to the next query's */
        for (uint32_t k = 0; k < nt;) {
            uint32_t j = k;
            while (j < nt && htq[j].x == htq[k].x) ++j;
            const uint32_t b = htq[k].y, e = j < nt ? htq[j].y : total;
            out[htq[k].x].assign(hres.begin() + b, hres.begin() + e);
            k = j;
        }
        return;
    }
    throw std::runtime_error("ball search: node frontier exceeds the work capacity");
}

void dev_ball(const DevTree &t, uint32_t nq, const double4 *centre_r, uint32_t *counts, const uint32_t *offsets,
              uint32_t *out, const double *rho, double rho_min, int filter, hipStream_t s)
{
    if (nq == 0) return;
    plan_derived d{};
    d.cell_length[0] = d.cell_length[1] = d.cell_length[2] = 1;
    d.cell_volume = 1;
    TreeView tv = tree_view(t, d);
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

