// boundness.hip — per-clump gravitational boundness with a pairwise
// potential, in two forms:
//
//   FP64 per clump: every clump on its own, one launch sequence per clump and
//     iteration, binary64 pair arithmetic;
//   FP32 batched: all clumps in each launch, the pair sums in binary32 with
//     compensated (Kahan) accumulation on positions relative to the clump's
//     starting centre of mass, binary64 for centre-of-mass and energy totals,
//     binary64 re-evaluation of every particle whose energy is within 1e-5 of
//     zero (relative to its terms), and of every clump whose total energy ends
//     within 1e-3 of zero (relative to its potential energy).  Iterations are
//     captured in a HIP graph and replayed; the pair kernel is persistent and
//     its tile shape is picked by timing the candidates on the batch.
//
// The reference decides boundness by a two-body test of each particle against
// the rest of the clump (src/tree.hpp:4615-4674) and has no pairwise potential
// at all; this is an extension of the path (the boundness clumps/s metric
// compares its FP32 and FP64 forms), with its own FP64
// oracle (oracle/plan_oracle.cpp, orc_boundness).  Definition, per clump with
// bound set B (initially every member):
//   M = sum_B m, X = sum_B m x / M, V = sum_B m v / M            (binary64)
//   phi_i = -G sum_{j in B, x_j != x_i} m_j / |x_i - x_j|
//   dv_i = v_i - V + s (X_x - x_i,x),  s = (0, q Omega, 0)  (the reference's
//          shear-corrected relative velocity, src/tree.hpp:4656-4659)
//   e_i = |dv_i|^2 / 2 + phi_i;  particles with e_i > 0 leave B;
//   repeat until none leaves.
//   E_kin = sum_B m |dv|^2 / 2,  E_pot = sum_B m phi / 2 of the final B.
// After a removal the potentials are updated by the removed particles'
// contributions when fewer than half remain to be subtracted, else
// recomputed.
#include "boundness.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace plan {

static constexpr int RED = 256;  /* threads of the per-clump reduction blocks */

enum : uint32_t { B_OUT = 0, B_IN = 1, B_LEFT = 2 };   /* member flags */
enum : uint32_t { C_ACTIVE = 0, C_DONE = 1 };         /* clump state */

/* ------------------------------------------------------------ set-up */

/* starting centre of mass per clump (reference point of the relative
 * binary32 positions) */
__global__ __launch_bounds__(RED) void k_bnd_ref(const double4 *__restrict__ pos, const uint32_t *__restrict__ off,
                                                 uint32_t c0, double4 *__restrict__ ref)
{
    __shared__ double sm[4][RED];
    const uint32_t c = c0 + blockIdx.x;
    const uint32_t b = off[c], e = off[c + 1];
    double M = 0, x = 0, y = 0, z = 0;
    for (uint32_t i = b + threadIdx.x; i < e; i += RED) {
        const double4 p = pos[i];
        M += p.w;
        x += p.w * p.x;
        y += p.w * p.y;
        z += p.w * p.z;
    }
    sm[0][threadIdx.x] = M;
    sm[1][threadIdx.x] = x;
    sm[2][threadIdx.x] = y;
    sm[3][threadIdx.x] = z;
    __syncthreads();
    for (int s = RED / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 4; ++k) sm[k][threadIdx.x] += sm[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = sm[0][0];
        ref[c] = m > 0 ? make_double4(sm[1][0] / m, sm[2][0] / m, sm[3][0] / m, 0.0) : make_double4(0, 0, 0, 0);
    }
}

__global__ void k_bnd_init(const double4 *__restrict__ pos, const uint32_t *__restrict__ cid, uint32_t m0,
                           uint32_t m1, const double4 *__restrict__ ref, float4 *__restrict__ posf,
                           uint32_t *__restrict__ flag, double *__restrict__ phi)
{
    const uint32_t i = m0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m1) return;
    const double4 p = pos[i];
    const double4 r = ref[cid[i]];
    posf[i] = make_float4((float)(p.x - r.x), (float)(p.y - r.y), (float)(p.z - r.z), (float)p.w);
    flag[i] = B_IN;
    phi[i] = 0.0;
}

__global__ void k_bnd_reset_clumps(uint32_t c0, uint32_t c1, uint32_t *__restrict__ state,
                                   uint32_t *__restrict__ iters, uint32_t *__restrict__ nearz)
{
    const uint32_t c = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= c1) return;
    state[c] = C_ACTIVE;
    iters[c] = 0;
    nearz[c] = 0;
}

/* ------------------------------------------------------------ lists */

__global__ void k_bnd_marks(const uint32_t *__restrict__ flag, uint32_t m0, uint32_t m1, uint32_t *__restrict__ inb,
                            uint32_t *__restrict__ inr)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m1 - m0) return;
    const uint32_t f = flag[m0 + k];
    inb[k] = f == B_IN;
    inr[k] = f == B_LEFT;
}

/* compacted bound and just-removed lists (member ids and positions, in
 * member order); particles removed last iteration become B_OUT */
template <typename R>
__global__ void k_bnd_compact(const double4 *__restrict__ pos, const float4 *__restrict__ posf, uint32_t m0,
                              uint32_t m1, uint32_t *__restrict__ flag, const uint32_t *__restrict__ sb,
                              const uint32_t *__restrict__ sr, uint32_t *__restrict__ blist,
                              uint32_t *__restrict__ rlist, void *__restrict__ pj, void *__restrict__ pr)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m1 - m0) return;
    const uint32_t i = m0 + k;
    const uint32_t f = flag[i];
    if (f == B_IN) {
        blist[sb[k]] = i;
        if (sizeof(R) == 8) ((double4 *)pj)[sb[k]] = pos[i];
        else ((float4 *)pj)[sb[k]] = posf[i];
    } else if (f == B_LEFT) {
        rlist[sr[k]] = i;
        if (sizeof(R) == 8) ((double4 *)pr)[sr[k]] = pos[i];
        else ((float4 *)pr)[sr[k]] = posf[i];
        flag[i] = B_OUT;
    }
}

/* per clump: segment of the lists, pair-kernel mode and tile count */
__global__ void k_bnd_segments(const uint32_t *__restrict__ off, uint32_t c0, uint32_t c1, uint32_t m0,
                               const uint32_t *__restrict__ sb, const uint32_t *__restrict__ sr,
                               const uint32_t *__restrict__ state, const uint32_t *__restrict__ iters, uint32_t ti,
                               uint4 *__restrict__ seg, uint32_t *__restrict__ mode, uint32_t *__restrict__ tiles)
{
    const uint32_t c = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= c1) return;
    const uint32_t b0 = sb[off[c] - m0], b1 = sb[off[c + 1] - m0];
    const uint32_t r0 = sr[off[c] - m0], r1 = sr[off[c + 1] - m0];
    seg[c] = make_uint4(b0, b1 - b0, r0, r1 - r0);
    const uint32_t nb = b1 - b0, nr = r1 - r0;
    /* 1 = recompute from B, 2 = subtract the removed particles' terms */
    const uint32_t md = (state[c] == C_DONE || nb == 0) ? 0u : ((iters[c] == 0 || 2 * nr > nb) ? 1u : (nr ? 2u : 0u));
    mode[c] = md;
    tiles[c - c0] = md ? (nb + ti - 1) / ti : 0u;
}

/* ------------------------------------------------------------ totals */

/* M, X, V of the bound set (binary64, fixed reduction order) */
__global__ __launch_bounds__(RED) void k_bnd_totals(const double4 *__restrict__ pos, const double4 *__restrict__ vel,
                                                    const uint32_t *__restrict__ blist, const uint4 *__restrict__ seg,
                                                    const uint32_t *__restrict__ state, uint32_t c0,
                                                    double *__restrict__ tot)
{
    __shared__ double sm[7][RED];
    const uint32_t c = c0 + blockIdx.x;
    if (state[c] == C_DONE) return;
    const uint4 sg = seg[c];
    double a[7] = {0, 0, 0, 0, 0, 0, 0};
    for (uint32_t k = threadIdx.x; k < sg.y; k += RED) {
        const uint32_t i = blist[sg.x + k];
        const double4 p = pos[i], v = vel[i];
        a[0] += p.w;
        a[1] += p.w * p.x;
        a[2] += p.w * p.y;
        a[3] += p.w * p.z;
        a[4] += p.w * v.x;
        a[5] += p.w * v.y;
        a[6] += p.w * v.z;
    }
    for (int k = 0; k < 7; ++k) sm[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int s = RED / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 7; ++k) sm[k][threadIdx.x] += sm[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double M = sm[0][0];
        double *t = tot + (size_t)c * 8;
        t[0] = M;
        for (int k = 0; k < 3; ++k) {
            t[1 + k] = M > 0 ? sm[1 + k][0] / M : 0.0;
            t[4 + k] = M > 0 ? sm[4 + k][0] / M : 0.0;
        }
    }
}

/* ------------------------------------------------------------ pair sums */

struct PairArgs {
    const uint32_t *off;
    uint32_t c0, c1;
    const uint4 *seg;
    const uint32_t *mode;
    const uint32_t *tile_off;   /* exclusive scan of tiles over [c0, c1), total at [c1 - c0] */
    const uint32_t *blist;
    const void *pj;             /* bound positions (compacted) */
    const void *pr;             /* removed positions (compacted) */
    double *phi;
    double G;
    uint32_t *counter;
    unsigned long long *pairs;
};

template <typename R> struct V4;
template <> struct V4<float> { typedef float4 T; };
template <> struct V4<double> { typedef double4 T; };

/* Persistent: each block takes (clump, tile of TI bound members) items from a
 * counter; each thread owns RPT rows and sums over the clump's bound (mode 1)
 * or just-removed (mode 2) positions staged in LDS TJ at a time (padded to a
 * multiple of 8 with massless far-away points).  binary32: 8 terms are summed
 * plainly, then added to the row's Kahan pair (sum, compensation). */
template <typename R, int TI, int RPT, int TJ>
__global__ __launch_bounds__(TI / RPT) void k_bnd_pairs(PairArgs a)
{
    typedef typename V4<R>::T P;
    __shared__ P tile[TJ];
    __shared__ uint32_t item;
    const int NT = TI / RPT;
    const uint32_t total = a.tile_off[a.c1 - a.c0];
    P pad;
    pad.x = pad.y = pad.z = (R)1e19;
    pad.w = 0;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) item = atomicAdd(a.counter, 1u);
        __syncthreads();
        const uint32_t it = item;
        if (it >= total) break;
        /* clump of this item */
        uint32_t lo = 0, hi = a.c1 - a.c0;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (a.tile_off[mid] <= it) lo = mid;
            else hi = mid;
        }
        const uint32_t c = a.c0 + lo;
        const uint4 sg = a.seg[c];
        const uint32_t md = a.mode[c];
        const uint32_t row0 = (it - a.tile_off[lo]) * TI;
        const P *jl = md == 1 ? (const P *)a.pj + sg.x : (const P *)a.pr + sg.z;
        const uint32_t nj = md == 1 ? sg.y : sg.w;
        P me[RPT];
        R s[RPT], cmp[RPT];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const uint32_t row = row0 + threadIdx.x + r * NT;
            me[r] = row < sg.y ? ((const P *)a.pj)[sg.x + row] : ((const P *)a.pj)[sg.x];
            s[r] = 0;
            cmp[r] = 0;
        }
        for (uint32_t j0 = 0; j0 < nj; j0 += TJ) {
            const uint32_t cnt = min((uint32_t)TJ, nj - j0);
            const uint32_t cnt8 = (cnt + 7u) & ~7u;
            __syncthreads();
            for (uint32_t k = threadIdx.x; k < cnt8; k += NT) tile[k] = k < cnt ? jl[j0 + k] : pad;
            __syncthreads();
            for (uint32_t k = 0; k < cnt8; k += 8) {
                R acc[RPT];
#pragma unroll
                for (int r = 0; r < RPT; ++r) acc[r] = 0;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const P q = tile[k + u];
#pragma unroll
                    for (int r = 0; r < RPT; ++r) {
                        const R dx = q.x - me[r].x, dy = q.y - me[r].y, dz = q.z - me[r].z;
                        if (sizeof(R) == 4) {
                            const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                            const float inv = r2 > 0.f ? __builtin_amdgcn_rsqf(r2) : 0.f;
                            acc[r] = __builtin_fmaf(q.w, inv, acc[r]);
                        } else {
                            const double r2 = (dx * dx + dy * dy) + dz * dz;
                            if (r2 > 0) acc[r] += q.w / sqrt(r2);
                        }
                    }
                }
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
                    if (sizeof(R) == 4) {
                        const R y = acc[r] - cmp[r];
                        const R t = s[r] + y;
                        cmp[r] = (t - s[r]) - y;
                        s[r] = t;
                    } else {
                        s[r] += acc[r];
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const uint32_t row = row0 + threadIdx.x + r * NT;
            if (row >= sg.y) continue;
            const uint32_t i = a.blist[sg.x + row];
            const double sum = (double)s[r] - (double)cmp[r];
            if (md == 1) a.phi[i] = -a.G * sum;
            else a.phi[i] += a.G * sum;
        }
        if (threadIdx.x == 0) {
            const uint32_t rows = min((uint32_t)TI, sg.y - row0);
            atomicAdd(a.pairs, (unsigned long long)rows * nj);
        }
    }
}

/* ------------------------------------------------------------ energies */

struct EnergyArgs {
    const double4 *pos, *vel;
    const uint32_t *blist;
    const uint4 *seg;
    const uint32_t *state;
    const double *tot;
    double *phi;
    uint32_t *flag;
    double qO;   /* q Omega */
    uint32_t c0;
    double near_rel;            /* re-check |e| <= near_rel (k + |phi|) in binary64 (0: off) */
    uint32_t *near_list, *near_count;
    double *eout;               /* per clump {E_kin, E_pot, removed, nb} */
};

__device__ __forceinline__ double spec_kin(const double4 &p, const double4 &v, const double *t, double qO)
{
    const double dvx = v.x - t[4];
    const double dvy = (v.y - t[5]) + qO * (t[1] - p.x);
    const double dvz = v.z - t[6];
    return 0.5 * ((dvx * dvx + dvy * dvy) + dvz * dvz);
}

/* members whose energy is too close to zero for the binary32 sums */
__global__ __launch_bounds__(RED) void k_bnd_near(EnergyArgs a)
{
    const uint32_t c = a.c0 + blockIdx.x;
    if (a.state[c] == C_DONE) return;
    const uint4 sg = a.seg[c];
    const double *t = a.tot + (size_t)c * 8;
    for (uint32_t k = threadIdx.x; k < sg.y; k += RED) {
        const uint32_t i = a.blist[sg.x + k];
        const double kin = spec_kin(a.pos[i], a.vel[i], t, a.qO);
        const double e = kin + a.phi[i];
        if (fabs(e) <= a.near_rel * (kin + fabs(a.phi[i]))) {
            const uint32_t slot = atomicAdd(a.near_count, 1u);
            a.near_list[2 * slot] = i;
            a.near_list[2 * slot + 1] = c;
        }
    }
}

/* binary64 potential of the listed members over their clump's bound set
 * (one wave per member; lanes stride, then a fixed-order reduction) */
__global__ __launch_bounds__(64) void k_bnd_fix64(const double4 *__restrict__ pos, const uint32_t *__restrict__ blist,
                                                  const uint4 *__restrict__ seg, const uint32_t *__restrict__ near_list,
                                                  const uint32_t *__restrict__ near_count, double G,
                                                  double *__restrict__ phi)
{
    const uint32_t n = *near_count;
    for (uint32_t w = blockIdx.x; w < n; w += gridDim.x) {
        const uint32_t i = near_list[2 * w], c = near_list[2 * w + 1];
        const uint4 sg = seg[c];
        const double4 p = pos[i];
        double s = 0;
        for (uint32_t k = threadIdx.x; k < sg.y; k += 64) {
            const double4 q = pos[blist[sg.x + k]];
            const double dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
            const double r2 = (dx * dx + dy * dy) + dz * dz;
            if (r2 > 0) s += q.w / sqrt(r2);
        }
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (threadIdx.x == 0) phi[i] = -G * s;
    }
}

/* decisions and energy sums of the bound set (block per clump, fixed order) */
__global__ __launch_bounds__(RED) void k_bnd_energy(EnergyArgs a)
{
    __shared__ double sm[3][RED];
    const uint32_t c = a.c0 + blockIdx.x;
    if (a.state[c] == C_DONE) return;
    const uint4 sg = a.seg[c];
    const double *t = a.tot + (size_t)c * 8;
    double ek = 0, ep = 0, nrem = 0;
    for (uint32_t k = threadIdx.x; k < sg.y; k += RED) {
        const uint32_t i = a.blist[sg.x + k];
        const double4 p = a.pos[i];
        const double kin = spec_kin(p, a.vel[i], t, a.qO);
        const double ph = a.phi[i];
        ek += p.w * kin;
        ep += 0.5 * p.w * ph;
        if (kin + ph > 0) {
            a.flag[i] = B_LEFT;
            nrem += 1;
        }
    }
    sm[0][threadIdx.x] = ek;
    sm[1][threadIdx.x] = ep;
    sm[2][threadIdx.x] = nrem;
    __syncthreads();
    for (int s = RED / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 3; ++k) sm[k][threadIdx.x] += sm[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double *e = a.eout + (size_t)c * 4;
        e[0] = sm[0][0];
        e[1] = sm[1][0];
        e[2] = sm[2][0];
        e[3] = sg.y;
    }
}

/* count the iteration; a clump stops when nothing left it.  nearz[c] records
 * an iteration whose total energy was within 1e-3 of zero (relative to the
 * potential energy): such a clump is re-decided in binary64 */
__global__ void k_bnd_finalize(uint32_t c0, uint32_t c1, const double *__restrict__ eout,
                               uint32_t *__restrict__ state, uint32_t *__restrict__ iters, uint32_t max_iter,
                               uint32_t *__restrict__ active, uint32_t *__restrict__ nearz)
{
    const uint32_t c = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= c1 || state[c] == C_DONE) return;
    iters[c] += 1;
    const double ek = eout[(size_t)c * 4], ep = eout[(size_t)c * 4 + 1];
    if (fabs(ek + ep) < 1e-3 * fabs(ep)) nearz[c] = 1;
    if (eout[(size_t)c * 4 + 2] == 0 || iters[c] >= max_iter) state[c] = C_DONE;
    else atomicAdd(active, 1u);
}

__global__ void k_bnd_zero(uint32_t *__restrict__ a, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

/* ------------------------------------------------------------ host side */

struct PairShape {
    int ti, rpt, tj;
};
static const PairShape kShapes[] = {{256, 2, 256}, {512, 4, 256}, {256, 4, 512}, {1024, 4, 512}};

template <typename R>
static void launch_pairs(const PairShape &sh, const PairArgs &pa, int blocks, hipStream_t s)
{
#define PLAN_PAIRS(TI, RPT, TJ)                                                                           \
    if (sh.ti == TI && sh.rpt == RPT && sh.tj == TJ) {                                                   \
        hipLaunchKernelGGL((k_bnd_pairs<R, TI, RPT, TJ>), dim3(blocks), dim3(TI / RPT), 0, s, pa);     \
        LAUNCHCHK();                                                                                      \
        return;                                                                                           \
    }
    PLAN_PAIRS(256, 2, 256)
    PLAN_PAIRS(512, 4, 256)
    PLAN_PAIRS(256, 4, 512)
    PLAN_PAIRS(1024, 4, 512)
#undef PLAN_PAIRS
    throw std::runtime_error("boundness: unknown tile shape");
}

__global__ void k_bnd_gather_tree(const uint32_t *__restrict__ members, uint32_t nm, const double4 *__restrict__ pm,
                                  const double4 *__restrict__ va, double4 *__restrict__ pos, double4 *__restrict__ vel)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nm) return;
    const uint32_t t = members[i];
    pos[i] = pm[t];
    const double4 v = va[t];
    vel[i] = make_double4(v.x, v.y, v.z, 0.0);
}

static void clump_ids(const std::vector<uint32_t> &off, std::vector<uint32_t> &cid)
{
    const uint32_t nc = (uint32_t)off.size() - 1;
    cid.resize(off[nc]);
    for (uint32_t c = 0; c < nc; ++c)
        for (uint32_t i = off[c]; i < off[c + 1]; ++i) cid[i] = c;
}

void BoundnessEngine::upload(const plan_boundness_in &in, hipStream_t s)
{
    nc_ = in.num_clumps;
    nm_ = in.offsets[nc_];
    G_ = in.G;
    qO_ = in.q * in.Omega;
    off_h_.assign(in.offsets, in.offsets + nc_ + 1);
    std::vector<double4> hp(nm_), hv(nm_);
    for (uint32_t i = 0; i < nm_; ++i) {
        hp[i] = make_double4(in.pos[3 * (size_t)i], in.pos[3 * (size_t)i + 1], in.pos[3 * (size_t)i + 2], in.mass[i]);
        hv[i] = make_double4(in.vel[3 * (size_t)i], in.vel[3 * (size_t)i + 1], in.vel[3 * (size_t)i + 2], 0.0);
    }
    pos_ = a_.get_t<double4>("b_pos", nm_ + 1);
    vel_ = a_.get_t<double4>("b_vel", nm_ + 1);
    HIPCHK(hipMemcpyAsync(pos_, hp.data(), nm_ * sizeof(double4), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(vel_, hv.data(), nm_ * sizeof(double4), hipMemcpyHostToDevice, s));
    alloc_run_buffers(s);
}

void BoundnessEngine::upload_tree(uint32_t nc, const uint32_t *offsets_h, const uint32_t *members_h,
                                  const double4 *pm, const double4 *va, double G, double Omega, double q,
                                  hipStream_t s)
{
    nc_ = nc;
    nm_ = offsets_h[nc];
    G_ = G;
    qO_ = q * Omega;
    off_h_.assign(offsets_h, offsets_h + nc + 1);
    pos_ = a_.get_t<double4>("b_pos", nm_ + 1);
    vel_ = a_.get_t<double4>("b_vel", nm_ + 1);
    uint32_t *mem = a_.get_t<uint32_t>("b_members", nm_ + 1);
    HIPCHK(hipMemcpyAsync(mem, members_h, nm_ * 4, hipMemcpyHostToDevice, s));
    if (nm_) {
        hipLaunchKernelGGL(k_bnd_gather_tree, dim3(ceil_div(nm_, 256)), dim3(256), 0, s, mem, nm_, pm, va, pos_, vel_);
        LAUNCHCHK();
    }
    alloc_run_buffers(s);
}

void BoundnessEngine::alloc_run_buffers(hipStream_t s)
{
    std::vector<uint32_t> cid;
    clump_ids(off_h_, cid);
    posf_ = a_.get_t<float4>("b_posf", nm_ + 1);
    cid_ = a_.get_t<uint32_t>("b_cid", nm_ + 1);
    off_ = a_.get_t<uint32_t>("b_off", nc_ + 1);
    flag_ = a_.get_t<uint32_t>("b_flag", nm_ + 1);
    phi_ = a_.get_t<double>("b_phi", nm_ + 1);
    ref_ = a_.get_t<double4>("b_ref", nc_ + 1);
    if (nm_) HIPCHK(hipMemcpyAsync(cid_, cid.data(), nm_ * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(off_, off_h_.data(), (nc_ + 1) * 4, hipMemcpyHostToDevice, s));
    if (nc_) {
        hipLaunchKernelGGL(k_bnd_ref, dim3(nc_), dim3(RED), 0, s, pos_, off_, 0u, ref_);
        LAUNCHCHK();
    }
    /* per-run buffers, sized once so that graph capture never allocates */
    inb_ = a_.get_t<uint32_t>("b_inb", nm_ + 1);
    inr_ = a_.get_t<uint32_t>("b_inr", nm_ + 1);
    sb_ = a_.get_t<uint32_t>("b_sb", nm_ + 1);
    sr_ = a_.get_t<uint32_t>("b_sr", nm_ + 1);
    blist_ = a_.get_t<uint32_t>("b_blist", nm_ + 1);
    rlist_ = a_.get_t<uint32_t>("b_rlist", nm_ + 1);
    pj_ = a_.get_t<double4>("b_pj", nm_ + 1);
    pr_ = a_.get_t<double4>("b_pr", nm_ + 1);
    seg_ = a_.get_t<uint4>("b_seg", nc_ + 1);
    mode_ = a_.get_t<uint32_t>("b_mode", nc_ + 1);
    tiles_ = a_.get_t<uint32_t>("b_tiles", nc_ + 1);
    toff_ = a_.get_t<uint32_t>("b_toff", nc_ + 1);
    state_ = a_.get_t<uint32_t>("b_state", nc_ + 1);
    iters_ = a_.get_t<uint32_t>("b_iters", nc_ + 1);
    nearz_ = a_.get_t<uint32_t>("b_nearz", nc_ + 1);
    tot_ = a_.get_t<double>("b_tot", (size_t)nc_ * 8 + 8);
    eout_ = a_.get_t<double>("b_eout", (size_t)nc_ * 4 + 4);
    misc_ = a_.get_t<uint32_t>("b_misc", 8);
    pairs_ = a_.get_t<unsigned long long>("b_pairs", 1);
    near_ = a_.get_t<uint32_t>("b_near", 2 * (size_t)nm_ + 2);
    HIPCHK(hipMemsetAsync(eout_, 0, ((size_t)nc_ * 4 + 4) * 8, s));
    HIPCHK(hipMemsetAsync(tot_, 0, ((size_t)nc_ * 8 + 8) * 8, s));
    HIPCHK(hipMemsetAsync(iters_, 0, (nc_ + 1) * 4, s));
    /* scratch of the scans, at the largest size this batch uses */
    scan_exclusive_u32(a_, inb_, sb_, nm_ ? nm_ : 1, s, nullptr);
    scan_exclusive_u32(a_, tiles_, toff_, nc_ ? nc_ : 1, s, nullptr);
    shape_idx_ = -1;
    HIPCHK(plan::stream_wait(s));
}

/* reset members [m0,m1) and clumps [c0,c1) to "all bound" */
void BoundnessEngine::reset(uint32_t c0, uint32_t c1, hipStream_t s)
{
    const uint32_t m0 = off_h_[c0], m1 = off_h_[c1];
    if (m1 > m0) {
        hipLaunchKernelGGL(k_bnd_init, dim3(ceil_div(m1 - m0, 256)), dim3(256), 0, s, pos_, cid_, m0, m1, ref_, posf_,
                           flag_, phi_);
        LAUNCHCHK();
    }
    if (c1 > c0) {
        hipLaunchKernelGGL(k_bnd_reset_clumps, dim3(ceil_div(c1 - c0, 256)), dim3(256), 0, s, c0, c1, state_, iters_,
                           nearz_);
        LAUNCHCHK();
    }
}

/* one unbinding iteration for clumps [c0,c1); kernels only (capturable) */
template <typename R>
void BoundnessEngine::iteration(uint32_t c0, uint32_t c1, const PairShape &sh, bool near, uint32_t max_iter,
                                hipStream_t s)
{
    const uint32_t m0 = off_h_[c0], m1 = off_h_[c1], nm = m1 - m0, nc = c1 - c0;
    if (nm == 0 || nc == 0) return;
    hipLaunchKernelGGL(k_bnd_marks, dim3(ceil_div(nm, 256)), dim3(256), 0, s, flag_, m0, m1, inb_, inr_);
    LAUNCHCHK();
    scan_exclusive_u32(a_, inb_, sb_, nm, s, sb_ + nm);
    scan_exclusive_u32(a_, inr_, sr_, nm, s, sr_ + nm);
    hipLaunchKernelGGL(k_bnd_compact<R>, dim3(ceil_div(nm, 256)), dim3(256), 0, s, pos_, posf_, m0, m1, flag_, sb_,
                       sr_, blist_, rlist_, (void *)pj_, (void *)pr_);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_bnd_segments, dim3(ceil_div(nc, 256)), dim3(256), 0, s, off_, c0, c1, m0, sb_, sr_, state_,
                       iters_, (uint32_t)sh.ti, seg_, mode_, tiles_);
    LAUNCHCHK();
    scan_exclusive_u32(a_, tiles_, toff_, nc, s, toff_ + nc);
    hipLaunchKernelGGL(k_bnd_totals, dim3(nc), dim3(RED), 0, s, pos_, vel_, blist_, seg_, state_, c0, tot_);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_bnd_zero, dim3(1), dim3(64), 0, s, misc_, 8u);
    LAUNCHCHK();
    PairArgs pa;
    pa.off = off_;
    pa.c0 = c0;
    pa.c1 = c1;
    pa.seg = seg_;
    pa.mode = mode_;
    pa.tile_off = toff_;
    pa.blist = blist_;
    pa.pj = pj_;
    pa.pr = pr_;
    pa.phi = phi_;
    pa.G = G_;
    pa.counter = misc_;
    pa.pairs = pairs_;
    launch_pairs<R>(sh, pa, persistent_blocks_, s);
    EnergyArgs ea;
    ea.pos = pos_;
    ea.vel = vel_;
    ea.blist = blist_;
    ea.seg = seg_;
    ea.state = state_;
    ea.tot = tot_;
    ea.phi = phi_;
    ea.flag = flag_;
    ea.qO = qO_;
    ea.c0 = c0;
    ea.near_rel = near ? 1e-5 : 0.0;
    ea.near_list = near_;
    ea.near_count = misc_ + 1;
    ea.eout = eout_;
    if (near) {
        hipLaunchKernelGGL(k_bnd_near, dim3(nc), dim3(RED), 0, s, ea);
        LAUNCHCHK();
        hipLaunchKernelGGL(k_bnd_fix64, dim3(1024), dim3(64), 0, s, pos_, blist_, seg_, near_, misc_ + 1, G_, phi_);
        LAUNCHCHK();
    }
    hipLaunchKernelGGL(k_bnd_energy, dim3(nc), dim3(RED), 0, s, ea);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_bnd_finalize, dim3(ceil_div(nc, 256)), dim3(256), 0, s, c0, c1, eout_, state_, iters_,
                       max_iter, misc_ + 2, nearz_);
    LAUNCHCHK();
}

uint32_t BoundnessEngine::read_active(hipStream_t s)
{
    uint32_t v = 0;
    HIPCHK(hipMemcpyAsync(&v, misc_ + 2, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(plan::stream_wait(s));
    return v;
}

/* FP64 per clump: every clump on its own until it stops losing members */
void BoundnessEngine::run_fp64_per_clump(const std::vector<uint32_t> &clumps, uint32_t max_iter, hipStream_t s)
{
    const PairShape sh{256, 2, 256};
    for (uint32_t c : clumps) {
        reset(c, c + 1, s);
        for (uint32_t it = 0; it < max_iter; ++it) {
            iteration<double>(c, c + 1, sh, false, max_iter, s);
            if (read_active(s) == 0) break;
        }
    }
}

/* FP32 batched: all clumps per launch, ITER iterations per graph replay */
void BoundnessEngine::run_fp32_batched(uint32_t max_iter, hipStream_t s)
{
    if (nc_ == 0) return;
    if (shape_idx_ < 0) autotune(s);
    const PairShape sh = kShapes[shape_idx_];
    reset(0, nc_, s);
    /* first iteration eagerly (sizes every buffer), the rest from a graph */
    iteration<float>(0, nc_, sh, true, max_iter, s);
    if (read_active(s) == 0) return;
    const int ITER = 4;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < ITER; ++k) iteration<float>(0, nc_, sh, true, max_iter, s);
    HIPCHK(hipStreamEndCapture(s, &g));
    HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    uint32_t done_iters = 1;
    while (done_iters < max_iter) {
        HIPCHK(hipGraphLaunch(ge, s));
        done_iters += ITER;
        if (read_active(s) == 0) break;
    }
    HIPCHK(hipGraphExecDestroy(ge));
    HIPCHK(hipGraphDestroy(g));
    graph_replays_ = (done_iters - 1) / ITER;
}

/* time each tile shape on the first full evaluation of a subset of the
 * clumps (every 4th) and keep the fastest */
void BoundnessEngine::autotune(hipStream_t s)
{
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreateWithFlags(&e0, hipEventBlockingSync));
    HIPCHK(hipEventCreateWithFlags(&e1, hipEventBlockingSync));
    float best = 1e30f;
    int bi = 0;
    const uint32_t cend = std::min<uint32_t>(nc_, std::max<uint32_t>(1, nc_ / 4));
    tune_ms_.clear();
    for (int k = 0; k < (int)(sizeof(kShapes) / sizeof(kShapes[0])); ++k) {
        reset(0, cend, s);
        HIPCHK(hipEventRecord(e0, s));
        iteration<float>(0, cend, kShapes[k], false, 1, s);
        HIPCHK(hipEventRecord(e1, s));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        tune_ms_.push_back(ms);
        if (ms < best) {
            best = ms;
            bi = k;
        }
    }
    shape_idx_ = bi;
    HIPCHK(hipEventDestroy(e0));
    HIPCHK(hipEventDestroy(e1));
}

void BoundnessEngine::run(int form, uint32_t max_iter, hipStream_t s, plan_boundness_stats *st)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    persistent_blocks_ = prop.multiProcessorCount * 4;
    unsigned long long zero = 0;
    HIPCHK(hipMemcpyAsync(pairs_, &zero, 8, hipMemcpyHostToDevice, s));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreateWithFlags(&e0, hipEventBlockingSync));
    HIPCHK(hipEventCreateWithFlags(&e1, hipEventBlockingSync));
    if (form == 1 && shape_idx_ < 0) autotune(s);
    HIPCHK(hipMemcpyAsync(pairs_, &zero, 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipEventRecord(e0, s));
    std::vector<uint32_t> refit;
    graph_replays_ = 0;
    if (form == 0) {
        std::vector<uint32_t> all(nc_);
        for (uint32_t c = 0; c < nc_; ++c) all[c] = c;
        run_fp64_per_clump(all, max_iter, s);
    } else {
        run_fp32_batched(max_iter, s);
        /* clumps whose total energy came within 1e-3 of zero: redo in binary64 */
        std::vector<uint32_t> nz(nc_);
        HIPCHK(hipMemcpyAsync(nz.data(), nearz_, nc_ * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(plan::stream_wait(s));
        for (uint32_t c = 0; c < nc_; ++c)
            if (nz[c]) refit.push_back(c);
        if (!refit.empty()) run_fp64_per_clump(refit, max_iter, s);
    }
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    unsigned long long pairs = 0;
    HIPCHK(hipMemcpy(&pairs, pairs_, 8, hipMemcpyDeviceToHost));
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->ms = ms;
        st->pair_evaluations = (double)pairs;
        st->refit_clumps = (uint32_t)refit.size();
        st->graph_replays = graph_replays_;
        st->tile_shape = shape_idx_ >= 0 && form == 1 ? shape_idx_ : -1;
        for (size_t k = 0; k < tune_ms_.size() && k < 4; ++k) st->tune_ms[k] = tune_ms_[k];
    }
    HIPCHK(hipEventDestroy(e0));
    HIPCHK(hipEventDestroy(e1));
}

void BoundnessEngine::download(uint8_t *bound, plan_boundness_clump *out, hipStream_t s)
{
    std::vector<uint32_t> fl(nm_), it(nc_);
    std::vector<double> eo((size_t)nc_ * 4), tot((size_t)nc_ * 8);
    if (nm_) HIPCHK(hipMemcpyAsync(fl.data(), flag_, nm_ * 4, hipMemcpyDeviceToHost, s));
    if (nc_) {
        HIPCHK(hipMemcpyAsync(it.data(), iters_, nc_ * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(eo.data(), eout_, eo.size() * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(tot.data(), tot_, tot.size() * 8, hipMemcpyDeviceToHost, s));
    }
    HIPCHK(plan::stream_wait(s));
    if (bound)
        for (uint32_t i = 0; i < nm_; ++i) bound[i] = fl[i] == B_IN ? 1 : 0;
    if (out)
        for (uint32_t c = 0; c < nc_; ++c) {
            plan_boundness_clump &o = out[c];
            o.n_in = off_h_[c + 1] - off_h_[c];
            o.n_bound = (uint32_t)eo[4 * (size_t)c + 3];
            o.iterations = it[c];
            o.mass = tot[8 * (size_t)c];
            for (int k = 0; k < 3; ++k) {
                o.center_of_mass[k] = tot[8 * (size_t)c + 1 + k];
                o.vel_com[k] = tot[8 * (size_t)c + 4 + k];
            }
            o.E_kin = eo[4 * (size_t)c];
            o.E_pot = eo[4 * (size_t)c + 1];
        }
}

template void BoundnessEngine::iteration<float>(uint32_t, uint32_t, const PairShape &, bool, uint32_t, hipStream_t);
template void BoundnessEngine::iteration<double>(uint32_t, uint32_t, const PairShape &, bool, uint32_t, hipStream_t);

}  // namespace plan
