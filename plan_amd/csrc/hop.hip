// hop.hip — HOP on the device: particles above the density threshold, their
// densest neighbours, pointer jumping to the peaks, and the reference's member
// discovery order.
//
// The reference walks densest-neighbour chains serially in tree-index order
// (src/tree.hpp:3766-3809).  The pointers form a forest (each step strictly
// raises (density, -index)), so the peak of every particle is its forest root
// and pointer jumping finds it in O(log depth) rounds.  The order in which the
// serial walk appends members to a clump is also recovered in parallel: the
// walk that first reaches particle p starts at f(p) = the smallest index in
// p's subtree, and within one walk members appear from the deepest up.  So
// sorting members by (peak, f(p), -depth(p)) reproduces the reference's
// member lists exactly.
#include "ctx.h"
#include "hop.h"

namespace plan {

static constexpr int BLK = 256;

__global__ void k_above(const double *__restrict__ rho, uint32_t n, double thr, uint32_t *__restrict__ flag)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = rho[i] >= thr ? 1u : 0u;
}

__global__ void k_compact(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pos, uint32_t n,
                          uint32_t *__restrict__ a_idx, uint32_t *__restrict__ apos)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (flag[i]) {
        a_idx[pos[i]] = i;
        apos[i] = pos[i];
    } else {
        apos[i] = PLAN_NONE;
    }
}

__global__ void k_hop_init(const uint32_t *__restrict__ densest, const uint32_t *__restrict__ apos, uint32_t na,
                           const uint32_t *__restrict__ a_idx, uint32_t *__restrict__ par, uint32_t *__restrict__ dep,
                           uint32_t *__restrict__ fmin)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= na) return;
    uint32_t p = apos[densest[j]];
    par[j] = p;
    dep[j] = (p == j) ? 0u : 1u;
    fmin[j] = 0xffffffffu;
    (void)a_idx;
}

__global__ void k_jump(const uint32_t *__restrict__ par, const uint32_t *__restrict__ dep, uint32_t na,
                       uint32_t *__restrict__ par2, uint32_t *__restrict__ dep2, uint32_t *__restrict__ changed)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= na) return;
    uint32_t p = par[j];
    uint32_t pp = par[p];
    par2[j] = pp;
    dep2[j] = dep[j] + (p == j ? 0u : dep[p]);
    if (pp != p) *changed = 1u;
}

/* subtree minimum of A positions along the densest-neighbour pointers */
__global__ void k_subtree_min(const uint32_t *__restrict__ up, uint32_t na, uint32_t *__restrict__ fmin)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= na) return;
    uint32_t x = j;
    for (uint32_t steps = 0; steps <= na; ++steps) { /* a chain has at most na links */
        uint32_t old = atomicMin(&fmin[x], j);
        if (old <= j) break;
        uint32_t nx = up[x];
        if (nx == x) break;
        x = nx;
    }
}

__global__ void k_group_keys(const uint32_t *__restrict__ root, const uint32_t *__restrict__ fmin,
                             const uint32_t *__restrict__ depth, uint32_t na, uint4 *__restrict__ keys)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= na) return;
    keys[j] = make_uint4(0xffffffffu - depth[j], fmin[j], root[j], j);
}

__global__ void k_clump_flags(const uint4 *__restrict__ keys, uint32_t na, uint32_t *__restrict__ flag)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= na) return;
    flag[j] = (j == 0 || keys[j].z != keys[j - 1].z) ? 1u : 0u;
}

__global__ void k_clump_starts(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ cidx, uint32_t na,
                               uint32_t *__restrict__ cstart)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= na) return;
    if (flag[j]) cstart[cidx[j]] = j;
}

/* gather member data in grouped order */
__global__ void k_gather_members(const uint4 *__restrict__ keys, uint32_t na, const uint32_t *__restrict__ a_idx,
                                 const double4 *__restrict__ pm, const double4 *__restrict__ va,
                                 const double *__restrict__ rho, const uint32_t *__restrict__ oid,
                                 const uint32_t *__restrict__ sinkcnt, MemberRec *__restrict__ out,
                                 uint32_t *__restrict__ msink)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= na) return;
    uint32_t ap = keys[j].w;
    uint32_t ti = a_idx[ap];
    MemberRec m;
    double4 p = pm[ti];
    double4 v = va[ti];
    m.x = p.x;
    m.y = p.y;
    m.z = p.z;
    m.m = p.w;
    m.vx = v.x;
    m.vy = v.y;
    m.vz = v.z;
    m.ath = v.w;
    m.rho = rho[ti];
    m.tidx = ti;
    m.oid = oid[ti];
    out[j] = m;
    msink[j] = sinkcnt[ti];
}

static uint32_t read_u32(const uint32_t *dp, hipStream_t s)
{
    uint32_t v;
    HIPCHK(hipMemcpyAsync(&v, dp, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(plan::stream_wait(s));
    return v;
}

void dev_hop(const DevTree &t, Arena &a, const plan_derived &d, double thr, uint32_t H, HopResult &res,
             hipStream_t s)
{
    const uint32_t n = t.n;
    uint32_t *flag = a.get_t<uint32_t>("h_flag", n + 1);
    uint32_t *pos = a.get_t<uint32_t>("h_pos", n + 1);
    uint32_t *tot = a.get_t<uint32_t>("h_tot", 8);
    hipLaunchKernelGGL(k_above, dim3(ceil_div(n, BLK)), dim3(BLK), 0, s, t.rho, n, thr, flag);
    LAUNCHCHK();
    scan_exclusive_u32(a, flag, pos, n, s, tot);
    const uint32_t na = read_u32(tot, s);
    res.na = na;
    res.a_idx = a.get_t<uint32_t>("h_aidx", na + 1);
    res.apos = a.get_t<uint32_t>("h_apos", n + 1);
    hipLaunchKernelGGL(k_compact, dim3(ceil_div(n, BLK)), dim3(BLK), 0, s, flag, pos, n, res.a_idx, res.apos);
    LAUNCHCHK();
    res.nclumps = 0;
    if (na == 0) return;
    /* densest neighbour of each A particle (H nearest, reference candidate sets) */
    uint32_t *dn = a.get_t<uint32_t>("h_dn", na);
    KnnOut ko;
    ko.densest = dn;
    res.knn_overflow_waves = 0;
    res.knn_res_queries = 0;
    ko.novf = &res.knn_overflow_waves;
    ko.nres = &res.knn_res_queries;
    dev_knn(t, H, 1, 0, na, res.a_idx, nullptr, nullptr, t.rho, d, ko, s);
    res.densest = dn;
    /* pointer jumping */
    uint32_t *par = a.get_t<uint32_t>("h_par", na);
    uint32_t *dep = a.get_t<uint32_t>("h_dep", na);
    uint32_t *par2 = a.get_t<uint32_t>("h_par2", na);
    uint32_t *dep2 = a.get_t<uint32_t>("h_dep2", na);
    uint32_t *up = a.get_t<uint32_t>("h_up", na);
    uint32_t *fmin = a.get_t<uint32_t>("h_fmin", na);
    uint32_t *chg = tot + 4;
    hipLaunchKernelGGL(k_hop_init, dim3(ceil_div(na, BLK)), dim3(BLK), 0, s, dn, res.apos, na, res.a_idx, par, dep,
                       fmin);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(up, par, (size_t)na * 4, hipMemcpyDeviceToDevice, s));
    /* two jumps per check of the change flag: a jump after convergence
     * changes nothing, so the extra one is harmless and halves the syncs */
    for (int it = 0; it < 20; ++it) {
        HIPCHK(hipMemsetAsync(chg, 0, 4, s));
        for (int h = 0; h < 2; ++h) {
            hipLaunchKernelGGL(k_jump, dim3(ceil_div(na, BLK)), dim3(BLK), 0, s, par, dep, na, par2, dep2, chg);
            LAUNCHCHK();
            std::swap(par, par2);
            std::swap(dep, dep2);
        }
        if (!read_u32(chg, s)) break;
    }
    hipLaunchKernelGGL(k_subtree_min, dim3(ceil_div(na, BLK)), dim3(BLK), 0, s, up, na, fmin);
    LAUNCHCHK();
    res.root = par;
    res.depth = dep;
    res.up = up;
    res.fmin = fmin;
    /* member order */
    uint4 *k0 = a.get_t<uint4>("h_keys_a", na);
    uint4 *k1 = a.get_t<uint4>("h_keys_b", na);
    hipLaunchKernelGGL(k_group_keys, dim3(ceil_div(na, BLK)), dim3(BLK), 0, s, par, fmin, dep, na, k0);
    LAUNCHCHK();
    radix_sort_96(a, &k0, &k1, na, 3, s);
    res.order = k0;
    uint32_t *cflag = a.get_t<uint32_t>("h_cflag", na + 1);
    uint32_t *cidx = a.get_t<uint32_t>("h_cidx", na + 1);
    hipLaunchKernelGGL(k_clump_flags, dim3(ceil_div(na, BLK)), dim3(BLK), 0, s, k0, na, cflag);
    LAUNCHCHK();
    scan_exclusive_u32(a, cflag, cidx, na, s, tot + 1);
    res.nclumps = read_u32(tot + 1, s);
    res.cstart = a.get_t<uint32_t>("h_cstart", res.nclumps + 1);
    hipLaunchKernelGGL(k_clump_starts, dim3(ceil_div(na, BLK)), dim3(BLK), 0, s, cflag, cidx, na, res.cstart);
    LAUNCHCHK();
    res.members = a.get_t<MemberRec>("h_members", na);
    res.msink = a.get_t<uint32_t>("h_msink", na);
    hipLaunchKernelGGL(k_gather_members, dim3(ceil_div(na, BLK)), dim3(BLK), 0, s, k0, na, res.a_idx, t.pm, t.va,
                       t.rho, t.oid, t.sinkcnt, res.members, res.msink);
    LAUNCHCHK();
}

}  // namespace plan
