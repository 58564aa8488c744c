// analyses.hip — the density-against-scale pass of BasicAnalysesWithTree
// (src/analyses.hpp:196-333) on the device tree.  For every cell centre of
// the grid between z_min and z_max, and for each radius r_i = (cell/2) 2^i,
// the reference counts the particles in a sheared-periodic ball
// (RecursiveBallSearchCountWithShear, src/tree.hpp:3383-3402); levels 0-2
// also search at the 8 half-cell offsets and level 0 at the quarter-cell
// offsets.  The largest count of each level becomes a density.
//
// Every search is one lane: a stackless walk over the skip pointers (the
// pre-order successor of a node's subtree) makes the same contain /
// intersect / leaf decisions as the recursion, in the same node order.
// Lanes of a wave hold neighbouring grid points at the same level and
// offset, so their walks stay close.  Each wave reduces its lanes' counts
// (max and sum) before one atomic per wave.

#include "ctx.h"

#include <cmath>

namespace plan {

struct ScaleGrid {
    const double *xs, *ys, *zs; /* grid coordinates, as the reference's loops accumulate them */
    uint32_t nx, ny, nz;
    double sub[3], subsub[3];   /* half- and quarter-cell steps */
    double boxmin[3], boxmax[3], boxlen[3];
    double shear;               /* shear_speed * time */
};

/* ShearedCenter2Center (src/tree.hpp:3307-3330): once the x offset wraps,
 * the node centre's y moves by the shear and is folded back into the box,
 * and the moved y feeds the y offset */
__device__ __forceinline__ void sheared_c2c(const ScaleGrid &g, const double c[3], const NodeGeo &n,
                                            const double maxd[3], double c2c[3])
{
    double ny = n.cy;
    c2c[0] = __dsub_rn(c[0], n.cx);
    if (fabs(c2c[0]) > maxd[0]) {
        c2c[0] = c2c[0] > 0 ? __dsub_rn(c2c[0], g.boxlen[0]) : __dadd_rn(c2c[0], g.boxlen[0]);
        if (c[0] < 0) {
            ny = __dsub_rn(ny, g.shear);
            ny = __dadd_rn(ny, __dmul_rn((double)(int)__ddiv_rn(__dsub_rn(g.boxmax[1], ny), g.boxlen[1]),
                                         g.boxlen[1]));
        } else {
            ny = __dadd_rn(ny, g.shear);
            ny = __dsub_rn(ny, __dmul_rn((double)(int)__ddiv_rn(__dsub_rn(ny, g.boxmin[1]), g.boxlen[1]),
                                         g.boxlen[1]));
        }
    }
    c2c[1] = __dsub_rn(c[1], ny);
    if (fabs(c2c[1]) > maxd[1]) c2c[1] = c2c[1] > 0 ? __dsub_rn(c2c[1], g.boxlen[1]) : __dadd_rn(c2c[1], g.boxlen[1]);
    c2c[2] = __dsub_rn(c[2], n.cz);
    if (fabs(c2c[2]) > maxd[2]) c2c[2] = c2c[2] > 0 ? __dsub_rn(c2c[2], g.boxlen[2]) : __dadd_rn(c2c[2], g.boxlen[2]);
}

__device__ __forceinline__ double norm2_3(double a, double b, double c)
{
    return __dadd_rn(__dadd_rn(__dmul_rn(a, a), __dmul_rn(b, b)), __dmul_rn(c, c));
}

/* RecursiveBallSearchCountWithShear (src/tree.hpp:3383-3402) with
 * SphereContainNodeWithShear (3334-3345) and SphereNodeIntersectWithShear
 * (3349-3378); leaf particles are tested at their own positions */
__device__ uint32_t sheared_count(const TreeView &tv, const ScaleGrid &g, const double c[3], double r,
                                  const double maxd[3])
{
    const double r2 = __dmul_rn(r, r);
    uint32_t count = 0, node = 0;
    while (node != PLAN_NONE) {
        const NodeGeo n = tv.ngeo[node];
        double c2c[3];
        sheared_c2c(g, c, n, maxd, c2c);
        const double diag = __dmul_rn(n.hw, tv.to_diag);
        const double t0 = __dadd_rn(fabs(c2c[0]), diag), t1 = __dadd_rn(fabs(c2c[1]), diag),
                     t2 = __dadd_rn(fabs(c2c[2]), diag);
        if (norm2_3(t0, t1, t2) < r2) {
            const NodeTop nt = tv.ntop[node];
            count += nt.end - nt.begin;
            node = tv.skip[node];
            continue;
        }
        const double c2 = norm2_3(c2c[0], c2c[1], c2c[2]);
        const double td = __dadd_rn(diag, r);
        bool hit = !(c2 > __dmul_rn(td, td));
        if (hit && !(c2 < r || c2 < n.hw)) {
            double md = 0;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const double base = __dsub_rn(c[i], c2c[i]);
                const double lo = __dsub_rn(base, n.hw), hi = __dadd_rn(base, n.hw);
                if (c[i] < lo) md = __dadd_rn(md, __dmul_rn(__dsub_rn(c[i], lo), __dsub_rn(c[i], lo)));
                else if (c[i] > hi) md = __dadd_rn(md, __dmul_rn(__dsub_rn(c[i], hi), __dsub_rn(c[i], hi)));
            }
            hit = md <= r2;
        }
        if (!hit) {
            node = tv.skip[node];
            continue;
        }
        const NodeTop nt = tv.ntop[node];
        if ((tv.ninfo[node] & 0xffu) == 0) {
            for (uint32_t p = nt.begin; p < nt.end; ++p) {
                const double4 q = tv.pm[p];
                if (norm2_3(__dsub_rn(c[0], q.x), __dsub_rn(c[1], q.y), __dsub_rn(c[2], q.z)) <= r2) ++count;
            }
            node = tv.skip[node];
        } else {
            node = nt.first_daughter;
        }
    }
    return count;
}

/* One level's searches: sample s of point p is search s * npts + p (s = 0
 * the cell centre, 1-8 the half-cell offsets, 9-16 the quarter-cell ones).
 * out[0] = max count (atomicMax), out[1..2] = 64-bit sum of counts. */
__global__ __launch_bounds__(256) void k_scale_level(TreeView tv, ScaleGrid g, uint32_t nsamp, double r, double mx,
                                                     double my, double mz, uint32_t *__restrict__ out)
{
    const uint32_t npts = g.nx * g.ny * g.nz;
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t cnt = 0;
    if (q < (uint64_t)npts * nsamp) {
        const uint32_t s = (uint32_t)(q / npts), p = (uint32_t)(q % npts);
        const uint32_t ix = p % g.nx, iy = (p / g.nx) % g.ny, iz = p / (g.nx * g.ny);
        double c[3] = {g.xs[ix], g.ys[iy], g.zs[iz]};
        if (s >= 1) {
            const uint32_t o = (s - 1) & 7u; /* Orthant<3>::orthants (src/tree.cpp:21-30): bit k = +1 along axis k */
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double sg = ((o >> k) & 1u) ? 1.0 : -1.0;
                c[k] = __dadd_rn(c[k], __dmul_rn(g.sub[k], sg));
                if (s >= 9) c[k] = __dadd_rn(c[k], __dmul_rn(g.subsub[k], sg));
            }
        }
        const double maxd[3] = {mx, my, mz};
        cnt = sheared_count(tv, g, c, r, maxd);
    }
    uint32_t m = cnt;
    uint64_t sum = cnt;
    for (int off = 32; off > 0; off >>= 1) {
        m = max(m, (uint32_t)__shfl_xor((int)m, off, 64));
        sum += (uint64_t)__shfl_xor((long long)sum, off, 64);
    }
    if (__lane_id() == 0) {
        atomicMax(out, m);
        atomicAdd(reinterpret_cast<unsigned long long *>(out + 2), (unsigned long long)sum);
    }
}

/* RecursiveFindExtreme along z (src/analyses.hpp:171-191), one lane: the
 * same depth-first order and the same pruning (a daughter on the far side
 * of its parent's centre is entered only while the parent's centre beats
 * the current extreme), so the result is the reference's even where the
 * pruning would miss a particle. */
__global__ void k_find_extreme_z(TreeView tv, int smaller, uint32_t *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t ext = 0;
    uint32_t stk_node[40], stk_next[40];
    int sp = 0;
    stk_node[0] = 0;
    stk_next[0] = 0;
    auto better = [smaller](double a, double b) { return smaller ? a < b : a > b; };
    while (sp >= 0) {
        const uint32_t node = stk_node[sp];
        const NodeTop nt = tv.ntop[node];
        const uint32_t nd = tv.ninfo[node] & 0xffu;
        if (nd == 0) {
            for (uint32_t p = nt.begin; p < nt.end; ++p)
                if (better(tv.pm[p].z, tv.pm[ext].z)) ext = p;
            --sp;
            continue;
        }
        bool pushed = false;
        while (stk_next[sp] < nd) {
            const uint32_t d = nt.first_daughter + stk_next[sp]++;
            const double cz = tv.ngeo[node].cz;
            if (better(tv.ngeo[d].cz, cz) || better(cz, tv.pm[ext].z)) {
                if (sp + 1 >= 40) { /* deeper than any tree (32 levels): report, do not overflow */
                    out[1] = 1u;
                    return;
                }
                ++sp;
                stk_node[sp] = d;
                stk_next[sp] = 0;
                pushed = true;
                break;
            }
        }
        if (!pushed) --sp;
    }
    out[0] = ext;
}

void dev_density_vs_scale(const DevTree &t, const plan_params &pp, const plan_derived &d, Arena &a, hipStream_t s,
                          ScaleResult &res)
{
    const TreeView tv = tree_view(t, d);
    uint32_t *ex = a.get_t<uint32_t>("sc_ext", 4);
    HIPCHK(hipMemsetAsync(ex, 0, 16, s));
    hipLaunchKernelGGL(k_find_extreme_z, dim3(1), dim3(64), 0, s, tv, 1, ex);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_find_extreme_z, dim3(1), dim3(64), 0, s, tv, 0, ex + 2);
    LAUNCHCHK();
    uint32_t he[4];
    HIPCHK(hipMemcpyAsync(he, ex, 16, hipMemcpyDeviceToHost, s));
    HIPCHK(plan::stream_wait(s));
    if (he[1] || he[3]) throw std::runtime_error("density_vs_scale: tree deeper than 40 levels");
    double zlo, zhi;
    HIPCHK(hipMemcpy(&zlo, reinterpret_cast<const char *>(t.pm + he[0]) + 16, 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&zhi, reinterpret_cast<const char *>(t.pm + he[2]) + 16, 8, hipMemcpyDeviceToHost));

    /* src/analyses.hpp:230-236: start at the central cell's lower face */
    double z_min = d.box_center[2] - d.cell_length[2] / 2.0;
    const double z_max = zhi;
    while (z_min - d.cell_length[2] > zlo) z_min -= d.cell_length[2];
    res.z_min = z_min;
    res.z_max = z_max;
    const int L = (int)std::round(std::log10(d.box_half_width[0] * 2 / d.cell_length[0]) / std::log10(2.0)) + 1;
    res.radii.assign(L, d.cell_length[0] / 2.0);
    res.max_rhop.assign(L, 0.0);
    res.sums.assign(L, 0);

    /* the grid as the reference's loops accumulate it (src/analyses.hpp:277-281) */
    std::vector<double> xs, ys, zs;
    const double cx0 = pp.box_min[0] + 0.5 * d.cell_length[0], cy0 = pp.box_min[1] + 0.5 * d.cell_length[1];
    for (double z = z_min; z < z_max; z += d.cell_length[2]) zs.push_back(z);
    for (double y = cy0; y < pp.box_max[1]; y += d.cell_length[1]) ys.push_back(y);
    for (double x = cx0; x < pp.box_max[0]; x += d.cell_length[0]) xs.push_back(x);
    res.points = (uint64_t)xs.size() * ys.size() * zs.size();
    if (res.points == 0) return;
    if (res.points * 17 > 0xffffffffull) throw std::runtime_error("density_vs_scale: grid too large");
    double *dx = a.get_t<double>("sc_xs", xs.size()), *dy = a.get_t<double>("sc_ys", ys.size()),
           *dz = a.get_t<double>("sc_zs", zs.size());
    HIPCHK(hipMemcpyAsync(dx, xs.data(), xs.size() * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dy, ys.data(), ys.size() * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dz, zs.data(), zs.size() * 8, hipMemcpyHostToDevice, s));
    ScaleGrid g;
    g.xs = dx;
    g.ys = dy;
    g.zs = dz;
    g.nx = (uint32_t)xs.size();
    g.ny = (uint32_t)ys.size();
    g.nz = (uint32_t)zs.size();
    for (int k = 0; k < 3; ++k) {
        g.sub[k] = d.cell_length[k] / 2.0;
        g.subsub[k] = d.cell_length[k] / 4.0;
        g.boxmin[k] = pp.box_min[k];
        g.boxmax[k] = pp.box_max[k];
        g.boxlen[k] = d.box_length[k];
    }
    g.shear = d.shear_speed * pp.time;
    uint32_t *acc = a.get_t<uint32_t>("sc_acc", 4 * (size_t)L);
    HIPCHK(hipMemsetAsync(acc, 0, 16 * (size_t)L, s));
    for (int i = 0; i < L; ++i) {
        res.radii[i] *= std::pow(2.0, i);
        const double r = res.radii[i];
        const uint32_t nsamp = i == 0 ? 17u : (i < 3 ? 9u : 1u);
        const uint64_t nq = res.points * nsamp;
        hipLaunchKernelGGL(k_scale_level, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, tv, g, nsamp, r,
                           d.box_length[0] - r, d.box_length[1] - r, 100 * d.max_half_width, acc + 4 * i);
        LAUNCHCHK();
    }
    std::vector<uint32_t> h(4 * (size_t)L);
    HIPCHK(hipMemcpyAsync(h.data(), acc, 16 * (size_t)L, hipMemcpyDeviceToHost, s));
    HIPCHK(plan::stream_wait(s));
    for (int i = 0; i < L; ++i) {
        const double best = (double)h[4 * i];
        res.sums[i] = (uint64_t)h[4 * i + 2] | ((uint64_t)h[4 * i + 3] << 32);
        res.max_rhop[i] = best * (d.mass_per_particle[0] / (4.1887902047863910 * std::pow(res.radii[i], 3)));
    }
}

}  // namespace plan
