// build.hip — ghost particles, Morton keys, key sort, sink merging and the
// level-major octree, all on the device.
//
// Reference behaviour restated: ParticleSet::MakeGhostParticles
// (src/tree.hpp:1626-1754); BHtree::BuildTree (src/tree.hpp:2759-2908) with
// SortPoints (2608-2651), MakeSinkParticle (2591-2604), CountNodesLeaves /
// FillTree (2655-2755).  The recursive count/fill passes become one
// level-synchronous pass per tree level: at level L every particle of a
// non-leaf node flags the start of its octant run; a prefix sum over the
// flags numbers the level-(L+1) nodes in key order, which is exactly the
// reference's level-major, DFS-within-level numbering.
#include "ctx.h"
#include "morton.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace plan {

static constexpr int BLK = 256;

/* ---------------------------------------------------------------- ghosts */

struct GhostGeom {
    double ngmin[3], ngmax[3];
    double Lx, Ly, ymin, ymax, shear_t;
};

__global__ void k_ghost_count(const double4 *__restrict__ pos, uint32_t N, GhostGeom g, uint32_t *__restrict__ cx,
                              uint32_t *__restrict__ cy)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double4 p = pos[i];
    cx[i] = (uint32_t)(p.x > g.ngmax[0]) + (uint32_t)(p.x < g.ngmin[0]);
    cy[i] = (uint32_t)(p.y < g.ngmin[1]) + (uint32_t)(p.y > g.ngmax[1]);
}

__device__ __forceinline__ void copy_particle(double4 *pos, double4 *vel, int32_t *prop, uint32_t *id, uint32_t dst,
                                              uint32_t src, double4 p)
{
    pos[dst] = p;
    vel[dst] = vel[src];
    prop[dst] = prop[src];
    id[dst] = dst;
}

/* x images with the shearing-box y offset (src/tree.hpp:1660-1681) */
__global__ void k_ghost_x(double4 *pos, double4 *vel, int32_t *prop, uint32_t *id, uint32_t N,
                          const uint32_t *__restrict__ ox, GhostGeom g)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double4 p = pos[i];
    uint32_t slot = N + ox[i];
    if (p.x > g.ngmax[0]) {
        double4 q = p;
        q.x = __dsub_rn(q.x, g.Lx);
        double ny = __dadd_rn(q.y, g.shear_t);
        double k = (double)(int)(__ddiv_rn(__dsub_rn(ny, g.ymin), g.Ly));
        q.y = __dsub_rn(ny, __dmul_rn(k, g.Ly));
        copy_particle(pos, vel, prop, id, slot++, i, q);
    }
    if (p.x < g.ngmin[0]) {
        double4 q = p;
        q.x = __dadd_rn(q.x, g.Lx);
        double ny = __dsub_rn(q.y, g.shear_t);
        double k = (double)(int)(__ddiv_rn(__dsub_rn(g.ymax, ny), g.Ly));
        q.y = __dadd_rn(ny, __dmul_rn(k, g.Ly));
        copy_particle(pos, vel, prop, id, slot, i, q);
    }
}

__global__ void k_ghost_count_y(const double4 *__restrict__ pos, uint32_t first, uint32_t cnt, GhostGeom g,
                                uint32_t *__restrict__ cy)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cnt) return;
    double y = pos[first + j].y;
    cy[j] = (uint32_t)(y < g.ngmin[1]) + (uint32_t)(y > g.ngmax[1]);
}

/* y images (src/tree.hpp:1685-1720) of particles [first, first+cnt) */
__global__ void k_ghost_y(double4 *pos, double4 *vel, int32_t *prop, uint32_t *id, uint32_t first, uint32_t cnt,
                          uint32_t out_base, const uint32_t *__restrict__ oy, GhostGeom g)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cnt) return;
    uint32_t src = first + j;
    double4 p = pos[src];
    uint32_t slot = out_base + oy[j];
    if (p.y < g.ngmin[1]) {
        double4 q = p;
        q.y = __dadd_rn(q.y, g.Ly);
        copy_particle(pos, vel, prop, id, slot++, src, q);
    }
    if (p.y > g.ngmax[1]) {
        double4 q = p;
        q.y = __dsub_rn(q.y, g.Ly);
        copy_particle(pos, vel, prop, id, slot, src, q);
    }
}

static uint32_t read_u32(const uint32_t *dp, hipStream_t s)
{
    uint32_t v;
    HIPCHK(hipMemcpyAsync(&v, dp, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(plan::stream_wait(s));
    return v;
}

void dev_make_ghosts(DevParticles &ps, const plan_params &pp, const plan_derived &d, hipStream_t s)
{
    const uint32_t N = ps.N;
    GhostGeom g;
    for (int k = 0; k < 3; ++k) {
        double w = d.box_half_width[k] - d.ghost_zone_width[k];
        if (w < 0) w = 0;
        g.ngmin[k] = d.box_center[k] - w;
        g.ngmax[k] = d.box_center[k] + w;
    }
    g.Lx = d.box_length[0];
    g.Ly = d.box_length[1];
    g.ymin = pp.box_min[1];
    g.ymax = pp.box_max[1];
    g.shear_t = d.shear_speed * pp.time;
    Arena &a = ps.arena;
    uint32_t *cx = a.get_t<uint32_t>("g_cx", N + 1);
    uint32_t *cy = a.get_t<uint32_t>("g_cy", N + 1);
    uint32_t *ox = a.get_t<uint32_t>("g_ox", N + 1);
    uint32_t *oy = a.get_t<uint32_t>("g_oy", N + 1);
    uint32_t *tot = a.get_t<uint32_t>("g_tot", 4);
    hipLaunchKernelGGL(k_ghost_count, dim3(ceil_div(N, BLK)), dim3(BLK), 0, s, ps.pos, N, g, cx, cy);
    LAUNCHCHK();
    scan_exclusive_u32(a, cx, ox, N, s, tot);
    scan_exclusive_u32(a, cy, oy, N, s, tot + 1);
    uint32_t tt[2];
    HIPCHK(hipMemcpyAsync(tt, tot, 8, hipMemcpyDeviceToHost, s));
    HIPCHK(plan::stream_wait(s));
    const uint32_t GX = tt[0], GYO = tt[1];
    /* capacity: the y images of the x images are bounded by 2 GX */
    uint64_t need = (uint64_t)N + GX + 2ull * GX + GYO;
    if (need >= 0xfffffffeull) throw std::runtime_error("too many particles for 32-bit tree indices");
    if (need > ps.cap) {
        /* grow, preserving the N originals */
        double4 *np = static_cast<double4 *>(nullptr);
        double4 *nv = nullptr;
        int32_t *npr = nullptr;
        uint32_t *nid = nullptr;
        HIPCHK(hipMalloc(&np, need * sizeof(double4)));
        HIPCHK(hipMalloc(&nv, need * sizeof(double4)));
        HIPCHK(hipMalloc(&npr, need * sizeof(int32_t)));
        HIPCHK(hipMalloc(&nid, need * sizeof(uint32_t)));
        HIPCHK(hipMemcpyAsync(np, ps.pos, (size_t)N * sizeof(double4), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(nv, ps.vel, (size_t)N * sizeof(double4), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(npr, ps.prop, (size_t)N * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(nid, ps.id, (size_t)N * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        HIPCHK(plan::stream_wait(s));
        HIPCHK(hipFree(ps.pos));
        HIPCHK(hipFree(ps.vel));
        HIPCHK(hipFree(ps.prop));
        HIPCHK(hipFree(ps.id));
        ps.pos = np;
        ps.vel = nv;
        ps.prop = npr;
        ps.id = nid;
        ps.cap = (uint32_t)need;
    }
    if (GX) {
        hipLaunchKernelGGL(k_ghost_x, dim3(ceil_div(N, BLK)), dim3(BLK), 0, s, ps.pos, ps.vel, ps.prop, ps.id, N, ox,
                           g);
        LAUNCHCHK();
    }
    uint32_t *cyx = a.get_t<uint32_t>("g_cyx", GX + 1);
    uint32_t *oyx = a.get_t<uint32_t>("g_oyx", GX + 1);
    uint32_t GYX = 0;
    if (GX) {
        hipLaunchKernelGGL(k_ghost_count_y, dim3(ceil_div(GX, BLK)), dim3(BLK), 0, s, ps.pos, N, GX, g, cyx);
        LAUNCHCHK();
        scan_exclusive_u32(a, cyx, oyx, GX, s, tot + 2);
        GYX = read_u32(tot + 2, s);
        hipLaunchKernelGGL(k_ghost_y, dim3(ceil_div(GX, BLK)), dim3(BLK), 0, s, ps.pos, ps.vel, ps.prop, ps.id, N,
                           GX, N + GX, oyx, g);
        LAUNCHCHK();
    }
    if (GYO) {
        hipLaunchKernelGGL(k_ghost_y, dim3(ceil_div(N, BLK)), dim3(BLK), 0, s, ps.pos, ps.vel, ps.prop, ps.id, 0u, N,
                           N + GX + GYX, oy, g);
        LAUNCHCHK();
    }
    ps.T = N + GX + GYX + GYO;
}

/* ------------------------------------------------------------ tree build */

__global__ void k_bbox(const double4 *__restrict__ pos, uint32_t T, double *__restrict__ out)
{
    __shared__ double sm[6][BLK];
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < T; i += gridDim.x * blockDim.x) {
        double4 p = pos[i];
        mn[0] = fmin(mn[0], p.x);
        mn[1] = fmin(mn[1], p.y);
        mn[2] = fmin(mn[2], p.z);
        mx[0] = fmax(mx[0], p.x);
        mx[1] = fmax(mx[1], p.y);
        mx[2] = fmax(mx[2], p.z);
    }
    for (int k = 0; k < 3; ++k) {
        sm[k][threadIdx.x] = mn[k];
        sm[3 + k][threadIdx.x] = mx[k];
    }
    __syncthreads();
    for (int o = BLK / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int k = 0; k < 3; ++k) {
                sm[k][threadIdx.x] = fmin(sm[k][threadIdx.x], sm[k][threadIdx.x + o]);
                sm[3 + k][threadIdx.x] = fmax(sm[3 + k][threadIdx.x], sm[3 + k][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[blockIdx.x * 6 + threadIdx.x] = sm[threadIdx.x][0];
}

struct MortonFrame {
    double boxmin[3], scale[3];
};

/* A particle's fields in one 64-byte record, so that the gather in key
 * order (random positions) reads one cache line per particle instead of one
 * per field array. */
struct __align__(64) GatherRec {
    double x, y, z, w;
    double vx, vy, vz;
    int32_t prop;
    uint32_t id;
};
static_assert(sizeof(GatherRec) == 64, "one 64-byte line per particle");

/* keys, and the particles packed into gather records (sequential pass) */
__global__ void k_morton(const double4 *__restrict__ pos, const double4 *__restrict__ vel,
                         const int32_t *__restrict__ prop, const uint32_t *__restrict__ pid, uint32_t T, MortonFrame f,
                         uint4 *__restrict__ keys, GatherRec *__restrict__ rec)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T) return;
    double4 p = pos[i];
    uint32_t ix = double_to_grid(__dmul_rn(__dsub_rn(p.x, f.boxmin[0]), f.scale[0]));
    uint32_t iy = double_to_grid(__dmul_rn(__dsub_rn(p.y, f.boxmin[1]), f.scale[1]));
    uint32_t iz = double_to_grid(__dmul_rn(__dsub_rn(p.z, f.boxmin[2]), f.scale[2]));
    keys[i] = morton_key(ix, iy, iz, i);
    const double4 v = vel[i];
    GatherRec r;
    r.x = p.x;
    r.y = p.y;
    r.z = p.z;
    r.w = p.w;
    r.vx = v.x;
    r.vy = v.y;
    r.vz = v.z;
    r.prop = prop[i];
    r.id = pid[i];
    rec[i] = r;
}

/* Keys sorted stably by their high 48 bits (word z and the top half of
 * word y, the first 16 octree levels) are in input order inside each run of
 * equal high parts; sorting every such run by the low 48 bits, stably, gives
 * the order of a stable sort by all 96 bits.  Runs hold particles within
 * 2^-16 of the root box width of each other and are short (about 0.2% of the
 * particles of the densest benchmark clump core share a cell); a run longer
 * than RUN_FIX_MAX is reported (err) and the caller sorts by the full key
 * instead. */
constexpr uint32_t RUN_FIX_MAX = 64;
__global__ void k_high_run_starts(const uint4 *__restrict__ keys, uint32_t T, uint32_t *__restrict__ start)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T) return;
    start[i] = (i == 0) ? 1u : (((keys[i].y >> 16) != (keys[i - 1].y >> 16) || keys[i].z != keys[i - 1].z) ? 1u : 0u);
}
__global__ void k_high_run_fix(uint4 *__restrict__ keys, uint32_t T, const uint32_t *__restrict__ start,
                               uint32_t *__restrict__ err)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T || !start[i]) return;
    uint32_t e = i + 1;
    while (e < T && !start[e] && e - i <= RUN_FIX_MAX) ++e;
    if (e - i == 1) return;
    if (e - i > RUN_FIX_MAX) {
        atomicOr(err, 1u);
        return;
    }
    /* low 48 bits: the bottom half of word y, then word x */
    auto low = [](const uint4 &k) { return ((uint64_t)(k.y & 0xffffu) << 32) | k.x; };
    for (uint32_t a = i + 1; a < e; ++a) {
        const uint4 v = keys[a];
        const uint64_t lv = low(v);
        uint32_t b = a;
        while (b > i && low(keys[b - 1]) > lv) {
            keys[b] = keys[b - 1];
            --b;
        }
        keys[b] = v;
    }
}

__global__ void k_run_flags(const uint4 *__restrict__ keys, uint32_t T, uint32_t *__restrict__ flags)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T) return;
    flags[i] = (i == 0) ? 1u : (key_eq(keys[i], keys[i - 1]) ? 0u : 1u);
}

__global__ void k_run_starts(const uint32_t *__restrict__ flags, const uint32_t *__restrict__ ridx, uint32_t T,
                             uint32_t *__restrict__ start)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T) return;
    if (flags[i]) start[ridx[i]] = i;
}

/* number of sink particles (runs of more than one equal key) */
__global__ void k_count_sinks(const uint32_t *__restrict__ sinkcnt, uint32_t n, uint32_t *__restrict__ count)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t m = __ballot(i < n && sinkcnt[i] > 1u);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(count, (uint32_t)__popcll(m));
}

/* gather the unique particles in key order; runs of equal 96-bit keys become
 * one sink particle (first of the run represents; mass and momentum summed
 * in run order) */
__global__ void k_gather_particles(const uint4 *__restrict__ skeys, const uint32_t *__restrict__ start, uint32_t n,
                                   uint32_t T, const GatherRec *__restrict__ rec, const double *__restrict__ mass_of_type, uint32_t ntypes, double4 *__restrict__ pm,
                                   double4 *__restrict__ va, uint32_t *__restrict__ oid, uint32_t *__restrict__ sinkcnt,
                                   uint32_t *__restrict__ sinkoff, uint4 *__restrict__ ukeys, double m0,
                                   uint32_t *__restrict__ nonuniform)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    uint32_t s = start[r];
    uint32_t e = (r + 1 < n) ? start[r + 1] : T;
    uint4 k = skeys[s];
    uint32_t src = k.w;
    const GatherRec g = rec[src];
    double4 p = make_double4(g.x, g.y, g.z, g.w);
    double4 v = make_double4(g.vx, g.vy, g.vz, 0.0);
    int32_t ty = g.prop;
    double m = mass_of_type[(ty >= 0 && (uint32_t)ty < ntypes) ? ty : 0];
    if (e - s > 1) {
        double ms = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
        for (uint32_t j = s; j < e; ++j) {
            const GatherRec &gq = rec[skeys[j].w];
            int32_t tq = gq.prop;
            double mq = mass_of_type[(tq >= 0 && (uint32_t)tq < ntypes) ? tq : 0];
            vx = __dadd_rn(vx, __dmul_rn(mq, gq.vx));
            vy = __dadd_rn(vy, __dmul_rn(mq, gq.vy));
            vz = __dadd_rn(vz, __dmul_rn(mq, gq.vz));
            ms = __dadd_rn(ms, mq);
        }
        m = ms;
        v.x = __ddiv_rn(vx, ms);
        v.y = __ddiv_rn(vy, ms);
        v.z = __ddiv_rn(vz, ms);
    }
    pm[r] = make_double4(p.x, p.y, p.z, m);
    if (m != m0) *nonuniform = 1u; /* any thread may set it; all write 1 */
    va[r] = make_double4(v.x, v.y, v.z, p.w);
    oid[r] = g.id;
    sinkcnt[r] = e - s;
    sinkoff[r] = s;
    ukeys[r] = make_uint4(k.x, k.y, k.z, s);
}

/* ------------------------------------------------ node array from the keys
 *
 * The level-major node array of CountNodesLeaves/FillTree
 * (src/tree.hpp:2655-2755) follows from the sorted unique keys without a
 * pass per level.  With lcp(a) the number of leading octal digits keys[a-1]
 * and keys[a] share, position a starts a level-l range (the particles that
 * share digits 1..l-1) iff a == 0 or lcp(a) < l - 1.  A level-l range holds
 * more than 8 particles iff 9 consecutive keys in it share l - 1 digits, so
 * the leaf holding particle i is at level
 *     lam(i) = min(32, 2 + max(0, max W(s))),   W(s) = lcp(keys[s], keys[s + 8])
 * over the windows s <= i <= s + 8 (the root is always split and a level-32
 * node is always a leaf).  Position a starts the nodes of levels lo(a)..lam(a)
 * (lo(0) = 1, lo(a) = lcp(a) + 2), and a node's id is the first id of its
 * level plus the rank of a among the starts of that level.  Sibling groups
 * are contiguous, so a node that is not its position's lowest takes its
 * parent from that position, and the others from the nearest preceding
 * first daughter of their level (at most 7 ids back). */
static constexpr int NB_TILE = 256;
static constexpr uint32_t NODE_PENDING = 0xfffffffeu;

/* number of leading octal digits (levels 1..32) two keys share */
__device__ __forceinline__ int key_lcp(const uint4 &a, const uint4 &b)
{
    const uint32_t z = a.z ^ b.z, y = a.y ^ b.y, x = a.x ^ b.x;
    int hb;
    if (z) hb = 95 - __clz((int)z);
    else if (y) hb = 63 - __clz((int)y);
    else if (x) hb = 31 - __clz((int)x);
    else return 32;
    return (95 - hb) / 3;
}

/* per position (lo, lam) and, per level, the number of node starts of each
 * tile of NB_TILE positions (counts laid out level-major: cnt[(l-1) nb + b]) */
__global__ __launch_bounds__(NB_TILE) void k_node_owners(const uint4 *__restrict__ keys, uint32_t n,
                                                         uint16_t *__restrict__ lolam, uint32_t *__restrict__ cnt,
                                                         uint32_t nb)
{
    __shared__ uint4 sk[NB_TILE + 16];
    __shared__ uint8_t sw[NB_TILE + 8];
    __shared__ uint32_t wc[NB_TILE / 64][33];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int64_t b0 = (int64_t)blockIdx.x * NB_TILE;
    for (int j = t; j < NB_TILE + 16; j += NB_TILE) {
        const int64_t g = b0 - 8 + j;
        sk[j] = (g >= 0 && g < (int64_t)n) ? keys[g] : make_uint4(0, 0, 0, 0);
    }
    __syncthreads();
    for (int j = t; j < NB_TILE + 8; j += NB_TILE) {
        const int64_t w = b0 - 8 + j;
        sw[j] = (w >= 0 && w + 8 < (int64_t)n) ? (uint8_t)key_lcp(sk[j], sk[j + 8]) : (uint8_t)0xff;
    }
    __syncthreads();
    const int64_t i = b0 + t;
    const bool valid = i < (int64_t)n;
    int lo = 33, lam = 0;
    if (valid) {
        int m = 0;
        for (int d = 0; d <= 8; ++d) {
            const int w = sw[t + d]; /* window s = i - 8 + d */
            if (w != 0xff) m = max(m, w);
        }
        lam = min(32, m + 2);
        lo = (i == 0) ? 1 : key_lcp(sk[t + 7], sk[t + 8]) + 2;
        lolam[i] = (uint16_t)(lo | (lam << 8));
    }
    for (int l = 1; l <= 32; ++l) {
        const uint64_t m = __ballot(valid && lo <= l && l <= lam);
        if (lane == 0) wc[wave][l] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (t < 32) {
        uint32_t c = 0;
        for (int w = 0; w < NB_TILE / 64; ++w) c += wc[w][t + 1];
        cnt[(uint64_t)t * nb + blockIdx.x] = c;
    }
}

/* first id of each level (out[l - 1]) from the scanned counts; out[32], the
 * node count, is written by the scan */
__global__ void k_level_ptrs(const uint32_t *__restrict__ off, uint32_t nb, uint32_t *__restrict__ out)
{
    const uint32_t l = threadIdx.x;
    if (l < 32) out[l] = off[(uint64_t)l * nb];
}

/* write every node: centre by the parent-to-child chain of binary64 adds from
 * the root, begin, level, orthant, first daughter, and the parent where the
 * position knows it */
__global__ __launch_bounds__(NB_TILE) void k_node_emit(const uint4 *__restrict__ keys, uint32_t n,
                                                       const uint16_t *__restrict__ lolam,
                                                       const uint32_t *__restrict__ off, uint32_t nb, NodeGeo root,
                                                       NodeGeo *__restrict__ ngeo, NodeTop *__restrict__ ntop,
                                                       uint32_t *__restrict__ ninfo)
{
    __shared__ uint32_t wc[NB_TILE / 64][33];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const uint64_t i = (uint64_t)blockIdx.x * NB_TILE + t;
    const bool valid = i < n;
    int lo = 33, lam = 0;
    uint4 key = make_uint4(0, 0, 0, 0);
    if (valid) {
        const uint32_t v = lolam[i];
        lo = (int)(v & 0xffu);
        lam = (int)(v >> 8);
        key = keys[i];
    }
    for (int l = 1; l <= 32; ++l) {
        const uint64_t m = __ballot(valid && lo <= l && l <= lam);
        if (lane == 0) wc[wave][l] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    const uint64_t lt = lanemask_lt();
    double cx = root.cx, cy = root.cy, cz = root.cz, hw = root.hw;
    uint32_t prev = PLAN_NONE;
    for (int l = 1; l <= 32; ++l) {
        uint32_t o = 0;
        if (l >= 2 && l <= lam) {
            hw = __dmul_rn(0.5, hw);
            o = key_digit(key, l - 1);
            cx = __dadd_rn(cx, (o & 1u) ? hw : -hw);
            cy = __dadd_rn(cy, (o & 2u) ? hw : -hw);
            cz = __dadd_rn(cz, (o & 4u) ? hw : -hw);
        }
        const bool own = valid && lo <= l && l <= lam;
        const uint64_t m = __ballot(own);
        if (own) {
            uint32_t pre = 0;
            for (int w = 0; w < wave; ++w) pre += wc[w][l];
            const uint32_t id = off[(uint64_t)(l - 1) * nb + blockIdx.x] + pre + (uint32_t)__popcll(m & lt);
            const uint32_t par = (l > lo) ? prev : (i == 0 ? PLAN_NONE : NODE_PENDING);
            ngeo[id] = NodeGeo{cx, cy, cz, hw};
            ntop[id] = NodeTop{(uint32_t)i, 0u, 0u, par};
            ninfo[id] = ((uint32_t)l << 8) | (o << 16);
            if (l > lo) ntop[prev].first_daughter = id;
            prev = id;
        }
    }
}

/* parent of every node: a position's lowest node takes the parent of the
 * nearest preceding first daughter of its level */
__global__ void k_node_parent(const NodeTop *__restrict__ ntop, uint32_t nn, uint32_t *__restrict__ par,
                              uint32_t *__restrict__ err)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    uint32_t p = ntop[v].parent;
    for (uint32_t u = v, k = 0; p == NODE_PENDING && u > 0 && k < 8; ++k) p = ntop[--u].parent;
    if (p == NODE_PENDING || (v != 0 && p >= nn)) {
        atomicAdd(err, 1u);
        p = 0;
    }
    par[v] = p;
}

/* end of every node (the next sibling's begin, else its parent's end) and,
 * from each last child, its parent's number of daughters */
__global__ void k_node_end(NodeTop *__restrict__ ntop, const uint32_t *__restrict__ par, uint32_t nn, uint32_t n,
                           uint32_t *__restrict__ nd)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    uint32_t x = v, e = n;
    for (int k = 0; k < 40 && x != 0; ++k) {
        if (x + 1 < nn && par[x + 1] == par[x]) {
            e = ntop[x + 1].begin;
            break;
        }
        x = par[x];
    }
    const uint32_t pv = par[v];
    ntop[v].end = e;
    ntop[v].parent = pv;
    if (v != 0 && (v + 1 >= nn || par[v + 1] != pv)) nd[pv] = v + 1 - ntop[pv].first_daughter;
}

__global__ void k_node_daughters(const NodeTop *__restrict__ ntop, const uint32_t *__restrict__ nd, uint32_t nn,
                                 uint32_t *__restrict__ ninfo)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    if (ntop[v].first_daughter != 0) ninfo[v] |= nd[v];
}

/* skip pointer: first node after the subtree in pre-order */
__global__ void k_skip(const NodeTop *__restrict__ ntop, const uint32_t *__restrict__ ninfo, uint32_t nn,
                       uint32_t *__restrict__ skip)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    uint32_t x = v, res = PLAN_NONE;
    for (int up = 0; x != 0 && up < 40; ++up) { /* at most 32 levels */
        uint32_t P = ntop[x].parent;
        uint32_t fd = ntop[P].first_daughter, nd = ninfo[P] & 0xffu;
        if (x + 1 < fd + nd) {
            res = x + 1;
            break;
        }
        x = P;
    }
    skip[v] = res;
}

__global__ void k_leaf_mark(const NodeTop *__restrict__ ntop, const uint32_t *__restrict__ ninfo, uint32_t nn,
                            uint32_t *__restrict__ leaf_of, uint32_t *__restrict__ lflag)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn || v == 0) return;
    if ((ninfo[v] & 0xffu) != 0) return;
    NodeTop t = ntop[v];
    lflag[t.begin] = 1u;
    for (uint32_t i = t.begin; i < t.end; ++i) leaf_of[i] = v;
}

__global__ void k_leaf_index(const NodeTop *__restrict__ ntop, const uint32_t *__restrict__ ninfo, uint32_t nn,
                             const uint32_t *__restrict__ lidx, int32_t *__restrict__ node2leaf,
                             int32_t *__restrict__ leaf_nodes)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    bool leaf = v != 0 && (ninfo[v] & 0xffu) == 0;
    if (!leaf) {
        node2leaf[v] = -1;
        return;
    }
    uint32_t k = lidx[ntop[v].begin];
    node2leaf[v] = (int32_t)k;
    leaf_nodes[k] = (int32_t)v;
}

void dev_build_tree(DevTree &t, const DevParticles &ps, const plan_derived &d, bool check_pos,
                    const double *root_center_in, double half_width_in, hipStream_t s)
{
    const uint32_t T = ps.T;
    if (T == 0) throw std::runtime_error("BuildTree: no particles");
    Arena &a = t.arena;
    /* root box (src/tree.hpp:2763-2803) */
    if (check_pos) {
        unsigned nb = std::min<unsigned>(1024, ceil_div(T, BLK));
        double *bb = a.get_t<double>("bbox", (size_t)nb * 6);
        hipLaunchKernelGGL(k_bbox, dim3(nb), dim3(BLK), 0, s, ps.pos, T, bb);
        LAUNCHCHK();
        std::vector<double> hb((size_t)nb * 6);
        HIPCHK(hipMemcpyAsync(hb.data(), bb, hb.size() * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(plan::stream_wait(s));
        double pmin[3], pmax[3];
        for (int k = 0; k < 3; ++k) {
            pmin[k] = hb[k];
            pmax[k] = hb[3 + k];
        }
        for (unsigned b = 1; b < nb; ++b)
            for (int k = 0; k < 3; ++k) {
                pmin[k] = std::min(hb[b * 6 + k], pmin[k]);
                pmax[k] = std::max(hb[b * 6 + 3 + k], pmax[k]);
            }
        double extent = 0.0;
        for (int k = 0; k < 3; ++k) extent = std::max(extent, pmax[k] - pmin[k]);
        t.half_width = d.max_half_width + d.max_ghost_zone_width;
        for (int k = 0; k < 3; ++k) t.root_center[k] = d.box_center[k];
        if (extent < t.half_width) {
            t.half_width = (std::ceil(extent * 100.) + 1) / 100.;
            for (int k = 0; k < 3; ++k) {
                if (pmax[k] - pmin[k] < 0.5 * d.box_length[k]) {
                    t.root_center[k] = (pmax[k] + pmin[k]) / 2.0;
                    t.root_center[k] = std::round(t.root_center[k] * 1000.) / 1000.;
                } else {
                    t.root_center[k] = d.box_center[k];
                }
            }
        }
    } else {
        t.half_width = half_width_in;
        for (int k = 0; k < 3; ++k) t.root_center[k] = root_center_in[k];
        if (t.half_width < 1e-16) {
            t.half_width = d.max_half_width + d.max_ghost_zone_width;
            for (int k = 0; k < 3; ++k) t.root_center[k] = d.box_center[k];
        }
    }
    MortonFrame f;
    for (int k = 0; k < 3; ++k) {
        volatile double lo = t.root_center[k] - t.half_width;
        volatile double hi = t.root_center[k] + t.half_width;
        t.boxmin[k] = lo;
        t.scale[k] = 1.0 / (hi - lo);
        f.boxmin[k] = t.boxmin[k];
        f.scale[k] = t.scale[k];
    }
    t.epsilon = 2 * t.half_width / std::pow(2, 31);

    /* keys and sort */
    uint4 *k0 = a.get_t<uint4>("keys_a", T);
    uint4 *k1 = a.get_t<uint4>("keys_b", T);
    GatherRec *grec = a.get_t<GatherRec>("gather_rec", T);
    hipLaunchKernelGGL(k_morton, dim3(ceil_div(T, BLK)), dim3(BLK), 0, s, ps.pos, ps.vel, ps.prop, ps.id, T, f, k0,
                       grec);
    LAUNCHCHK();
    /* sort by the high 48 bits, then each run of equal high parts by the
     * low 48 bits */
    uint32_t *flags = a.get_t<uint32_t>("flags", T + 1);
    uint32_t *ridx = a.get_t<uint32_t>("ridx", T + 1);
    uint32_t *tot = a.get_t<uint32_t>("tot", 8);
    HIPCHK(hipMemsetAsync(tot + 1, 0, 4, s));
    radix_sort_96(a, &k0, &k1, T, 3, s, 6);
    hipLaunchKernelGGL(k_high_run_starts, dim3(ceil_div(T, BLK)), dim3(BLK), 0, s, k0, T, flags);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_high_run_fix, dim3(ceil_div(T, BLK)), dim3(BLK), 0, s, k0, T, flags, tot + 1);
    LAUNCHCHK();
    uint4 *skeys = k0;

    /* sink runs */
    hipLaunchKernelGGL(k_run_flags, dim3(ceil_div(T, BLK)), dim3(BLK), 0, s, skeys, T, flags);
    LAUNCHCHK();
    scan_exclusive_u32(a, flags, ridx, T, s, tot);
    uint32_t h2[2];
    HIPCHK(hipMemcpyAsync(h2, tot, sizeof(h2), hipMemcpyDeviceToHost, s));
    HIPCHK(plan::stream_wait(s));
    if (h2[1]) {
        /* a long run of equal high halves: stable sort by the full key (equal
         * full keys are in input order already) */
        radix_sort_96(a, &k0, &k1, T, 3, s);
        skeys = k0;
        hipLaunchKernelGGL(k_run_flags, dim3(ceil_div(T, BLK)), dim3(BLK), 0, s, skeys, T, flags);
        LAUNCHCHK();
        scan_exclusive_u32(a, flags, ridx, T, s, tot);
        HIPCHK(hipMemcpyAsync(h2, tot, sizeof(h2), hipMemcpyDeviceToHost, s));
        HIPCHK(plan::stream_wait(s));
    }
    const uint32_t n = h2[0];
    t.n = n;
    t.skeys = skeys;
    t.num_sinks = 0;
    uint32_t *start = a.get_t<uint32_t>("rstart", n + 1);
    hipLaunchKernelGGL(k_run_starts, dim3(ceil_div(T, BLK)), dim3(BLK), 0, s, flags, ridx, T, start);
    LAUNCHCHK();
    double *mot = a.get_t<double>("mass_of_type", PLAN_MAX_TYPES);
    HIPCHK(hipMemcpyAsync(mot, d.mass_per_particle, sizeof(double) * PLAN_MAX_TYPES, hipMemcpyHostToDevice, s));
    t.pm = a.get_t<double4>("pm", n);
    t.va = a.get_t<double4>("va", n);
    t.rho = a.get_t<double>("rho", n);
    t.oid = a.get_t<uint32_t>("oid", n);
    t.sinkcnt = a.get_t<uint32_t>("sinkcnt", n);
    t.sinkoff = a.get_t<uint32_t>("sinkoff", n);
    t.keys = a.get_t<uint4>("ukeys", n);
    HIPCHK(hipMemsetAsync(tot + 1, 0, 8, s));
    hipLaunchKernelGGL(k_gather_particles, dim3(ceil_div(n, BLK)), dim3(BLK), 0, s, skeys, start, n, T, grec, mot,
                       (uint32_t)PLAN_MAX_TYPES, t.pm, t.va, t.oid, t.sinkcnt,
                       t.sinkoff, t.keys, d.mass_per_particle[0], tot + 2);
    LAUNCHCHK();
    HIPCHK(hipMemsetAsync(t.rho, 0, (size_t)n * 8, s));
    if (n != T) {
        hipLaunchKernelGGL(k_count_sinks, dim3(ceil_div(n, BLK)), dim3(BLK), 0, s, t.sinkcnt, n, tot + 1);
        LAUNCHCHK();
    }
    {
        uint32_t h[2];
        HIPCHK(hipMemcpyAsync(h, tot + 1, sizeof(h), hipMemcpyDeviceToHost, s));
        HIPCHK(plan::stream_wait(s));
        t.num_sinks = h[0];
        t.mass_uniform = h[1] == 0u;
    }

    /* node array straight from the sorted keys (grow-only node arrays) */
    const uint32_t nbn = ceil_div(n, NB_TILE);
    uint16_t *lolam = a.get_t<uint16_t>("lolam", n);
    uint32_t *ncnt = a.get_t<uint32_t>("node_cnt", (size_t)32 * nbn);
    uint32_t *noff = a.get_t<uint32_t>("node_off", (size_t)32 * nbn);
    uint32_t *lptr = a.get_t<uint32_t>("node_lptr", 40);
    hipLaunchKernelGGL(k_node_owners, dim3(nbn), dim3(NB_TILE), 0, s, t.keys, n, lolam, ncnt, nbn);
    LAUNCHCHK();
    scan_exclusive_u32(a, ncnt, noff, 32 * nbn, s, lptr + 32);
    hipLaunchKernelGGL(k_level_ptrs, dim3(1), dim3(64), 0, s, noff, nbn, lptr);
    LAUNCHCHK();
    HIPCHK(hipMemsetAsync(lptr + 33, 0, 4, s));
    uint32_t hl[33];
    HIPCHK(hipMemcpyAsync(hl, lptr, sizeof(hl), hipMemcpyDeviceToHost, s));
    HIPCHK(plan::stream_wait(s));
    const uint32_t nn = hl[32];
    if (t.node_cap < (size_t)nn) {
        if (t.ngeo) HIPCHK(hipFree(t.ngeo));
        if (t.ntop) HIPCHK(hipFree(t.ntop));
        if (t.ninfo) HIPCHK(hipFree(t.ninfo));
        t.ngeo = nullptr;
        t.ntop = nullptr;
        t.ninfo = nullptr;
        const size_t cap = std::max<size_t>(nn, (size_t)n + n / 2 + 64);
        HIPCHK(hipMalloc(&t.ngeo, cap * sizeof(NodeGeo)));
        HIPCHK(hipMalloc(&t.ntop, cap * sizeof(NodeTop)));
        HIPCHK(hipMalloc(&t.ninfo, cap * sizeof(uint32_t)));
        t.node_cap = cap;
    }
    NodeGeo *ngeo = t.ngeo;
    NodeTop *ntop = t.ntop;
    uint32_t *ninfo = t.ninfo;
    const NodeGeo rg{t.root_center[0], t.root_center[1], t.root_center[2], t.half_width};
    hipLaunchKernelGGL(k_node_emit, dim3(nbn), dim3(NB_TILE), 0, s, t.keys, n, lolam, noff, nbn, rg, ngeo, ntop,
                       ninfo);
    LAUNCHCHK();
    uint32_t *npar = a.get_t<uint32_t>("node_par", nn);
    uint32_t *ndau = a.get_t<uint32_t>("node_nd", nn);
    hipLaunchKernelGGL(k_node_parent, dim3(ceil_div(nn, BLK)), dim3(BLK), 0, s, ntop, nn, npar, lptr + 33);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_node_end, dim3(ceil_div(nn, BLK)), dim3(BLK), 0, s, ntop, npar, nn, n, ndau);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_node_daughters, dim3(ceil_div(nn, BLK)), dim3(BLK), 0, s, ntop, ndau, nn, ninfo);
    LAUNCHCHK();
    t.depth = 1;
    for (int l = 0; l < 32; ++l) {
        t.level_ptr[l] = (int)hl[l];
        t.level_count[l] = (int)((l + 1 < 32 ? hl[l + 1] : nn) - hl[l]);
        if (t.level_count[l] > 0) t.depth = l + 1;
    }
    t.num_nodes = nn;
    uint32_t *cidx = a.get_t<uint32_t>("cidx", n + 1);
    t.skip = a.get_t<uint32_t>("skip", nn);
    t.leaf_of = a.get_t<uint32_t>("leaf_of", n);
    t.node2leaf = a.get_t<int32_t>("node2leaf", nn);
    hipLaunchKernelGGL(k_skip, dim3(ceil_div(nn, BLK)), dim3(BLK), 0, s, ntop, ninfo, nn, t.skip);
    LAUNCHCHK();
    HIPCHK(hipMemsetAsync(flags, 0, (size_t)n * 4, s));
    hipLaunchKernelGGL(k_leaf_mark, dim3(ceil_div(nn, BLK)), dim3(BLK), 0, s, ntop, ninfo, nn, t.leaf_of, flags);
    LAUNCHCHK();
    scan_exclusive_u32(a, flags, cidx, n, s, tot);
    t.num_leaves = read_u32(tot, s);
    t.leaf_nodes = a.get_t<int32_t>("leaf_nodes", t.num_leaves + 1);
    hipLaunchKernelGGL(k_leaf_index, dim3(ceil_div(nn, BLK)), dim3(BLK), 0, s, ntop, ninfo, nn, cidx, t.node2leaf,
                       t.leaf_nodes);
    LAUNCHCHK();
    if (read_u32(lptr + 33, s) != 0) throw std::runtime_error("BuildTree: inconsistent node array");
}

}  // namespace plan
