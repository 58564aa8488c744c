// scan_sort.hip — device prefix sum and stable LSD radix sort (gfx950, wave64).
#include "devmem.h"

namespace plan {

static constexpr int SCAN_BLOCK = 256;
static constexpr int SCAN_ITEMS = 8;
static constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

/* exclusive block scan of one value per thread (256 threads); returns the
 * block total in *total */
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *lds4, uint32_t *total)
{
    const int lane = __lane_id();
    const int wave = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) lds4[wave] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < SCAN_BLOCK / 64; ++w) {
        uint32_t s = lds4[w];
        if (w < wave) off += s;
        tot += s;
    }
    *total = tot;
    __syncthreads();
    return off + inc - v;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(const uint32_t *__restrict__ in, uint32_t n,
                                                            uint32_t *__restrict__ sums)
{
    __shared__ uint32_t lds4[4];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        uint64_t j = base + i;
        if (j < n) s += in[j];
    }
    uint32_t tot;
    (void)block_excl_scan(s, lds4, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_apply(const uint32_t *__restrict__ in,
                                                           uint32_t *__restrict__ out, uint32_t n,
                                                           const uint32_t *__restrict__ block_off,
                                                           uint32_t *__restrict__ total)
{
    __shared__ uint32_t lds4[4];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        uint64_t j = base + i;
        v[i] = (j < n) ? in[j] : 0u;
        s += v[i];
    }
    uint32_t tot;
    uint32_t off = block_excl_scan(s, lds4, &tot);
    if (block_off) off += block_off[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        uint64_t j = base + i;
        if (j < n) out[j] = off;
        off += v[i];
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_BLOCK - 1) *total = off;
}

static void scan_rec(Arena &a, const uint32_t *in, uint32_t *out, uint32_t n, hipStream_t s, uint32_t *total,
                     int depth)
{
    unsigned nb = ceil_div(n, SCAN_TILE);
    if (nb <= 1) {
        hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(SCAN_BLOCK), 0, s, in, out, n, nullptr, total);
        LAUNCHCHK();
        return;
    }
    std::string tag = "scan_sums" + std::to_string(depth);
    uint32_t *sums = a.get_t<uint32_t>(tag, nb);
    uint32_t *soff = a.get_t<uint32_t>(tag + "o", nb);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_BLOCK), 0, s, in, n, sums);
    LAUNCHCHK();
    scan_rec(a, sums, soff, nb, s, nullptr, depth + 1);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_BLOCK), 0, s, in, out, n, soff, total);
    LAUNCHCHK();
}

void scan_exclusive_u32(Arena &a, const uint32_t *in, uint32_t *out, uint32_t n, hipStream_t s,
                        uint32_t *total_dev)
{
    if (n == 0) {
        if (total_dev) HIPCHK(hipMemsetAsync(total_dev, 0, 4, s));
        return;
    }
    scan_rec(a, in, out, n, s, total_dev, 0);
}

/* ----------------------------------------------------------- radix sort */

static constexpr int RS_BLOCK = 256;
static constexpr int RS_ROUNDS = 8;
static constexpr int RS_TILE = RS_BLOCK * RS_ROUNDS;

__device__ __forceinline__ uint32_t rs_digit(const uint4 &k, int d)
{
    uint32_t w = (d < 4) ? k.x : (d < 8 ? k.y : k.z);
    return (w >> ((d & 3) * 8)) & 0xffu;
}

__global__ __launch_bounds__(RS_BLOCK) void k_rs_hist(const uint4 *__restrict__ keys, uint32_t n, int d,
                                                      uint32_t *__restrict__ counts, uint32_t nb)
{
    __shared__ uint32_t hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int r = 0; r < RS_ROUNDS; ++r) {
        uint64_t j = base + (uint64_t)r * RS_BLOCK + threadIdx.x;
        if (j < n) atomicAdd(&hist[rs_digit(keys[j], d)], 1u);
    }
    __syncthreads();
    counts[(uint64_t)threadIdx.x * nb + blockIdx.x] = hist[threadIdx.x];
}

/* One stable counting-sort pass over a tile of RS_TILE keys: the keys are
 * ranked by digit in LDS (tile order kept within a digit), then written out
 * digit run by digit run, so consecutive threads write consecutive global
 * positions. */
__global__ __launch_bounds__(RS_BLOCK) void k_rs_scatter(const uint4 *__restrict__ keys, uint4 *__restrict__ out,
                                                         uint32_t n, int d, const uint32_t *__restrict__ offs,
                                                         uint32_t nb)
{
    __shared__ uint4 tile[RS_TILE];
    __shared__ uint32_t lbase[256];
    __shared__ uint32_t gbase[256];
    __shared__ uint32_t running[256];
    __shared__ uint32_t wsum[RS_BLOCK / 64];
    __shared__ uint32_t wcount[RS_BLOCK / 64][256];
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63;
    const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
    const uint32_t cnt = (uint32_t)((uint64_t)n - base < (uint64_t)RS_TILE ? (uint64_t)n - base : (uint64_t)RS_TILE);
    running[t] = 0;
#pragma unroll
    for (int w = 0; w < RS_BLOCK / 64; ++w) wcount[w][t] = 0;
    __syncthreads();
    uint4 kr[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const uint32_t i = (uint32_t)r * RS_BLOCK + t;
        kr[r] = i < cnt ? keys[base + i] : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r)
        if ((uint32_t)r * RS_BLOCK + t < cnt) atomicAdd(&running[rs_digit(kr[r], d)], 1u);
    __syncthreads();
    /* exclusive prefix of the tile's digit counts */
    const uint32_t c = running[t];
    uint32_t incl = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t wpre = 0;
    for (int w = 0; w < wave; ++w) wpre += wsum[w];
    lbase[t] = wpre + incl - c;
    gbase[t] = offs[(uint64_t)t * nb + blockIdx.x];
    running[t] = 0;
    __syncthreads();
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const bool valid = (uint32_t)r * RS_BLOCK + t < cnt;
        const uint4 k = kr[r];
        const uint32_t dig = valid ? rs_digit(k, d) : 256u;
        uint64_t peers = __ballot(1);
#pragma unroll
        for (int b = 0; b < 9; ++b) {
            const uint64_t m = __ballot((dig >> b) & 1u);
            peers &= ((dig >> b) & 1u) ? m : ~m;
        }
        const uint32_t rank = __popcll(peers & lt);
        if (valid && rank == 0) wcount[wave][dig] = __popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pre = 0;
            for (int w = 0; w < wave; ++w) pre += wcount[w][dig];
            tile[lbase[dig] + running[dig] + pre + rank] = k;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < RS_BLOCK / 64; ++w) {
            add += wcount[w][t];
            wcount[w][t] = 0;
        }
        running[t] += add;
        __syncthreads();
    }
    for (uint32_t i = t; i < cnt; i += RS_BLOCK) {
        const uint4 k = tile[i];
        const uint32_t v = rs_digit(k, d);
        out[gbase[v] + (i - lbase[v])] = k;
    }
}

/* One-workgroup bitonic sort of up to SMALL96_N records by (w2, w1, w0,
 * payload).  Every caller's payload is the record's input position, so this
 * order is the stable LSD sort's, for any key words the caller sorts on:
 * ordering by the further words too only fixes what a later pass or fix-up
 * would.  Small sorts (clump trees, small snapshots) then cost one launch
 * instead of three per digit. */
constexpr uint32_t SMALL96_N = 2048, SMALL96_T = 1024;
__device__ __forceinline__ bool key96_less(const uint4 &a, const uint4 &b)
{
    if (a.z != b.z) return a.z < b.z;
    if (a.y != b.y) return a.y < b.y;
    if (a.x != b.x) return a.x < b.x;
    return a.w < b.w;
}
__global__ __launch_bounds__(SMALL96_T) void k_small_sort96(uint4 *__restrict__ keys, uint32_t n, uint32_t p)
{
    __shared__ uint4 sk[SMALL96_N];
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < p; i += SMALL96_T)
        sk[i] = i < n ? keys[i] : make_uint4(0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu);
    __syncthreads();
    for (uint32_t k = 2; k <= p; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = t; i < p; i += SMALL96_T) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const uint4 x = sk[i], y = sk[l];
                    if (key96_less(y, x) == ((i & k) == 0)) {
                        sk[i] = y;
                        sk[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    for (uint32_t i = t; i < n; i += SMALL96_T) keys[i] = sk[i];
}

void radix_sort_96(Arena &a, uint4 **keys, uint4 **tmp, uint32_t n, int nwords, hipStream_t s, int first_digit)
{
    if (n <= 1) return;
    if (n <= SMALL96_N) {
        uint32_t p = 2;
        while (p < n) p <<= 1;
        hipLaunchKernelGGL(k_small_sort96, dim3(1), dim3(SMALL96_T), 0, s, *keys, n, p);
        LAUNCHCHK();
        return;
    }
    unsigned nb = ceil_div(n, RS_TILE);
    uint32_t *counts = a.get_t<uint32_t>("rs_counts", (size_t)256 * nb);
    uint32_t *offs = a.get_t<uint32_t>("rs_offs", (size_t)256 * nb);
    int nd = nwords * 4;
    for (int d = first_digit; d < nd; ++d) {
        hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(RS_BLOCK), 0, s, *keys, n, d, counts, nb);
        LAUNCHCHK();
        scan_exclusive_u32(a, counts, offs, 256 * nb, s);
        hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(RS_BLOCK), 0, s, *keys, *tmp, n, d, offs, nb);
        LAUNCHCHK();
        std::swap(*keys, *tmp);
    }
}

}  // namespace plan
