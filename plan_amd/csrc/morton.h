// morton.h — 96-bit Morton keys (host and device).
//
// Key layout follows the reference (src/tree.hpp:2315-2336, src/tree.cpp:106-174):
// coordinate d of the integer position contributes its bit b at key bit
// 3b + d; the scaled coordinate maps to 32 bits by rounding
// d * (2^32 - 2) to nearest-even (the reference adds 1.5 * 2^52 and keeps the
// low word).  Here the 96-bit key is carried as three 32-bit words w0 (bits
// 0-31), w1 (32-63), w2 (64-95).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plan {

__host__ __device__ __forceinline__ uint32_t double_to_grid(double s)
{
    /* two roundings, as the reference: product to binary64, then the sum with
     * 1.5 * 2^52 rounds the product to an integer (ties to even) */
#if defined(__HIP_DEVICE_COMPILE__)
    double t = __dadd_rn(__dmul_rn(s, 4294967294.0), 6755399441055744.0);
    return (uint32_t)(uint64_t)__double_as_longlong(t);
#else
    volatile double p = s * 4294967294.0;
    double t = p + 6755399441055744.0;
    uint64_t b;
    __builtin_memcpy(&b, &t, 8);
    return (uint32_t)b;
#endif
}

/* spread the low 21 bits of x to every third bit (bit b -> bit 3b) */
__host__ __device__ __forceinline__ uint64_t spread3_21(uint64_t x)
{
    x &= 0x1fffffull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

/* dilate a 32-bit value to 96 bits: returns bits 0-63 in lo, 64-95 in hi */
__host__ __device__ __forceinline__ void dilate3_96(uint32_t v, uint64_t &lo, uint64_t &hi)
{
    lo = spread3_21(v) | ((uint64_t)((v >> 21) & 1u) << 63);
    hi = spread3_21(v >> 22) << 2;
}

__host__ __device__ __forceinline__ uint4 morton_key(uint32_t ix, uint32_t iy, uint32_t iz, uint32_t payload)
{
    uint64_t lo = 0, hi = 0, l, h;
    dilate3_96(ix, l, h);
    lo |= l;
    hi |= h;
    dilate3_96(iy, l, h);
    lo |= l << 1;
    hi |= (h << 1) | (l >> 63);
    dilate3_96(iz, l, h);
    lo |= l << 2;
    hi |= (h << 2) | (l >> 62);
    return make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, payload);
}

/* octant digit of a key at tree level L (1..32): bits 93-3(L-1) .. +2
 * (src/tree.hpp:2242-2245) */
__host__ __device__ __forceinline__ uint32_t key_digit(const uint4 &k, int level)
{
    int s = 93 - 3 * (level - 1);
    uint64_t lo = ((uint64_t)k.y << 32) | k.x;
    uint64_t hi = k.z;
    uint64_t v;
    if (s >= 64) v = hi >> (s - 64);
    else if (s == 0) v = lo;
    else v = (lo >> s) | (hi << (64 - s));
    return (uint32_t)(v & 7u);
}

__host__ __device__ __forceinline__ bool key_eq(const uint4 &a, const uint4 &b)
{
    return a.x == b.x && a.y == b.y && a.z == b.z;
}

}  // namespace plan
