// common.h — shared definitions for the gfx950 clump-finding library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <stdexcept>

#define PLAN_NONE 0xffffffffu

namespace plan {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPCHK(expr)                                                                    \
    do {                                                                                \
        hipError_t e__ = (expr);                                                        \
        if (e__ != hipSuccess)                                                          \
            throw ::plan::HipError(std::string(#expr) + ": " + hipGetErrorString(e__)); \
    } while (0)

#define LAUNCHCHK() HIPCHK(hipGetLastError())

static inline unsigned ceil_div(uint64_t a, uint64_t b) { return (unsigned)((a + b - 1) / b); }

/* Node record used by the traversal kernels: geometry in one 32-byte load
 * and topology in one 16-byte load. */
struct NodeGeo {
    double cx, cy, cz, hw;
};
struct NodeTop {
    uint32_t begin, end, first_daughter, parent;
};

}  // namespace plan

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)
namespace plan {
__device__ __forceinline__ uint64_t lanemask_lt()
{
    uint32_t lane = __lane_id();
    return lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
}
__device__ __forceinline__ int wave_first_lane(uint64_t mask) { return __ffsll((unsigned long long)mask) - 1; }
__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double bcast_f64(double v, int lane)
{
    return __shfl(v, lane, 64);
}
}  // namespace plan
#endif
