// common.h — shared definitions for the gfx950 clump-finding library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <chrono>
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

/* Waits for the work queued on s.  The wait is on an event (one per thread
 * and device) recorded behind that work and created with
 * hipEventBlockingSync: it polls for the first SPIN_US microseconds (the
 * short kernels of ball searches and scans finish within them, and a
 * sleeping thread wakes late when every host core is busy), then sleeps
 * until the device signals.  hipStreamSynchronize spins for the whole wait,
 * a host core taken from the clump logic that another context runs
 * meanwhile.  A stream of another device than the thread's current one
 * takes hipStreamSynchronize. */
inline hipError_t stream_wait(hipStream_t s)
{
    constexpr int MAX_DEV = 16;
    constexpr double SPIN_US = 200.0;
    thread_local hipEvent_t ev[MAX_DEV] = {};
    int cur = 0;
    hipDevice_t dev = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    e = hipStreamGetDevice(s, &dev);
    if (e != hipSuccess) return e;
    if (dev != cur || cur < 0 || cur >= MAX_DEV) return hipStreamSynchronize(s);
    if (!ev[cur]) {
        e = hipEventCreateWithFlags(&ev[cur], hipEventBlockingSync | hipEventDisableTiming);
        if (e != hipSuccess) {
            ev[cur] = nullptr;
            return e;
        }
    }
    e = hipEventRecord(ev[cur], s);
    if (e != hipSuccess) return e;
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        e = hipEventQuery(ev[cur]);
        if (e != hipErrorNotReady) return e;
        if (std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count() > SPIN_US) break;
    }
    return hipEventSynchronize(ev[cur]);
}

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
