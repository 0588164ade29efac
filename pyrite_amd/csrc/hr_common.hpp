// hr_common.hpp -- shared helpers of libhiprite: error state, launch checks,
// counter-addressed SplitMix64 streams on the device.
#pragma once
#include <rocprofiler-sdk-roctx/roctx.h>

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>

#include "hiprite.h"

namespace hr {

// Thread-local text of the last failure, returned by hr_last_error().
void set_error(const char *fmt, ...);
const char *last_error();

#define HR_CHECK_HIP(expr)                                                           \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) {                                                      \
            ::hr::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),   \
                            __FILE__, __LINE__);                                     \
            return _e == hipErrorOutOfMemory ? HR_ERR_OOM : HR_ERR_HIP;              \
        }                                                                            \
    } while (0)

#define HR_REQUIRE(cond, ...)                 \
    do {                                      \
        if (!(cond)) {                        \
            ::hr::set_error(__VA_ARGS__);     \
            return HR_ERR_INVALID;            \
        }                                     \
    } while (0)

// A roctx range over an entry point's host work (validation and the launches
// it enqueues), named after the entry point: `rocprofv3 --marker-trace` shows
// it beside the kernels of the call.
struct TraceRange {
    explicit TraceRange(const char *name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange &) = delete;
    TraceRange &operator=(const TraceRange &) = delete;
};
#define HR_TRACE(name) ::hr::TraceRange hr_trace_range_(name)

// Constants of the counter-addressed streams (kinematics.py:181-193).
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;
constexpr uint64_t kMix1 = 0xBF58476D1CE4E5B9ULL;
constexpr uint64_t kMix2 = 0x94D049BB133111EBULL;
constexpr uint64_t kUrbanSalt = 0xD6E8FEB86659FD93ULL;  // straggling.py:33
constexpr double kU53 = 1.0 / 9007199254740992.0;

__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x = (x ^ (x >> 30)) * kMix1;
    x = (x ^ (x >> 27)) * kMix2;
    return x ^ (x >> 31);
}

// Draw `counter` of stream `key`, in [0, 1); the 53-bit value converts exactly.
__host__ __device__ __forceinline__ double stream_uniform(uint64_t key, uint64_t counter) {
    const uint64_t z = splitmix64(key + kGolden * (counter + 1ULL));
    return (double)(z >> 11) * kU53;
}

// Particle owning compact row r of slot rows laid out by an exclusive prefix
// sum of per-particle row counts: the last i with row_offset[i] <= r
// (particles without rows share the next one's offset and are passed over).
// row_offset[0] <= r < row_offset[n-1] + count[n-1] is the caller's range.
__device__ __forceinline__ int64_t row_owner(const int64_t *__restrict__ row_offset, int64_t n, int64_t r) {
    int64_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (row_offset[mid] <= r) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Compute units of the current device (a host-side property cache; no
// allocation, no launch).
hr_status device_cu(int &cu);

// A caller-owned workspace holds at least `need` bytes, 256-byte aligned.
#define HR_REQUIRE_WS(ws, need, who)                                                              \
    HR_REQUIRE((need) == 0 || ((ws).ptr != nullptr && (ws).bytes >= (size_t)(need) &&             \
                               ((uintptr_t)(ws).ptr & 255u) == 0),                                \
               "%s: workspace of %zu bytes at %p is smaller than the %zu it needs or misaligned "  \
               "(query the *_workspace_bytes entry point)",                                      \
               who, (size_t)(ws).bytes, (ws).ptr, (size_t)(need))

inline int pick_block(int requested, int fallback) {
    if (requested <= 0) return fallback;
    if (requested > 1024) return 1024;
    return (requested + 63) / 64 * 64;  // whole wave64 wavefronts
}

}  // namespace hr
