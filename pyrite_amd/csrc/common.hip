// common.hip -- error state, version and device-side stream keys of libhiprite.
#include <cstdarg>
#include <cstdio>

#include "hr_common.hpp"

namespace hr {
namespace {
thread_local char g_error[1024] = "";
}

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

const char *last_error() { return g_error; }

namespace {
int g_cu[64];  // compute units per device, filled on first use
}

hr_status device_cu(int &cu) {
    int dev = 0;
    HR_CHECK_HIP(hipGetDevice(&dev));
    HR_REQUIRE(dev >= 0 && dev < 64, "device index %d out of range", dev);
    if (g_cu[dev] == 0) {
        int n = 0;
        HR_CHECK_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
        HR_REQUIRE(n > 0, "device %d reports %d compute units", dev, n);
        g_cu[dev] = n;
    }
    cu = g_cu[dev];
    return HR_OK;
}

namespace {
// key[e] = splitmix64(seed + PHI*(start + e + 1))  (kinematics.py:226-245)
__global__ void __launch_bounds__(256) stream_keys_kernel(uint64_t seed, int64_t start, int64_t n,
                                                          uint64_t *__restrict__ keys) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = splitmix64(seed + kGolden * ((uint64_t)(start + i) + 1ULL));
}
}  // namespace
}  // namespace hr

extern "C" {

const char *hr_last_error(void) { return hr::last_error(); }

const char *hr_version(void) { return "hiprite 0.1.0 gfx950"; }

int32_t hr_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

hr_status hr_stream_keys(uint64_t seed, int64_t start, int64_t n, uint64_t *keys_d, void *stream) {
    HR_TRACE("hr_stream_keys");
    HR_REQUIRE(n >= 0 && start >= 0, "hr_stream_keys: negative start or count");
    if (n == 0) return HR_OK;
    HR_REQUIRE(keys_d != nullptr, "hr_stream_keys: null output");
    const int block = 256;
    const int64_t grid = (n + block - 1) / block;
    hipLaunchKernelGGL(hr::stream_keys_kernel, dim3((unsigned)grid), dim3(block), 0,
                       (hipStream_t)stream, seed, start, n, keys_d);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}
}
