// tally.hip -- depth-dose tally of transport rows on gfx950.
//
// Every row loses E - E_end along its flight; the loss is spread over the
// depth bins the row crosses in proportion to the depth it covers there (the
// continuous-slowing-down picture of the condensed-history step), and a
// history that ends at the cutoff deposits its residual E_end at the row's
// end.  With the energy carried out by escaping electrons this closes the
// energy balance exactly: sum(dose) + sum(escaped E) = Ne E0.
//
// One thread per slot (particle i = t / rows, row s = t % rows), reading the
// SoA row columns coalesced; a block accumulates its rows into an LDS
// histogram (FP64 LDS atomics) and adds it to the output with one global
// atomic per bin.
#include "hr_common.hpp"

namespace hr {
namespace {

constexpr int kDoseThreads = 256;
constexpr int kDoseMaxBins = 8192;  // 64 KB of FP64 LDS
constexpr int8_t kEventCutoff = 7;  // montecarlo/transport/events.py:43

__global__ void __launch_bounds__(kDoseThreads) depth_dose_kernel(hr_dose_args p) {
    __shared__ double hist[kDoseMaxBins];
    const int nb = p.n_bins;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) hist[b] = 0.0;
    __syncthreads();
    const int64_t total = p.n * p.rows;
    const double inv_dz = 1.0 / p.dz;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t / p.rows, s = t - i * p.rows;
        if (p.seg_count) {
            const int64_t m = p.seg_count[i] < p.rows ? p.seg_count[i] : p.rows;
            if (s >= m) continue;
        }
        const double half = 0.5 * p.seg_len[t] * p.seg_dir[3 * t + 2];
        const double zm = p.seg_mid[3 * t + 2];
        const double za = zm - half, zb = zm + half;
        const double dE = p.seg_E[t] - p.seg_E_end[t];
        const double w = p.weight ? p.weight[i] * p.scale : p.scale;
        const double lo = fmin(za, zb), hi = fmax(za, zb);
        double fb_lo = floor((lo - p.z0) * inv_dz), fb_hi = floor((hi - p.z0) * inv_dz);
        fb_lo = fmin(fmax(fb_lo, 0.0), (double)(nb - 1));
        fb_hi = fmin(fmax(fb_hi, 0.0), (double)(nb - 1));
        const int b0 = (int)fb_lo, b1 = (int)fb_hi;
        if (b0 == b1 || !(hi > lo)) {
            atomicAdd(&hist[b0], w * dE);
        } else {
            const double inv_span = 1.0 / (hi - lo);
            for (int b = b0; b <= b1; ++b) {
                // first and last bins also take what lies outside the grid
                const double edge_lo = b == b0 ? lo : p.z0 + (double)b * p.dz;
                const double edge_hi = b == b1 ? hi : p.z0 + (double)(b + 1) * p.dz;
                atomicAdd(&hist[b], w * dE * ((edge_hi - edge_lo) * inv_span));
            }
        }
        if (p.seg_event && p.seg_event[t] == kEventCutoff) {
            double fe = floor((zb - p.z0) * inv_dz);
            fe = fmin(fmax(fe, 0.0), (double)(nb - 1));
            atomicAdd(&hist[(int)fe], w * p.seg_E_end[t]);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        if (hist[b] != 0.0) atomicAdd(&p.out[b], hist[b]);
}

}  // namespace
}  // namespace hr

extern "C" hr_status hr_depth_dose(const hr_dose_args *args, void *stream) {
    HR_TRACE("hr_depth_dose");
    using namespace hr;
    HR_REQUIRE(args != nullptr, "hr_depth_dose: null args");
    const hr_dose_args &p = *args;
    HR_REQUIRE(p.n >= 0 && p.rows >= 1, "hr_depth_dose: bad sizes (n=%lld rows=%lld)", (long long)p.n,
               (long long)p.rows);
    HR_REQUIRE(p.n_bins >= 1 && p.n_bins <= kDoseMaxBins, "hr_depth_dose: n_bins %d outside [1, %d]", p.n_bins,
               kDoseMaxBins);
    HR_REQUIRE(p.dz > 0.0, "hr_depth_dose: dz must be positive");
    if (p.n == 0) return HR_OK;
    HR_REQUIRE(p.seg_mid && p.seg_dir && p.seg_len && p.seg_E && p.seg_E_end && p.out,
               "hr_depth_dose: null row or output pointer");
    HR_REQUIRE(p.n <= INT64_MAX / p.rows, "hr_depth_dose: too many slots");
    const int64_t total = p.n * p.rows;
    int64_t blocks = (total + kDoseThreads - 1) / kDoseThreads;
    blocks = blocks < 1024 ? blocks : 1024;
    hipLaunchKernelGGL(depth_dose_kernel, dim3((unsigned)blocks), dim3(kDoseThreads), 0, (hipStream_t)stream, p);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}
