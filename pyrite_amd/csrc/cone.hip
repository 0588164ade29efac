// cone.hip -- copy records of detector-cone splitting.
//
// hr_cone_collect compacts the per-electron copy slots a cone-split
// hr_transport_run wrote into one launch state in (electron, event) order,
// after the power-of-two weight ladder (cone_split.py:140-179 weight_ladder
// and copy_stream_keys, :412-444 _ConeSplitPass.collect).  Three passes over
// electrons: per-block kept counts and largest slot count, one block scanning
// the block totals, then each block scanning its electrons' kept counts and
// writing its copies.  The ladder draw and the copy stream key depend only on
// (split key, flight), so the output is the same for any launch shape.
#include <cmath>

#include "cone.hpp"

namespace hr {
namespace {

constexpr int kCollectThreads = 256;

// Power-of-two ladder weight of w > 0: w itself when w is a power of two, else
// the next power of two above it (frexp is exact: w = m 2^e, m in [1/2, 1)).
__device__ __forceinline__ double ladder_weight(double w) {
    int e = 0;
    const double m = frexp(w, &e);
    return m == 0.5 ? w : ldexp(1.0, e);
}

// Whether copy slot r of electron i survives the ladder (always, without it).
__device__ __forceinline__ bool copy_kept(const hr_cone_split &c, uint64_t split_key, int64_t r, bool ladder,
                                          double *W) {
    const double w = c.weight[r];
    if (!ladder) {
        *W = w;
        return true;
    }
    const double L = ladder_weight(w);
    *W = L;
    const uint64_t key = cone::event_key(split_key, c.flight[r]);
    return stream_uniform(key, 4ULL) < w / L;
}

__device__ __forceinline__ int64_t electron_kept(const hr_cone_split &c, int64_t e_start, int64_t i, bool ladder) {
    const int64_t n = c.count[i] < c.cap ? c.count[i] : c.cap;
    const uint64_t sk = c.split_key[e_start + i];
    int64_t kept = 0;
    for (int64_t k = 0; k < n; ++k) {
        double W;
        kept += copy_kept(c, sk, i * c.cap + k, ladder, &W) ? 1 : 0;
    }
    return kept;
}

// Block-wide inclusive scan of one value per thread (kCollectThreads threads).
__device__ __forceinline__ int64_t block_inclusive_scan(int64_t v, int64_t *s_buf) {
    const int t = threadIdx.x;
    s_buf[t] = v;
    __syncthreads();
    for (int off = 1; off < kCollectThreads; off <<= 1) {
        const int64_t add = t >= off ? s_buf[t - off] : 0;
        __syncthreads();
        s_buf[t] += add;
        __syncthreads();
    }
    return s_buf[t];
}

__global__ void __launch_bounds__(kCollectThreads) collect_count_kernel(hr_cone_split c, int64_t e_start,
                                                                        int64_t e_count, int ladder,
                                                                        int64_t *__restrict__ block_kept,
                                                                        int64_t *__restrict__ block_max) {
    __shared__ int64_t s_buf[kCollectThreads];
    const int64_t i = (int64_t)blockIdx.x * kCollectThreads + threadIdx.x;
    const int64_t kept = i < e_count ? electron_kept(c, e_start, i, ladder != 0) : 0;
    const int64_t cnt = i < e_count ? c.count[i] : 0;
    const int64_t total = block_inclusive_scan(kept, s_buf);
    if (threadIdx.x == kCollectThreads - 1) block_kept[blockIdx.x] = total;
    __syncthreads();
    // largest slot count of the block
    s_buf[threadIdx.x] = cnt;
    __syncthreads();
    for (int off = kCollectThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off && s_buf[threadIdx.x + off] > s_buf[threadIdx.x])
            s_buf[threadIdx.x] = s_buf[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) block_max[blockIdx.x] = s_buf[0];
}

// One block: exclusive offsets of the block totals, the kept total and the
// largest slot count.
__global__ void __launch_bounds__(kCollectThreads) collect_scan_kernel(int64_t nb, const int64_t *__restrict__ block_kept,
                                                                       const int64_t *__restrict__ block_max,
                                                                       int64_t *__restrict__ block_off,
                                                                       int64_t *__restrict__ n_out) {
    __shared__ int64_t s_buf[kCollectThreads];
    int64_t carry = 0, biggest = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += kCollectThreads) {
        const int64_t b = b0 + threadIdx.x;
        const int64_t v = b < nb ? block_kept[b] : 0;
        const int64_t incl = block_inclusive_scan(v, s_buf);
        if (b < nb) block_off[b] = carry + incl - v;
        __syncthreads();
        carry += s_buf[kCollectThreads - 1];
        __syncthreads();
        const int64_t m = b < nb ? block_max[b] : 0;
        s_buf[threadIdx.x] = m;
        __syncthreads();
        for (int off = kCollectThreads / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off && s_buf[threadIdx.x + off] > s_buf[threadIdx.x])
                s_buf[threadIdx.x] = s_buf[threadIdx.x + off];
            __syncthreads();
        }
        if (s_buf[0] > biggest) biggest = s_buf[0];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        n_out[0] = carry;
        n_out[1] = biggest;
    }
}

__global__ void __launch_bounds__(kCollectThreads) collect_write_kernel(hr_cone_collect_args p, hr_cone_split c,
                                                                        const int64_t *__restrict__ block_off) {
    __shared__ int64_t s_buf[kCollectThreads];
    const int64_t i = (int64_t)blockIdx.x * kCollectThreads + threadIdx.x;
    const bool ladder = p.ladder != 0;
    const int64_t kept = i < p.e_count ? electron_kept(c, p.e_start, i, ladder) : 0;
    const int64_t incl = block_inclusive_scan(kept, s_buf);
    if (i >= p.e_count || kept == 0) return;
    int64_t o = block_off[blockIdx.x] + incl - kept;
    const int64_t n = c.count[i] < c.cap ? c.count[i] : c.cap;
    const uint64_t sk = c.split_key[p.e_start + i];
    for (int64_t k = 0; k < n; ++k) {
        const int64_t r = i * c.cap + k;
        double W;
        if (!copy_kept(c, sk, r, ladder, &W)) continue;
        if (o < p.out_cap) {
            p.parent[o] = i;
            p.flight[o] = c.flight[r];
            for (int d = 0; d < 3; ++d) {
                p.pos[3 * o + d] = c.pos[3 * r + d];
                p.dir[3 * o + d] = c.dir[3 * r + d];
            }
            p.E_keV[o] = c.E_keV[r];
            p.clock[o] = c.clock[r];
            p.weight[o] = W;
            p.stream_key[o] = splitmix64(cone::event_key(sk, c.flight[r]) ^ cone::kCopySalt);
        }
        ++o;
    }
}

constexpr uint64_t kRouletteSalt = 0x61C8864680B583EBULL;

__global__ void cone_roulette_kernel(int64_t n, const float *__restrict__ log_imp,
                                     const uint64_t *__restrict__ key, double log_T, double p_min,
                                     double *__restrict__ weight, uint8_t *__restrict__ keep) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double w = weight[i];
    const double s = log(w) + (double)log_imp[i] - log_T;  // log(w I / T)
    const double p = s >= 0.0 ? 1.0 : fmax(p_min, exp(s));
    const bool k = p >= 1.0 || stream_uniform(splitmix64(key[i] ^ kRouletteSalt), 0ULL) < p;
    keep[i] = k ? 1 : 0;
    weight[i] = k ? w / p : 0.0;
}

__global__ void zero_pair_kernel(int64_t *__restrict__ v) {
    v[0] = 0;
    v[1] = 0;
}

__global__ void split_keys_kernel(const uint64_t *__restrict__ key, int64_t n, uint64_t *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = splitmix64(key[i] ^ cone::kSplitSalt);
}

__global__ void cone_event_kernel(hr_cone_split c, int64_t n, const double *__restrict__ d,
                                  const double *__restrict__ alpha, const double *__restrict__ E,
                                  const uint64_t *__restrict__ key, double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const cone::Copy cp = cone::event(d[3 * i], d[3 * i + 1], d[3 * i + 2], c.nx, c.ny, c.nz, c.cos_c, c.sin_c,
                                      c.target_weight, c.angle_scale, alpha[i], E[i], key[i]);
    out[5 * i] = cp.made ? 1.0 : 0.0;
    out[5 * i + 1] = cp.ox;
    out[5 * i + 2] = cp.oy;
    out[5 * i + 3] = cp.oz;
    out[5 * i + 4] = cp.w;
}

__global__ void cone_event_tab_kernel(hr_cone_split c, hr_elsepa t, int64_t n, const double *__restrict__ d,
                                      const int64_t *__restrict__ el, const double *__restrict__ E,
                                      const uint64_t *__restrict__ key, double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cone::TabLaw law(&t, (size_t)el[i], E[i]);
    const cone::Copy cp = cone::event_law(d[3 * i], d[3 * i + 1], d[3 * i + 2], c.nx, c.ny, c.nz, c.cos_c, c.sin_c,
                                          c.target_weight, c.angle_scale, law, E[i], key[i]);
    out[5 * i] = cp.made ? 1.0 : 0.0;
    out[5 * i + 1] = cp.ox;
    out[5 * i + 2] = cp.oy;
    out[5 * i + 3] = cp.oz;
    out[5 * i + 4] = cp.w;
}

int64_t collect_blocks(int64_t e_count) { return (e_count + kCollectThreads - 1) / kCollectThreads; }

}  // namespace
}  // namespace hr

extern "C" int64_t hr_cone_collect_workspace_bytes(int64_t e_count) {
    if (e_count <= 0) return 0;
    const int64_t nb = hr::collect_blocks(e_count);
    return (3 * nb * (int64_t)sizeof(int64_t) + 255) / 256 * 256;
}

extern "C" hr_status hr_cone_collect(const hr_cone_collect_args *args, void *stream) {
    HR_TRACE("hr_cone_collect");
    using namespace hr;
    HR_REQUIRE(args != nullptr && args->cone != nullptr, "hr_cone_collect: null args");
    const hr_cone_collect_args &p = *args;
    const hr_cone_split &c = *p.cone;
    HR_REQUIRE(p.e_start >= 0 && p.e_count >= 0 && p.out_cap >= 0, "hr_cone_collect: negative range or capacity");
    HR_REQUIRE(p.ladder == 0 || p.ladder == 1, "hr_cone_collect: ladder must be 0 or 1");
    HR_REQUIRE(p.n_out != nullptr, "hr_cone_collect: null n_out");
    hipStream_t s = (hipStream_t)stream;
    if (p.e_count == 0) {
        hipLaunchKernelGGL(zero_pair_kernel, dim3(1), dim3(1), 0, s, p.n_out);
        HR_CHECK_HIP(hipGetLastError());
        return HR_OK;
    }
    HR_REQUIRE(c.cap >= 1 && c.count && c.split_key && c.pos && c.dir && c.E_keV && c.clock && c.weight && c.flight,
               "hr_cone_collect: the cone record needs its slots, counts and split keys");
    HR_REQUIRE(p.out_cap == 0 || (p.parent && p.flight && p.pos && p.dir && p.E_keV && p.clock && p.weight &&
                                  p.stream_key),
               "hr_cone_collect: null output");
    const int64_t need = hr_cone_collect_workspace_bytes(p.e_count);
    HR_REQUIRE_WS(p.ws, need, "hr_cone_collect");
    const int64_t nb = collect_blocks(p.e_count);
    HR_REQUIRE(nb <= (int64_t)0x7fffffff, "hr_cone_collect: too many electrons");
    int64_t *block_kept = (int64_t *)p.ws.ptr;
    int64_t *block_max = block_kept + nb;
    int64_t *block_off = block_max + nb;
    hipLaunchKernelGGL(collect_count_kernel, dim3((unsigned)nb), dim3(kCollectThreads), 0, s, c, p.e_start, p.e_count,
                       p.ladder, block_kept, block_max);
    HR_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(collect_scan_kernel, dim3(1), dim3(kCollectThreads), 0, s, nb, block_kept, block_max, block_off,
                       p.n_out);
    HR_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(collect_write_kernel, dim3((unsigned)nb), dim3(kCollectThreads), 0, s, p, c, block_off);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

extern "C" hr_status hr_cone_split_keys(const uint64_t *stream_key, int64_t n, uint64_t *out, void *stream) {
    HR_TRACE("hr_cone_split_keys");
    using namespace hr;
    HR_REQUIRE(n >= 0, "hr_cone_split_keys: negative n");
    if (n == 0) return HR_OK;
    HR_REQUIRE(stream_key && out, "hr_cone_split_keys: null pointer");
    hipLaunchKernelGGL(split_keys_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       stream_key, n, out);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

extern "C" hr_status hr_cone_event(const hr_cone_split *cone, int64_t n, const double *d, const double *alpha,
                                   const double *E_keV, const uint64_t *event_key, double *out, void *stream) {
    HR_TRACE("hr_cone_event");
    using namespace hr;
    HR_REQUIRE(cone != nullptr && n >= 0, "hr_cone_event: null cone or negative n");
    if (n == 0) return HR_OK;
    HR_REQUIRE(d && alpha && E_keV && event_key && out, "hr_cone_event: null pointer");
    HR_REQUIRE(cone->sin_c > 0.0 && cone->cos_c > 0.0 && cone->target_weight > 0.0 && cone->angle_scale >= 0.0,
               "hr_cone_event: bad cone parameters");
    hipLaunchKernelGGL(cone_event_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       *cone, n, d, alpha, E_keV, event_key, out);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

extern "C" hr_status hr_cone_event_tab(const hr_cone_split *cone, const hr_elsepa *tables, int64_t n,
                                       const double *d, const int64_t *el, const double *E_keV,
                                       const uint64_t *event_key, double *out, void *stream) {
    HR_TRACE("hr_cone_event_tab");
    using namespace hr;
    HR_REQUIRE(cone != nullptr && tables != nullptr && n >= 0, "hr_cone_event_tab: null cone/tables or negative n");
    if (n == 0) return HR_OK;
    HR_REQUIRE(d && el && E_keV && event_key && out, "hr_cone_event_tab: null pointer");
    HR_REQUIRE(cone->sin_c > 0.0 && cone->cos_c > 0.0 && cone->target_weight > 0.0 && cone->angle_scale >= 0.0,
               "hr_cone_event_tab: bad cone parameters");
    const hr_elsepa &t = *tables;
    HR_REQUIRE(t.on == 1 && t.start && t.length && t.logE && t.mu && t.cdf && t.pdf && t.guide && t.n_mu >= 2 &&
                   t.guide_cells >= 1 && (t.guide_cells & (t.guide_cells - 1)) == 0,
               "hr_cone_event_tab: incomplete tables");
    hipLaunchKernelGGL(cone_event_tab_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       *cone, t, n, d, el, E_keV, event_key, out);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

extern "C" hr_status hr_cone_roulette(int64_t n, const float *log_imp, const uint64_t *stream_key, double log_T,
                                      double p_min, double *weight, uint8_t *keep, void *stream) {
    HR_TRACE("hr_cone_roulette");
    using namespace hr;
    HR_REQUIRE(n >= 0, "hr_cone_roulette: negative n");
    if (n == 0) return HR_OK;
    HR_REQUIRE(log_imp && stream_key && weight && keep, "hr_cone_roulette: null pointer");
    HR_REQUIRE(p_min > 0.0 && p_min <= 1.0 && std::isfinite(log_T), "hr_cone_roulette: p_min in (0, 1], finite log_T");
    hipLaunchKernelGGL(cone_roulette_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n,
                       log_imp, stream_key, log_T, p_min, weight, keep);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}
