// score.hip -- per-segment line prologue, sinc^2 reduction and BH+Elwert
// bremsstrahlung on gfx950.
//
// Algorithms follow PyRITE's batched incoherent line route and brem reducer
// (src/pyrite/montecarlo/spectrum/cxr/_batched.py:461-728, cxr/_fused.py,
// cxr/_recoil.py:35-76, cxr_jit_kernel.py:32-613, brem_jit_kernel.py:76-153,
// segment_escape.py:286-300).  Precision policy (MI355X-first):
//  * the prologue (resonance root, couplings, amplitudes, weights) runs in FP64:
//    it is once per line, and keeping the resonance energy E_r and the sinc slope
//    a in FP64 keeps the phase a(E - E_r) exact to ~1e-11 turns;
//  * the reduction cell evaluates the phase u = a(E - E_r)/pi in FP64 (one FMA
//    on a uniform grid), reduces it to r = u - rint(u) in FP64, and takes
//    [sin(pi r) / (pi u)]^2 in FP32 (polynomial sine, hardware reciprocal); cells
//    are summed in FP32 over groups of 4 lines (all terms are positive) and in
//    FP64 beyond;
//  * the brem cell is FP32 (as the reference's), its sums FP64.
// Reductions never use atomics: each workgroup writes a partial spectrum and a
// second pass adds the partials in a fixed order, so results are bitwise
// reproducible for a fixed launch shape.
#include <math.h>

#include <mutex>

#include "hr_common.hpp"
#include "line_core.hpp"

namespace hr {
namespace {

using namespace line;

constexpr double kReCm2 = 7.9407877e-26;   // brem.py:63
constexpr double kMc2 = 510.99895;

// Lines of one (segment j, direction d) pair: emit(g, E_r, a, w) for every
// reflection g (w = 0 and E_r = a = 0 for a rejected pair).
template <class Emit>
__device__ __forceinline__ void segment_lines(const hr_line_args &p, int64_t j, int d, Emit &&emit) {
    const double nx = p.n_hat[3 * d], ny = p.n_hat[3 * d + 1], nz = p.n_hat[3 * d + 2];
    const double dir_w = p.dir_weight[d] * p.scale * (p.row_weight ? (double)p.row_weight[j] : 1.0);
    const SegKin k = seg_kin(p.r_mid[3 * j + 2], p.v_hat[3 * j], p.v_hat[3 * j + 1], p.v_hat[3 * j + 2], p.L[j],
                             p.E[j], nx, ny, nz, p.thickness);
    const GridGuess guess = grid_guess(p.E_tab, p.n_tab);
    for (int g = 0; g < p.n_g; ++g) {
        const LineCore o = line_core(p, k, nx, ny, nz, d, g, guess);
        double w = 0.0, a_width = 0.0, Er = 0.0;
        if (o.keep) {
            double A2 = 0.0;
#pragma unroll
            for (int pol = 0; pol < 2; ++pol) A2 += o.A_re[pol] * o.A_re[pol] + o.A_im[pol] * o.A_im[pol];
            // _batched.py:702-728 + _fused.py:267-281, one piece, fraction 1
            const double T_abs = mean_transmission(k.path0 * o.mu, k.path1 * o.mu);
            const double pref = kAlpha * o.omega / (4.0 * kPi * kPi * kHbarc) * (k.t_L * k.t_L) * T_abs;
            const double wt = pref * A2 * p.wm[g];
            if (isfinite(wt) && wt > 0.0) {
                w = wt * dir_w;
                a_width = o.slope * k.t_L / (2.0 * kHbarc);
                Er = o.E_res;
            }
        }
        emit(g, Er, a_width, w);
    }
}

// One (segment, direction) pair per thread: writes its n_g lines.
__global__ void __launch_bounds__(256) cxr_lines_kernel(hr_line_args p) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = p.n_seg * (int64_t)p.n_dir;
    if (tid >= total) return;
    const int64_t j = tid % p.n_seg;
    const int d = (int)(tid / p.n_seg);
    const int64_t out0 = ((int64_t)d * p.n_seg + j) * p.n_g;
    segment_lines(p, j, d, [&](int g, double Er, double a_width, double w) {
        p.E_r[out0 + g] = Er;
        p.a[out0 + g] = a_width;
        p.w[out0 + g] = (float)w;
    });
}

// sin(pi r)/pi for |r| <= 1/2 as r P(r^2): an odd polynomial of degree 9 fitted
// to sin(pi r) for minimax relative error on [0, 1/2] (Lawson's iteration of
// weighted least squares on Chebyshev-clustered nodes, fit error 5.3e-9), its
// coefficients divided by pi and rounded to FP32.
constexpr float kSinPiOverPi[5] = {1.0f, -1.6449331045150757f, 0.8117124438285828f, -0.19042645394802094f,
                                   0.024688366800546646f};

// [sin(pi u)/(pi u)]^2 from r = u - rint(u) (|r| <= 1/2) and u in FP32:
// x = r P(r^2) / u with one hardware reciprocal (v_rcp_f32, 1 ulp); maximum
// relative error 5.4e-7 over |u| <= 5e4.  |x| <= 1 in exact arithmetic (the
// central period has u == r, elsewhere |u| >= 1/2), so fminf(x, 1) only trims
// rounding, and at u == 0 (r == 0) it turns 0 * inf = NaN into the limit 1,
// where the reference substitutes x = 1e-20.
__device__ __forceinline__ float sinc2_cell(float r, float uf) {
    const float r2 = r * r;
    float s = kSinPiOverPi[4];
    s = fmaf(s, r2, kSinPiOverPi[3]);
    s = fmaf(s, r2, kSinPiOverPi[2]);
    s = fmaf(s, r2, kSinPiOverPi[1]);
    s = fmaf(s, r2, kSinPiOverPi[0]);
    const float x = fminf(r * __builtin_amdgcn_rcpf(uf) * s, 1.0f);
    return x * x;
}

constexpr int kRedThreads = 256;
constexpr int kLineTile = 256;  // lines staged in LDS per step
constexpr int kSincGroup = 4;   // cells summed in FP32 before each FP64 add (all terms >= 0)

// Dense sum over a uniform grid E_k = E0 + k dE.  Block b takes line tiles
// b, b + gridDim.x, ...; thread t owns bins t + kRedThreads*m (m < BINS_PER_THREAD)
// of the bin window starting at blockIdx.y * kRedThreads * BINS.  A staged line
// is (u0, du) in one 16-byte LDS slot (one ds_read_b128) with the weights of
// kSincGroup lines in one float4; a group whose lines all have zero weight
// (copies that do not carry this tally) costs one LDS read.  The phase
// u = u0 + k du is formed and reduced in FP64 per cell: on this device an FP64
// FMA/add issues at the rate of a non-packed FP32 one, and FP32 stepping of the
// reduced phase would take more instructions per cell.
template <int BINS>
__global__ void __launch_bounds__(kRedThreads) sincsq_uniform_kernel(
    const double *__restrict__ E_r, const double *__restrict__ a, const float *__restrict__ w,
    int64_t n_lines, double E0, double dE, int64_t n_E, double *__restrict__ partial) {
    static_assert(kSincGroup == 4 && kLineTile % kSincGroup == 0, "weights are read as float4 groups");
    __shared__ double2 s_ud[kLineTile];
    __shared__ float4 s_w4[kLineTile / kSincGroup];
    float *s_w = reinterpret_cast<float *>(s_w4);
    const int t = threadIdx.x;
    const int64_t bin0 = (int64_t)blockIdx.y * kRedThreads * BINS;
    double acc[BINS], kd[BINS];
#pragma unroll
    for (int m = 0; m < BINS; ++m) {
        acc[m] = 0.0;
        kd[m] = (double)(bin0 + t + (int64_t)m * kRedThreads);
    }
    const double inv_pi = 1.0 / kPi;
    const int64_t n_tiles = (n_lines + kLineTile - 1) / kLineTile;
    for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const int64_t l = tile * kLineTile + t;
        // u(E_k) = a (E_k - E_r)/pi = u0 + k du; the weight is read first and a
        // zero-weight line's E_r and a are never fetched (rejected resonances
        // come in whole waves: a ray whose in-medium root never settles)
        const float wl = l < n_lines ? w[l] : 0.0f;
        s_w[t] = wl;
        if (wl != 0.0f) {
            const double al = a[l] * inv_pi;
            s_ud[t] = make_double2(al * (E0 - E_r[l]), al * dE);
        } else {
            s_ud[t] = make_double2(0.0, 0.0);
        }
        __syncthreads();
        for (int q0 = 0; q0 < kLineTile; q0 += kSincGroup) {
            const float4 w4 = s_w4[q0 / kSincGroup];
            // uniform across the block: every thread reads the same slots
            if (w4.x == 0.0f && w4.y == 0.0f && w4.z == 0.0f && w4.w == 0.0f) continue;
            const float wg[kSincGroup] = {w4.x, w4.y, w4.z, w4.w};
            float g[BINS];
#pragma unroll
            for (int m = 0; m < BINS; ++m) g[m] = 0.0f;
#pragma unroll
            for (int j = 0; j < kSincGroup; ++j) {
                const float wq = wg[j];
                if (wq == 0.0f) continue;
                const double2 ud = s_ud[q0 + j];
#pragma unroll
                for (int m = 0; m < BINS; ++m) {
                    const double u = fma(kd[m], ud.y, ud.x);
                    const float r = (float)(u - rint(u));
                    g[m] = fmaf(wq, sinc2_cell(r, (float)u), g[m]);
                }
            }
#pragma unroll
            for (int m = 0; m < BINS; ++m) acc[m] += (double)g[m];
        }
        __syncthreads();
    }
    double *out = partial + (int64_t)blockIdx.x * n_E;
#pragma unroll
    for (int m = 0; m < BINS; ++m) {
        const int64_t k = bin0 + t + (int64_t)m * kRedThreads;
        if (k < n_E) out[k] = acc[m];
    }
}

// General grid: x = a (E_k - E_r) formed in FP64 per cell.
template <int BINS>
__global__ void __launch_bounds__(kRedThreads) sincsq_general_kernel(
    const double *__restrict__ E_r, const double *__restrict__ a, const float *__restrict__ w,
    int64_t n_lines, const double *__restrict__ E_grid, int64_t n_E, double *__restrict__ partial) {
    __shared__ double s_er[kLineTile];
    __shared__ double s_a[kLineTile];
    __shared__ float s_w[kLineTile];
    const int t = threadIdx.x;
    const int64_t bin0 = (int64_t)blockIdx.y * kRedThreads * BINS;
    double acc[BINS], Ek[BINS];
#pragma unroll
    for (int m = 0; m < BINS; ++m) {
        acc[m] = 0.0;
        const int64_t k = bin0 + t + (int64_t)m * kRedThreads;
        Ek[m] = k < n_E ? E_grid[k] : 0.0;
    }
    const double inv_pi = 1.0 / kPi;
    const int64_t n_tiles = (n_lines + kLineTile - 1) / kLineTile;
    for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const int64_t l = tile * kLineTile + t;
        const float wl = l < n_lines ? w[l] : 0.0f;  // zero-weight lines: E_r and a not fetched
        s_w[t] = wl;
        s_er[t] = wl != 0.0f ? E_r[l] : 0.0;
        s_a[t] = wl != 0.0f ? a[l] * inv_pi : 0.0;
        __syncthreads();
        for (int q0 = 0; q0 < kLineTile; q0 += kSincGroup) {
            float g[BINS];
#pragma unroll
            for (int m = 0; m < BINS; ++m) g[m] = 0.0f;
            for (int q = q0; q < q0 + kSincGroup; ++q) {
                const float wq = s_w[q];
                if (wq == 0.0f) continue;
                const double er = s_er[q], aq = s_a[q];
#pragma unroll
                for (int m = 0; m < BINS; ++m) {
                    const double u = aq * (Ek[m] - er);
                    const float r = (float)(u - rint(u));
                    g[m] = fmaf(wq, sinc2_cell(r, (float)u), g[m]);
                }
            }
#pragma unroll
            for (int m = 0; m < BINS; ++m) acc[m] += (double)g[m];
        }
        __syncthreads();
    }
    double *out = partial + (int64_t)blockIdx.x * n_E;
#pragma unroll
    for (int m = 0; m < BINS; ++m) {
        const int64_t k = bin0 + t + (int64_t)m * kRedThreads;
        if (k < n_E) out[k] = acc[m];
    }
}

// out[k] += sum_b partial[b][k], b ascending (fixed order).
__global__ void __launch_bounds__(256) sum_partials_kernel(const double *__restrict__ partial,
                                                           int64_t n_part, int64_t n_E,
                                                           double *__restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_E) return;
    double s = 0.0;
    for (int64_t b = 0; b < n_part; ++b) s += partial[b * n_E + k];
    out[k] += s;
}

// First level of the fixed-order partial sum: split s adds rows
// [s*chunk, min((s+1)*chunk, n_part)) of every bin into level[s][k].
__global__ void __launch_bounds__(256) sum_partials_split_kernel(const double *__restrict__ partial,
                                                                 int64_t n_part, int64_t chunk, int64_t n_E,
                                                                 double *__restrict__ level) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_E) return;
    const int64_t b0 = (int64_t)blockIdx.y * chunk;
    const int64_t b1 = b0 + chunk < n_part ? b0 + chunk : n_part;
    double s = 0.0;
    for (int64_t b = b0; b < b1; ++b) s += partial[b * n_E + k];
    level[(int64_t)blockIdx.y * n_E + k] = s;
}

// Splits of the two-level partial sum (a fixed tree: deterministic, and wide
// enough that the first level reads the partials at full bandwidth).
constexpr int64_t kPartialSplits = 64;

constexpr float kMc2f = 510.99895f;
constexpr float kInvMc2f = (float)(1.0 / 510.99895);
constexpr double kLog2e = 1.4426950408889634;

// Hardware single-instruction forms (v_rcp_f32, v_sqrt_f32, v_exp_f32): about
// 1 ulp each, against the multi-instruction correctly rounded sequences.
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fast_log2(float x) { return __builtin_amdgcn_logf(x); }
constexpr float kLn2f = 0.69314718055994531f;
constexpr double kLn2d = 0.6931471805599453;

// BH+Elwert brem over segment tiles (brem.py:396-423, brem_jit_kernel.py:76-153).
constexpr int kBremTile = 128;
// NEL: element count of the composition (1..4), a template parameter so the
// per-element loop unrolls with its constants in registers.
template <int BINS, int NEL>
__global__ void __launch_bounds__(kRedThreads) brem_kernel(hr_brem_args p, double *__restrict__ partial) {
    // per staged segment: kinetic energy, momentum, the shorter and longer
    // escape paths from its two ends along n_hat, their mean and difference,
    // the per-element prefactors and whether any of them is nonzero
    __shared__ float s_T[kBremTile], s_pi[kBremTile], s_plo[kBremTile], s_phi[kBremTile], s_pmid[kBremTile],
        s_dp[kBremTile];
    __shared__ float s_pref[kBremTile * 4];
    __shared__ int s_live[kBremTile];
    const int t = threadIdx.x;
    const int64_t bin0 = (int64_t)blockIdx.y * kRedThreads * BINS;
    constexpr int n_el = NEL;
    double acc[BINS];
    // per-bin constants: k [keV], k / m^2 [1/keV], ln 2 / k [1/eV], mu, mu log2(e)
    float keV[BINS], k_m2[BINS], ln2_inv_k[BINS], mu[BINS], mu2[BINS];
#pragma unroll
    for (int m = 0; m < BINS; ++m) {
        acc[m] = 0.0;
        const int64_t k = bin0 + t + (int64_t)m * kRedThreads;
        const double kev = k < p.n_E ? p.E_grid_eV[k] : 0.0;
        keV[m] = (float)(kev * 1.0e-3);
        k_m2[m] = (float)(kev * 1.0e-3 / (kMc2 * kMc2));
        ln2_inv_k[m] = (float)(kLn2d / fmax(kev, 1.0e-30));
        mu[m] = k < p.n_E ? p.mu[k] : 0.0f;
        mu2[m] = (float)((double)mu[m] * kLog2e);
    }
    float zi2[4];
    for (int el = 0; el < 4; ++el)  // exp(-2 pi Z alpha / beta) = exp2(-(2 pi Z alpha log2 e) / beta)
        zi2[el] = el < n_el ? (float)(2.0 * kPi * p.Z[el] * kAlpha * kLog2e) : 0.0f;
    const int64_t n_tiles = (p.n_seg + kBremTile - 1) / kBremTile;
    for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        if (t < kBremTile) {
            const int64_t j = tile * kBremTile + t;
            if (j < p.n_seg) {
                const float T = p.E_keV[j];
                const float pi_ = sqrtf(T * (T + 2.0f * (float)kMc2)) / (float)kMc2;
                const float bi = pi_ / (1.0f + T / (float)kMc2);
                s_T[t] = T;
                s_pi[t] = pi_;
                // escape paths from both ends along n_hat (segment_escape.py:190-211)
                const double half = 0.5 * (double)p.L[j];
                const double rz = p.r_mid[3 * j + 2], vz = p.v_hat[3 * j + 2];
                const double p0 = slab_path(rz - half * vz, p.thickness, p.n_hat[2]);
                const double p1 = slab_path(rz + half * vz, p.thickness, p.n_hat[2]);
                const double plo = fmin(p0, p1), phi = fmax(p0, p1);
                s_plo[t] = (float)plo;
                s_phi[t] = (float)phi;
                s_pmid[t] = (float)(0.5 * (plo + phi));
                s_dp[t] = (float)(phi - plo);
                int live = 0;
                for (int el = 0; el < 4; ++el) {
                    float pref = 0.0f;
                    if (el < n_el) {
                        // n L 1e-8 (16/3) alpha r_e^2 Z^2 beta_i (1 - exp(-2 pi Z alpha/beta_i)) / p_i^2
                        const double den_i = 1.0 - exp(-2.0 * kPi * p.Z[el] * kAlpha / (double)bi);
                        const double rw = p.row_weight ? (double)p.row_weight[j] : 1.0;
                        pref = (float)(rw * p.n_cm3[el] * (double)p.L[j] * 1.0e-8 * (16.0 / 3.0) * kAlpha * kReCm2 *
                                       p.Z[el] * p.Z[el] * (double)bi * den_i / ((double)pi_ * (double)pi_));
                    }
                    s_pref[t * 4 + el] = pref;
                    live |= pref != 0.0f;
                }
                s_live[t] = live;
            } else {
                s_T[t] = 0.0f;
                s_pi[t] = 0.0f;
                s_plo[t] = 0.0f;
                s_phi[t] = 0.0f;
                s_pmid[t] = 0.0f;
                s_dp[t] = 0.0f;
                for (int el = 0; el < 4; ++el) s_pref[t * 4 + el] = 0.0f;
                s_live[t] = 0;
            }
        }
        __syncthreads();
        const int64_t left = p.n_seg - tile * kBremTile;
        const int cnt = left < kBremTile ? (int)left : kBremTile;
        // cells of kGroup segments are summed in FP32 (8 positive terms: relative
        // rounding below 5e-7) before each FP64 accumulate
        constexpr int kGroup = 8;
        float part[BINS];
#pragma unroll
        for (int m = 0; m < BINS; ++m) part[m] = 0.0f;
        for (int q = 0; q < cnt; ++q) {
            if ((q & (kGroup - 1)) == 0 && q > 0) {
#pragma unroll
                for (int m = 0; m < BINS; ++m) {
                    acc[m] += (double)part[m];
                    part[m] = 0.0f;
                }
            }
            // a zero-weight row (a copy that does not carry this tally) adds nothing
            if (!s_live[q]) continue;
            const float T = s_T[q], pi_ = s_pi[q];
            const float plo = s_plo[q], phi = s_phi[q], pmid = s_pmid[q], dp = s_dp[q];
#pragma unroll
            for (int m = 0; m < BINS; ++m) {
                const float Tf = T - keV[m];
                if (!(Tf > 1.0e-6f) || !(keV[m] > 0.0f)) continue;
                const float Tm = Tf + 2.0f * kMc2f;
                // p_f m = sqrt(T_f (T_f + 2m)) and 1/beta_f = (m + T_f) / (p_f m), from one v_rsq_f32
                const float TT = Tf * Tm;
                const float rs = fast_rsq(TT);
                const float pf = TT * rs * kInvMc2f;
                const float rcp_bf = rs * (kMc2f + Tf);
                // ln((p_i + p_f)^2 m^2 / (k (T_i + T_f + 2m))) = ln((p_i + p_f)/(p_i - p_f)): the
                // difference form has no cancellation for k << T_i; the argument is >= 1 up to
                // rounding (no denormal path in the hardware log2); ln 2 / k is folded per bin
                const float spf = pi_ + pf;
                const float born_k = fast_log2(spf * spf * fast_rcp(k_m2[m] * (T + Tm))) * ln2_inv_k[m];
                // Elwert factors sum_el pref_el / (1 - exp(-2 pi Z_el alpha / beta_f)) as one
                // fraction num / den
                float num = s_pref[q * 4];
                float den = 1.0f - fast_exp2(-zi2[0] * rcp_bf);
#pragma unroll
                for (int el = 1; el < n_el; ++el) {
                    const float d_el = 1.0f - fast_exp2(-zi2[el] * rcp_bf);
                    num = fmaf(num, d_el, s_pref[q * 4 + el] * den);
                    den *= d_el;
                }
                // segment-mean escape (exp(-p_lo mu) - exp(-p_hi mu)) / d, d = (p_hi - p_lo) mu;
                // for d <= 1 as exp(-p_mid mu) sinh(d/2)/(d/2), the series truncated after
                // (d/2)^6 / 7! (remainder < 3e-8 relative)
                const float d = dp * mu[m];
                float esc;
                if (d > 1.0f) {
                    esc = fast_exp2(-plo * mu2[m]) - fast_exp2(-phi * mu2[m]);
                    den *= d;
                } else {
                    const float x2 = 0.25f * d * d;
                    const float shc = fmaf(x2, fmaf(x2, fmaf(x2, 1.0f / 5040, 1.0f / 120), 1.0f / 6), 1.0f);
                    esc = fast_exp2(-pmid * mu2[m]) * shc;
                }
                part[m] = fmaf(num * born_k * rcp_bf * esc, fast_rcp(den), part[m]);
            }
        }
#pragma unroll
        for (int m = 0; m < BINS; ++m) acc[m] += (double)part[m];
        __syncthreads();
    }
    double *out = partial + (int64_t)blockIdx.x * p.n_E;
#pragma unroll
    for (int m = 0; m < BINS; ++m) {
        const int64_t k = bin0 + t + (int64_t)m * kRedThreads;
        if (k < p.n_E) out[k] = acc[m];
    }
}

// ---- support-windowed reduction with the cycle-averaged far tail
// (spectrum/cxr/_node_window.py:86-250 NodeTail / row_window_plan,
// :454-522 _sincsq_window_reduce).  A line deposits its exact sinc^2 on the
// bins within |a (E - E_r)| <= x_c (minus the tail there) and three tail
// moments (m2, m3, m4) = w (1, 2 eps, 3 eps^2) / (2 a^2) at its node; a second
// pass adds sum_node m2/x^2 + m3/x^3 + m4/x^4 everywhere.  Broad windows (a
// quarter of the grid or more) and nodes outside [-n, 2n) are summed exactly
// over the whole grid.  Per-direction spectra and in-grid moments live in LDS
// (FP64 atomics); out-of-grid moments go to global memory.
// Threads per window workgroup (one workgroup per CU: the spectrum and moments
// fill its LDS); a tuning build may set HR_WINDOW_THREADS.
#ifndef HR_WINDOW_THREADS
#define HR_WINDOW_THREADS 1024
#endif
constexpr int kWinThreads = HR_WINDOW_THREADS;
constexpr int kWinMaxBins = 4096;

struct WindowShared {
    double spec[kWinMaxBins];
    double dep[3][kWinMaxBins];
};
static_assert(sizeof(WindowShared) <= 160 * 1024, "window LDS");

__device__ __forceinline__ double tail_value(double m2, double m3, double m4, double x) {
    const double inv = 1.0 / x;
    return (m2 + (m3 + m4 * inv) * inv) * inv * inv;
}

// Exact cell of the dense reducer at u = a (E - E_r) / pi (FP64 phase, FP32 profile).
__device__ __forceinline__ double sincsq_cell(double u) {
    return (double)sinc2_cell((float)(u - rint(u)), (float)u);
}

struct WindowGrid {
    int64_t n;
    double E0, dE, x_c;
};

__device__ __forceinline__ void window_line(WindowShared &sh, double *__restrict__ dep_far, const WindowGrid &G,
                                            double Er, double a, double w) {
    const int64_t n = G.n;
    const double nd = rint((Er - G.E0) / G.dE);
    const double eps = Er - (G.E0 + nd * G.dE);
    double R = ceil((fabs(eps) + G.x_c / a) / G.dE);
    R = fmin(R, 4503599627370496.0);
    const int64_t node = (int64_t)nd;
    const bool conv_range = node + n >= 0 && node + n < 3 * n;
    int64_t lo = 0, hi = n - 1;
    bool conv = false;
    if (conv_range) {
        const int64_t RR = (int64_t)R;
        const int64_t l = node - RR < 0 ? 0 : node - RR;
        const int64_t h = node + RR > n - 1 ? n - 1 : node + RR;
        if ((double)(h - l + 1) < 0.25 * (double)n) {
            lo = l;
            hi = h;
            conv = true;
        }
    }
    const double m2 = 0.5 / (a * a);
    const double m3 = 2.0 * eps * m2, m4 = 3.0 * eps * eps * m2;
    const double inv_pi = 1.0 / kPi;
    const double al = a * inv_pi;
    for (int64_t e = lo; e <= hi; ++e) {
        double v = sincsq_cell(al * ((G.E0 + (double)e * G.dE) - Er));
        if (conv && e != node) v -= tail_value(m2, m3, m4, (double)(e - node) * G.dE);
        atomicAdd(&sh.spec[e], w * v);
    }
    if (conv) {
        if (node >= 0 && node < n) {
            atomicAdd(&sh.dep[0][node], w * m2);
            atomicAdd(&sh.dep[1][node], w * m3);
            atomicAdd(&sh.dep[2][node], w * m4);
        } else {
            const int64_t L = 3 * n;
            atomicAdd(&dep_far[node + n], w * m2);
            atomicAdd(&dep_far[L + node + n], w * m3);
            atomicAdd(&dep_far[2 * L + node + n], w * m4);
        }
    }
}

__device__ __forceinline__ void window_zero(WindowShared &sh, int64_t n) {
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        sh.spec[i] = 0.0;
        sh.dep[0][i] = 0.0;
        sh.dep[1][i] = 0.0;
        sh.dep[2][i] = 0.0;
    }
}

// Block partial: [spec | m2 | m3 | m4] over the grid.
__device__ __forceinline__ void window_flush(const WindowShared &sh, int64_t n, double *__restrict__ out) {
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        out[i] = sh.spec[i];
        out[n + i] = sh.dep[0][i];
        out[2 * n + i] = sh.dep[1][i];
        out[3 * n + i] = sh.dep[2][i];
    }
}

// Segments [chunk share) x direction blockIdx.y: prologue fused with the window deposit.
__global__ void __launch_bounds__(kWinThreads) window_lines_kernel(hr_line_args p, WindowGrid G,
                                                                   double *__restrict__ partial,
                                                                   double *__restrict__ dep_far) {
    __shared__ WindowShared sh;
    const int d = blockIdx.y;
    window_zero(sh, G.n);
    __syncthreads();
    const int64_t per = (p.n_seg + gridDim.x - 1) / gridDim.x;
    const int64_t j0 = (int64_t)blockIdx.x * per;
    const int64_t j1 = j0 + per < p.n_seg ? j0 + per : p.n_seg;
    double *far = dep_far + (int64_t)d * 9 * G.n;
    for (int64_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
        segment_lines(p, j, d, [&](int, double Er, double a_width, double w64) {
            const float wf = (float)w64;  // the line weight the dense route stores
            if (wf == 0.0f || !(a_width > 0.0)) return;
            window_line(sh, far, G, Er, a_width, (double)wf);
        });
    }
    __syncthreads();
    window_flush(sh, G.n, partial + ((int64_t)d * gridDim.x + blockIdx.x) * 4 * G.n);
}

// Given lines (one target): the windowed counterpart of hr_sincsq_reduce.
__global__ void __launch_bounds__(kWinThreads) window_given_kernel(const double *__restrict__ E_r,
                                                                   const double *__restrict__ a,
                                                                   const float *__restrict__ w, int64_t n_lines,
                                                                   WindowGrid G, double *__restrict__ partial,
                                                                   double *__restrict__ dep_far) {
    __shared__ WindowShared sh;
    window_zero(sh, G.n);
    __syncthreads();
    for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n_lines;
         l += (int64_t)gridDim.x * blockDim.x) {
        const float wf = w[l];
        if (wf == 0.0f || !(a[l] > 0.0)) continue;
        window_line(sh, dep_far, G, E_r[l], a[l], (double)wf);
    }
    __syncthreads();
    window_flush(sh, G.n, partial + (int64_t)blockIdx.x * 4 * G.n);
}

// Sum the block partials of each direction in a fixed order; spectra to spec,
// in-grid moments to the node slots [n, 2n) of the moment array.
__global__ void __launch_bounds__(256) window_sum_kernel(const double *__restrict__ partial, int64_t n_blocks,
                                                         int64_t n, int n_dir, double *__restrict__ spec,
                                                         double *__restrict__ dep) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n_dir * 4 * n) return;
    const int64_t d = i / (4 * n), r = i % (4 * n);
    double s = 0.0;
    for (int64_t b = 0; b < n_blocks; ++b) s += partial[(d * n_blocks + b) * 4 * n + r];
    const int64_t part = r / n, e = r % n;
    if (part == 0) spec[d * n + e] = s;
    else dep[d * 9 * n + (part - 1) * 3 * n + n + e] = s;
}

// out[d][e] += spec[d][e] + sum_q tail(dep[d][.][q], (e - (q - n)) dE) over node slots q in [q_lo, q_hi).
constexpr int kConvTile = 256;
__global__ void __launch_bounds__(kConvTile) window_tail_kernel(const double *__restrict__ spec,
                                                                const double *__restrict__ dep, int64_t n,
                                                                double dE, int64_t q_lo, int64_t q_hi,
                                                                double *__restrict__ out) {
    __shared__ double s_m[3][kConvTile];
    __shared__ double s_inv[2 * kConvTile];
    const int d = blockIdx.y;
    const int64_t e0 = (int64_t)blockIdx.x * kConvTile;
    const int t = threadIdx.x;
    const int64_t e = e0 + t;
    const double *D = dep + (int64_t)d * 9 * n;
    const int64_t L = 3 * n;
    double acc = 0.0;
    for (int64_t q0 = q_lo; q0 < q_hi; q0 += kConvTile) {
        const int64_t q = q0 + t;
        const bool in = q < q_hi;
        s_m[0][t] = in ? D[q] : 0.0;
        s_m[1][t] = in ? D[L + q] : 0.0;
        s_m[2][t] = in ? D[2 * L + q] : 0.0;
        // offsets e - (q - n) over the tile: o = (e0 - q0 + n) + (t_e - t_q), t_e - t_q in [-255, 255]
        for (int k = t; k < 2 * kConvTile; k += kConvTile) {
            const int64_t o = (e0 - q0 + n) + (int64_t)(k - (kConvTile - 1));
            s_inv[k] = o == 0 ? 0.0 : 1.0 / ((double)o * dE);
        }
        __syncthreads();
        const int m = q_hi - q0 < kConvTile ? (int)(q_hi - q0) : kConvTile;
        for (int j = 0; j < m; ++j) {
            const double inv = s_inv[t - j + (kConvTile - 1)];
            acc += (s_m[0][j] + (s_m[1][j] + s_m[2][j] * inv) * inv) * inv * inv;
        }
        __syncthreads();
    }
    if (e < n) out[(int64_t)d * n + e] += spec[(int64_t)d * n + e] + acc;
}

__global__ void fill_zero_kernel(double *__restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Workspace bytes for gx partial rows of n_E bins plus the second level.
size_t partial_bytes(int64_t gx, int64_t n_E) {
    const int64_t splits = gx < kPartialSplits ? gx : kPartialSplits;
    return (size_t)(gx + splits) * (size_t)n_E * sizeof(double);
}

// out[k] += the fixed two-level sum of gx partial rows held at ws.
hr_status reduce_partials(void *ws, int64_t gx, int64_t n_E, double *out, hipStream_t s) {
    const double *partial = (const double *)ws;
    const int64_t splits = gx < kPartialSplits ? gx : kPartialSplits;
    const int64_t chunk = (gx + splits - 1) / splits;
    double *level = (double *)ws + gx * n_E;
    const unsigned bx = (unsigned)((n_E + 255) / 256);
    hipLaunchKernelGGL(sum_partials_split_kernel, dim3(bx, (unsigned)splits), dim3(256), 0, s, partial, gx, chunk,
                       n_E, level);
    HR_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(sum_partials_kernel, dim3(bx), dim3(256), 0, s, (const double *)level, splits, n_E, out);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

// Blocks along the line axis: enough to fill the chip several times over, but
// never more than there are line tiles.
int64_t line_blocks(int cu, int64_t n_tiles, int64_t bin_tiles) {
    int64_t want = (int64_t)cu * 8 / (bin_tiles > 0 ? bin_tiles : 1);
    if (want < 1) want = 1;
    return n_tiles < want ? (n_tiles > 0 ? n_tiles : 1) : want;
}

constexpr int kDenseBins = 4;  // bins per thread of the dense reducers

// Launch shape and workspace bytes of a dense reducer over n_items items in
// tiles of `tile`, on a grid of n_E bins, for the current device.
struct DensePlan {
    int64_t gx = 0, bin_tiles = 0;
    size_t bytes = 0;
};
hr_status dense_plan(int64_t n_items, int64_t tile, int64_t n_E, DensePlan &P) {
    int cu = 0;
    hr_status st = device_cu(cu);
    if (st != HR_OK) return st;
    P.bin_tiles = (n_E + kRedThreads * kDenseBins - 1) / (kRedThreads * kDenseBins);
    const int64_t n_tiles = (n_items + tile - 1) / tile;
    P.gx = line_blocks(cu, n_tiles, P.bin_tiles);
    P.bytes = (n_items == 0 || n_E == 0) ? 0 : partial_bytes(P.gx, n_E);
    return HR_OK;
}

// Blocks of the standalone windowed reduction and its workspace bytes:
// [partials blocks*4*n | spec n | node moments 9n].
int64_t window_reduce_blocks(int cu, int64_t n_lines) {
    int64_t blocks = (n_lines + kWinThreads - 1) / kWinThreads;
    return blocks < 2 * (int64_t)cu ? blocks : 2 * (int64_t)cu;
}

// Line-axis chunks per direction of the fused pixel route (about two resident
// blocks per CU in total; one block fills a CU's LDS).
int64_t window_chunks(int cu, int64_t n_seg, int32_t n_dir) {
    int64_t chunks = (2 * (int64_t)cu + n_dir - 1) / n_dir;
    const int64_t max_chunks = (n_seg + kWinThreads - 1) / kWinThreads;
    chunks = chunks < max_chunks ? chunks : max_chunks;
    return chunks < 1 ? 1 : chunks;
}

}  // namespace
}  // namespace hr

extern "C" hr_status hr_cxr_lines(const hr_line_args *args, void *stream) {
    HR_TRACE("hr_cxr_lines");
    using namespace hr;
    HR_REQUIRE(args != nullptr, "hr_cxr_lines: null args");
    const hr_line_args &p = *args;
    HR_REQUIRE(p.n_seg >= 0 && p.n_dir >= 1 && p.n_g >= 1 && p.n_tab >= 2 && p.n_mu >= 1,
               "hr_cxr_lines: bad sizes (n_seg=%lld n_dir=%d n_g=%d n_tab=%lld n_mu=%d)",
               (long long)p.n_seg, p.n_dir, p.n_g, (long long)p.n_tab, p.n_mu);
    if (p.n_seg == 0) return HR_OK;
    HR_REQUIRE(p.r_mid && p.v_hat && p.L && p.E && p.n_hat && p.dir_weight && p.g && p.wm && p.row &&
                   p.e_s && p.e_p && p.E_tab && p.delta_tab && p.chi_re && p.chi_im && p.u_re && p.u_im &&
                   p.log_mu && p.E_r && p.a && p.w,
               "hr_cxr_lines: null pointer");
    const int64_t total = p.n_seg * (int64_t)p.n_dir;
    const int block = 256;
    const int64_t grid = (total + block - 1) / block;
    HR_REQUIRE(grid < (1LL << 31), "hr_cxr_lines: %lld segment-direction pairs exceed one launch",
               (long long)total);
    hipLaunchKernelGGL(cxr_lines_kernel, dim3((unsigned)grid), dim3(block), 0, (hipStream_t)stream, p);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

extern "C" hr_status hr_sincsq_reduce(const hr_sincsq_args *args, void *stream) {
    HR_TRACE("hr_sincsq_reduce");
    using namespace hr;
    HR_REQUIRE(args != nullptr, "hr_sincsq_reduce: null args");
    const hr_sincsq_args &p = *args;
    HR_REQUIRE(p.n_lines >= 0 && p.n_E >= 0, "hr_sincsq_reduce: negative size");
    if (p.n_lines == 0 || p.n_E == 0) return HR_OK;
    HR_REQUIRE(p.E_r && p.a && p.w && p.out, "hr_sincsq_reduce: null pointer");
    HR_REQUIRE(p.uniform || p.E_grid, "hr_sincsq_reduce: a non-uniform grid needs E_grid");
    constexpr int BINS = kDenseBins;
    DensePlan P;
    hr_status st = dense_plan(p.n_lines, kLineTile, p.n_E, P);
    if (st != HR_OK) return st;
    HR_REQUIRE_WS(p.ws, P.bytes, "hr_sincsq_reduce");
    const int64_t gx = P.gx;
    void *ws = p.ws.ptr;
    hipStream_t s = (hipStream_t)stream;
    dim3 grid((unsigned)gx, (unsigned)P.bin_tiles);
    if (p.uniform) {
        hipLaunchKernelGGL(sincsq_uniform_kernel<BINS>, grid, dim3(kRedThreads), 0, s, p.E_r, p.a, p.w,
                           p.n_lines, p.E0, p.dE, p.n_E, (double *)ws);
    } else {
        hipLaunchKernelGGL(sincsq_general_kernel<BINS>, grid, dim3(kRedThreads), 0, s, p.E_r, p.a, p.w,
                           p.n_lines, p.E_grid, p.n_E, (double *)ws);
    }
    HR_CHECK_HIP(hipGetLastError());
    return reduce_partials(ws, gx, p.n_E, p.out, s);
}

// Scratch of hr_sincsq_reduce_static: the partial spectra of one call live in
// this file-scope device array instead of a caller workspace.  Each device
// holds its own copy at its own address, found once per device with
// hipGetSymbolAddress; host code uses that address, never the symbol itself.
static __device__ __align__(256) unsigned char scratch[HR_SINCSQ_STATIC_SCRATCH_BYTES];

namespace hr {
namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[64];

hr_status static_scratch(void **out) {
    int dev = 0;
    HR_CHECK_HIP(hipGetDevice(&dev));
    HR_REQUIRE(dev >= 0 && dev < 64, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        HR_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        HR_REQUIRE(p != nullptr && ((uintptr_t)p & 255u) == 0, "static scratch at %p is null or misaligned", p);
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return HR_OK;
}
}  // namespace
}  // namespace hr

extern "C" hr_status hr_sincsq_reduce_static(const hr_sincsq_args *args, void *stream) {
    HR_TRACE("hr_sincsq_reduce_static");
    using namespace hr;
    HR_REQUIRE(args != nullptr, "hr_sincsq_reduce_static: null args");
    HR_REQUIRE(args->n_lines >= 0 && args->n_E >= 0, "hr_sincsq_reduce_static: negative size");
    if (args->n_lines == 0 || args->n_E == 0) return HR_OK;
    DensePlan P;
    hr_status st = dense_plan(args->n_lines, kLineTile, args->n_E, P);
    if (st != HR_OK) return st;
    HR_REQUIRE(P.bytes <= (size_t)HR_SINCSQ_STATIC_SCRATCH_BYTES,
               "hr_sincsq_reduce_static: %lld bins need %zu bytes of partials, more than the %zu-byte static "
               "scratch (use hr_sincsq_reduce with a workspace)",
               (long long)args->n_E, P.bytes, (size_t)HR_SINCSQ_STATIC_SCRATCH_BYTES);
    void *ptr = nullptr;
    st = static_scratch(&ptr);
    if (st != HR_OK) return st;
    hr_sincsq_args q = *args;
    q.ws.ptr = ptr;
    q.ws.bytes = (size_t)HR_SINCSQ_STATIC_SCRATCH_BYTES;
    return hr_sincsq_reduce(&q, stream);
}

extern "C" hr_status hr_sincsq_workspace_bytes(const hr_sincsq_args *args, size_t *bytes) {
    using namespace hr;
    HR_REQUIRE(args != nullptr && bytes != nullptr, "hr_sincsq_workspace_bytes: null argument");
    HR_REQUIRE(args->n_lines >= 0 && args->n_E >= 0, "hr_sincsq_workspace_bytes: negative size");
    DensePlan P;
    hr_status st = dense_plan(args->n_lines, kLineTile, args->n_E, P);
    if (st != HR_OK) return st;
    *bytes = P.bytes;
    return HR_OK;
}

extern "C" hr_status hr_brem_workspace_bytes(const hr_brem_args *args, size_t *bytes) {
    using namespace hr;
    HR_REQUIRE(args != nullptr && bytes != nullptr, "hr_brem_workspace_bytes: null argument");
    HR_REQUIRE(args->n_seg >= 0 && args->n_E >= 0, "hr_brem_workspace_bytes: negative size");
    DensePlan P;
    hr_status st = dense_plan(args->n_seg, kBremTile, args->n_E, P);
    if (st != HR_OK) return st;
    *bytes = P.bytes;
    return HR_OK;
}

extern "C" hr_status hr_brem_reduce(const hr_brem_args *args, void *stream) {
    HR_TRACE("hr_brem_reduce");
    using namespace hr;
    HR_REQUIRE(args != nullptr, "hr_brem_reduce: null args");
    const hr_brem_args &p = *args;
    HR_REQUIRE(p.n_seg >= 0 && p.n_E >= 0, "hr_brem_reduce: negative size");
    HR_REQUIRE(p.n_el >= 1 && p.n_el <= 4, "hr_brem_reduce: 1..4 elements supported, got %d", p.n_el);
    if (p.n_seg == 0 || p.n_E == 0) return HR_OK;
    HR_REQUIRE(p.r_mid && p.v_hat && p.L && p.E_keV && p.E_grid_eV && p.mu && p.out,
               "hr_brem_reduce: null pointer");
    constexpr int BINS = kDenseBins;
    DensePlan P;
    hr_status st = dense_plan(p.n_seg, kBremTile, p.n_E, P);
    if (st != HR_OK) return st;
    HR_REQUIRE_WS(p.ws, P.bytes, "hr_brem_reduce");
    const int64_t gx = P.gx;
    void *ws = p.ws.ptr;
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid((unsigned)gx, (unsigned)P.bin_tiles);
    switch (p.n_el) {
    case 1: hipLaunchKernelGGL((brem_kernel<BINS, 1>), grid, dim3(kRedThreads), 0, s, p, (double *)ws); break;
    case 2: hipLaunchKernelGGL((brem_kernel<BINS, 2>), grid, dim3(kRedThreads), 0, s, p, (double *)ws); break;
    case 3: hipLaunchKernelGGL((brem_kernel<BINS, 3>), grid, dim3(kRedThreads), 0, s, p, (double *)ws); break;
    default: hipLaunchKernelGGL((brem_kernel<BINS, 4>), grid, dim3(kRedThreads), 0, s, p, (double *)ws); break;
    }
    HR_CHECK_HIP(hipGetLastError());
    return reduce_partials(ws, gx, p.n_E, p.out, s);
}

namespace hr {
namespace {

hr_status window_setup(int64_t n_E, double E0, double dE, double x_c, WindowGrid &G) {
    HR_REQUIRE(n_E >= 1 && n_E <= kWinMaxBins, "windowed reduction: n_E %lld outside [1, %d]", (long long)n_E,
               kWinMaxBins);
    HR_REQUIRE(dE > 0.0 && isfinite(E0), "windowed reduction: needs a uniform ascending grid");
    HR_REQUIRE(x_c > 0.0, "windowed reduction: x_c must be positive");
    G.n = n_E;
    G.E0 = E0;
    G.dE = dE;
    G.x_c = x_c;
    return HR_OK;
}

// Launch window_sum + window_tail over `blocks` partials per direction; the
// workspace holds [partials | spec | dep] and dep was zeroed before the deposit.
hr_status window_finish(double *partial, int64_t blocks, int n_dir, const WindowGrid &G, double *spec, double *dep,
                        int64_t q_lo, int64_t q_hi, double *out, hipStream_t s) {
    const int64_t n = G.n;
    const int64_t tot = (int64_t)n_dir * 4 * n;
    hipLaunchKernelGGL(window_sum_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, partial, blocks, n,
                       n_dir, spec, dep);
    HR_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(window_tail_kernel, dim3((unsigned)((n + kConvTile - 1) / kConvTile), (unsigned)n_dir),
                       dim3(kConvTile), 0, s, spec, dep, n, G.dE, q_lo, q_hi, out);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

}  // namespace
}  // namespace hr

extern "C" hr_status hr_window_reduce_workspace_bytes(int64_t n_lines, int64_t n_E, size_t *bytes) {
    using namespace hr;
    HR_REQUIRE(bytes != nullptr && n_lines >= 0 && n_E >= 0, "hr_window_reduce_workspace_bytes: bad arguments");
    *bytes = 0;
    if (n_lines == 0 || n_E == 0) return HR_OK;
    int cu = 0;
    hr_status st = device_cu(cu);
    if (st != HR_OK) return st;
    const int64_t blocks = window_reduce_blocks(cu, n_lines);
    *bytes = (size_t)(blocks * 4 * n_E + n_E + 9 * n_E) * sizeof(double);
    return HR_OK;
}

extern "C" hr_status hr_window_reduce(const double *E_r, const double *a, const float *w, int64_t n_lines,
                                      int64_t n_E, double E0, double dE, double x_c, double *out, hr_workspace ws,
                                      void *stream) {
    HR_TRACE("hr_window_reduce");
    using namespace hr;
    HR_REQUIRE(n_lines >= 0, "hr_window_reduce: negative n_lines");
    if (n_lines == 0 || n_E == 0) return HR_OK;
    HR_REQUIRE(E_r && a && w && out, "hr_window_reduce: null pointer");
    WindowGrid G;
    hr_status st = window_setup(n_E, E0, dE, x_c, G);
    if (st != HR_OK) return st;
    int cu = 0;
    st = device_cu(cu);
    if (st != HR_OK) return st;
    const int64_t blocks = window_reduce_blocks(cu, n_lines);
    const int64_t n = n_E;
    const size_t need = (size_t)(blocks * 4 * n + n + 9 * n) * sizeof(double);
    HR_REQUIRE_WS(ws, need, "hr_window_reduce");
    double *partial = (double *)ws.ptr, *spec = partial + blocks * 4 * n, *dep = spec + n;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((9 * n + 255) / 256)), dim3(256), 0, s, dep, 9 * n);
    HR_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(window_given_kernel, dim3((unsigned)blocks), dim3(kWinThreads), 0, s, E_r, a, w, n_lines, G,
                       partial, dep);
    HR_CHECK_HIP(hipGetLastError());
    return window_finish(partial, blocks, 1, G, spec, dep, 0, 3 * n, out, s);
}

extern "C" hr_status hr_window_spectra_workspace_bytes(const hr_window_args *args, size_t *bytes) {
    using namespace hr;
    HR_REQUIRE(args != nullptr && args->lines != nullptr && bytes != nullptr,
               "hr_window_spectra_workspace_bytes: null argument");
    const hr_line_args &p = *args->lines;
    HR_REQUIRE(p.n_seg >= 0 && p.n_dir >= 1 && args->n_E >= 0, "hr_window_spectra_workspace_bytes: bad sizes");
    *bytes = 0;
    if (p.n_seg == 0 || args->n_E == 0) return HR_OK;
    int cu = 0;
    hr_status st = device_cu(cu);
    if (st != HR_OK) return st;
    const int64_t n = args->n_E;
    const int64_t chunks = window_chunks(cu, p.n_seg, p.n_dir);
    *bytes = (size_t)((int64_t)p.n_dir * (chunks * 4 * n + n + 9 * n)) * sizeof(double);
    return HR_OK;
}

extern "C" hr_status hr_window_spectra(const hr_window_args *args, void *stream) {
    HR_TRACE("hr_window_spectra");
    using namespace hr;
    HR_REQUIRE(args != nullptr && args->lines != nullptr, "hr_window_spectra: null args");
    const hr_line_args &p = *args->lines;
    HR_REQUIRE(p.n_seg >= 0 && p.n_dir >= 1 && p.n_g >= 1, "hr_window_spectra: bad sizes");
    if (p.n_seg == 0 || args->n_E == 0) return HR_OK;
    HR_REQUIRE(p.r_mid && p.v_hat && p.L && p.E && p.n_hat && p.dir_weight && p.g && p.wm && p.row && p.e_s &&
                   p.e_p && p.E_tab && p.delta_tab && p.chi_re && p.chi_im && p.u_re && p.u_im && p.log_mu &&
                   args->out,
               "hr_window_spectra: null pointer");
    HR_REQUIRE(p.n_tab >= 2 && p.n_mu >= 1, "hr_window_spectra: tables too small");
    HR_REQUIRE(p.n_dir <= 65535, "hr_window_spectra: at most 65535 directions");
    WindowGrid G;
    hr_status st = window_setup(args->n_E, args->E0, args->dE, args->x_c, G);
    if (st != HR_OK) return st;
    int cu = 0;
    st = device_cu(cu);
    if (st != HR_OK) return st;
    const int64_t n = G.n;
    const int64_t chunks = window_chunks(cu, p.n_seg, p.n_dir);
    const size_t need = (size_t)((int64_t)p.n_dir * (chunks * 4 * n + n + 9 * n)) * sizeof(double);
    HR_REQUIRE_WS(args->ws, need, "hr_window_spectra");
    double *partial = (double *)args->ws.ptr;
    double *spec = partial + (int64_t)p.n_dir * chunks * 4 * n;
    double *dep = spec + (int64_t)p.n_dir * n;
    hipStream_t s = (hipStream_t)stream;
    const int64_t nd9 = (int64_t)p.n_dir * 9 * n;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((nd9 + 255) / 256)), dim3(256), 0, s, dep, nd9);
    HR_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(window_lines_kernel, dim3((unsigned)chunks, (unsigned)p.n_dir), dim3(kWinThreads), 0, s, p,
                       G, partial, dep);
    HR_CHECK_HIP(hipGetLastError());
    // node slots reachable from the keep window
    int64_t q_lo = (int64_t)floor((p.lo_keep - G.E0) / G.dE) - 1 + n;
    int64_t q_hi = (int64_t)ceil((p.hi_keep - G.E0) / G.dE) + 2 + n;
    q_lo = q_lo < 0 ? 0 : (q_lo > 3 * n ? 3 * n : q_lo);
    q_hi = q_hi < q_lo ? q_lo : (q_hi > 3 * n ? 3 * n : q_hi);
    return window_finish(partial, chunks, p.n_dir, G, spec, dep, q_lo, q_hi, args->out, s);
}
