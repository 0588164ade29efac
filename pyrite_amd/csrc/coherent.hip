// coherent.hip -- the coherent line route on gfx950: per-line complex
// records, phased field accumulation, and the squaring finalize.
//
// Follows the reference's coherent stream pipeline (prologue
// coherent_prologue_jit_kernel.py; field coherent_stream_jit_kernel.py:81-450;
// finalize :452-477) and its FP64 CPU route (cxr/_batched.py:561-690,
// cxr/_formation.py:66-115, cxr/_recoil.py:118-147,
// cxr/_batched_finalize.py:201-235).  The fields of every segment of a
// reflection row are summed before squaring, so the planes persist across
// batches and are squared once at the end.
//
// Precision policy (MI355X-first):
//  * the prologue is FP64 (shared line kinematics, line_core.hpp); positions
//    and ages are read from the FP64 transport rows, since they enter phases
//    of up to ~1e7 rad;
//  * the field cell evaluates both phases -- the formation argument v and the
//    propagation phase phi -- in FP64 turns, reduces them to half a turn in
//    FP64, and takes sin/cos from the FP32 hardware units; the formation factor
//    and the complex products are FP32, summed in FP32 over groups of 8 lines
//    and in FP64 beyond;
//  * no atomics: a launch splits each row's lines over S workgroups, each
//    writes its partial field, and a second pass adds the partials in split
//    order, so results are bitwise reproducible for a fixed launch shape.
#include <math.h>

#include "hr_common.hpp"
#include "line_core.hpp"

namespace hr {
namespace {

using namespace line;

constexpr int kRecFields = 13;
constexpr double kFormationSmallQ = 0.5;  // cxr/_formation.py:62
constexpr float kSeriesW2 = 1.0e-6f;      // cxr/_formation.py:59
constexpr double kInv2Pi = 0.15915494309189535;

inline int64_t blocks_of(int64_t n, int t) { return (n + t - 1) / t; }

// ---------------------------------------------------------------- prologue
// One thread per written row (its particle found by binary search over the
// row offsets; with n_lines = 0, one thread per slot); for each direction
// and reflection it writes the 13 record planes at column
// first + row_offset[i] + s (coalesced along the rows).
__global__ void __launch_bounds__(256) coherent_lines_kernel(hr_coherent_args a, hr_line_args p) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t i, s;
    if (a.n_lines > 0) {
        if (tid >= a.n_lines) return;
        const int64_t r = a.row_offset[0] + tid;
        i = row_owner(a.row_offset, a.n, r);
        s = r - a.row_offset[i];
    } else {
        if (tid >= a.n * a.rows) return;
        i = tid / a.rows;
        s = tid - i * a.rows;
    }
    const int64_t m = a.seg_count[i] < a.rows ? a.seg_count[i] : a.rows;
    if (s >= m) return;
    const int64_t t = i * a.rows + s;
    const int64_t col = a.first + a.row_offset[i] + s;
    const int64_t plane = (int64_t)p.n_dir * p.n_g * a.ld;
    const double rx = a.seg_mid[3 * t], ry = a.seg_mid[3 * t + 1], rz = a.seg_mid[3 * t + 2];
    const double vhx = a.seg_dir[3 * t], vhy = a.seg_dir[3 * t + 1], vhz = a.seg_dir[3 * t + 2];
    const double L = a.seg_len[t];
    const double E = 0.5 * (a.seg_E[t] + a.seg_E_end[t]);
    const GridGuess guess = grid_guess(p.E_tab, p.n_tab);
    for (int d = 0; d < p.n_dir; ++d) {
        const double nx = p.n_hat[3 * d], ny = p.n_hat[3 * d + 1], nz = p.n_hat[3 * d + 2];
        const SegKin k = seg_kin(rz, vhx, vhy, vhz, L, E, nx, ny, nz, p.thickness);
        // midpoint age and offset-free retardation (_setup.py:1025, :721-722)
        const double age = a.seg_t0[t] + 0.5 * k.t_L;
        const double d_geom = age - (rx * nx + ry * ny + rz * nz);
        for (int g = 0; g < p.n_g; ++g) {
            const LineCore o = line_core(p, k, nx, ny, nz, d, g, guess);
            double r[kRecFields];
#pragma unroll
            for (int f = 0; f < kRecFields; ++f) r[f] = 0.0;
            if (o.keep) {
                const double amp = sqrt(kAlpha * o.omega / (4.0 * kPi * kPi * kHbarc));
                // formation_coefficients(L_start mu, L_end mu)
                const double tau0 = k.path0 * o.mu, tau1 = k.path1 * o.mu;
                const double q = 0.25 * (tau1 - tau0);
                double apb, bma;
                if (fabs(q) < kFormationSmallQ) {
                    const double c0 = exp(-0.25 * (tau0 + tau1));
                    apb = 2.0 * c0 * cosh(q);
                    bma = -2.0 * c0 * sinh(q);
                } else {
                    const double ea = exp(-0.5 * tau0), eb = exp(-0.5 * tau1);
                    apb = ea + eb;
                    bma = eb - ea;
                }
                // _recoil_formation_line: exact vacuum centre, width, recoil phase shifts
                const double numerator = o.vdg - o.r_num;
                const double denom = k.denom_row + o.r_den;
                const double E_c = kHbarc * numerator / denom;
                const double x = numerator / denom * k.inv_E;
                const double kappa = 1.0 / (1.0 - x);
                const double a_vac = kappa * denom * k.t_L / (2.0 * kHbarc);
                const double d_shift = kappa * (x * k.denom_row + o.r_den) * age;
                const double g_shift = kappa * (x * o.vdg - o.r_num) * age;
                const double Gx = p.g[3 * g], Gy = p.g[3 * g + 1], Gz = p.g[3 * g + 2];
                const double ct = amp * k.t_L;
                const double cs_re = ct * o.A_re[0], cs_im = ct * o.A_im[0];
                const double cp_re = ct * o.A_re[1], cp_im = ct * o.A_im[1];
                const bool good = isfinite(amp) && amp > 0.0 && k.t_L > 0.0 && isfinite(apb) && apb > 0.0 &&
                                  isfinite(bma) && isfinite(cs_re) && isfinite(cs_im) && isfinite(cp_re) &&
                                  isfinite(cp_im);
                if (good) {
                    r[0] = E_c;
                    r[1] = a_vac;
                    r[2] = (d_geom + d_shift) / kHbarc;
                    r[3] = rx * Gx + ry * Gy + rz * Gz + g_shift;
                    r[4] = cs_re;
                    r[5] = cs_im;
                    r[6] = cp_re;
                    r[7] = cp_im;
                    r[8] = 0.5 * (k.path0 + k.path1);
                    r[9] = 0.5 * (k.path1 - k.path0);
                    r[10] = apb;
                    r[11] = bma;
                    r[12] = q;
                }
            }
            double *out = a.rec + ((int64_t)d * p.n_g + g) * a.ld + col;
#pragma unroll
            for (int f = 0; f < kRecFields; ++f) out[f * plane] = r[f];
        }
    }
}

// ------------------------------------------------------------------ field
// sin and cos of 2 pi u for u in turns: u is reduced in FP64 to
// r = u - rint(u) in [-1/2, 1/2] (exact) and the hardware's v_sin_f32 /
// v_cos_f32, which take their argument in turns, do the rest (measured on
// gfx950 over 4M points of [-1/2, 1/2]: max abs error 1.25e-7 against FP64).
__device__ __forceinline__ void sincos_turns(double u, float &s, float &c) {
    const float r = (float)(u - rint(u));
    s = __builtin_amdgcn_sinf(r);
    c = __builtin_amdgcn_cosf(r);
}

constexpr int kFieldThreads = 256;
constexpr int kFieldTile = 256;
constexpr int kFieldGroup = 8;  // lines summed in FP32 before each FP64 add

// Line constants staged in LDS: FP64 terms of the two phases (in turns); FP32
// coefficients and the formation factor's constants with its 1/2 folded in
// (h_apb = apb/2, h_bma = bma/2, h_qapb = q apb/2, h_qbma = q bma/2).
struct FieldTile {
    double a_t[kFieldTile], aEc_t[kFieldTile], hdL_t[kFieldTile];
    double ps_t[kFieldTile], gp_t[kFieldTile], Le_t[kFieldTile];
    float c[4][kFieldTile], h_apb[kFieldTile], h_bma[kFieldTile], h_qapb[kFieldTile], h_qbma[kFieldTile];
    float q[kFieldTile];
    int flag[kFieldTile];  // 0 rejected, 1 live, 2 live and |q| small enough for the series branch
};

// Stage record column j (row planes R, plane stride `plane`) into tile slot t.
__device__ __forceinline__ void stage_line(FieldTile &T, const double *__restrict__ R, int64_t plane, int64_t j,
                                           int tid) {
        const double cs_re = R[4 * plane + j], cs_im = R[5 * plane + j];
        const double cp_re = R[6 * plane + j], cp_im = R[7 * plane + j];
        const bool live = cs_re != 0.0 || cs_im != 0.0 || cp_re != 0.0 || cp_im != 0.0;
        const double a = live ? R[plane + j] : 0.0;
        const double Ec = live ? R[j] : 0.0;
        const double apb = live ? R[10 * plane + j] : 0.0, bma = live ? R[11 * plane + j] : 0.0;
        const double q = live ? R[12 * plane + j] : 0.0;
        T.a_t[tid] = a * kInv2Pi;
        T.aEc_t[tid] = a * Ec * kInv2Pi;
        T.ps_t[tid] = live ? R[2 * plane + j] * kInv2Pi : 0.0;
        T.gp_t[tid] = live ? R[3 * plane + j] * kInv2Pi : 0.0;
        T.Le_t[tid] = live ? R[8 * plane + j] * kInv2Pi : 0.0;
        T.hdL_t[tid] = live ? R[9 * plane + j] * kInv2Pi : 0.0;
        T.c[0][tid] = (float)cs_re;
        T.c[1][tid] = (float)cs_im;
        T.c[2][tid] = (float)cp_re;
        T.c[3][tid] = (float)cp_im;
        T.h_apb[tid] = (float)(0.5 * apb);
        T.h_bma[tid] = (float)(0.5 * bma);
        T.h_qapb[tid] = (float)(0.5 * q * apb);
        T.h_qbma[tid] = (float)(0.5 * q * bma);
        T.q[tid] = (float)q;
        T.flag[tid] = live ? ((float)q * (float)q < kSeriesW2 ? 2 : 1) : 0;
}

// One (line l, bin m) cell added to the FP32 group sums.
template <bool kSeries, int BINS>
__device__ __forceinline__ void field_cell(const FieldTile &T, int l, const double (&E)[BINS],
                                           const double (&D)[BINS], float (&g)[BINS][4]) {
    const double a_t = T.a_t[l], aEc_t = T.aEc_t[l], hdL_t = T.hdL_t[l];
    const double ps_t = T.ps_t[l], gp_t = T.gp_t[l], Le_t = T.Le_t[l];
    const float csr = T.c[0][l], csi = T.c[1][l], cpr = T.c[2][l], cpi = T.c[3][l];
    const float h_apb = T.h_apb[l], h_bma = T.h_bma[l], h_qapb = T.h_qapb[l], h_qbma = T.h_qbma[l];
    const float q = T.q[l], qq = q * q;
#pragma unroll
    for (int m = 0; m < BINS; ++m) {
        const double uv = fma(-hdL_t, D[m], fma(a_t, E[m], -aEc_t));
        const double uph = fma(-Le_t, D[m], fma(ps_t, E[m], -gp_t));
        float sv, cv, sp, cp;
        sincos_turns(uv, sv, cv);
        sincos_turns(uph, sp, cp);
        const float v = (float)uv * 6.2831853071795865f;
        const float w2 = fmaf(v, v, qq);
        const float inv = __builtin_amdgcn_rcpf(w2);
        const float vs = v * sv, vc = v * cv;
        float fr = fmaf(vs, h_apb, -cv * h_qbma) * inv;
        float fi = -fmaf(vc, h_bma, sv * h_qapb) * inv;
        if (kSeries && w2 < kSeriesW2) {  // second-order series about w = 0
            fr = h_apb * (1.0f - qq * (1.0f / 3.0f) - v * v * (1.0f / 6.0f));
            fi = -h_apb * q * v * (1.0f / 3.0f);
        }
        const float xr = fmaf(fr, cp, -fi * sp), xi = fmaf(fr, sp, fi * cp);
        g[m][0] = fmaf(csr, xr, fmaf(-csi, xi, g[m][0]));
        g[m][1] = fmaf(csr, xi, fmaf(csi, xr, g[m][1]));
        g[m][2] = fmaf(cpr, xr, fmaf(-cpi, xi, g[m][2]));
        g[m][3] = fmaf(cpr, xi, fmaf(cpi, xr, g[m][3]));
    }
}

// acc += the fields of tile lines [0, nt), in FP32 groups of kFieldGroup lines.
template <int BINS>
__device__ __forceinline__ void tile_fields(const FieldTile &T, int nt, const double (&E)[BINS],
                                            const double (&D)[BINS], double (&acc)[BINS][4]) {
    for (int l0 = 0; l0 < nt; l0 += kFieldGroup) {
        float g[BINS][4];
#pragma unroll
        for (int m = 0; m < BINS; ++m)
#pragma unroll
            for (int c = 0; c < 4; ++c) g[m][c] = 0.0f;
        const int l1 = l0 + kFieldGroup < nt ? l0 + kFieldGroup : nt;
        for (int l = l0; l < l1; ++l) {
            const int f = T.flag[l];  // uniform over the block
            if (f == 1) field_cell<false, BINS>(T, l, E, D, g);
            else if (f == 2) field_cell<true, BINS>(T, l, E, D, g);
        }
#pragma unroll
        for (int m = 0; m < BINS; ++m)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[m][c] += (double)g[m][c];
    }
}

// Block (split s, row r, bin window w): lines of row r in tiles s, s+S, ...;
// thread t owns bins w*256*BINS + t + 256*m.  Writes its partial field
// (or adds it when S == 1).
template <int BINS>
__global__ void __launch_bounds__(kFieldThreads) coherent_field_kernel(hr_coherent_field_args p, int W, int S,
                                                                       double *__restrict__ out) {
    __shared__ FieldTile T;
    const int tid = threadIdx.x;
    const int row = (int)(blockIdx.y / W), win = (int)(blockIdx.y % W);
    const int split = (int)blockIdx.x;
    const int64_t n_E = p.n_E;
    double E[BINS], D[BINS];
    bool valid[BINS];
#pragma unroll
    for (int m = 0; m < BINS; ++m) {
        const int64_t k = ((int64_t)win * BINS + m) * kFieldThreads + tid;
        valid[m] = k < n_E;
        E[m] = valid[m] ? p.E_grid[k] : 0.0;
        D[m] = valid[m] ? p.dom[k] : 0.0;
    }
    double acc[BINS][4];
#pragma unroll
    for (int m = 0; m < BINS; ++m)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[m][c] = 0.0;
    const int64_t plane = (int64_t)p.n_rows * p.ld;
    const double *R = p.rec + (int64_t)row * p.ld;
    const int64_t n_tiles = (p.n_lines + kFieldTile - 1) / kFieldTile;
    for (int64_t tile = split; tile < n_tiles; tile += S) {
        const int64_t j0 = tile * kFieldTile;
        const int nt = (int)(p.n_lines - j0 < kFieldTile ? p.n_lines - j0 : kFieldTile);
        __syncthreads();
        if (tid < nt) stage_line(T, R, plane, j0 + tid, tid);
        __syncthreads();
        tile_fields<BINS>(T, nt, E, D, acc);
    }
#pragma unroll
    for (int m = 0; m < BINS; ++m) {
        if (!valid[m]) continue;
        const int64_t k = ((int64_t)win * BINS + m) * kFieldThreads + tid;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (S == 1) out[((int64_t)row * 4 + c) * n_E + k] += acc[m][c];
            else out[(((int64_t)split * p.n_rows + row) * 4 + c) * n_E + k] = acc[m][c];
        }
    }
}

// Block (chunk s, row r, bin window w): groups [n_groups s / S, n_groups (s+1) / S)
// of row r, each group's columns [gs[e], gs[e+1]) summed to one FP64 field
// whose |F_s|^2 + |F_p|^2 is added to the thread's bins; writes its partial
// (or adds it when S == 1).
template <int BINS>
__global__ void __launch_bounds__(kFieldThreads) coherent_grouped_kernel(hr_coherent_field_args p,
                                                                         const int64_t *__restrict__ gs,
                                                                         int64_t n_groups, int W, int S,
                                                                         double *__restrict__ out) {
    __shared__ FieldTile T;
    const int tid = threadIdx.x;
    const int row = (int)(blockIdx.y / W), win = (int)(blockIdx.y % W);
    const int split = (int)blockIdx.x;
    const int64_t n_E = p.n_E;
    double E[BINS], D[BINS];
    bool valid[BINS];
#pragma unroll
    for (int m = 0; m < BINS; ++m) {
        const int64_t k = ((int64_t)win * BINS + m) * kFieldThreads + tid;
        valid[m] = k < n_E;
        E[m] = valid[m] ? p.E_grid[k] : 0.0;
        D[m] = valid[m] ? p.dom[k] : 0.0;
    }
    double G[BINS];
#pragma unroll
    for (int m = 0; m < BINS; ++m) G[m] = 0.0;
    const int64_t plane = (int64_t)p.n_rows * p.ld;
    const double *R = p.rec + (int64_t)row * p.ld;
    const int64_t g_lo = n_groups * split / S, g_hi = n_groups * (split + 1) / S;
    for (int64_t grp = g_lo; grp < g_hi; ++grp) {
        int64_t b = gs[grp], e = gs[grp + 1];
        b = b < 0 ? 0 : (b > p.n_lines ? p.n_lines : b);
        e = e < b ? b : (e > p.n_lines ? p.n_lines : e);
        double acc[BINS][4];
#pragma unroll
        for (int m = 0; m < BINS; ++m)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[m][c] = 0.0;
        for (int64_t j0 = b; j0 < e; j0 += kFieldTile) {
            const int nt = (int)(e - j0 < kFieldTile ? e - j0 : kFieldTile);
            __syncthreads();
            if (tid < nt) stage_line(T, R, plane, j0 + tid, tid);
            __syncthreads();
            tile_fields<BINS>(T, nt, E, D, acc);
        }
#pragma unroll
        for (int m = 0; m < BINS; ++m)
            G[m] += acc[m][0] * acc[m][0] + acc[m][1] * acc[m][1] + acc[m][2] * acc[m][2] + acc[m][3] * acc[m][3];
    }
#pragma unroll
    for (int m = 0; m < BINS; ++m) {
        if (!valid[m]) continue;
        const int64_t k = ((int64_t)win * BINS + m) * kFieldThreads + tid;
        if (S == 1) out[(int64_t)row * n_E + k] += G[m];
        else out[((int64_t)split * p.n_rows + row) * n_E + k] = G[m];
    }
}

// field[i] += sum_s partial[s][i] in split order, i over n_rows*4*n_E cells.
__global__ void __launch_bounds__(256) field_partials_kernel(const double *__restrict__ partial, int S,
                                                             int64_t cells, double *__restrict__ field) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells) return;
    double s = 0.0;
    for (int k = 0; k < S; ++k) s += partial[(int64_t)k * cells + i];
    field[i] += s;
}

__global__ void __launch_bounds__(256) coherent_finalize_kernel(const double *__restrict__ field, int n_dir, int n_g,
                                                                const double *__restrict__ dir_weight,
                                                                const double *__restrict__ wm, int64_t n_E,
                                                                double scale, double *__restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_E) return;
    double total = 0.0;
    for (int d = 0; d < n_dir; ++d) {
        double dir = 0.0;
        for (int g = 0; g < n_g; ++g) {
            const double *F = field + ((int64_t)d * n_g + g) * 4 * n_E + k;
            const double sr = F[0], si = F[n_E], pr = F[2 * n_E], pi_ = F[3 * n_E];
            dir += wm[g] * (sr * sr + si * si + pr * pr + pi_ * pi_);
        }
        total += dir_weight[d] * dir;
    }
    out[k] += scale * total;
}

// Decoherence blend per bin (coherent_population.py:83-96): with P the
// squared all-line field and G the grouped term of each row,
// pair < 0: (1 - F) G + F P; otherwise G + pair F (P - G).
__global__ void __launch_bounds__(256) coherent_blend_kernel(const double *__restrict__ field,
                                                             const double *__restrict__ grouped, int n_dir, int n_g,
                                                             const double *__restrict__ dir_weight,
                                                             const double *__restrict__ wm,
                                                             const double *__restrict__ factor, double pair,
                                                             int64_t n_E, double scale, double *__restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_E) return;
    const double F = factor[k];
    double total = 0.0;
    for (int d = 0; d < n_dir; ++d) {
        double dir = 0.0;
        for (int g = 0; g < n_g; ++g) {
            const int64_t r = (int64_t)d * n_g + g;
            const double *A = field + r * 4 * n_E + k;
            const double P = A[0] * A[0] + A[n_E] * A[n_E] + A[2 * n_E] * A[2 * n_E] + A[3 * n_E] * A[3 * n_E];
            const double Gr = grouped[r * n_E + k];
            const double mixed = pair < 0.0 ? (1.0 - F) * Gr + F * P : Gr + pair * F * (P - Gr);
            dir += wm[g] * mixed;
        }
        total += dir_weight[d] * dir;
    }
    out[k] += scale * total;
}

constexpr int kFieldBins = 4;
constexpr size_t kMaxPartialBytes = size_t(1) << 30;

// Launch shape of hr_coherent_field: S splits along the line axis (fill the
// chip ~8 workgroups per CU, never more than there are tiles, partials under
// kMaxPartialBytes) over n_rows*W row windows; workspace bytes for S > 1.
struct FieldPlan {
    int64_t W = 0, S = 1;
    size_t bytes = 0;
};
hr_status field_plan(const hr_coherent_field_args &p, FieldPlan &P) {
    int cu = 0;
    hr_status st = device_cu(cu);
    if (st != HR_OK) return st;
    const int64_t per_block = (int64_t)kFieldThreads * kFieldBins;
    P.W = (p.n_E + per_block - 1) / per_block;
    const int64_t n_tiles = (p.n_lines + kFieldTile - 1) / kFieldTile;
    const int64_t cells = (int64_t)p.n_rows * 4 * p.n_E;
    int64_t S = (int64_t)cu * 8 / ((int64_t)p.n_rows * P.W);
    if (S < 1) S = 1;
    if (S > n_tiles) S = n_tiles;
    const int64_t s_cap = (int64_t)(kMaxPartialBytes / ((size_t)cells * sizeof(double)));
    if (S > s_cap) S = s_cap > 1 ? s_cap : 1;
    if (S < 1) S = 1;
    P.S = S;
    P.bytes = (S > 1 && p.n_lines > 0) ? (size_t)S * (size_t)cells * sizeof(double) : 0;
    return HR_OK;
}

// Launch shape of hr_coherent_grouped: S chunks of groups (as the field plan,
// never more than there are groups), partials of n_rows*n_E cells.
hr_status grouped_plan(const hr_coherent_field_args &p, int64_t n_groups, FieldPlan &P) {
    int cu = 0;
    hr_status st = device_cu(cu);
    if (st != HR_OK) return st;
    const int64_t per_block = (int64_t)kFieldThreads * kFieldBins;
    P.W = (p.n_E + per_block - 1) / per_block;
    const int64_t cells = (int64_t)p.n_rows * p.n_E;
    int64_t S = (int64_t)cu * 8 / ((int64_t)p.n_rows * P.W);
    if (S > n_groups) S = n_groups;
    const int64_t s_cap = (int64_t)(kMaxPartialBytes / ((size_t)cells * sizeof(double)));
    if (S > s_cap) S = s_cap;
    if (S < 1) S = 1;
    P.S = S;
    P.bytes = (S > 1 && n_groups > 0) ? (size_t)S * (size_t)cells * sizeof(double) : 0;
    return HR_OK;
}

}  // namespace
}  // namespace hr

extern "C" hr_status hr_coherent_lines(const hr_coherent_args *args, void *stream) {
    HR_TRACE("hr_coherent_lines");
    using namespace hr;
    HR_REQUIRE(args != nullptr && args->lines != nullptr, "hr_coherent_lines: null args");
    const hr_coherent_args &a = *args;
    const hr_line_args &p = *a.lines;
    HR_REQUIRE(a.n >= 0 && a.rows >= 1, "hr_coherent_lines: bad sizes (n=%lld rows=%lld)", (long long)a.n,
               (long long)a.rows);
    HR_REQUIRE(p.n_dir >= 1 && p.n_g >= 1 && p.n_tab >= 2 && p.n_mu >= 1,
               "hr_coherent_lines: bad table sizes (n_dir=%d n_g=%d n_tab=%lld n_mu=%d)", p.n_dir, p.n_g,
               (long long)p.n_tab, p.n_mu);
    HR_REQUIRE(a.first >= 0 && a.ld >= a.first, "hr_coherent_lines: bad record columns (first=%lld ld=%lld)",
               (long long)a.first, (long long)a.ld);
    if (a.n == 0) return HR_OK;
    HR_REQUIRE(a.n <= INT64_MAX / a.rows, "hr_coherent_lines: too many slots");
    HR_REQUIRE(a.seg_count && a.row_offset && a.seg_mid && a.seg_dir && a.seg_len && a.seg_E && a.seg_E_end &&
                   a.seg_t0 && a.rec,
               "hr_coherent_lines: null row or record pointer");
    HR_REQUIRE(p.n_hat && p.g && p.row && p.e_s && p.e_p && p.E_tab && p.delta_tab && p.chi_re && p.chi_im &&
                   p.u_re && p.u_im && p.log_mu,
               "hr_coherent_lines: null table pointer");
    HR_REQUIRE(a.n_lines >= 0 && a.n_lines <= a.n * a.rows, "hr_coherent_lines: n_lines %lld outside [0, %lld]",
               (long long)a.n_lines, (long long)(a.n * a.rows));
    const int64_t total = a.n_lines > 0 ? a.n_lines : a.n * a.rows;
    hipLaunchKernelGGL(coherent_lines_kernel, dim3((unsigned)blocks_of(total, 256)), dim3(256), 0,
                       (hipStream_t)stream, a, p);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

extern "C" hr_status hr_coherent_field_workspace_bytes(const hr_coherent_field_args *args, size_t *bytes) {
    using namespace hr;
    HR_REQUIRE(args != nullptr && bytes != nullptr, "hr_coherent_field_workspace_bytes: null argument");
    HR_REQUIRE(args->n_lines >= 0 && args->n_rows >= 1 && args->n_E >= 1,
               "hr_coherent_field_workspace_bytes: bad sizes");
    FieldPlan P;
    hr_status st = field_plan(*args, P);
    if (st != HR_OK) return st;
    *bytes = P.bytes;
    return HR_OK;
}

extern "C" hr_status hr_coherent_field(const hr_coherent_field_args *args, void *stream) {
    HR_TRACE("hr_coherent_field");
    using namespace hr;
    HR_REQUIRE(args != nullptr, "hr_coherent_field: null args");
    const hr_coherent_field_args &p = *args;
    HR_REQUIRE(p.n_lines >= 0 && p.ld >= p.n_lines && p.n_rows >= 1 && p.n_E >= 1,
               "hr_coherent_field: bad sizes (n_lines=%lld ld=%lld n_rows=%d n_E=%lld)", (long long)p.n_lines,
               (long long)p.ld, p.n_rows, (long long)p.n_E);
    if (p.n_lines == 0) return HR_OK;
    HR_REQUIRE(p.rec && p.E_grid && p.dom && p.field, "hr_coherent_field: null pointer");
    FieldPlan P;
    hr_status st = field_plan(p, P);
    if (st != HR_OK) return st;
    const int64_t W = P.W, S = P.S;
    const int64_t cells = (int64_t)p.n_rows * 4 * p.n_E;
    HR_REQUIRE((int64_t)p.n_rows * W <= 65535, "hr_coherent_field: %lld row windows exceed the grid",
               (long long)(p.n_rows * W));
    HR_REQUIRE_WS(p.ws, P.bytes, "hr_coherent_field");
    void *ws = p.ws.ptr;
    hipStream_t s = (hipStream_t)stream;
    double *out = p.field;
    if (S > 1) out = (double *)ws;
    hipLaunchKernelGGL(coherent_field_kernel<kFieldBins>, dim3((unsigned)S, (unsigned)(p.n_rows * W)),
                       dim3(kFieldThreads), 0, s, p, (int)W, (int)S, out);
    HR_CHECK_HIP(hipGetLastError());
    if (S > 1) {
        hipLaunchKernelGGL(field_partials_kernel, dim3((unsigned)blocks_of(cells, 256)), dim3(256), 0, s,
                           (const double *)ws, (int)S, cells, p.field);
        HR_CHECK_HIP(hipGetLastError());
    }
    return HR_OK;
}

extern "C" hr_status hr_coherent_finalize(const double *field, int32_t n_dir, int32_t n_g, const double *dir_weight,
                                          const double *wm, int64_t n_E, double scale, double *out, void *stream) {
    HR_TRACE("hr_coherent_finalize");
    using namespace hr;
    HR_REQUIRE(n_dir >= 1 && n_g >= 1 && n_E >= 0, "hr_coherent_finalize: bad sizes (n_dir=%d n_g=%d n_E=%lld)",
               n_dir, n_g, (long long)n_E);
    if (n_E == 0) return HR_OK;
    HR_REQUIRE(field && dir_weight && wm && out, "hr_coherent_finalize: null pointer");
    hipLaunchKernelGGL(coherent_finalize_kernel, dim3((unsigned)blocks_of(n_E, 256)), dim3(256), 0,
                       (hipStream_t)stream, field, n_dir, n_g, dir_weight, wm, n_E, scale, out);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

extern "C" hr_status hr_coherent_grouped_workspace_bytes(const hr_coherent_grouped_args *args, size_t *bytes) {
    using namespace hr;
    HR_REQUIRE(args != nullptr && bytes != nullptr, "hr_coherent_grouped_workspace_bytes: null argument");
    const hr_coherent_field_args &p = args->field;
    HR_REQUIRE(p.n_lines >= 0 && p.n_rows >= 1 && p.n_E >= 1 && args->n_groups >= 0,
               "hr_coherent_grouped_workspace_bytes: bad sizes");
    FieldPlan P;
    hr_status st = grouped_plan(p, args->n_groups, P);
    if (st != HR_OK) return st;
    *bytes = P.bytes;
    return HR_OK;
}

extern "C" hr_status hr_coherent_grouped(const hr_coherent_grouped_args *args, void *stream) {
    HR_TRACE("hr_coherent_grouped");
    using namespace hr;
    HR_REQUIRE(args != nullptr, "hr_coherent_grouped: null args");
    const hr_coherent_field_args &p = args->field;
    const int64_t n_groups = args->n_groups;
    HR_REQUIRE(p.n_lines >= 0 && p.ld >= p.n_lines && p.n_rows >= 1 && p.n_E >= 1 && n_groups >= 0,
               "hr_coherent_grouped: bad sizes (n_lines=%lld ld=%lld n_rows=%d n_E=%lld n_groups=%lld)",
               (long long)p.n_lines, (long long)p.ld, p.n_rows, (long long)p.n_E, (long long)n_groups);
    if (p.n_lines == 0 || n_groups == 0) return HR_OK;
    HR_REQUIRE(p.rec && p.E_grid && p.dom && args->group_start && args->grouped,
               "hr_coherent_grouped: null pointer");
    FieldPlan P;
    hr_status st = grouped_plan(p, n_groups, P);
    if (st != HR_OK) return st;
    const int64_t W = P.W, S = P.S;
    const int64_t cells = (int64_t)p.n_rows * p.n_E;
    HR_REQUIRE((int64_t)p.n_rows * W <= 65535, "hr_coherent_grouped: %lld row windows exceed the grid",
               (long long)(p.n_rows * W));
    HR_REQUIRE_WS(p.ws, P.bytes, "hr_coherent_grouped");
    hipStream_t s = (hipStream_t)stream;
    double *out = S > 1 ? (double *)p.ws.ptr : args->grouped;
    hipLaunchKernelGGL(coherent_grouped_kernel<kFieldBins>, dim3((unsigned)S, (unsigned)(p.n_rows * W)),
                       dim3(kFieldThreads), 0, s, p, args->group_start, n_groups, (int)W, (int)S, out);
    HR_CHECK_HIP(hipGetLastError());
    if (S > 1) {
        hipLaunchKernelGGL(field_partials_kernel, dim3((unsigned)blocks_of(cells, 256)), dim3(256), 0, s,
                           (const double *)p.ws.ptr, (int)S, cells, args->grouped);
        HR_CHECK_HIP(hipGetLastError());
    }
    return HR_OK;
}

extern "C" hr_status hr_coherent_blend_finalize(const double *field, const double *grouped, int32_t n_dir,
                                                int32_t n_g, const double *dir_weight, const double *wm,
                                                int64_t n_E, const double *factor, double pair_scale, double scale,
                                                double *out, void *stream) {
    HR_TRACE("hr_coherent_blend_finalize");
    using namespace hr;
    HR_REQUIRE(n_dir >= 1 && n_g >= 1 && n_E >= 0,
               "hr_coherent_blend_finalize: bad sizes (n_dir=%d n_g=%d n_E=%lld)", n_dir, n_g, (long long)n_E);
    HR_REQUIRE(pair_scale < 0.0 || (pair_scale >= 0.0 && pair_scale <= 1.0e300),
               "hr_coherent_blend_finalize: pair_scale must be negative (sampled bunch) or finite");
    if (n_E == 0) return HR_OK;
    HR_REQUIRE(field && grouped && dir_weight && wm && factor && out, "hr_coherent_blend_finalize: null pointer");
    hipLaunchKernelGGL(coherent_blend_kernel, dim3((unsigned)blocks_of(n_E, 256)), dim3(256), 0,
                       (hipStream_t)stream, field, grouped, n_dir, n_g, dir_weight, wm, factor, pair_scale, n_E, scale,
                       out);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}
