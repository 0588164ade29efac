// vr.hip -- round-based transport and weight-window kernels of the learned
// variance reduction on gfx950.
//
// There is no counterpart in the reference: PyRITE's variance reduction is
// likelihood-ratio detector-cone splitting on the CPU core
// (src/pyrite/_cone_splitting.py:57, transport/cone_split.py:223) and line
// roulette (src/pyrite/_line_roulette.py:29).  Here an importance network
// (mlp.hip) scores particle states between transport rounds and a weight window
// splits or roulettes them.  Every decision keeps the expected weight:
//   split:    n copies of weight w/n (total weight exactly w),
//   roulette: survive with probability w/w_t at weight w_t,
// so every tally stays unbiased whatever the network predicts.
//
// Particles live in a structure-of-arrays bank (pos/dir/E/clock/weight/key/
// primary id).  A round advances every particle by at most `rows` transport
// rows with the analog step (transport_step.hpp), then the particle is
// suspended.  Resuming a suspended particle from a fresh stream is exact:
// the pending flight is exponential in optical depth (memoryless), so a new
// optical depth is drawn, and the Urban key domain comes from the new key.
#include "transport_step.hpp"

namespace hr {
namespace {

constexpr uint64_t kDecisionCounter = 1ULL << 40;       // draw of a particle's split/roulette decision
constexpr uint64_t kSpawnSalt = 0x6A09E667F3BCC909ULL;  // child-key domain (sqrt(2) fraction bits)

__host__ __device__ __forceinline__ uint64_t child_key(uint64_t key, int64_t j) {
    return splitmix64(splitmix64(key ^ kSpawnSalt) + kGolden * (uint64_t)(j + 1));
}

// One particle per thread: at most `rows` steps, then suspend.
template <bool kMixed>
__global__ void __launch_bounds__(256) vr_round_kernel(hr_transport_args a, int32_t rows,
                                                       const int32_t *__restrict__ rows_each) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.e_count) return;
    const int64_t e = a.e_start + i;
    int32_t mine = rows;
    if (rows_each) {
        const int32_t r = rows_each[e];
        mine = r < 1 ? 1 : (r < rows ? r : rows);
    }
    Lane s;
    load_lane(a, e, s);
    int code = -1;
    while (code < 0 && s.nseg < mine) code = step_lane<true, kMixed>(a, s);
    store_lane(a, s, code < 0 ? (int8_t)EXIT_SUSPENDED : (int8_t)code);
}

// Network inputs x = (E/E0, u_x, u_y, u_z, z/thickness, tilt/3 deg).  With a
// row length array the position is the row midpoint and the row start is used.
__global__ void __launch_bounds__(256) vr_features_kernel(int64_t n, const double *__restrict__ pos,
                                                          const double *__restrict__ dirs,
                                                          const double *__restrict__ E,
                                                          const double *__restrict__ seg_len, double inv_E0,
                                                          double inv_thickness, float tilt_feature,
                                                          float *__restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double dz = dirs[3 * i + 2];
    double z = pos[3 * i + 2];
    if (seg_len) z -= 0.5 * seg_len[i] * dz;
    x[6 * i + 0] = (float)(E[i] * inv_E0);
    x[6 * i + 1] = (float)dirs[3 * i];
    x[6 * i + 2] = (float)dirs[3 * i + 1];
    x[6 * i + 3] = (float)dz;
    x[6 * i + 4] = (float)(z * inv_thickness);
    x[6 * i + 5] = tilt_feature;
}

// Weight window on the predicted log-importance y: target weight
// w_t = clamp(exp(log_I_ref - y), w_min, w_max); ratio r = w / w_t.
__global__ void __launch_bounds__(256) vr_window_kernel(hr_vr_window_args p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    int32_t count = 0;
    double wc = 0.0, w2c = 0.0;
    const double w2 = p.weight2 ? p.weight2[i] : 0.0;
    if (p.exit_code[i] == EXIT_SUSPENDED) {
        const double w = p.weight[i];
        double wt = exp(p.log_I_ref - (double)p.log_importance[i]);
        wt = fmin(fmax(wt, p.w_min), p.w_max);
        const double r = w / wt;
        const double u = stream_uniform(p.key[i], kDecisionCounter);
        if (r > p.ratio_hi) {
            double fl = floor(r);
            if (u < r - fl) fl += 1.0;
            const double cap = (double)p.n_max;
            count = (int32_t)(fl < cap ? fl : cap);
            wc = w / (double)count;
            w2c = w2;
        } else if (r >= p.ratio_lo) {
            count = 1;
            wc = w;
            w2c = w2;
        } else if (u < r) {
            count = 1;
            wc = wt;
            w2c = w2 / r;
        }
    }
    p.count[i] = count;
    p.w_child[i] = wc;
    if (p.w2_child) p.w2_child[i] = w2c;
    if (p.rows_next) {
        // far below the split threshold: the next round may take rows_far rows
        const double s = p.weight[i] * exp((double)p.log_importance[i] - p.log_I_ref);
        p.rows_next[i] = (count > 0 && s < p.far_ratio) ? p.rows_far : p.rows_near;
    }
}

// Children of parent i occupy [offset[i], offset[i] + count[i]) of the output bank.
__global__ void __launch_bounds__(256) vr_spawn_kernel(hr_vr_spawn_args p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const int32_t c = p.count[i];
    if (c <= 0) return;
    const int64_t o = p.offset[i];
    const double px = p.pos[3 * i], py = p.pos[3 * i + 1], pz = p.pos[3 * i + 2];
    const double dx = p.dirs[3 * i], dy = p.dirs[3 * i + 1], dz = p.dirs[3 * i + 2];
    const double E = p.E[i], clock = p.clock[i], wc = p.w_child[i];
    const uint64_t key = p.key[i];
    const int64_t prim = p.prim[i];
    const double w2c = p.w2_child ? p.w2_child[i] : 0.0;
    const int32_t rn = p.rows_next ? p.rows_next[i] : 0;
    for (int32_t j = 0; j < c; ++j) {
        const int64_t k = o + j;
        p.out_pos[3 * k] = px;
        p.out_pos[3 * k + 1] = py;
        p.out_pos[3 * k + 2] = pz;
        p.out_dirs[3 * k] = dx;
        p.out_dirs[3 * k + 1] = dy;
        p.out_dirs[3 * k + 2] = dz;
        p.out_E[k] = E;
        p.out_clock[k] = clock;
        p.out_weight[k] = wc;
        p.out_key[k] = child_key(key, j);
        p.out_prim[k] = prim;
        if (p.out_weight2) p.out_weight2[k] = j == 0 ? w2c : 0.0;
        if (p.out_rows) p.out_rows[k] = rn;
    }
}

// Rows of a round (particle i, slot i*rows + s, s < seg_count[i]) gathered to
// [row_offset[i], ...) as FP32 scoring records with the particle's weight.
// One thread per slot (particle i = t / rows, row s = t % rows): consecutive
// threads read consecutive slot rows and write consecutive records.
__global__ void __launch_bounds__(256) vr_gather_kernel(hr_vr_gather_args p) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t i, s;
    if (p.n_out > 0) {  // one thread per written row
        if (t >= p.n_out) return;
        const int64_t r = p.row_offset[0] + t;
        i = row_owner(p.row_offset, p.n, r);
        s = r - p.row_offset[i];
    } else {  // one thread per slot
        if (t >= p.n * p.rows) return;
        i = t / p.rows;
        s = t - i * p.rows;
    }
    const int64_t m = p.seg_count[i] < p.rows ? p.seg_count[i] : p.rows;
    if (s >= m) return;
    const int64_t src = i * p.rows + s, dst = p.row_offset[i] + s;
    for (int c = 0; c < 3; ++c) {
        p.r_mid[3 * dst + c] = (float)p.seg_mid[3 * src + c];
        p.v_hat[3 * dst + c] = (float)p.seg_dir[3 * src + c];
    }
    p.L[dst] = (float)p.seg_len[src];
    p.E[dst] = (float)(0.5 * (p.seg_E[src] + p.seg_E_end[src]));
    p.row_weight[dst] = (float)p.weight[i];
    if (p.row_weight2) p.row_weight2[dst] = p.weight2 ? (float)p.weight2[i] : 0.0f;
    if (p.row_prim) p.row_prim[dst] = p.prim ? p.prim[i] : i;
}

// Expected window content of each segment's lines: a narrow sinc^2 line of
// weight w and slope a deposits w * pi / (a * dE) on a uniform grid of
// spacing dE, on average over its offset to the bins.  Lines are laid out
// (d * n_seg + j) * n_g + g.
__global__ void __launch_bounds__(256) line_mass_kernel(const double *__restrict__ E_r, const double *__restrict__ a,
                                                        const float *__restrict__ w, int64_t n_seg, int32_t n_dir,
                                                        int32_t n_g, double lo, double hi, double inv_dE,
                                                        double *__restrict__ mass) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_seg) return;
    double m = 0.0;
    for (int d = 0; d < n_dir; ++d) {
        for (int g = 0; g < n_g; ++g) {
            const int64_t k = ((int64_t)d * n_seg + j) * n_g + g;
            const double e = E_r[k], ak = a[k], wk = (double)w[k];
            if (wk != 0.0 && ak > 0.0 && e >= lo && e < hi) m += wk * kPi * inv_dE / ak;
        }
    }
    mass[j] = m;
}

// Segment j's lines survive with probability min(1, mass/m_thr), weight 1/p.
__global__ void __launch_bounds__(256) line_roulette_kernel(float *__restrict__ w, int64_t n_seg, int32_t n_dir,
                                                            int32_t n_g, const double *__restrict__ mass,
                                                            double m_thr, uint64_t key, int64_t row0) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_seg) return;
    const double m = mass[j];
    if (m >= m_thr) return;
    const double pkeep = m > 0.0 ? m / m_thr : 0.0;
    const double u = stream_uniform(key, (uint64_t)(row0 + j));
    const float scale = u < pkeep ? (float)(1.0 / pkeep) : 0.0f;
    for (int d = 0; d < n_dir; ++d)
        for (int g = 0; g < n_g; ++g) {
            const int64_t k = ((int64_t)d * n_seg + j) * n_g + g;
            w[k] = w[k] * scale;
        }
}

// Per-primary line roulette (montecarlo/runner/line_roulette.py): the draw of
// primary `id` is counter `id` of the stream keyed by seed ^ "LINEROUL"
// (:33-45), apart from every transport stream; pilot primaries and those at or
// above the threshold are scored surely, positive masses below it with
// probability `floor` (:114-126).
constexpr uint64_t kLineRouletteSalt = 0x4C494E45524F554CULL;

__global__ void __launch_bounds__(256) roulette_uniforms_kernel(uint64_t seed, int64_t first_id, int64_t n,
                                                                double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = stream_uniform(seed ^ kLineRouletteSalt, (uint64_t)(first_id + i));
}

__global__ void __launch_bounds__(256) line_screen_kernel(const double *__restrict__ mass, int64_t n,
                                                          int64_t first_id, int64_t pilot, double threshold,
                                                          double floor_p, uint64_t seed,
                                                          double *__restrict__ probability,
                                                          uint8_t *__restrict__ kept) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t id = first_id + i;
    const double m = mass[i];
    double p = 1.0;
    bool k = m > 0.0;
    if (id >= pilot && m > 0.0 && m < threshold) {
        p = floor_p;
        k = stream_uniform(seed ^ kLineRouletteSalt, (uint64_t)id) < floor_p;
    }
    probability[i] = p;
    kept[i] = k ? 1 : 0;
}

unsigned blocks_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

}  // namespace
}  // namespace hr

extern "C" {

hr_status hr_vr_round(const hr_transport_args *a, int32_t rows, const int32_t *rows_each, void *stream) {
    HR_TRACE("hr_vr_round");
    using namespace hr;
    HR_REQUIRE(a != nullptr, "hr_vr_round: null args");
    HR_REQUIRE(rows >= 1 && rows <= a->cap, "hr_vr_round: rows %d outside [1, cap=%lld]", rows, (long long)a->cap);
    HR_REQUIRE(a->e_count >= 0 && a->e_start >= 0, "hr_vr_round: bad electron range");
    if (a->e_count == 0) return HR_OK;
    HR_REQUIRE(a->lut_n_energy >= 2 && a->n_layers >= 1 && a->max_el >= 1, "hr_vr_round: bad tables");
    HR_REQUIRE(a->pos && a->dirs && a->E_keV && a->clock && a->E_cut && a->stream_key, "hr_vr_round: null state");
    HR_REQUIRE(a->seg_dir && a->seg_mid && a->seg_len && a->seg_E && a->seg_t0 && a->seg_id && a->seg_lay,
               "hr_vr_round: null row buffer");
    HR_REQUIRE(a->energy_model != 1 || (a->seg_E_end && a->seg_t_end && a->seg_flight && a->seg_substep &&
                                        a->seg_event),
               "hr_vr_round: midpoint rows need E_end/t_end/flight/substep/event");
    HR_REQUIRE(!a->straggle_on || a->stragg_dE, "hr_vr_round: straggling needs stragg_dE");
    HR_REQUIRE(a->seg_count && a->exit_code, "hr_vr_round: null seg_count/exit_code");
    HR_REQUIRE(a->e_count < (1LL << 40), "hr_vr_round: too many particles");
    HR_REQUIRE(!a->hard.rad_on && !a->hard.shell_on, "hr_vr_round: the hard-event modes run on hr_transport_run only");
    HR_REQUIRE(!a->cone.on, "hr_vr_round: cone splitting runs on hr_transport_run only");
    const hr_status est = elsepa::check(*a, "hr_vr_round");
    if (est != HR_OK) return est;
    HR_REQUIRE(a->precision == 0 || (a->precision == 1 && a->lut_total_rate_f && a->lut_dEds_f &&
                                      a->lut_inv_beta_f && a->lut_cdf_f && a->lut_alpha_f),
               "hr_vr_round: precision 0, or 1 with the FP32 table copies");
    if (a->precision == 1)
        hipLaunchKernelGGL(vr_round_kernel<true>, dim3(blocks_for(a->e_count, 256)), dim3(256), 0,
                           (hipStream_t)stream, *a, rows, rows_each);
    else
        hipLaunchKernelGGL(vr_round_kernel<false>, dim3(blocks_for(a->e_count, 256)), dim3(256), 0,
                           (hipStream_t)stream, *a, rows, rows_each);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

hr_status hr_vr_features(int64_t n, const double *pos, const double *dirs, const double *E, const double *seg_len,
                         double E0_keV, double thickness_ang, double tilt_deg, float *x, void *stream) {
    HR_TRACE("hr_vr_features");
    using namespace hr;
    HR_REQUIRE(n >= 0, "hr_vr_features: negative n");
    if (n == 0) return HR_OK;
    HR_REQUIRE(pos && dirs && E && x, "hr_vr_features: null pointer");
    HR_REQUIRE(E0_keV > 0.0 && thickness_ang > 0.0, "hr_vr_features: E0 and thickness must be positive");
    hipLaunchKernelGGL(vr_features_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, n, pos,
                       dirs, E, seg_len, 1.0 / E0_keV, 1.0 / thickness_ang, (float)(tilt_deg / 3.0), x);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

hr_status hr_vr_window(const hr_vr_window_args *p, void *stream) {
    HR_TRACE("hr_vr_window");
    using namespace hr;
    HR_REQUIRE(p != nullptr && p->n >= 0, "hr_vr_window: bad args");
    if (p->n == 0) return HR_OK;
    HR_REQUIRE(p->exit_code && p->log_importance && p->weight && p->key && p->count && p->w_child,
               "hr_vr_window: null pointer");
    HR_REQUIRE(p->w_min > 0.0 && p->w_min <= p->w_max, "hr_vr_window: need 0 < w_min <= w_max");
    HR_REQUIRE(p->ratio_lo > 0.0 && p->ratio_lo <= 1.0 && p->ratio_hi >= 1.0,
               "hr_vr_window: need 0 < ratio_lo <= 1 <= ratio_hi");
    HR_REQUIRE(p->n_max >= 1 && p->n_max <= 1024, "hr_vr_window: n_max %d outside [1, 1024]", p->n_max);
    HR_REQUIRE(!p->rows_next || (p->rows_near >= 1 && p->rows_far >= 1), "hr_vr_window: round lengths must be >= 1");
    hipLaunchKernelGGL(vr_window_kernel, dim3(blocks_for(p->n, 256)), dim3(256), 0, (hipStream_t)stream, *p);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

hr_status hr_vr_spawn(const hr_vr_spawn_args *p, void *stream) {
    HR_TRACE("hr_vr_spawn");
    using namespace hr;
    HR_REQUIRE(p != nullptr && p->n >= 0, "hr_vr_spawn: bad args");
    if (p->n == 0) return HR_OK;
    HR_REQUIRE(p->count && p->offset && p->w_child && p->pos && p->dirs && p->E && p->clock && p->key && p->prim,
               "hr_vr_spawn: null input");
    HR_REQUIRE(p->out_pos && p->out_dirs && p->out_E && p->out_clock && p->out_weight && p->out_key &&
                   p->out_prim,
               "hr_vr_spawn: null output");
    hipLaunchKernelGGL(vr_spawn_kernel, dim3(blocks_for(p->n, 256)), dim3(256), 0, (hipStream_t)stream, *p);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

hr_status hr_vr_gather(const hr_vr_gather_args *p, void *stream) {
    HR_TRACE("hr_vr_gather");
    using namespace hr;
    HR_REQUIRE(p != nullptr && p->n >= 0 && p->rows >= 1, "hr_vr_gather: bad args");
    if (p->n == 0) return HR_OK;
    HR_REQUIRE(p->seg_count && p->row_offset && p->weight && p->seg_mid && p->seg_dir && p->seg_len && p->seg_E &&
                   p->seg_E_end,
               "hr_vr_gather: null input");
    HR_REQUIRE(p->r_mid && p->v_hat && p->L && p->E && p->row_weight, "hr_vr_gather: null output");
    HR_REQUIRE(p->n <= (INT64_MAX / p->rows) && blocks_for(p->n * p->rows, 256) < (1ULL << 31),
               "hr_vr_gather: %lld slots exceed one launch", (long long)p->n * (long long)p->rows);
    HR_REQUIRE(p->n_out >= 0 && p->n_out <= p->n * p->rows, "hr_vr_gather: n_out %lld outside [0, %lld]",
               (long long)p->n_out, (long long)(p->n * p->rows));
    const int64_t threads = p->n_out > 0 ? p->n_out : p->n * p->rows;
    hipLaunchKernelGGL(vr_gather_kernel, dim3(blocks_for(threads, 256)), dim3(256), 0, (hipStream_t)stream, *p);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

hr_status hr_line_mass(const double *E_r, const double *a, const float *w, int64_t n_seg, int32_t n_dir,
                       int32_t n_g, double lo_eV, double hi_eV, double dE_eV, double *mass, void *stream) {
    HR_TRACE("hr_line_mass");
    using namespace hr;
    HR_REQUIRE(n_seg >= 0 && n_dir >= 1 && n_g >= 1, "hr_line_mass: bad sizes");
    if (n_seg == 0) return HR_OK;
    HR_REQUIRE(E_r && a && w && mass, "hr_line_mass: null pointer");
    HR_REQUIRE(dE_eV > 0.0, "hr_line_mass: dE must be positive");
    hipLaunchKernelGGL(line_mass_kernel, dim3(blocks_for(n_seg, 256)), dim3(256), 0, (hipStream_t)stream, E_r, a,
                       w, n_seg, n_dir, n_g, lo_eV, hi_eV, 1.0 / dE_eV, mass);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

hr_status hr_line_roulette(float *w, int64_t n_seg, int32_t n_dir, int32_t n_g, const double *mass, double m_thr,
                           uint64_t key, int64_t row0, void *stream) {
    HR_TRACE("hr_line_roulette");
    using namespace hr;
    HR_REQUIRE(n_seg >= 0 && n_dir >= 1 && n_g >= 1 && row0 >= 0, "hr_line_roulette: bad sizes");
    if (n_seg == 0) return HR_OK;
    HR_REQUIRE(w && mass, "hr_line_roulette: null pointer");
    HR_REQUIRE(m_thr > 0.0, "hr_line_roulette: m_thr must be positive");
    hipLaunchKernelGGL(line_roulette_kernel, dim3(blocks_for(n_seg, 256)), dim3(256), 0, (hipStream_t)stream, w,
                       n_seg, n_dir, n_g, mass, m_thr, key, row0);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

hr_status hr_roulette_uniforms(uint64_t seed, int64_t first_id, int64_t n, double *out, void *stream) {
    HR_TRACE("hr_roulette_uniforms");
    using namespace hr;
    HR_REQUIRE(n >= 0 && first_id >= 0, "hr_roulette_uniforms: bad range");
    if (n == 0) return HR_OK;
    HR_REQUIRE(out != nullptr, "hr_roulette_uniforms: null output");
    hipLaunchKernelGGL(roulette_uniforms_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, seed,
                       first_id, n, out);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

hr_status hr_line_screen(const double *mass, int64_t n, int64_t first_id, int64_t pilot, double threshold,
                         double floor_p, uint64_t seed, double *probability, uint8_t *kept, void *stream) {
    HR_TRACE("hr_line_screen");
    using namespace hr;
    HR_REQUIRE(n >= 0 && first_id >= 0 && pilot >= 0, "hr_line_screen: bad range");
    HR_REQUIRE(floor_p > 0.0 && floor_p <= 1.0, "hr_line_screen: floor must lie in (0, 1]");
    if (n == 0) return HR_OK;
    HR_REQUIRE(mass && probability && kept, "hr_line_screen: null pointer");
    hipLaunchKernelGGL(line_screen_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, mass, n,
                       first_id, pilot, threshold, floor_p, seed, probability, kept);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}
}
