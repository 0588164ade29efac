/*
 * hiprite.h -- C ABI of libhiprite.so, the MI355X (gfx950) hot path of PyRITE.
 *
 * Plain C: fixed-width integers, plain pointers and sizes, no C++ or torch
 * types.  Every pointer named *_d is a device pointer (hipMalloc memory or a
 * torch.cuda tensor's data_ptr()); host pointers are named *_h.  Every entry
 * point returns an hr_status; no exception crosses the ABI.  `stream` is a
 * hipStream_t passed as void* (NULL = the null stream); launches are
 * asynchronous on it unless the comment says otherwise.
 *
 * Each entry point cites the reference interface it replaces (paths relative
 * to the PyRITE tree, src/pyrite/...).
 */
#ifndef HIPRITE_H
#define HIPRITE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t hr_status;
#define HR_OK 0
#define HR_ERR_INVALID 1     /* bad argument (shape, null pointer, range) */
#define HR_ERR_HIP 2         /* a HIP runtime call failed; see hr_last_error() */
#define HR_ERR_OOM 3         /* device allocation failed */
#define HR_ERR_UNSUPPORTED 4 /* combination the kernels do not implement */

/* Human-readable text of the last error on the calling thread. */
const char *hr_last_error(void);
/* Library version string, e.g. "hiprite 0.1.0 gfx950". */
const char *hr_version(void);
/* Number of visible HIP devices (0 when none); never launches a kernel. */
int32_t hr_device_count(void);

/* ------------------------------------------------------------------------
 * Counter-addressed SplitMix64 stream keys.
 * Replaces montecarlo/transport/kinematics.py:226 stream_keys(seed, Ne, start=)
 * (host NumPy there): key[e] = splitmix64(seed + PHI*(start+e+1)).  Derived on
 * the device so 1e8-electron runs never stage 0.8 GB of keys over PCIe.
 * ---------------------------------------------------------------------- */
hr_status hr_stream_keys(uint64_t seed, int64_t start, int64_t n, uint64_t *keys_d, void *stream);

/* ------------------------------------------------------------------------
 * Per-electron LUT transport core.
 * Replaces the device core the reference batch driver calls,
 *   core((e_start, e_count, cap, keys), control, geometry, materials, mott,
 *        state, scratch, (seg_count, exit_code), (straggle_on, stragg_dE), ...)
 * montecarlo/transport/batching.py:547-555, positional signature
 * montecarlo/transport/_jit_launch.py:234-252, kernel
 * montecarlo/transport/_jit_lut_kernel.py:64-574, CPU twin
 * montecarlo/transport/cores.py:96 make_cpu_transport_core(per_electron=True, lut=True).
 *
 * Semantics kept from the reference contract:
 *  - electron e = e_start + i (i local); the state arrays are indexed by e;
 *  - row s of local electron i goes to slot i*cap + s only while s < cap, but
 *    transport continues and seg_count[i] reports the true count (capacity
 *    replay is the caller's job);
 *  - pos, dirs, E_keV and clock are updated in place;
 *  - exit_code[i] in {0 cutoff, 1 backscattered, 2 transmitted, 3 side,
 *    4 step-limited, 5 not-entered};
 *  - draws are counter-addressed by (key[e], counter), so results do not
 *    depend on launch geometry or batch split.
 * All pointers are device pointers.  Rows are structure-of-arrays.
 * ---------------------------------------------------------------------- */
typedef struct hr_transport_args {
    /* run */
    int64_t e_start, e_count, cap;
    const uint64_t *stream_key; /* [Ne] */
    /* control */
    int32_t max_steps;
    int32_t energy_model; /* 1 midpoint, 0 frozen */
    double max_dE_frac;   /* 0 disables the energy-loss substep cap */
    int32_t straggle_on;  /* Urban straggling on its own key domain */
    /* geometry */
    int32_t n_layers;
    const double *internal_bounds; /* [max(n_layers-1,1)] */
    double z_total;
    int32_t finite_footprint;
    double width_ang, height_ang;
    const int32_t *L_nel; /* [n_layers] */
    const double *L_top;  /* [n_layers] */
    const double *L_bot;  /* [n_layers] */
    /* materials: padded (n_layers, max_el) rows */
    int32_t max_el;
    const double *L_Z, *L_J, *L_k, *L_coeff, *L_E_cross;
    /* LUT, uniform in ln E */
    double lut_log_E_min, lut_inv_dlogE;
    int32_t lut_n_energy;
    const double *lut_total_rate; /* [n_layers, n_E]          1/cm   */
    const double *lut_dEds;       /* [n_layers, n_E]          keV/A  */
    const double *lut_inv_beta;   /* [n_E]                           */
    const double *lut_cdf;        /* [n_layers, max_el, n_E]         */
    const double *lut_alpha;      /* [n_layers, max_el, n_E]         */
    /* state, indexed by global electron id */
    const uint8_t *alive;
    double *clock, *pos, *dirs, *E_keV;
    const double *E_cut;
    double *stragg_dE;
    /* rows (midpoint schema, montecarlo/transport/batching.py:401-425) */
    double *seg_dir, *seg_mid, *seg_len, *seg_E, *seg_t0;
    int64_t *seg_id;
    int16_t *seg_lay;
    double *seg_E_end, *seg_t_end;
    int64_t *seg_flight, *seg_substep;
    int8_t *seg_event;
    /* per local electron */
    int64_t *seg_count;
    int8_t *exit_code;
    /* launch shape: 0 picks the tuned default */
    int32_t block_size;
} hr_transport_args;

hr_status hr_transport_run(const hr_transport_args *args, void *stream);

/* ------------------------------------------------------------------------
 * Incoherent PXR/CBS line prologue over dumped segments.
 * Replaces the batched line route's per-(segment, reflection) steps 1-7,
 * montecarlo/spectrum/cxr/_batched.py:461-728 (_batched_block,
 * _line_amplitudes, _escape_weights) with the fused device bodies of
 * montecarlo/spectrum/cxr/_fused.py:22-389 and the exact line elements of
 * montecarlo/spectrum/cxr/_recoil.py:35-76, for a laterally infinite slab.
 * One call covers n_dir detector directions; line (d, j, g) is written to
 * index (d*n_seg + j)*n_g + g as (E_r [eV], a [1/eV], w).  The direction's
 * solid-angle weight and `scale` (1/Ne) are folded into w, so summing the
 * lines' sinc^2 profiles gives the solid-angle-integrated per-electron
 * spectrum (mc_cxr_spectrum_solid_angle, cxr/_spectrum.py:401).  Rejected
 * pairs are written as (0, 0, 0).  Segment records and tables are FP32 (the
 * reference's device REAL); the arithmetic is FP64.
 * ---------------------------------------------------------------------- */
typedef struct hr_line_args {
    int64_t n_seg;
    const float *r_mid; /* [n_seg,3] A */
    const float *v_hat; /* [n_seg,3] */
    const float *L;     /* [n_seg] A */
    const float *E;     /* [n_seg] representative energy keV */
    double thickness;   /* slab [A] */
    int32_t n_dir;
    const double *n_hat;      /* [n_dir,3] sample frame */
    const double *dir_weight; /* [n_dir] solid angle [sr] */
    int32_t n_g;
    const double *g;    /* [n_g,3] 1/A */
    const double *wm;   /* [n_g] orientation weight */
    const int32_t *row; /* [n_g] coupling-table row */
    const double *e_s;  /* [n_dir,n_g,3] polarization basis */
    const double *e_p;  /* [n_dir,n_g,3] */
    int64_t n_tab;
    const float *E_tab;     /* [n_tab] eV ascending */
    const float *delta_tab; /* [n_tab] 1 - Re n */
    const float *chi_re, *chi_im, *u_re, *u_im; /* [rows, n_tab] */
    int32_t n_mu;
    const float *log_mu; /* [n_mu, n_tab] log(1/A) per absorber element */
    double lo_keep, hi_keep; /* resonance keep window [eV] */
    double scale;            /* multiplies every weight (1/Ne) */
    double *E_r;             /* [n_dir*n_seg*n_g] */
    double *a;
    float *w;
} hr_line_args;

hr_status hr_cxr_lines(const hr_line_args *args, void *stream);

/* ------------------------------------------------------------------------
 * Dense sinc^2 line reduction: out[k] += sum_j w_j [sin x/x]^2,
 * x = a_j (E_k - E_r,j), x = 0 -> 1.
 * Replaces montecarlo/spectrum/cxr_jit_kernel.py:561 run_reduction_kernel
 * (and the CPU twin cxr/_cpu_reduce.py:26 _sincsq_reduce).  FP64 phase,
 * FP32 profile, FP64 accumulation; no atomics (fixed-order partial sums).
 * `uniform` selects the grid E_k = E0 + k*dE (E_grid may then be NULL).
 * Uses a library-held per-device workspace: concurrent calls on different
 * streams of one device must not overlap.
 * ---------------------------------------------------------------------- */
typedef struct hr_sincsq_args {
    int64_t n_lines;
    const double *E_r;
    const double *a;
    const float *w;
    int64_t n_E;
    const double *E_grid; /* [n_E] eV, or NULL when uniform */
    int32_t uniform;
    double E0, dE;
    double *out; /* [n_E] FP64, accumulated */
} hr_sincsq_args;

hr_status hr_sincsq_reduce(const hr_sincsq_args *args, void *stream);

/* ------------------------------------------------------------------------
 * Bethe-Heitler + Elwert bremsstrahlung track-length reduction:
 * out[k] += sum_seg sum_el n_el L dsigma/dk(T, k) * <exp(-mu_k path)>.
 * Replaces montecarlo/spectrum/brem_jit_kernel.py:779 run_brem_reduction_kernel
 * (cell :76-153; incident prefactor brem.py:396-423); the caller divides by
 * 4 pi Ne (brem.py:1169).  FP32 cells, FP64 sums, single absorber layer,
 * 1..4 elements.  Same workspace rule as hr_sincsq_reduce.
 * ---------------------------------------------------------------------- */
typedef struct hr_brem_args {
    int64_t n_seg;
    const float *r_mid; /* [n_seg,3] A (the escape path reads z) */
    const float *v_hat; /* [n_seg,3] */
    const float *L;     /* [n_seg] A */
    const float *E_keV; /* [n_seg] representative energy */
    double thickness;   /* slab [A]; escape paths from both segment ends */
    double n_hat[3];    /* observation direction (by value) */
    int32_t n_el;     /* 1..4 */
    double Z[4];      /* atomic numbers (by value) */
    double n_cm3[4];  /* number densities [1/cm^3] (by value) */
    int64_t n_E;
    const double *E_grid_eV; /* [n_E] */
    const float *mu;         /* [n_E] 1/A total attenuation */
    double *out;             /* [n_E] FP64, accumulated */
} hr_brem_args;

hr_status hr_brem_reduce(const hr_brem_args *args, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* HIPRITE_H */
