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

#ifdef __cplusplus
}
#endif
#endif /* HIPRITE_H */
