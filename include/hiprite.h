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

/* Caller-owned device scratch of one call (partial sums, counters).  The
 * library allocates nothing on its own: each entry point that needs scratch
 * takes one of these, 256-byte aligned and at least as large as its
 * *_workspace_bytes query says for the same arguments on the current device
 * (the answer depends on the device's compute-unit count).  A workspace is
 * written by the launches of the call it is passed to, so calls that may run
 * at the same time (different streams) need different workspaces. */
typedef struct hr_workspace {
    void *ptr;    /* device memory */
    size_t bytes; /* its size */
} hr_workspace;

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
 * All pointers are device pointers.  Rows are structure-of-arrays.  Lanes claim
 * electrons from the caller's claim counter `work` (8 bytes of device memory,
 * reset by the call itself with a kernel): calls that may run at the same time
 * need different counters.  The call allocates nothing.
 * ---------------------------------------------------------------------- */
/* ------------------------------------------------------------------------
 * Hard-event modes of the transport step (both switches 0: the plain step).
 *
 * rad_on -- coupled radiative mode (montecarlo/transport/_jit_kernel.py:253-378,
 * 554-558, 866-905; cores.py:465-480, 880-906; device helpers
 * _jit_radiative_device.py:40-481): bremsstrahlung below rad_cutoff_eV is a
 * continuous loss (lut_rad_soft, subtracted from dEds and added to each
 * straggled row as its mean); photons above it are explicit events on their
 * own optical depth and key (rad_key = stream keys of seed ^
 * 0xA54FF53A5F1D36F1, _jit_radiative.py:17-22), scheduled at the row-start
 * rate (lut_rad_rate) and sampled, element and photon energy, from the exact
 * per-element moments at the row-end energy; the electron keeps its direction.
 * Row event 9 (HARD_RADIATIVE), photon energy seg_rad_k [eV] and emitting Z
 * seg_rad_Z (0 on other rows).
 *
 * shell_on -- shell soft/hard inelastic mode (cores.py:456-463, 616-650,
 * 850-880; hard_inelastic.py:120-554): lut_dEds holds the soft stopping,
 * soft losses of a straggled row come from the two-moment sampler (mean
 * -dEds s, variance lut_sh_omega2 s) on the straggling key domain, and hard
 * collisions are explicit events on their own optical depth and key
 * (splitmix64(stream_key ^ 0x3C6EF372FE94F82B), hard_inelastic.py:40,77-80):
 * channel by cumulative rate at the row start, loss by inverting the
 * channel's restricted CDF, primary deflection from its recoil kinematics.
 * Row event 8 (HARD_INELASTIC), loss seg_hard_W [keV] and channel code
 * seg_hard_ch (-1 on other rows).
 *
 * Both modes need the midpoint energy model.  Tables: radiative slot
 * s = layer*max_el + i (the layer's element order); shell channel c of layer
 * L at L*sh_max_ch + c; LUT columns on the transport LUT's ln-E grid.
 * ---------------------------------------------------------------------- */
typedef struct hr_hard_events {
    int32_t rad_on, shell_on;
    /* radiative */
    const uint64_t *rad_key;     /* [Ne] by global electron id */
    const double *lut_rad_soft;  /* [n_layers, n_E] keV/A */
    const double *lut_rad_rate;  /* [n_layers, n_E] 1/A */
    const double *rad_incident;  /* [slots, rad_max_t] keV */
    const double *rad_nominal;   /* [slots, rad_n_red] reduced photon energies */
    const double *rad_top;       /* [slots, rad_max_t] tip of each incident row */
    const double *rad_chi;       /* [slots, rad_max_t, rad_n_red] scaled SDCS, mb */
    const double *rad_z;         /* [slots] */
    const double *rad_ncm3;      /* [slots] atoms/cm^3 */
    const int32_t *rad_nt;       /* [slots] incident rows in use */
    const double *rad_plasma;    /* [n_layers] plasma energy, eV (0: no dielectric suppression) */
    int32_t rad_max_t, rad_n_red;
    double rad_cutoff_eV;
    /* shell */
    const double *lut_sh_rate;    /* [n_layers, n_E] total hard rate 1/A */
    const double *lut_sh_omega2;  /* [n_layers, n_E] soft variance per length keV^2/A */
    const double *lut_sh_ch_rate; /* [n_layers, sh_max_ch, n_E] 1/A */
    const double *sh_U, *sh_W;    /* [n_layers, sh_max_ch] ionisation and resonance energies, eV */
    const int32_t *sh_branch;     /* [n_layers, sh_max_ch] 0 distant longitudinal, 1 distant transverse, 2 close */
    const int16_t *sh_code;       /* [n_layers, sh_max_ch] channel code written to seg_hard_ch */
    const int32_t *sh_nch;        /* [n_layers] */
    int32_t sh_max_ch;
    double sh_cutoff_eV;
    /* mode rows (NULL: not written) */
    double *seg_hard_W;   /* keV */
    int16_t *seg_hard_ch;
    double *seg_rad_k;    /* eV */
    int16_t *seg_rad_Z;
} hr_hard_events;

/* Radiative LUT columns: soft loss [keV/A] and hard photon rate [1/A] of each
 * layer at each node of E_grid [keV] (row-major [n_layers, n_E]), from the
 * exact per-element moments (the reference's prepared moment lookup,
 * _jit_moment_lookup_device.py:10-70, on the transport LUT's grid). */
hr_status hr_radiative_lut(const hr_hard_events *h, int32_t n_layers, int32_t max_el, const int32_t *L_nel,
                           const double *E_grid_keV, int64_t n_E, double *soft_out, double *rate_out, void *stream);
/* Hard photon energies [eV] of element slot `slot` at energies E_eV[i] and
 * uniforms u[i], with the plasma energy of the slot's layer (the event
 * sampler, _jit_radiative_device.py:359-481). */
hr_status hr_radiative_sample(const hr_hard_events *h, int32_t slot, double plasma_eV, int64_t n, const double *E_eV,
                              const double *u, double *k_out, void *stream);

/* ------------------------------------------------------------------------
 * Detector-cone splitting of elastic events (the reference's opt-in VR:
 * src/pyrite/montecarlo/transport/cone_split.py:223-335 cone_split_event,
 * cores.py:1189-1226 its call site, _cone_splitting.py:57-141 the policy),
 * here on this transport's screened-Rutherford angular law.  At an elastic
 * event with E >= min_energy_keV, a copy may represent the outcomes inside
 * the cone C = {Omega : Omega . n >= cos_c}: its polar angle comes from the
 * analog law restricted to the band |theta - theta_dn| <= theta_c (exact:
 * the law's CDF (1 + alpha) mu / (alpha + mu) is inverted in closed form),
 * its azimuth uniformly from the arc inside C, and it carries the likelihood
 * ratio P_band dphi/pi.  A roulette against target_weight (and, with
 * angle_scale = u_ref > 0, a second one against target max(1, (u/u_ref)^2),
 * u = gamma angle(copy, n)) bounds the copy count; draws are counters 0-3 of
 * the event key splitmix64(split_key + G (flight + 1)), G the SplitMix64
 * increment, split_key = splitmix64(stream_key ^ 0xC2B2AE3D27D4EB4F).  The
 * primary keeps its analog draw and ends with exit code 6 (cone split) when
 * that draw lands in C.  Copy k of local electron i goes to slot i*cap + k
 * while k < cap; count[i] counts every copy made (count > cap: the caller
 * must rerun with a larger cap).  Not combined with the hard-event modes or
 * the continuation/round entry points (exit code 6 is "suspended" there).
 * ---------------------------------------------------------------------- */
#define HR_EXIT_CONE_SPLIT 6
typedef struct hr_cone_split {
    int32_t on, pad_;
    double nx, ny, nz;       /* detector direction (unit) */
    double cos_c, sin_c;     /* cone half-angle */
    double target_weight;    /* roulette target copy weight */
    double min_energy_keV;   /* events below this kinetic energy stay analog */
    double angle_scale;      /* u_ref of the angle roulette; 0 = off */
    const uint64_t *split_key; /* [Ne] by global electron id */
    int64_t cap;             /* copy slots per electron */
    int64_t *count;          /* [e_count] copies made */
    uint8_t *killed;         /* [e_count] 1: ended at a cone-entering event */
    double *pos, *dir;       /* [e_count*cap, 3] copy state at the event */
    double *E_keV, *clock, *weight; /* [e_count*cap] */
    int64_t *flight;         /* [e_count*cap] flight index of the event */
} hr_cone_split;

/* Tabulated (ELSEPA) elastic angles.  Replaces the exact core's
 * elastic_model_code == 2 branch (montecarlo/transport/_jit_kernel.py:1130-
 * 1160; device inversion _jit_device.py:493-620): at an elastic event the
 * chosen element's two energy rows bracketing ln E[eV] are inverted at the
 * angle uniform (guided bisection for the panel, then the exact root of the
 * piecewise-linear density's quadratic CDF) and the quantiles mixed linearly
 * in ln E; cos theta = 1 - 2 mu.  The rate and element choice stay on the LUT
 * (its columns built from the tables' log-log rates).  on = 0: the analytic
 * screened-Rutherford angle from lut_alpha. */
typedef struct hr_elsepa {
    int32_t on;
    int32_t n_mu;          /* angular nodes, shared by every row */
    int32_t guide_cells;   /* power of two */
    int32_t pad_;
    const int64_t *start;  /* [n_layers, max_el] first row of each element */
    const int64_t *length; /* [n_layers, max_el] its energy rows (>= 2) */
    const double *logE;    /* [rows] ln E[eV] */
    const double *mu;      /* [n_mu] */
    const double *cdf;     /* [rows, n_mu] */
    const double *pdf;     /* [rows, n_mu] */
    const int32_t *guide;  /* [rows, guide_cells + 1] */
} hr_elsepa;

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
    /* transport precision (transport/_jit_precision.py): 0 FP64; 1 mixed --
     * FP32 copies of the LUT columns below and FP32 table queries, ln E,
     * sqrt(1 - cos^2) and azimuth sin/cos, FP64 state, draws and straggling */
    int32_t precision;
    const float *lut_total_rate_f, *lut_dEds_f, *lut_inv_beta_f, *lut_cdf_f, *lut_alpha_f;
    /* claim counter of hr_transport_run (8 bytes, caller-owned; unused by the
     * continuation and round entry points, which run one lane per thread) */
    unsigned long long *work;
    /* hard-event modes (hr_hard_events; both switches 0 = the plain step) */
    hr_hard_events hard;
    /* detector-cone splitting (hr_cone_split; on = 0: analog) */
    hr_cone_split cone;
    /* tabulated elastic angles (hr_elsepa; on = 0: screened Rutherford) */
    hr_elsepa elsepa;
} hr_transport_args;

hr_status hr_transport_run(const hr_transport_args *args, void *stream);

/* Copy records of one cone-split launch, compacted in (electron, event) order
 * (cone_split.py:412-444 _ConeSplitPass.collect).  With ladder = 1 each copy
 * of weight w is kept with probability w/W, W = 2^ceil(log2 w), and carries W
 * (cone_split.py:140-165 weight_ladder; the draw is counter 4 of the copy's
 * event key).  Copy stream keys are splitmix64(event_key ^ 0x165667B19E3779F9)
 * (cone_split.py:168-179).  Writes n_out[0] = copies kept and n_out[1] = the
 * largest count[i]; nothing is written past out_cap copies (n_out[0] may
 * exceed out_cap: rerun with room).  parent is the local electron index.
 * Workspace: hr_cone_collect_workspace_bytes(e_count). */
typedef struct hr_cone_collect_args {
    const hr_cone_split *cone; /* the launch's copy slots, count and split keys */
    int64_t e_start, e_count;
    int32_t ladder, pad_;
    int64_t out_cap;
    int64_t *parent, *flight;
    double *pos, *dir, *E_keV, *clock, *weight;
    uint64_t *stream_key;
    int64_t *n_out;            /* [2], device */
    hr_workspace ws;
} hr_cone_collect_args;
int64_t hr_cone_collect_workspace_bytes(int64_t e_count);
hr_status hr_cone_collect(const hr_cone_collect_args *args, void *stream);
/* Split-decision keys splitmix64(stream_key[i] ^ 0xC2B2AE3D27D4EB4F) of n electrons. */
hr_status hr_cone_split_keys(const uint64_t *stream_key, int64_t n, uint64_t *out, void *stream);
/* Importance roulette of collected copies (no counterpart in the reference's
 * cone splitting, whose copy roulettes use the event state only): copy i of
 * weight w[i] and predicted log-importance log_imp[i] (per unit weight)
 * survives with p = clamp(w exp(log_imp) / T, p_min, 1), T = exp(log_T), and
 * then carries w/p; its draw is counter 0 of splitmix64(stream_key[i] ^
 * 0x61C8864680B583EB), disjoint from the copy's transport stream.  Writes
 * keep[i] (0/1) and the new weight in place (0 for a dropped copy). */
hr_status hr_cone_roulette(int64_t n, const float *log_imp, const uint64_t *stream_key, double log_T, double p_min,
                           double *weight, uint8_t *keep, void *stream);
/* One elastic event's copy, for tests: incident direction d[3], screening
 * alpha, kinetic energy E_keV, event key; out = (made, ox, oy, oz, weight). */
hr_status hr_cone_event(const hr_cone_split *cone, int64_t n, const double *d, const double *alpha,
                        const double *E_keV, const uint64_t *event_key, double *out, void *stream);
/* The same on a tabulated law, the reference's cone_split_event
 * (montecarlo/transport/cone_split.py:223-335): element row block el[i]
 * (layer * max_el + element) of `tables`, kinetic energy E_keV[i]. */
hr_status hr_cone_event_tab(const hr_cone_split *cone, const hr_elsepa *tables, int64_t n, const double *d,
                            const int64_t *el, const double *E_keV, const uint64_t *event_key, double *out,
                            void *stream);

/* ------------------------------------------------------------------------
 * Bounded continuation (montecarlo/transport/continuation.py:1-27 and the
 * round driver _drive_continuation_rounds, :506): a fixed population of
 * n = e_count lanes, each holding one history, advanced by at most `cap` rows
 * per launch.  Between launches the caller retires finished lanes and refills
 * them with pending primaries (fresh record: tau -1, every counter 0,
 * stragg_dE 0, and the primary's entry state and stream key in the lane).
 * The record carries every value the step loop holds between steps, so a
 * history's rows, exit code and final state are bit-identical to one
 * uninterrupted hr_transport_run, whatever the schedule.
 * State arrays, stream_key, alive, seg_count and exit_code are indexed by
 * lane (e_start must be 0); a launch's rows go to slot lane*cap + k for the
 * lane's k-th row of this launch, seg_id carries history[lane], and
 * seg_count[lane] counts this launch's rows.  exit_code is 0..5 for a history
 * that ended in this launch, 6 (suspended) for one still alive after `cap`
 * rows, and HR_LANE_EMPTY for a lane with history < 0.  alive[lane] is read
 * only at a history's first launch.
 * ---------------------------------------------------------------------- */
#define HR_LANE_EMPTY (-1)
typedef struct hr_continuation {
    const int64_t *history; /* [n] global electron id of each lane's history; < 0: empty lane */
    double *tau;            /* [n] open optical depth; < 0 draws a fresh one at the next step */
    uint64_t *draws;        /* [n] per-history stream counter */
    int64_t *nseg;          /* [n] rows the history has emitted before this launch (updated) */
    int64_t *flight, *substep;
    int32_t *steps;         /* [n] steps taken (max_steps applies to the whole history) */
} hr_continuation;

hr_status hr_transport_continue(const hr_transport_args *args, const hr_continuation *cont, void *stream);

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
    const float *row_weight; /* [n_seg] statistical weight of each row, or NULL (= 1) */
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
 * FP32 profile, FP64 accumulation; no atomics (fixed-order partial sums in
 * the caller's workspace, hr_sincsq_workspace_bytes).
 * `uniform` selects the grid E_k = E0 + k*dE (E_grid may then be NULL).
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
    hr_workspace ws;
} hr_sincsq_args;

hr_status hr_sincsq_reduce(const hr_sincsq_args *args, void *stream);
/* Workspace bytes hr_sincsq_reduce needs for these sizes on the current device. */
hr_status hr_sincsq_workspace_bytes(const hr_sincsq_args *args, size_t *bytes);

/* hr_sincsq_reduce without a workspace and without any allocation: the
 * partial spectra go to a static device array of
 * HR_SINCSQ_STATIC_SCRATCH_BYTES that each device holds until the process
 * ends (args->ws is ignored).  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops).  Sizes whose partials exceed the array are refused
 * with HR_ERR_INVALID (at most about 1M bins on a 256-CU device).  Every call
 * on one device shares the array: calls that overlap on one device (two
 * streams, or two host threads) race.  Results equal hr_sincsq_reduce's bit
 * for bit. */
#define HR_SINCSQ_STATIC_SCRATCH_BYTES (32u << 20)
hr_status hr_sincsq_reduce_static(const hr_sincsq_args *args, void *stream);

/* ------------------------------------------------------------------------
 * Bethe-Heitler + Elwert bremsstrahlung track-length reduction:
 * out[k] += sum_seg sum_el n_el L dsigma/dk(T, k) * <exp(-mu_k path)>.
 * Replaces montecarlo/spectrum/brem_jit_kernel.py:779 run_brem_reduction_kernel
 * (cell :76-153; incident prefactor brem.py:396-423); the caller divides by
 * 4 pi Ne (brem.py:1169).  FP32 cells, FP64 sums, single absorber layer,
 * 1..4 elements.  Fixed-order partial sums in the caller's workspace
 * (hr_brem_workspace_bytes).
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
    const float *row_weight; /* [n_seg] statistical weight of each row, or NULL (= 1) */
    double *out;             /* [n_E] FP64, accumulated */
    hr_workspace ws;
} hr_brem_args;

hr_status hr_brem_reduce(const hr_brem_args *args, void *stream);
hr_status hr_brem_workspace_bytes(const hr_brem_args *args, size_t *bytes);

/* ------------------------------------------------------------------------
 * Importance network (learned variance reduction).
 * No counterpart in the reference: PyRITE has no learned component (its
 * closest variance reduction is detector-cone splitting, src/pyrite/
 * _cone_splitting.py:57, transport/cone_split.py:223).  The network predicts
 * the (log) detected-photon yield to go from an electron state
 * x = (E/E0, ux, uy, uz, depth/thickness, tilt): 6 -> 64 -> 64 -> 1 with SiLU.
 * Parameters: FP32 master vector of hr_mlp_param_count() = 4673 floats in the
 * order W1[64][6], b1[64], W2[64][64], b2[64], w3[64], b3, plus BF16 shadows
 * of W1 (K padded to 8) and W2 read by the MFMA operands
 * (v_mfma_f32_32x32x8bf16_1k, FP32 accumulation).  x, h1 and dz2 are rounded
 * to BF16 before each GEMM.  All pointers are device pointers.
 * ---------------------------------------------------------------------- */
typedef struct hr_mlp {
    float *params; /* [4673] FP32 master */
    uint16_t *w1b; /* [64*8]  BF16 shadow of W1 */
    uint16_t *w2b; /* [64*64] BF16 shadow of W2 */
} hr_mlp;

int32_t hr_mlp_param_count(void);
/* Bytes of the gradient workspace hr_mlp_grad needs for a batch of n. */
size_t hr_mlp_workspace_bytes(int64_t n);
/* Refresh the BF16 shadows from the master vector. */
hr_status hr_mlp_shadow(const float *params_d, uint16_t *w1b_d, uint16_t *w2b_d, void *stream);
/* y[s] = net(x[s]) for s < n; x is [n,6] FP32. */
hr_status hr_mlp_forward(const hr_mlp *net, const float *x_d, int64_t n, float *y_d, void *stream);
/* Gradient of L = mean_s (net(x_s) - t_s)^2 w.r.t. the master vector (BF16
 * roundings straight-through); fixed-order reduction over per-wave partials. */
hr_status hr_mlp_grad(const hr_mlp *net, const float *x_d, const float *t_d, int64_t n,
                      float *workspace_d, float *grad_d, float *loss_d, void *stream);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on the master vector; refreshes shadows. */
hr_status hr_mlp_adam(hr_mlp *net, const float *grad_d, float *m1_d, float *m2_d, float lr,
                      int32_t step, void *stream);

/* ---- many configurations in one persistent work-queue transport kernel
 * Replaces a host loop of per-case transports (montecarlo/runner/case_driver.py:
 * 142-152 over transport/batching.py:477-628 per case) for tally-only
 * campaigns: every history of every configuration is a queue item; item i of
 * configuration c (offsets[c] <= i < offsets[c+1]) is its electron
 * e = i - offsets[c] with stream key splitmix64(seed + PHI*(e+1))
 * (transport/kinematics.py:220), a pencil beam from the origin along beam_dir.
 * No rows are stored.  Tally columns per configuration (FP64):
 *   0..6  histories ending with exit code 0..6
 *   7     rows, 8 rows^2
 *   9     sum of exit energies of transmitted histories [keV]
 *   10    sum of exit energies of backscattered histories [keV]
 *   11    energy deposited (E0 minus the energy carried out through a face) [keV] */
#define HR_QUEUE_TALLY 12
typedef struct hr_queue_config {
    hr_transport_args t; /* control, geometry, materials, LUT; run, state and row fields unused */
    uint64_t seed;
    double E0_keV, E_cut_keV;
    double beam_dir[3];
} hr_queue_config;
typedef struct hr_queue_launch {
    const hr_queue_config *configs; /* device [n_cfg], n_cfg <= 1024 */
    unsigned long long *counter;    /* caller-owned claim counter (8 bytes; reset by the launch) */
    int32_t n_cfg;
    const int64_t *offsets;         /* device [n_cfg + 1] */
    int64_t total;                  /* offsets[n_cfg] */
    double *tally;                  /* device [n_cfg, HR_QUEUE_TALLY] */
    int32_t accumulate;             /* 0: the tally is zeroed first (by a kernel, graph-safe) */
    int32_t block_size;             /* 0 (256) or 64..256 in wave64 steps */
    int32_t per_claim;              /* consecutive items a lane claims at once, 1..1024 */
    int32_t precision;              /* 0 FP64, 1 mixed (every configuration carries FP32 table copies) */
} hr_queue_launch;
hr_status hr_queue_run(const hr_queue_launch *launch, void *stream);
/* The same launch captured once into a hipGraph (reset + queue kernel) and
 * replayed by hr_queue_graph_launch; the graph uses launch->counter, which must
 * outlive it (replays of one graph, or of graphs sharing a counter, must not
 * overlap). */
hr_status hr_queue_graph_create(const hr_queue_launch *launch, void **graph);
hr_status hr_queue_graph_launch(void *graph, void *stream);
hr_status hr_queue_graph_destroy(void *graph);

/* ---- support-windowed line reduction with the cycle-averaged far tail
 * Replaces the per-row node-window route: row_window_plan and
 * _sincsq_window_reduce (src/pyrite/montecarlo/spectrum/cxr/_node_window.py:225-250,
 * :469-522) with the NodeTail deposit/apply (:86-180).  Exact sinc^2 on the
 * bins within |a (E - E_r)| <= x_c of each line (the reference default is
 * x_c = 100, DEFAULT_NODE_SUPPORT_CUTOFF), the tail m2/x^2 + m3/x^3 + m4/x^4
 * elsewhere; broad windows (a quarter of the grid) are exact everywhere.
 * Uniform grids E0 + k dE with n_E <= 4096.  Partials and node moments live in
 * the caller's workspace (hr_window_reduce_workspace_bytes). */
hr_status hr_window_reduce(const double *E_r, const double *a, const float *w, int64_t n_lines,
                           int64_t n_E, double E0, double dE, double x_c, double *out, hr_workspace ws,
                           void *stream);
hr_status hr_window_reduce_workspace_bytes(int64_t n_lines, int64_t n_E, size_t *bytes);

typedef struct hr_window_args {
    const hr_line_args *lines; /* prologue inputs; its E_r/a/w outputs are not used */
    int64_t n_E;
    double E0, dE;             /* uniform grid [eV] */
    double x_c;                /* exact support |a (E - E_r)| <= x_c */
    double *out;               /* [n_dir, n_E] FP64 per-direction (pixel) spectra, accumulated */
    hr_workspace ws;
} hr_window_args;
/* Per-direction spectra: the line prologue of hr_cxr_lines fused with the
 * windowed reduction, one spectrum per detector direction (pixel), without
 * materialising the lines (n_seg * n_dir * n_g of them). */
hr_status hr_window_spectra(const hr_window_args *args, void *stream);
hr_status hr_window_spectra_workspace_bytes(const hr_window_args *args, size_t *bytes);

/* ---- coherent line route: phased fields, squared once every batch is summed
 * Replaces the coherent stream pipeline: the prologue
 * (src/pyrite/montecarlo/spectrum/coherent_prologue_jit_kernel.py, CPU twin
 * cxr/_batched.py:561-690 _batched_coherent_block), the field accumulation
 * (coherent_stream_jit_kernel.py:523 run_coherent_field_accumulation_kernel)
 * and the finalize (:649 finalize_coherent_fields; CPU twin
 * cxr/_batched_finalize.py:201-235).  Reflections and directions stay
 * incoherent (one field row each); the flat term only (no inter-electron
 * decoherence blend).
 *
 * Line records, per (direction d, reflection g, segment j), 13 FP64 planes
 * rec[f][d*n_g + g][j]: f = 0 E_c (exact vacuum centre, eV), 1 a (sinc width,
 * 1/eV), 2 phase slope (t_mid - n.r + d_shift)/hbar c (1/eV), 3 lattice phase
 * g.r + g_shift, 4-7 c_s re/im, c_p re/im with c = sqrt(alpha omega/(4 pi^2
 * hbar c)) t_L A, 8 L_esc (mean escape path), 9 half_dL (half its change
 * along the segment), 10-12 apb, bma, q (formation-factor attenuation).
 * Rejected lines carry zero coefficients (and zeros elsewhere). */
typedef struct hr_coherent_args {
    const hr_line_args *lines; /* crystal, tables, n_hat/n_dir and keep window; its segment
                                  inputs, scale, dir_weight, row_weight and outputs are not used */
    int64_t n, rows;           /* transport slot rows: particle i, row s at i*rows + s */
    const int64_t *seg_count;  /* [n] rows written per particle (clamped to rows) */
    const int64_t *row_offset; /* [n] exclusive prefix sum of the clamped counts */
    const double *seg_mid;     /* [n*rows,3] midpoint [A] */
    const double *seg_dir;     /* [n*rows,3] */
    const double *seg_len;     /* [n*rows] [A] */
    const double *seg_E, *seg_E_end; /* [n*rows] keV; the row's energy is their mean */
    const double *seg_t0;      /* [n*rows] age at the segment start [A] (c = 1) */
    int64_t first;             /* record column of the first row written */
    int64_t ld;                /* column stride of the record planes (>= first + rows written) */
    double *rec;               /* [13][n_dir*n_g][ld] */
    int64_t n_lines;           /* rows written (sum of the clamped counts): one thread per row;
                                  0 = one thread per slot */
} hr_coherent_args;
hr_status hr_coherent_lines(const hr_coherent_args *args, void *stream);

typedef struct hr_coherent_field_args {
    const double *rec;  /* record planes of hr_coherent_lines */
    int64_t ld;         /* their column stride */
    int64_t n_lines;    /* columns [0, n_lines) are summed */
    int32_t n_rows;     /* n_dir*n_g field rows */
    int64_t n_E;
    const double *E_grid; /* [n_E] eV */
    const double *dom;    /* [n_E] delta(E) E / hbar c [1/A]: the escape-path phase rate */
    double *field;      /* [n_rows][4][n_E] FP64 (sigma re, sigma im, pi re, pi im), accumulated */
    hr_workspace ws;
} hr_coherent_field_args;
/* field[r][.][k] += sum_j c_j F(v_jk) exp(i phi_jk), v = a (E - E_c) -
 * half_dL dom, phi = phase_slope E - g_phase - L_esc dom, F = exp(-tau_c/2)
 * sinh(w)/w with w = i v - q (cxr/_formation.py:86-115).  FP64 phase and
 * sums, FP32 profile; each field cell has one writer per call (fixed-order
 * partial sums in the caller's workspace, no atomics). */
hr_status hr_coherent_field(const hr_coherent_field_args *args, void *stream);
hr_status hr_coherent_field_workspace_bytes(const hr_coherent_field_args *args, size_t *bytes);
/* out[k] += scale sum_d dir_weight[d] sum_g wm[g] (|F_s|^2 + |F_p|^2) */
hr_status hr_coherent_finalize(const double *field, int32_t n_dir, int32_t n_g, const double *dir_weight,
                               const double *wm, int64_t n_E, double scale, double *out, void *stream);

/* ---- inter-electron decoherence blend
 * Replaces the grouped-intensity reduction _grouped_intensity_kernel
 * (src/pyrite/montecarlo/spectrum/coherent_grouped_jit_kernel.py:62-230) and
 * the row estimator mixed_row_power (coherent_population.py:83-96).
 * grouped[r][k] += sum_e (|F_s,e|^2 + |F_p,e|^2), F_.,e the field of record
 * columns [group_start[e], group_start[e+1]) of row r (the same cell as
 * hr_coherent_field); group_start is ascending from 0 to field.n_lines
 * (reads are clamped into [0, n_lines]).  One writer per cell per call:
 * fixed-order partial sums in field.ws (hr_coherent_grouped_workspace_bytes);
 * field.field is not read. */
typedef struct hr_coherent_grouped_args {
    hr_coherent_field_args field;
    const int64_t *group_start; /* [n_groups + 1] device */
    int64_t n_groups;
    double *grouped;            /* [n_rows][n_E] FP64, accumulated */
} hr_coherent_grouped_args;
hr_status hr_coherent_grouped(const hr_coherent_grouped_args *args, void *stream);
hr_status hr_coherent_grouped_workspace_bytes(const hr_coherent_grouped_args *args, size_t *bytes);
/* out[k] += scale sum_d dir_weight[d] sum_g wm[g] M_dg(k), with P = |F_s|^2 + |F_p|^2
 * of the all-line field, G the grouped term and F = factor[k]:
 * pair_scale < 0 (sampled bunch): M = (1 - F) G + F P;
 * pair_scale >= 0 (physical population, (N-1)/(M-1)): M = G + pair_scale F (P - G). */
hr_status hr_coherent_blend_finalize(const double *field, const double *grouped, int32_t n_dir, int32_t n_g,
                                     const double *dir_weight, const double *wm, int64_t n_E, const double *factor,
                                     double pair_scale, double scale, double *out, void *stream);

/* ---- incident states on the entrance face
 * Replaces initial_beam_positions + primary_directions for an ungrooved
 * analytic beam (src/pyrite/montecarlo/transport/beam_entry.py:23-126,
 * :347-359; transverse.py:197-224; geometry.py:330-394).  Draw c of electron e
 * is u = ((splitmix64(k_e + PHI*(c+1)) >> 12) + 1/2) 2^-52 with
 * k_e = splitmix64(root + PHI*(e+1)) (kinematics.py:267-312) and the normal is
 * Phi^-1(u).  root is the SeedSequence child root the reference uses:
 * spawn(2)[1] for the spot, spawn(5)[4] for the Twiss beam. */
typedef struct hr_beam_args {
    int64_t start, n;        /* global electron ids [start, start+n) */
    int32_t mode;            /* 0 pencil, 1 Gaussian spot, 2 Courant-Snyder (Twiss) */
    uint64_t root;
    double sigma_x, sigma_y; /* spot sigma [A] (mode 1) */
    double twiss_x[3];       /* (sqrt(eps beta) [mm], sqrt(eps/beta) [rad], alpha) (mode 2) */
    double twiss_y[3];
    int32_t projected;       /* nonzero tilt: project lab offsets onto the z = 0 face */
    double e_x[3], e_y[3];   /* sample-frame lab spot axes (columns of R^T) */
    double beam_dir[3];      /* nominal unit beam axis, sample frame */
    double basis[9];         /* beam-frame basis, row-major; columns 0, 1 carry (x', y') */
    double E0_keV;
    double *pos, *dirs, *E_keV, *clock; /* out, local index */
    /* energy spread (beam_entry.py:251-297): E = E0 (1 + f z), z draw 0 of
     * electron e's counter normals under energy_root (the spawn(6)[5] child)
     * clamped to |z| <= 6; f = 0 leaves every E at E0 bit for bit */
    double energy_spread_frac;
    uint64_t energy_root;
    /* axis of the collimated lab beam in the sample frame, along which lab
     * spot offsets are carried to the z = 0 face (projected): the tilt's b
     * (geometry.py:330-394), while beam_dir above sets the directions */
    double axis[3];
} hr_beam_args;
hr_status hr_beam_entry(const hr_beam_args *args, void *stream);

/* ---- learned variance reduction: round-based transport and weight window.
 * No counterpart in the reference (its variance reduction is likelihood-ratio
 * cone splitting, src/pyrite/_cone_splitting.py:57 and
 * transport/cone_split.py:223, plus line roulette, src/pyrite/_line_roulette.py:29);
 * these entry points let an importance network split and roulette particles
 * between transport rounds with weight-preserving decisions. */

/* Advances particles [e_start, e_start+e_count) of a bank by at most `rows`
 * analog steps (1 <= rows <= cap), or min(rows_each[e], rows) when rows_each
 * is non-NULL, each from a fresh state: optical depth, flight and substep
 * restart and draws come from stream_key[e] (counter 0).  Rows go to slot
 * (e-e_start)*cap + s as in hr_transport_run; exit_code is 6 (suspended) for
 * particles still alive after their rows.  Bank state is updated in place.
 * alive[] is not read. */
hr_status hr_vr_round(const hr_transport_args *args, int32_t rows, const int32_t *rows_each,
                      void *stream);

/* x[i] = (E/E0, u_x, u_y, u_z, z/thickness, tilt_deg/3) as FP32 [n,6].  With
 * seg_len non-NULL, pos holds row midpoints and z is taken at the row start. */
hr_status hr_vr_features(int64_t n, const double *pos, const double *dirs, const double *E,
                         const double *seg_len, double E0_keV, double thickness_ang,
                         double tilt_deg, float *x, void *stream);

typedef struct hr_vr_window_args {
    int64_t n;
    const int8_t *exit_code;      /* [n] only 6 (suspended) particles continue */
    const float *log_importance;  /* [n] network output y */
    const double *weight;         /* [n] */
    const uint64_t *key;          /* [n] decision draw: counter 2^40 of the key */
    double log_I_ref;             /* target weight w_t = clamp(exp(log_I_ref - y), w_min, w_max) */
    double w_min, w_max;
    double ratio_lo, ratio_hi;    /* window: split when w/w_t > ratio_hi, roulette when < ratio_lo */
    int32_t n_max;                /* copies per particle, 1..1024 */
    int32_t *count;               /* [n] out: copies (0 = ended or rouletted) */
    double *w_child;              /* [n] out: weight of each copy */
    const double *weight2;        /* [n] second tally's weight, or NULL */
    double *w2_child;             /* [n] out: second weight of copy 0 (other copies carry 0), or NULL */
    int32_t rows_near, rows_far;  /* next round length: rows_far when w I/I_ref < far_ratio */
    double far_ratio;
    int32_t *rows_next;           /* [n] out, or NULL */
} hr_vr_window_args;
/* r = w/w_t; r > ratio_hi: n = floor(r) + [u < frac(r)], capped at n_max,
 * copies of weight w/n; r < ratio_lo: one copy of weight w_t with probability
 * r; otherwise the particle continues unchanged (one copy of weight w).  The
 * band [ratio_lo, ratio_hi] (0 < ratio_lo <= 1 <= ratio_hi) keeps noisy
 * importance estimates from splitting and rouletting the same particle back
 * and forth.  The second
 * weight (a tally that is not split, e.g. brem) stays with copy 0 (w2) and is
 * divided by the survival probability r on roulette. */
hr_status hr_vr_window(const hr_vr_window_args *args, void *stream);

typedef struct hr_vr_spawn_args {
    int64_t n;
    const int32_t *count;    /* [n] */
    const int64_t *offset;   /* [n] exclusive prefix sum of count */
    const double *w_child;   /* [n] */
    const double *pos, *dirs, *E, *clock;
    const uint64_t *key;
    const int64_t *prim;     /* [n] primary (incident electron) id */
    double *out_pos, *out_dirs, *out_E, *out_clock, *out_weight;
    uint64_t *out_key;       /* copy j of key k: splitmix64(splitmix64(k ^ S) + PHI*(j+1)) */
    int64_t *out_prim;
    const double *w2_child;  /* [n] or NULL */
    double *out_weight2;     /* copy 0: w2_child, other copies 0 (NULL with w2_child) */
    const int32_t *rows_next; /* [n] or NULL: every copy inherits its parent's next round length */
    int32_t *out_rows;        /* or NULL */
} hr_vr_spawn_args;
hr_status hr_vr_spawn(const hr_vr_spawn_args *args, void *stream);

typedef struct hr_vr_gather_args {
    int64_t n, rows;               /* particles; slots per particle (the round's cap) */
    const int64_t *seg_count;      /* [n] */
    const int64_t *row_offset;     /* [n] exclusive prefix sum of min(seg_count, rows) */
    const double *weight;          /* [n] */
    const int64_t *prim;           /* [n] or NULL */
    const double *seg_mid, *seg_dir, *seg_len, *seg_E, *seg_E_end; /* slot-major rows */
    float *r_mid, *v_hat, *L, *E;  /* out: FP32 scoring records, E = (E + E_end)/2 */
    float *row_weight;             /* out */
    int64_t *row_prim;             /* out or NULL */
    const double *weight2;         /* [n] or NULL */
    float *row_weight2;            /* out, or NULL */
    int64_t n_out;                 /* rows to write (sum of the clamped counts): one thread per row;
                                      0 = one thread per slot */
} hr_vr_gather_args;
hr_status hr_vr_gather(const hr_vr_gather_args *args, void *stream);

/* mass[j] = sum over the segment's n_dir*n_g lines with lo <= E_r < hi of
 * w*pi/(a*dE): the offset-averaged content of a narrow sinc^2 line on a
 * uniform grid of spacing dE.  Lines at (d*n_seg + j)*n_g + g. */
hr_status hr_line_mass(const double *E_r, const double *a, const float *w, int64_t n_seg,
                       int32_t n_dir, int32_t n_g, double lo_eV, double hi_eV, double dE_eV,
                       double *mass, void *stream);

/* Line roulette by exact mass: segment j's lines are kept surely when
 * mass[j] >= m_thr, else with probability p = mass[j]/m_thr and weight 1/p
 * (u = counter draw row0 + j of `key`); dropped lines get w = 0.  The
 * per-primary two-stratum LineRoulette of the reference
 * (src/pyrite/_line_roulette.py:29-80) applied per segment with a
 * proportional keep probability. */
hr_status hr_line_roulette(float *w, int64_t n_seg, int32_t n_dir, int32_t n_g, const double *mass,
                           double m_thr, uint64_t key, int64_t row0, void *stream);

/* ------------------------------------------------------------------------
 * Depth-dose tally of transport rows: each row's loss E - E_end spread over
 * the depth bins its flight crosses in proportion to the depth covered there,
 * plus the residual E_end of a history that ends at the cutoff (event 7,
 * montecarlo/transport/events.py:43) at the row's end; deposits outside the
 * grid go to the edge bins.  With the energy of escaping electrons this closes
 * the energy balance: sum(out)/scale + sum(escaped E) = Ne E0.  The depth-dose
 * profile is one of the plain method's parity observables; the reference's
 * transport modules have no single function for it (its runners bin rows).
 * Rows are slot-major: particle i's rows are slots i*rows + s, s < seg_count[i]
 * (seg_count NULL: every slot is a row, e.g. rows = 1 for compacted rows).
 * ---------------------------------------------------------------------- */
typedef struct hr_dose_args {
    int64_t n, rows;
    const int64_t *seg_count; /* [n] or NULL */
    const double *seg_mid, *seg_dir, *seg_len, *seg_E, *seg_E_end;
    const int8_t *seg_event;  /* or NULL: no cutoff residual */
    const double *weight;     /* [n] particle weights, or NULL (= 1) */
    double scale;             /* multiplies every deposit (1/Ne for per-electron doses) */
    double z0, dz;            /* top edge of bin 0 and the bin width [A] */
    int32_t n_bins;           /* 1 .. 8192 */
    double *out;              /* [n_bins] keV, accumulated */
} hr_dose_args;
hr_status hr_depth_dose(const hr_dose_args *args, void *stream);

/* Per-primary line roulette, the reference's LineRoulette estimator
 * (src/pyrite/_line_roulette.py:29-80, montecarlo/runner/line_roulette.py).
 * hr_roulette_uniforms: out[i] = uniform of primary first_id + i keyed only on
 * (seed, primary id) (roulette_uniforms, line_roulette.py:33-45).
 * hr_line_screen: for primaries first_id + i with exact line masses mass[i]:
 * probability 1 and kept = mass > 0, except ids >= pilot with
 * 0 < mass < threshold, which get probability `floor` and kept = u < floor
 * (LineScreen._select, line_roulette.py:114-126). */
hr_status hr_roulette_uniforms(uint64_t seed, int64_t first_id, int64_t n, double *out, void *stream);
hr_status hr_line_screen(const double *mass, int64_t n, int64_t first_id, int64_t pilot, double threshold,
                         double floor, uint64_t seed, double *probability, uint8_t *kept, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* HIPRITE_H */
