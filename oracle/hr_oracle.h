/*
 * hr_oracle.h -- FP64 CPU restatement of PyRITE's transport and scoring hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * never as the thing measured or shipped.  The product path (pyrite_amd) must
 * not link, import or execute anything under oracle/.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the PyRITE source tree, src/pyrite/...).  Arithmetic follows the reference's
 * expression order and is compiled with -ffp-contract=off so no multiply-add is
 * fused that the reference (NumPy / Numba without fastmath) does not fuse.
 *
 * Pinning status per function is listed in oracle/README.md.
 */
#ifndef HR_ORACLE_H
#define HR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-addressed SplitMix64 streams (transport/kinematics.py:198-252) */
uint64_t or_splitmix64(uint64_t x);
uint64_t or_stream_key(uint64_t seed, uint64_t electron);
double or_stream_uniform(uint64_t key, uint64_t counter);
void or_stream_keys(uint64_t seed, int64_t start, int64_t n, uint64_t *out);
uint64_t or_urban_stream_key(uint64_t stream_key);
uint64_t or_urban_flight_key(uint64_t urban_key, int64_t flight, int64_t substep);

/* ---- stopping power (transport/stopping.py:384-411, straggling.py:305-329) */
double or_dEds_spliced_element(double J, double k, double coeff, double E_cross, double E);
double or_dEds_spliced_packed(const double *J, const double *k, const double *coeff,
                              const double *E_cross, int n_el, double E);

/* ---- Urban straggling (transport/straggling.py:131-363) */
/* Returns the positive sampled loss [keV]; *counter is advanced. */
double or_urban_sample_compound(const double *Z, const double *J, const double *k,
                                const double *coeff, const double *E_cross, int n_el, double E,
                                double s, double stopping_scale, uint64_t key, uint64_t *counter);
/* Analytic (mean, variance) of one element's Urban loss (straggling.py:204-226). */
void or_urban_moments_element(double Z, double J, double C, double E, double s, double *mean,
                              double *var);
int64_t or_urban_poisson(double lam, uint64_t key, uint64_t *counter);

/* ---- LUT interpolation (transport/lut.py:94-124) */
void or_lut_index_frac(double E, double log_E_min, double inv_dlogE, int n_energy, int *i,
                       double *f);

/* ---- direction rotation (transport/core_geometry.py:16-36), in place */
void or_rotate_direction(double *d, double cos_t, double phi);

/* ---- per-electron LUT transport core (transport/cores.py:365-1250,
 *      transport/_jit_lut_kernel.py:64-574) */
/* ---- hard-event modes (coupled radiative, shell soft/hard) ---------------- */
typedef struct or_rad_tables {
    const double *incident, *nominal, *top, *chi, *z, *ncm3; /* slot-major, as hr_hard_events */
    const int32_t *nt;
    const double *plasma; /* [n_layers] eV */
    int32_t max_t, n_red;
    double cutoff_eV;
} or_rad_tables;

typedef struct or_hard_modes {
    int32_t rad_on, shell_on;
    or_rad_tables rad;
    const uint64_t *rad_key;     /* [Ne] */
    const double *lut_rad_soft;  /* [n_layers, n_E] keV/A */
    const double *lut_rad_rate;  /* [n_layers, n_E] 1/A */
    const double *lut_sh_rate, *lut_sh_omega2, *lut_sh_ch_rate; /* [L,nE], [L,nE], [L,max_ch,nE] */
    const double *sh_U, *sh_W;
    const int32_t *sh_branch;
    const int16_t *sh_code;
    const int32_t *sh_nch;
    int32_t sh_max_ch;
    double sh_cutoff_eV;
    double *seg_hard_W, *seg_rad_k; /* rows (NULL: not written) */
    int16_t *seg_hard_ch, *seg_rad_Z;
} or_hard_modes;

void or_rad_moments(const or_rad_tables *t, int slot, double E_eV, double cutoff_eV, double plasma_eV, double *soft,
                    double *hard_rate, double *hard_first);
double or_rad_sample_photon(const or_rad_tables *t, int slot, double E_eV, double cutoff_eV, double u,
                            double plasma_eV);
void or_rad_layer_moments(const or_rad_tables *t, int layer, int n_el, int max_el, double E_eV, double *soft,
                          double *hard, double *rates);
void or_rad_layer_lut(const or_rad_tables *t, int n_layers, const int32_t *nel, int max_el, const double *E_keV,
                      int64_t n_E, double *soft_keV_per_ang, double *rate_per_ang);
void or_shell_bounds(double E, double U, double W, int branch, double Wc, double *lower, double *upper,
                     double *peak_end);
double or_moller_j0(double E, double prime, double w);
double or_triangle_cdf(double peak_end, double lower, double w);
double or_shell_sample_transfer(double E, double U, double W, int branch, double Wc, double u);
double or_shell_sample_transfer_bisection(double E, double U, double W, int branch, double Wc, double u);
double or_qmin(double E, double W);
double or_shell_primary_cosine(double E, double U, double W, int branch, double W_eV, double u);
double or_shell_secondary_cosine(double E, double U, double W, int branch, double W_eV, double u);
double or_shell_soft_loss(double mean, double var, uint64_t key, uint64_t *counter);

/* Tabulated (ELSEPA) elastic angles (or_elsepa.c; scattering.py:540-620,
 * _jit_device.py:493-620).  on = 0: the screened-Rutherford angle. */
typedef struct or_elsepa {
    int32_t on, n_mu, guide_cells;
    const int64_t *start, *length; /* [n_layers, max_el] */
    const double *logE;            /* [rows] ln E[eV] */
    const double *mu;              /* [n_mu] */
    const double *cdf, *pdf;       /* [rows, n_mu] */
    const int32_t *guide;          /* [rows, guide_cells + 1] */
} or_elsepa;

/* Lower node (relative to row) and fraction of x on g[row .. row+n), clamped. */
void or_elsepa_bracket(const double *g, int64_t row, int64_t n, double x, int64_t *lo, double *f);
/* Inversion of one angular row at xi: guided panel search / whole-row bisection. */
double or_elsepa_invert_row(const or_elsepa *t, int64_t row, double xi);
double or_elsepa_invert_row_full(const or_elsepa *t, int64_t row, double xi);
/* mu of element slot `el` (L*max_el + i) at E_keV and uniform xi. */
double or_elsepa_sample_mu(const or_elsepa *t, int64_t el, double E_keV, double xi);
/* Bracketing first row and fraction of slot `el` at E_keV; the ln-E quantile
 * mixture of rows row, row+1 at xi; the exact CDF of one row at mu = m. */
void or_elsepa_energy_rows(const or_elsepa *t, int64_t el, double E_keV, int64_t *row, double *f);
double or_elsepa_mix_mu(const or_elsepa *t, int64_t row, double f, double xi);
double or_elsepa_cdf_at(const or_elsepa *t, int64_t row, double m);
void or_elsepa_sample_mu_batch(const or_elsepa *t, int64_t el, int64_t n, const double *E_keV, const double *xi,
                               double *out);
void or_elsepa_invert_batch(const or_elsepa *t, int64_t n, const int64_t *row, const double *xi, int guided,
                            double *out);

/* Detector-cone splitting of elastic events (or_cone.c; cone_split.py:223-335,
 * cores.py:1189-1226) on the screened-Rutherford law.  on = 0: analog.  Copy
 * k of local electron i in slot i*cap + k while k < cap; count[i] counts all. */
typedef struct or_cone_split {
    int32_t on;
    double nx, ny, nz, cos_c, sin_c, target_weight, min_energy_keV, angle_scale;
    const uint64_t *split_key; /* [Ne] */
    int64_t cap;
    int64_t *count;  /* [e_count] */
    uint8_t *killed; /* [e_count] */
    double *pos, *dir, *E_keV, *clock, *weight;
    int64_t *flight;
} or_cone_split;

uint64_t or_cone_split_key(uint64_t stream_key);
uint64_t or_cone_event_key(uint64_t split_key, int64_t flight);
uint64_t or_cone_copy_key(uint64_t split_key, int64_t flight);
/* out = (ox, oy, oz, weight); returns 1 when a copy is made */
int or_cone_event(double dx, double dy, double dz, const or_cone_split *c, double alpha, double E_keV,
                  uint64_t key, double *out);
/* The same on the tabulated law of element slot `el` (cone_split.py:223-335). */
int or_cone_event_tab(double dx, double dy, double dz, const or_cone_split *c, const or_elsepa *t, int64_t el,
                      double E_keV, uint64_t key, double *out);
void or_cone_event_tab_batch(const or_cone_split *c, const or_elsepa *t, int64_t n, const double *d,
                             const int64_t *el, const double *E_keV, const uint64_t *key, int32_t *made,
                             double *out);
double or_cone_ladder_weight(double w);
int64_t or_cone_collect(const or_cone_split *c, int64_t e_start, int64_t e_count, int ladder, int64_t out_cap,
                        int64_t *parent, int64_t *flight, double *pos, double *dir, double *E_keV, double *clock,
                        double *weight, uint64_t *stream_key, int64_t *max_count);

typedef struct or_transport_args {
    /* run */
    int64_t e_start, e_count, cap;
    const uint64_t *stream_key; /* [Ne] */
    /* control */
    int32_t max_steps, energy_model; /* 1 = midpoint, 0 = frozen */
    double max_dE_frac;
    int32_t straggle_on;
    /* geometry */
    int32_t n_layers;
    const double *internal_bounds; /* [n_layers-1] */
    double z_total;
    int32_t finite_footprint;
    double width_ang, height_ang;
    const int32_t *L_nel;   /* [n_layers] */
    const double *L_top;    /* [n_layers] */
    const double *L_bot;    /* [n_layers] */
    /* materials (padded per layer, stride max_el) */
    int32_t max_el;
    const double *L_Z, *L_J, *L_k, *L_coeff, *L_E_cross;
    /* lut */
    double lut_log_E_min, lut_inv_dlogE;
    int32_t lut_n_energy;
    const double *lut_total_rate; /* [n_layers, n_E] */
    const double *lut_dEds;       /* [n_layers, n_E] */
    const double *lut_inv_beta;   /* [n_E] */
    const double *lut_cdf;        /* [n_layers, max_el, n_E] */
    const double *lut_alpha;      /* [n_layers, max_el, n_E] */
    /* state (global electron index) */
    const uint8_t *alive;
    double *clock, *pos, *dirs, *E_keV;
    const double *E_cut;
    double *stragg_dE;
    /* rows, slot = i*cap + s (batching.py:401-425 midpoint schema) */
    double *seg_dir, *seg_mid, *seg_len, *seg_E, *seg_t0;
    int64_t *seg_id;
    int16_t *seg_lay;
    double *seg_E_end, *seg_t_end;
    int64_t *seg_flight, *seg_substep;
    int8_t *seg_event;
    /* per local electron */
    int64_t *seg_count;
    int8_t *exit_code;
    /* hard-event modes (zeroed: off) */
    or_hard_modes hard;
    /* detector-cone splitting (zeroed: off) */
    or_cone_split cone;
    /* tabulated elastic angles (zeroed: screened Rutherford) */
    or_elsepa elsepa;
} or_transport_args;

/* Runs electrons [e_start, e_start+e_count) with `threads` OpenMP threads
 * (<=0: all).  Each electron is independent, so the result does not depend on
 * the thread count. */
void or_transport_lut(const or_transport_args *a, int threads);

/* ---- scoring: per (segment, reflection) line prologue for one detector
 *      direction (spectrum/cxr/_batched.py:461-728), dense sinc^2 reduction
 *      (cxr/_cpu_reduce.py:26-45) and BH+Elwert brem (brem.py:829-1239).
 *      Segment records and tables are the FP32 inputs the device reads;
 *      the oracle widens them and computes in FP64. */
typedef struct or_line_args {
    int64_t n_seg;
    const float *r_mid; /* [n,3] A */
    const float *v_hat; /* [n,3] */
    const float *L;     /* [n] A (parent length; one escape piece) */
    const float *E;     /* [n] representative energy, keV */
    double thickness;   /* slab [A] */
    double n_hat[3];    /* detector direction, sample frame */
    int32_t n_g;
    const double *g;    /* [n_g,3] 1/A */
    const double *wm;   /* [n_g] orientation weight */
    const int32_t *row; /* [n_g] coupling-table row */
    const double *e_s;  /* [n_g,3] polarization basis */
    const double *e_p;  /* [n_g,3] */
    int64_t n_tab;
    const float *E_tab;     /* [n_tab] eV ascending */
    const float *delta_tab; /* [n_tab] 1 - Re n */
    const float *chi_re, *chi_im, *u_re, *u_im; /* [n_rows, n_tab] */
    int32_t n_mu;
    const float *log_mu; /* [n_mu, n_tab] log(1/A) */
    double lo_keep, hi_keep;
} or_line_args;

/* hbar*c of the line-recoil terms (default hbar*c; 0 = the classical limit, _recoil.py:14-16). */
void or_set_recoil_hbarc(double hbarc);
/* Exact second-order (pxr_num, f_cbs) of one polarization (cxr/_recoil.py:35-76). */
void or_quantum_line_factors(double h, double gamma, double vdg, double k_dot_v, double k_dot_g,
                             double g2, double k2, double v_dot_e, double g_dot_e, double k_dot_e,
                             double *out2);
/* Line (E_r [eV], a [1/eV], w) per (segment j, reflection g) at j*n_g+g;
 * rejected pairs get w = 0 and good = 0. */
void or_cxr_lines(const or_line_args *p, double *E_r, double *a, double *w, uint8_t *good,
                  int threads);
/* Coherent line records per (segment j, reflection g) at 13*(j*n_g+g): E_c, a,
 * phase slope, g phase, cs_re, cs_im, cp_re, cp_im, L_esc, half_dL, apb, bma, q
 * (definition in or_score.c); FP64 segment inputs, tables from p. */
void or_coherent_lines(const or_line_args *p, const double *r_mid, const double *v_hat, const double *L,
                       const double *E, const double *t_start, double *out, uint8_t *good, int threads);
/* exp(-tau_c/2) sinh(w)/w, w = i v - q (formation factor of a line piece) */
void or_formation_factor(double v, double apb, double bma, double q, double *re, double *im);
/* F[4][n_E] += phased field of one row's coherent records (13 per line) */
void or_coherent_field(const double *rec, int64_t n_lines, const double *E, const double *dom, int64_t n_E,
                       double *F, double *bound, int threads);
/* out[k] += sum_j w_j [sin x/x]^2, x = a_j (E_k - E_r,j) (x = 0 -> 1e-20) */
void or_sincsq_reduce(const double *E_r, const double *a, const double *w, int64_t n,
                      const double *E_grid, int64_t n_E, double *out, int threads);
/* ---- importance MLP (no reference counterpart; definition in or_mlp.c) */
double or_bf16(double v);
void or_mlp_set_rounding(int on); /* 1: device bfloat16 roundings (default), 0: exact */
void or_mlp_forward(const float *params, const float *x, int64_t n, double *y);
double or_mlp_loss_grad(const float *params, const float *x, const float *t, int64_t n,
                        double *grad);
/* ---- support-windowed sinc^2 reduction with the far tail (or_window.c,
 *      spectrum/cxr/_node_window.py:86-250,454-522); uniform grid E0 + k dE */
int or_window_plan(double E_r, double a, double E0, double dE, int64_t n, double x_c, int64_t *node,
                   int64_t *lo, int64_t *hi, double *mu);
void or_window_reduce(const double *E_r, const double *a, const float *w, int64_t n_lines, double E0,
                      double dE, int64_t n, double x_c, double *out);
/* ---- learned variance reduction (no reference counterpart; or_vr.c) */
uint64_t or_vr_child_key(uint64_t key, int64_t j);
void or_vr_window(int64_t n, const int8_t *exit_code, const float *log_importance, const double *weight,
                  const uint64_t *key, double log_I_ref, double w_min, double w_max, double ratio_lo,
                  double ratio_hi, int32_t n_max, int32_t *count, double *w_child, const double *weight2,
                  double *w2_child);
void or_line_roulette(float *w, int64_t n_seg, int32_t n_dir, int32_t n_g, const double *mass, double m_thr,
                      uint64_t key, int64_t row0);
void or_line_mass(const double *E_r, const double *a, const float *w, int64_t n_seg, int32_t n_dir,
                  int32_t n_g, double lo, double hi, double dE, double *mass);
/* out[k] += sum_seg sum_el n L dsigma/dk (BH+Elwert) * <exp(-mu_k path)>;
 * caller divides by 4 pi Ne. */
void or_brem_reduce(const float *r_mid, const float *v_hat, const float *L_ang,
                    const float *E_keV, int64_t n_seg, double thickness, const double *n_hat,
                    const double *Z, const double *n_cm3, int n_el, const double *E_grid_eV,
                    const float *mu_inv_ang, int64_t n_E, double *out, int threads);

#ifdef __cplusplus
}
#endif
#endif
