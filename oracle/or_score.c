/*
 * or_score.c -- FP64 CPU restatement of PyRITE's per-segment line (PXR/CBS)
 * prologue, dense sinc^2 reduction and Bethe-Heitler+Elwert bremsstrahlung.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Follows:
 *   montecarlo/spectrum/cxr/_setup.py:794-1175  (segment kinematics, escape gradient)
 *   montecarlo/spectrum/cxr/_kernels.py:624-766 (_cxr_kin_core, _vacuum_doppler,
 *                                                _in_medium_kinematics)
 *   montecarlo/spectrum/cxr/_fused.py:22-389    (recoil pair, line bracket,
 *                                                slope/keep, transmission, weights)
 *   montecarlo/spectrum/cxr/_recoil.py:35-76    (_quantum_line_factors)
 *   montecarlo/spectrum/cxr/_kernels.py:201-255, 562-620 (amplitude, table gather)
 *   montecarlo/spectrum/cxr/_batched.py:702-728 (escape weights)
 *   montecarlo/spectrum/segment_escape.py:30-300 (slab escape paths, mean transmission)
 *   montecarlo/spectrum/cxr/_cpu_reduce.py:26-45, cxr_jit_kernel.py:32-100 (sinc^2 sum)
 *   montecarlo/spectrum/brem.py:380-423, brem_jit_kernel.py:76-153 (BH+Elwert cell)
 *   montecarlo/spectrum/cxr/_batched.py:561-690, cxr/_formation.py:66-115,
 *   cxr/_recoil.py:118-147, cxr/_batched_finalize.py:201-235 (coherent records,
 *                                                formation factor, phased field)
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#include "hr_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* materials/crystal.py:47-53 */
#define HBARC 1973.269804
#define ALPHA_FS (1.0 / 137.035999)
#define M_E_EV 510998.95
/* brem.py:63-64 */
#define R_E_CM2 7.9407877e-26
#define MC2_KEV 510.99895
/* _kernels.py:675 */
#define ROOT_RTOL 1e-3

static double beta_from_keV(double E) { /* kinematics.py:131-138 */
    double g = 1.0 + E / 510.99895;
    double gis = 1.0 / (g * g);
    return sqrt(1.0 - gis);
}

/* numpy searchsorted(side="left") then clip(., 1, n-1) (_kernels.py:328-349) */
static int64_t bracket(const float *grid, int64_t n, double x) {
    int64_t y;
    if (isnan(x)) {
        y = n;
    } else if (x > grid[n - 1]) {
        y = n;
    } else {
        int64_t lo = 0, hi = n - 1;
        while (lo < hi) {
            int64_t m = lo + (hi - lo) / 2;
            if (grid[m] < x) lo = m + 1;
            else hi = m;
        }
        y = hi;
    }
    return y < 1 ? 1 : (y > n - 1 ? n - 1 : y);
}

/* _interp_gather1d with a fresh bracket (_kernels.py:495-507) */
static double gather1d(const float *grid, const float *f, int64_t n, double x) {
    int64_t i = bracket(grid, n, x);
    double x0 = grid[i - 1];
    double frac = (x - x0) / ((double)grid[i] - x0);
    double f0 = f[i - 1];
    double y = f0 + frac * ((double)f[i] - f0);
    if (x <= grid[0]) y = f[0];
    if (x >= grid[n - 1]) y = f[n - 1];
    return y;
}

/* segment_escape.py:30-75 slab branch (no layers, no footprint) */
static double slab_path(double z, double thickness, double n_z) {
    double inv_nz = 1.0 / fmax(fabs(n_z), 1.0e-12);
    double depth = z < 0.0 ? 0.0 : (z > thickness ? thickness : z);
    return n_z < 0.0 ? depth * inv_nz : (thickness - depth) * inv_nz;
}

/* segment_escape.py:286-300 mean_transmission */
static double mean_transmission(double tau0, double tau1) {
    double lo = fmin(tau0, tau1);
    double d = fabs(tau1 - tau0);
    double ratio = d > 1.0e-4 ? -expm1(-d) / d : 1.0 - 0.5 * d + d * d / 6.0;
    return exp(-lo) * ratio;
}

/* hbar*c in the line-recoil terms only (_recoil.py:14-16): zero is the
 * reference's hbar -> 0 hook, under which every line reduces to its classical
 * root, width and amplitudes.  Set by or_set_recoil_hbarc (tests only). */
static double g_recoil_hbarc = HBARC;
void or_set_recoil_hbarc(double hbarc) { g_recoil_hbarc = hbarc; }

/* _recoil.py:35-76 */
static void quantum_line_factors(double h, double gamma, double vdg, double k_dot_v, double k_dot_g,
                                 double g2, double k2, double v_dot_e, double g_dot_e,
                                 double k_dot_e, double *pxr_num, double *f_cbs) {
    double gg = gamma * gamma;
    double n_1 = h * g2 - 2.0 * gamma * vdg;
    double n_2 = n_1 + 2.0 * h * k_dot_g;
    double n_f = h * (g2 + 2.0 * k_dot_g + k2) - 2.0 * gamma * (vdg + k_dot_v);
    double e_1 = sqrt(gg + h * n_1);
    double e_2 = sqrt(gg + h * (h * k2 - 2.0 * gamma * k_dot_v));
    double e_f = sqrt(gg + h * n_f);
    double s_if = gamma + e_f;
    double ge_ke = g_dot_e + k_dot_e;
    *pxr_num = (-n_f * ge_ke - k2 * (2.0 * gamma * v_dot_e - h * ge_ke)) / s_if;
    double s_1 = e_1 + gamma;
    double s_3 = e_1 + e_f;
    double s_4 = gamma + e_2;
    double x = 2.0 * gamma * v_dot_e - h * k_dot_e;
    double sigma = s_if + e_1 + e_2;
    *f_cbs = 2.0 * g_dot_e * s_1 / (s_3 * n_1) +
             x * (n_1 * (n_f / s_if) * sigma - 2.0 * s_1 * s_4 * k_dot_g) / (s_3 * s_4 * n_1 * n_2);
}

void or_quantum_line_factors(double h, double gamma, double vdg, double k_dot_v, double k_dot_g,
                             double g2, double k2, double v_dot_e, double g_dot_e, double k_dot_e,
                             double *out2) {
    quantum_line_factors(h, gamma, vdg, k_dot_v, k_dot_g, g2, k2, v_dot_e, g_dot_e, k_dot_e, &out2[0],
                         &out2[1]);
}

/* Kinematics of one segment shared by its reflections (_setup.py:981-985,
 * _kernels.py:678-699, segment_escape.py:78-117,190-211). */
typedef struct seg_kin {
    double rz, vx, vy, vz, beta, gamma, denom_row, t_L, gz, v_dot_grad, grad2, path0, path1, inv_E;
} seg_kin;

static void segment_kinematics(const or_line_args *p, double rz, double vhx, double vhy, double vhz, double L,
                               double E, seg_kin *k) {
    const double nx = p->n_hat[0], ny = p->n_hat[1], nz = p->n_hat[2];
    k->rz = rz;
    k->beta = beta_from_keV(E);
    k->vx = k->beta * vhx;
    k->vy = k->beta * vhy;
    k->vz = k->beta * vhz;
    k->gamma = 1.0 + E / (M_E_EV / 1e3);
    const double one_minus_beta = 1.0 / (k->gamma * k->gamma * (1.0 + k->beta));
    const double dx = vhx - nx, dy = vhy - ny, dz = vhz - nz;
    const double one_minus_cos = 0.5 * (dx * dx + dy * dy + dz * dz);
    k->denom_row = one_minus_beta + k->beta * one_minus_cos;
    k->t_L = L / k->beta;
    /* escape gradient: slab, -e/(e.n) with e = +z; zero outside the slab */
    k->gz = -1.0 / nz;
    if (rz < 0.0 || rz > p->thickness) k->gz = 0.0;
    k->v_dot_grad = k->vz * k->gz;
    k->grad2 = k->gz * k->gz;
    /* escape paths at the segment ends */
    const double half = 0.5 * L;
    k->path0 = slab_path(rz - half * vhz, p->thickness, nz);
    k->path1 = slab_path(rz + half * vhz, p->thickness, nz);
    k->inv_E = g_recoil_hbarc / (k->gamma * M_E_EV); /* _recoil.py:17-23 */
}

/* One (segment, reflection) line at the in-medium root: the resonance, the
 * width slope, mu at the resonance and both polarizations' complex amplitudes
 * (before squaring).  Returns keep. */
typedef struct line_core {
    double E_res, omega, slope, mu, A_re[2], A_im[2], vdg, r_num, r_den, g2, n_dot_g;
} line_core;

static int reflection_line(const or_line_args *p, const seg_kin *k, int g, line_core *o) {
    const double nx = p->n_hat[0], ny = p->n_hat[1], nz = p->n_hat[2];
    const double h = g_recoil_hbarc / M_E_EV; /* _recoil.py:26-32 */
    const int64_t n_tab = p->n_tab;
    const double Gx = p->g[3 * g], Gy = p->g[3 * g + 1], Gz = p->g[3 * g + 2];
    const double g2 = Gx * Gx + Gy * Gy + Gz * Gz;
    const double n_dot_g = Gx * nx + Gy * ny + Gz * nz;
    /* _fused.py:335-345 recoil pair */
    const double r_num = 0.5 * g2 * k->inv_E;
    const double r_den = n_dot_g * k->inv_E;
    const double denom_vac = k->denom_row + r_den;
    const double v_dot_g_raw = k->vx * Gx + k->vy * Gy + k->vz * Gz;
    const double v_dot_g = v_dot_g_raw - r_num;
    /* _kernels.py:702-766 in-medium root, three passes */
    double denom = denom_vac, previous = denom, delta = 0.0;
    for (int pass = 0; pass < 3; ++pass) {
        const double Eres = HBARC * (v_dot_g / denom);
        delta = gather1d(p->E_tab, p->delta_tab, n_tab, Eres);
        previous = denom;
        denom = denom_vac - delta * k->v_dot_grad;
    }
    int settled = fabs(denom - previous) <= ROOT_RTOL * fabs(denom);
    settled = settled && (2.0 * fabs(delta) * k->grad2 < 1.0);
    if (!settled) denom = NAN;
    /* _fused.py:365-372 block n.g */
    const double n_dot_g_blk = n_dot_g + delta * (k->gz * Gz);
    /* _kernels.py:624-652 kinematics */
    const double omega = (v_dot_g_raw - r_num) / denom;
    const double k_dot_v = omega * ((1.0 - denom) + r_den);
    const double k_dot_g = omega * n_dot_g_blk;
    const double detuning = g2 + 2.0 * k_dot_g;
    /* _fused.py:150-171 line bracket */
    const double radicand = k->grad2 == 0.0 ? 1.0 : 1.0 - 2.0 * delta + delta * delta * k->grad2;
    const double k_mag = omega * sqrt(radicand);
    const double E_res = HBARC * omega;
    int keep = (E_res > p->lo_keep) && (E_res > 10.0) && (E_res < p->hi_keep);
    /* _fused.py:380-386 slope / positive mass shell */
    const double oe = omega * k->inv_E;
    const double slope = denom * (1.0 / (1.0 - oe));
    keep = keep && (denom > 0.0) && (oe <= 1.0 - 1.0 / k->gamma);
    o->E_res = E_res;
    o->omega = omega;
    o->slope = slope;
    o->vdg = v_dot_g_raw;
    o->r_num = r_num;
    o->r_den = r_den;
    o->g2 = g2;
    o->n_dot_g = n_dot_g;
    o->mu = 0.0;
    o->A_re[0] = o->A_re[1] = o->A_im[0] = o->A_im[1] = 0.0;
    if (!keep) return 0;
    /* _kernels.py:562-620 gather at the shared bracket */
    const int64_t i = bracket(p->E_tab, n_tab, E_res);
    const double x0 = p->E_tab[i - 1], x1 = p->E_tab[i];
    const double frac = (E_res - x0) / (x1 - x0);
    const int below = E_res <= p->E_tab[0];
    const int above = E_res >= p->E_tab[n_tab - 1];
    const double bounded = fmin(fmax(E_res, x0), x1);
    const double log_frac = log1p((bounded - x0) / x0) / log1p((x1 - x0) / x0);
    const int64_t base = (int64_t)p->row[g] * n_tab;
    double tab[4];
    const float *tables[4] = {p->chi_re, p->chi_im, p->u_re, p->u_im};
    for (int t = 0; t < 4; ++t) {
        const float *T = tables[t] + base;
        const double f0 = T[i - 1];
        double y = f0 + frac * ((double)T[i] - f0);
        if (below) y = T[0];
        if (above) y = T[n_tab - 1];
        tab[t] = y;
    }
    for (int el = 0; el < p->n_mu; ++el) {
        const float *T = p->log_mu + (int64_t)el * n_tab;
        const double l0 = T[i - 1];
        double lm = l0 + log_frac * ((double)T[i] - l0);
        if (below) lm = T[0];
        if (above) lm = T[n_tab - 1];
        o->mu += exp(lm);
    }
    /* _batched.py:789-826, both polarizations */
    for (int pol = 0; pol < 2; ++pol) {
        const double *e = pol == 0 ? p->e_s + 3 * g : p->e_p + 3 * g;
        const double v_dot_e = k->vx * e[0] + k->vy * e[1] + k->vz * e[2];
        const double k_dot_e = omega * delta * (k->gz * e[2]);
        const double g_dot_e = Gx * e[0] + Gy * e[1] + Gz * e[2];
        double pxr_num, f_cbs;
        quantum_line_factors(h, k->gamma, v_dot_g_raw, k_dot_v, k_dot_g, g2, k_mag * k_mag, v_dot_e, g_dot_e,
                             k_dot_e, &pxr_num, &f_cbs);
        const double f_pxr = pxr_num / detuning;
        o->A_re[pol] = tab[0] * f_pxr + tab[2] * f_cbs;
        o->A_im[pol] = tab[1] * f_pxr + tab[3] * f_cbs;
    }
    return 1;
}

void or_cxr_lines(const or_line_args *p, double *E_r_out, double *a_out, double *w_out,
                  uint8_t *good_out, int threads) {
    const int n_g = p->n_g;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(threads)
#endif
    for (int64_t j = 0; j < p->n_seg; ++j) {
        /* widened FP32 segment record */
        seg_kin k;
        segment_kinematics(p, p->r_mid[3 * j + 2], p->v_hat[3 * j], p->v_hat[3 * j + 1], p->v_hat[3 * j + 2], p->L[j],
                           p->E[j], &k);
        for (int g = 0; g < n_g; ++g) {
            const int64_t line = j * n_g + g;
            line_core o;
            const int keep = reflection_line(p, &k, g, &o);
            double A2 = 0.0;
            if (keep)
                for (int pol = 0; pol < 2; ++pol) A2 += o.A_re[pol] * o.A_re[pol] + o.A_im[pol] * o.A_im[pol];
            /* _batched.py:702-728 + _fused.py:267-281, one piece, fraction 1 */
            const double T_abs = mean_transmission(k.path0 * o.mu, k.path1 * o.mu);
            const double pref = ALPHA_FS * o.omega / (4.0 * M_PI * M_PI * HBARC) * (k.t_L * k.t_L) * T_abs;
            const double a_width = o.slope * k.t_L / (2.0 * HBARC);
            const double w = pref * A2 * p->wm[g];
            const int good = keep && isfinite(w) && w > 0.0;
            E_r_out[line] = o.E_res;
            a_out[line] = a_width;
            w_out[line] = good ? w : 0.0;
            good_out[line] = (uint8_t)good;
        }
    }
}

/* Coherent line records (cxr/_batched.py:561-690 _batched_coherent_block,
 * cxr/_formation.py:66-83 formation_coefficients, cxr/_recoil.py:118-147
 * _recoil_formation_line): per (segment j, reflection g) at j*n_g+g,
 * 13 FP64 values
 *   E_c, a, phase_slope, g_phase, cs_re, cs_im, cp_re, cp_im, L_esc, half_dL, apb, bma, q
 * with c = sqrt(alpha omega / (4 pi^2 hbar c)) t_L A (no transmission: the
 * formation factor carries the attenuation), E_c and a the exact vacuum sinc
 * centre and width, phase slope (t_mid - n.r + d_shift)/hbar c and
 * g_phase g.r + g_shift over the electron's midpoint age t_mid = t + t_L/2.
 * Segment inputs are FP64 (positions and times enter the phase). */
#define FORMATION_SMALL_Q 0.5
void or_coherent_lines(const or_line_args *p, const double *r_mid, const double *v_hat, const double *L,
                       const double *E, const double *t_start, double *out, uint8_t *good_out, int threads) {
    const int n_g = p->n_g;
    const double nx = p->n_hat[0], ny = p->n_hat[1], nz = p->n_hat[2];
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(threads)
#endif
    for (int64_t j = 0; j < p->n_seg; ++j) {
        seg_kin k;
        segment_kinematics(p, r_mid[3 * j + 2], v_hat[3 * j], v_hat[3 * j + 1], v_hat[3 * j + 2], L[j], E[j], &k);
        const double age = t_start[j] + 0.5 * k.t_L;
        const double d_geom = age - (r_mid[3 * j] * nx + r_mid[3 * j + 1] * ny + r_mid[3 * j + 2] * nz);
        for (int g = 0; g < n_g; ++g) {
            const int64_t line = j * n_g + g;
            double *rec = out + 13 * line;
            line_core o;
            int good = reflection_line(p, &k, g, &o);
            const double amp = sqrt(ALPHA_FS * o.omega / (4.0 * M_PI * M_PI * HBARC));
            /* formation_coefficients(L_start mu, L_end mu) */
            const double tau0 = k.path0 * o.mu, tau1 = k.path1 * o.mu;
            const double q = 0.25 * (tau1 - tau0);
            double apb, bma;
            if (fabs(q) < FORMATION_SMALL_Q) {
                const double c0 = exp(-0.25 * (tau0 + tau1));
                apb = 2.0 * c0 * cosh(q);
                bma = -2.0 * c0 * sinh(q);
            } else {
                const double a = exp(-0.5 * tau0), b = exp(-0.5 * tau1);
                apb = a + b;
                bma = b - a;
            }
            /* _recoil_formation_line */
            const double numerator = o.vdg - o.r_num;
            const double denom = k.denom_row + o.r_den;
            const double E_c = HBARC * numerator / denom;
            const double x = numerator / denom * k.inv_E;
            const double kappa = 1.0 / (1.0 - x);
            const double a_vac = kappa * denom * k.t_L / (2.0 * HBARC);
            const double d_shift = kappa * (x * k.denom_row + o.r_den) * age;
            const double g_shift = kappa * (x * o.vdg - o.r_num) * age;
            const double Gx = p->g[3 * g], Gy = p->g[3 * g + 1], Gz = p->g[3 * g + 2];
            const double cs_re = amp * k.t_L * o.A_re[0], cs_im = amp * k.t_L * o.A_im[0];
            const double cp_re = amp * k.t_L * o.A_re[1], cp_im = amp * k.t_L * o.A_im[1];
            good = good && isfinite(amp) && amp > 0.0 && k.t_L > 0.0 && isfinite(apb) && apb > 0.0 && isfinite(bma) &&
                   isfinite(cs_re) && isfinite(cs_im) && isfinite(cp_re) && isfinite(cp_im);
            rec[0] = E_c;
            rec[1] = a_vac;
            rec[2] = (d_geom + d_shift) / HBARC;
            rec[3] = r_mid[3 * j] * Gx + r_mid[3 * j + 1] * Gy + r_mid[3 * j + 2] * Gz + g_shift;
            rec[4] = good ? cs_re : 0.0;
            rec[5] = good ? cs_im : 0.0;
            rec[6] = good ? cp_re : 0.0;
            rec[7] = good ? cp_im : 0.0;
            rec[8] = 0.5 * (k.path0 + k.path1);
            rec[9] = 0.5 * (k.path1 - k.path0);
            rec[10] = apb;
            rec[11] = bma;
            rec[12] = q;
            good_out[line] = (uint8_t)good;
        }
    }
}

/* cxr/_formation.py:86-115 formation_factor: exp(-tau_c/2) sinh(w)/w with
 * w = i v - q, from apb = a + b, bma = b - a; second-order series below
 * |w|^2 = 1e-6 (FORMATION_SERIES_W2, _formation.py:59). */
#define FORMATION_SERIES_W2 1.0e-6
void or_formation_factor(double v, double apb, double bma, double q, double *re, double *im) {
    const double w2 = q * q + v * v;
    if (w2 < FORMATION_SERIES_W2) {
        const double half = 0.5 * apb;
        *re = half * (1.0 - q * q / 3.0 - v * v / 6.0);
        *im = -half * q * v / 3.0;
        return;
    }
    const double sv = sin(v), cv = cos(v), d = 2.0 * w2;
    *re = (v * apb * sv - q * bma * cv) / d;
    *im = -(v * bma * cv + q * apb * sv) / d;
}

/* Phased field of one reflection row (cxr/_batched_finalize.py:201-235):
 * F_pol(E_k) += sum_j c_pol,j F_form(v_jk) exp(i phi_jk) with
 *   v = a (E - E_c) - half_dL dom(E),  phi = ps E - g_phase - L_esc dom(E),
 * over the 13-value records rec[13*j] of or_coherent_lines (lines whose four
 * coefficients are all zero are rejected and skipped).  F is [4][n_E]:
 * sigma re, sigma im, pi re, pi im.  bound (or NULL) gets
 * sum_j (|c_s,j| + |c_p,j|) |F_form(v_jk)|, the L1 size of the sum. */
void or_coherent_field(const double *rec, int64_t n_lines, const double *E, const double *dom, int64_t n_E,
                       double *F, double *bound, int threads) {
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(threads)
#endif
    for (int64_t k = 0; k < n_E; ++k) {
        double sr = 0.0, si = 0.0, pr = 0.0, pi_ = 0.0, b = 0.0;
        for (int64_t j = 0; j < n_lines; ++j) {
            const double *r = rec + 13 * j;
            if (r[4] == 0.0 && r[5] == 0.0 && r[6] == 0.0 && r[7] == 0.0) continue;
            const double v = r[1] * (E[k] - r[0]) - r[9] * dom[k];
            const double phi = r[2] * E[k] - r[3] - r[8] * dom[k];
            double fr, fi;
            or_formation_factor(v, r[10], r[11], r[12], &fr, &fi);
            const double c = cos(phi), s = sin(phi);
            const double xr = fr * c - fi * s, xi = fr * s + fi * c;
            sr += r[4] * xr - r[5] * xi;
            si += r[4] * xi + r[5] * xr;
            pr += r[6] * xr - r[7] * xi;
            pi_ += r[6] * xi + r[7] * xr;
            if (bound) b += (hypot(r[4], r[5]) + hypot(r[6], r[7])) * hypot(fr, fi);
        }
        F[k] += sr;
        F[n_E + k] += si;
        F[2 * n_E + k] += pr;
        F[3 * n_E + k] += pi_;
        if (bound) bound[k] += b;
    }
}

/* cxr/_cpu_reduce.py:26-45: out[e] += sum_j w_j sinc^2(a_j (E_e - E_r,j)) */
void or_sincsq_reduce(const double *E_r, const double *a, const double *w, int64_t n,
                      const double *E_grid, int64_t n_E, double *out, int threads) {
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(threads)
#endif
    for (int64_t k = 0; k < n_E; ++k) {
        const double Ek = E_grid[k];
        double s = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            if (w[j] == 0.0) continue;
            double x = a[j] * (Ek - E_r[j]);
            if (x == 0.0) x = 1.0e-20;
            const double sx = sin(x) / x;
            s += w[j] * (sx * sx);
        }
        out[k] += s;
    }
}

/* brem.py:396-423 incident prefactor; brem_jit_kernel.py:76-153 cell and escape
 * mean.  out[k] += sum_seg cell(seg, k) * T(seg, k)   (caller divides by 4 pi Ne) */
void or_brem_reduce(const float *r_mid, const float *v_hat, const float *L_ang,
                    const float *E_keV, int64_t n_seg, double thickness, const double *n_hat,
                    const double *Z, const double *n_cm3, int n_el, const double *E_grid_eV,
                    const float *mu_inv_ang, int64_t n_E, double *out, int threads) {
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(threads)
#endif
    for (int64_t k = 0; k < n_E; ++k) {
        const double k_eV = E_grid_eV[k];
        const double k_keV = k_eV * 1.0e-3;
        const double mu = mu_inv_ang[k];
        double s = 0.0;
        for (int64_t j = 0; j < n_seg; ++j) {
            const double T = E_keV[j];
            const double Tf = T - k_keV;
            if (Tf <= 1.0e-6 || k_keV <= 0.0) continue;
            const double p_i = sqrt(T * (T + 2.0 * MC2_KEV)) / MC2_KEV;
            const double beta_i = p_i / (1.0 + T / MC2_KEV);
            const double p_f = sqrt(Tf * (Tf + 2.0 * MC2_KEV)) / MC2_KEV;
            const double beta_f = p_f / (1.0 + Tf / MC2_KEV);
            double dp = p_i - p_f;
            if (dp < 1.0e-30) dp = 1.0e-30;
            const double born_log = log((p_i + p_f) / dp);
            const double half = 0.5 * (double)L_ang[j];
            const double rz = r_mid[3 * j + 2], vz = v_hat[3 * j + 2];
            const double path0 = slab_path(rz - half * vz, thickness, n_hat[2]);
            const double path1 = slab_path(rz + half * vz, thickness, n_hat[2]);
            const double trans = mean_transmission(path0 * mu, path1 * mu);
            double cell = 0.0;
            for (int el = 0; el < n_el; ++el) {
                const double zi = 2.0 * M_PI * Z[el] * ALPHA_FS;
                const double pref = n_cm3[el] * (double)L_ang[j] * 1.0e-8 * (16.0 / 3.0) * ALPHA_FS *
                                    R_E_CM2 * Z[el] * Z[el] * beta_i * (1.0 - exp(-zi / beta_i)) /
                                    (p_i * p_i);
                const double den_f = 1.0 - exp(-zi / beta_f);
                cell += pref * born_log / (fmax(k_eV, 1e-30) * beta_f * den_f);
            }
            s += cell * trans;
        }
        out[k] += s;
    }
}
