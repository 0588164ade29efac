/*
 * or_hard.c -- FP64 CPU restatement of the hard-event modes of the transport
 * step: the coupled radiative partition and photon sampler, and the shell
 * soft/hard inelastic loss laws, recoil kinematics and soft-loss sampler.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): loaded by tests/, the smoke check and
 * bench.py's cpu_baseline leg, never by the product path.
 *
 * Follows the reference's host twins (not its device forms):
 *   src/pyrite/montecarlo/transport/_jit_radiative.py:26-288
 *       (_incident_row, _scaled_sdcs_at, _photon_nodes as a sorted union,
 *        radiative_moments_scalar, sample_hard_photon_energy_scalar)
 *   src/pyrite/montecarlo/transport/dielectric_radiative.py:19-80
 *   src/pyrite/montecarlo/transport/hard_inelastic.py:120-554
 *       (_hard_loss_bounds, _moller_j0, _triangle_cdf, Newton inversion and
 *        its bisection oracle, _qmin_ev_scalar, _hard_primary_cosine,
 *        _hard_secondary_cosine, _soft_loss_sample_keV)
 */
#include "hr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define RAD_MB_CM2 1.0e-27
#define MC2_EV 510998.95
#define TRUNC_WIDTH 1.015387
#define NEWTON_STEPS 200
#define NEWTON_RTOL 8.881784197001252e-16 /* 2^-50 */
#define BISECT_STEPS 200

/* ---- radiative ------------------------------------------------------------ */

/* _incident_row: lower row and ln-energy fraction (clamped). */
static void incident_row(const double *inc, int n, double E_eV, int *row, double *frac) {
    double le = log(E_eV * 1e-3);
    if (le <= log(inc[0])) { *row = 0; *frac = 0.0; return; }
    if (le >= log(inc[n - 1])) { *row = n - 2; *frac = 1.0; return; }
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (log(inc[mid]) <= le) lo = mid; else hi = mid;
    }
    *row = lo;
    *frac = (le - log(inc[lo])) / (log(inc[lo + 1]) - log(inc[lo]));
}

/* _scaled_sdcs_at: linear scan for the cell of each bracketing row. */
static double sdcs_at(double x, int row, double frac, const double *nom, int n_red, const double *top,
                      const double *chi /* [max_t, n_red] of the slot */) {
    double v[2];
    for (int side = 0; side < 2; ++side) {
        int r = row + side;
        const double *c = chi + (size_t)r * n_red;
        if (x >= top[r]) { v[side] = c[n_red - 1]; continue; }
        int lower = 0;
        while (lower + 1 < n_red - 1 && x >= nom[lower + 1]) lower++;
        double x0 = nom[lower];
        double x1 = (lower == n_red - 2) ? top[r] : nom[lower + 1];
        double w = (x - x0) / (x1 - x0);
        v[side] = c[lower] + w * (c[lower + 1] - c[lower]);
    }
    return v[0] + frac * (v[1] - v[0]);
}

static int cmp_double(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* _photon_nodes: sorted unique union of the nominal nodes, both tips, kc/T and 1,
 * as photon energies with the cutoff node exact.  Returns the count. */
static int photon_nodes(double E_eV, double cutoff_eV, int row, const double *nom, int n_red, const double *top,
                        double *out /* n_red + 3 */) {
    int n = 0;
    for (int i = 0; i < n_red - 1; ++i) out[n++] = nom[i];
    out[n++] = top[row];
    out[n++] = top[row + 1];
    out[n++] = cutoff_eV / E_eV;
    out[n++] = 1.0;
    qsort(out, (size_t)n, sizeof(double), cmp_double);
    int u = 1;
    for (int i = 1; i < n; ++i)
        if (out[i] > out[u - 1]) out[u++] = out[i];
    double cut_red = cutoff_eV / E_eV;
    for (int i = 0; i < u; ++i) out[i] = (out[i] == cut_red) ? cutoff_eV : out[i] * E_eV;
    return u;
}

/* dielectric_cell_moments: (zeroth, first) of chi*F on [left, right] before Z^2 mb scaling. */
static void dielectric_cell(double left, double right, double cl, double cr, double s, double *zeroth,
                            double *first) {
    if (right <= left) { *zeroth = 0.0; *first = 0.0; return; }
    double slope = (cr - cl) / (right - left);
    double icpt = cl - slope * left;
    double u = left / s, v = right / s;
    if (v < 0.1) {
        double z = 0.0, f = 0.0, sign = 1.0;
        for (int n = 0; n < 6; ++n) {
            int p = 2 * n + 2;
            z += sign * (icpt * (pow(v, p) - pow(u, p)) / p + slope * s * (pow(v, p + 1) - pow(u, p + 1)) / (p + 1));
            f += sign * s * (icpt * (pow(v, p + 1) - pow(u, p + 1)) / (p + 1) +
                             slope * s * (pow(v, p + 2) - pow(u, p + 2)) / (p + 2));
            sign = -sign;
        }
        *zeroth = z;
        *first = f;
        return;
    }
    double dc = 0.5 * (log1p(v * v) - log1p(u * u));
    double da = (v - u) - (atan(v) - atan(u));
    double db = 0.5 * (v * v - u * u) - dc;
    *zeroth = icpt * dc + slope * s * da;
    *first = s * (icpt * da + slope * s * db);
}

/* _linear_cell: (zeroth, first) of a linear chi cell. */
static void linear_cell(double left, double right, double cl, double cr, double *rate, double *first) {
    if (right <= left) { *rate = 0.0; *first = 0.0; return; }
    double slope = (cr - cl) / (right - left);
    double icpt = cl - slope * left;
    *rate = icpt * log(right / left) + slope * (right - left);
    *first = 0.5 * (cl + cr) * (right - left);
}

/* radiative_moments_scalar: soft first [eV cm^2], hard zeroth [cm^2], hard first [eV cm^2] per atom. */
void or_rad_moments(const or_rad_tables *t, int slot, double E_eV, double cutoff_eV, double plasma_eV, double *soft,
                    double *hard_rate, double *hard_first) {
    const double *inc = t->incident + (size_t)slot * t->max_t;
    const double *top = t->top + (size_t)slot * t->max_t;
    const double *nom = t->nominal + (size_t)slot * t->n_red;
    const double *chi = t->chi + (size_t)slot * t->max_t * t->n_red;
    int row;
    double frac;
    incident_row(inc, t->nt[slot], E_eV, &row, &frac);
    double nodes[128];
    int nn = photon_nodes(E_eV, cutoff_eV, row, nom, t->n_red, top, nodes);
    double scale = RAD_MB_CM2 * t->z[slot] * t->z[slot];
    double sup = (1.0 + E_eV / MC2_EV) * plasma_eV;
    double s = 0.0, hr = 0.0, hf = 0.0;
    double cl = sdcs_at(nodes[0] / E_eV, row, frac, nom, t->n_red, top, chi);
    for (int i = 0; i < nn - 1; ++i) {
        double left = nodes[i], right = nodes[i + 1];
        double cr = sdcs_at(right / E_eV, row, frac, nom, t->n_red, top, chi);
        double rate, first;
        if (sup == 0.0) {
            if (left == 0.0) {
                s += scale * 0.5 * (cl + cr) * right;
                cl = cr;
                continue;
            }
            linear_cell(left, right, cl, cr, &rate, &first);
        } else {
            dielectric_cell(left, right, cl, cr, sup, &rate, &first);
        }
        if (right <= cutoff_eV) {
            s += scale * first;
        } else {
            hr += scale * rate;
            hf += scale * first;
        }
        cl = cr;
    }
    *soft = s;
    *hard_rate = hr;
    *hard_first = hf;
}

/* sample_hard_photon_energy_scalar: invert the zeroth-moment CDF (52 bisection steps in the cell). */
double or_rad_sample_photon(const or_rad_tables *t, int slot, double E_eV, double cutoff_eV, double u,
                            double plasma_eV) {
    double soft, rate, hf;
    or_rad_moments(t, slot, E_eV, cutoff_eV, plasma_eV, &soft, &rate, &hf);
    if (rate <= 0.0) return E_eV;
    const double *inc = t->incident + (size_t)slot * t->max_t;
    const double *top = t->top + (size_t)slot * t->max_t;
    const double *nom = t->nominal + (size_t)slot * t->n_red;
    const double *chi = t->chi + (size_t)slot * t->max_t * t->n_red;
    int row;
    double frac;
    incident_row(inc, t->nt[slot], E_eV, &row, &frac);
    double nodes[128];
    int nn = photon_nodes(E_eV, cutoff_eV, row, nom, t->n_red, top, nodes);
    double scale = RAD_MB_CM2 * t->z[slot] * t->z[slot];
    double sup = (1.0 + E_eV / MC2_EV) * plasma_eV;
    double target = u * rate, cum = 0.0;
    double cl = sdcs_at(nodes[0] / E_eV, row, frac, nom, t->n_red, top, chi);
    for (int i = 0; i < nn - 1; ++i) {
        double left = nodes[i], right = nodes[i + 1];
        double cr = sdcs_at(right / E_eV, row, frac, nom, t->n_red, top, chi);
        if (right > cutoff_eV) {
            double mass, first;
            if (sup == 0.0) linear_cell(left, right, cl, cr, &mass, &first);
            else dielectric_cell(left, right, cl, cr, sup, &mass, &first);
            mass *= scale;
            if (cum + mass >= target || i == nn - 2) {
                double local = target - cum, lo = left, hi = right;
                for (int k = 0; k < 52; ++k) {
                    double mid = 0.5 * (lo + hi);
                    double cm = cl + (cr - cl) * (mid - left) / (right - left);
                    double part, f2;
                    if (sup == 0.0) linear_cell(left, mid, cl, cm, &part, &f2);
                    else dielectric_cell(left, mid, cl, cm, sup, &part, &f2);
                    if (scale * part < local) lo = mid; else hi = mid;
                }
                return 0.5 * (lo + hi);
            }
            cum += mass;
        }
        cl = cr;
    }
    return E_eV;
}

/* radiative_layer_moments_scalar: soft [eV/A] and hard [1/A]; element rates into rates[]. */
void or_rad_layer_moments(const or_rad_tables *t, int layer, int n_el, int max_el, double E_eV, double *soft,
                          double *hard, double *rates) {
    double s = 0.0, h = 0.0;
    for (int i = 0; i < n_el; ++i) {
        int slot = layer * max_el + i;
        double sm, r, hf;
        or_rad_moments(t, slot, E_eV, t->cutoff_eV, t->plasma[layer], &sm, &r, &hf);
        double per_ang = t->ncm3[slot] * 1e-8;
        s += per_ang * sm;
        double ri = per_ang * r;
        if (rates) rates[i] = ri;
        h += ri;
    }
    *soft = s;
    *hard = h;
}

/* ---- shell ---------------------------------------------------------------- */

/* _hard_loss_bounds (electron branches 0, 1, 2). */
void or_shell_bounds(double E, double U, double W, int branch, double Wc, double *lower, double *upper,
                     double *peak_end) {
    double w_max = 0.5 * (E + U);
    if (branch == 2) {
        double base = U > 0.0 ? U : W;
        *lower = fmax(base, Wc);
        *upper = w_max;
        *peak_end = 0.0;
        return;
    }
    if (U == 0.0) { *lower = W; *upper = W; *peak_end = 0.0; return; }
    double pe = 3.0 * W - 2.0 * U;
    if (E <= pe) pe = E;
    *upper = fmin(pe, w_max);
    *lower = fmax(U, Wc);
    *peak_end = pe;
}

/* _moller_j0: antiderivative of F(-)(E', W)/W^2. */
double or_moller_j0(double E, double prime, double w) {
    double a = (E / (E + MC2_EV)) * (E / (E + MC2_EV));
    double rest = prime - w;
    return -1.0 / w + 1.0 / rest + (1.0 - a) / prime * log(rest / w) + a * w / (prime * prime);
}

/* _triangle_cdf: unnormalised int_lower^w p_dis(W)/W dW. */
double or_triangle_cdf(double peak_end, double lower, double w) { return peak_end * log(w / lower) - (w - lower); }

static void cdf_density(int kind, double E, double prime, double pe, double lower, double w, double *v, double *d) {
    if (kind == 1) {
        double a = (E / (E + MC2_EV)) * (E / (E + MC2_EV));
        double rest = prime - w;
        double c1 = (1.0 - a) / prime;
        *v = -1.0 / w + 1.0 / rest + c1 * log(rest / w) + a * w / (prime * prime);
        *d = 1.0 / (w * w) + 1.0 / (rest * rest) - c1 * (1.0 / rest + 1.0 / w) + a / (prime * prime);
        return;
    }
    *v = pe * log(w / lower) - (w - lower);
    *d = pe / w - 1.0;
}

/* _sample_hard_transfer_eV: safeguarded Newton inversion (eV). */
double or_shell_sample_transfer(double E, double U, double W, int branch, double Wc, double u) {
    double lower, upper, pe;
    or_shell_bounds(E, U, W, branch, Wc, &lower, &upper, &pe);
    if (branch != 2 && U == 0.0) return W;
    if (!(upper > lower)) return lower;
    double prime = E + U;
    int kind = branch == 2 ? 1 : 0;
    double base, bd, top, td;
    cdf_density(kind, E, prime, pe, lower, lower, &base, &bd);
    cdf_density(kind, E, prime, pe, lower, upper, &top, &td);
    double target = u * (top - base);
    if (!(target > 0.0)) return lower;
    double w = lower + target / bd;
    if (kind != 0) {
        double inv = 1.0 / lower - target;
        w = inv > 0.0 ? 1.0 / inv : upper;
    }
    double lo = lower, hi = upper;
    if (!(w > lo && w < hi)) w = 0.5 * (lo + hi);
    for (int k = 0; k < NEWTON_STEPS; ++k) {
        double v, d;
        cdf_density(kind, E, prime, pe, lower, w, &v, &d);
        double res = v - base - target;
        if (res == 0.0) return w;
        if (res < 0.0) lo = w; else hi = w;
        double wn = w - res / d;
        if (!(wn > lo && wn < hi)) wn = 0.5 * (lo + hi);
        if (fabs(wn - w) <= NEWTON_RTOL * w || !(wn > lo && wn < hi)) return wn;
        w = wn;
    }
    return w;
}

/* _sample_hard_transfer_eV_bisection: the reference's own oracle of the Newton form. */
double or_shell_sample_transfer_bisection(double E, double U, double W, int branch, double Wc, double u) {
    double lower, upper, pe;
    or_shell_bounds(E, U, W, branch, Wc, &lower, &upper, &pe);
    if (branch != 2 && U == 0.0) return W;
    if (!(upper > lower)) return lower;
    double prime = E + U, base = 0.0, target;
    if (branch == 2) {
        base = or_moller_j0(E, prime, lower);
        target = u * (or_moller_j0(E, prime, upper) - base);
    } else {
        target = u * or_triangle_cdf(pe, lower, upper);
    }
    double lo = lower, hi = upper;
    for (int k = 0; k < BISECT_STEPS; ++k) {
        double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        double value = branch == 2 ? or_moller_j0(E, prime, mid) - base : or_triangle_cdf(pe, lower, mid);
        if (value < target) lo = mid; else hi = mid;
    }
    return 0.5 * (lo + hi);
}

/* _qmin_ev_scalar */
double or_qmin(double E, double W) {
    double p0 = sqrt(E * (E + 2.0 * MC2_EV));
    double rem = E - W;
    double p1 = sqrt(rem * (rem + 2.0 * MC2_EV));
    double dp = W * (2.0 * (E + MC2_EV) - W) / (p0 + p1);
    return dp * dp / (sqrt(dp * dp + MC2_EV * MC2_EV) + MC2_EV);
}

static int distant_q(double E, double U, double W, double u, double *Q, double *res) {
    double two = 2.0 * MC2_EV, q_up;
    if (U > 0.0) {
        double full = 3.0 * W - 2.0 * U;
        if (E > full) { *res = W; q_up = U; }
        else { *res = (E + 2.0 * U) / 3.0; q_up = U * E / full; }
    } else {
        *res = W;
        q_up = W;
    }
    if (!(*res < E)) return 0;
    double q_lo = or_qmin(E, *res);
    if (!(0.0 < q_lo && q_lo < q_up)) return 0;
    double ll = log(q_lo / (q_lo + two)), lu = log(q_up / (q_up + two));
    double lr = (1.0 - u) * ll + u * lu;
    *Q = two / expm1(-lr);
    return 1;
}

/* _hard_primary_cosine */
double or_shell_primary_cosine(double E, double U, double W, int branch, double W_eV, double u) {
    double two = 2.0 * MC2_EV;
    if (branch == 1) return 1.0;
    if (branch == 2) {
        double rem = E - W_eV;
        return sqrt(rem / E * (E + two) / (rem + two));
    }
    double Q, res;
    if (!distant_q(E, U, W, u, &Q, &res)) return 1.0;
    double p0 = E * (E + two), rem = E - res, p1 = rem * (rem + two);
    double c = (p0 + p1 - Q * (Q + two)) / (2.0 * sqrt(p0 * p1));
    return fmin(1.0, fmax(-1.0, c));
}

/* _hard_secondary_cosine */
double or_shell_secondary_cosine(double E, double U, double W, int branch, double W_eV, double u) {
    double two = 2.0 * MC2_EV;
    if (branch == 1) return 0.5;
    if (branch == 2) return sqrt(W_eV / E * (E + two) / (W_eV + two));
    double Q, res;
    if (!distant_q(E, U, W, u, &Q, &res)) return 1.0;
    double p0 = E * (E + two), rem = E - res, p1 = rem * (rem + two), q2 = Q * (Q + two);
    double c = (p0 + q2 - p1) / (2.0 * sqrt(p0 * q2));
    return fmin(1.0, fmax(-1.0, c));
}

/* _soft_loss_sample_keV: loss [keV]; advances *counter. */
double or_shell_soft_loss(double mean, double var, uint64_t key, uint64_t *counter) {
    if (mean <= 0.0) return 0.0;
    if (var <= 0.0) return mean;
    double sigma = sqrt(var), m2 = mean * mean;
    if (m2 > 9.0 * var) {
        double width = TRUNC_WIDTH * sigma;
        for (;;) {
            double u1 = or_stream_uniform(key, *counter);
            *counter += 1;
            double u2 = or_stream_uniform(key, *counter);
            *counter += 1;
            double z = sqrt(-2.0 * log(1.0 - u1)) * cos(2.0 * M_PI * u2);
            double off = width * z;
            if (fabs(off) < 3.0 * sigma) return mean + off;
        }
    }
    double u = or_stream_uniform(key, *counter);
    *counter += 1;
    if (m2 > 3.0 * var) {
        double half = sqrt(3.0) * sigma;
        return mean - half + 2.0 * half * u;
    }
    double weight = (3.0 * var - m2) / (3.0 * var + 3.0 * m2);
    double top = (3.0 * var + 3.0 * m2) / (2.0 * mean);
    if (u < weight) return 0.0;
    return top * (u - weight) / (1.0 - weight);
}

/* ---- batched helpers for the tests ------------------------------------------ */
void or_rad_layer_lut(const or_rad_tables *t, int n_layers, const int32_t *nel, int max_el, const double *E_keV,
                      int64_t n_E, double *soft_keV_per_ang, double *rate_per_ang) {
    for (int L = 0; L < n_layers; ++L)
        for (int64_t j = 0; j < n_E; ++j) {
            double s, h;
            or_rad_layer_moments(t, L, nel[L], max_el, E_keV[j] * 1e3, &s, &h, NULL);
            soft_keV_per_ang[(size_t)L * n_E + j] = s * 1e-3;
            rate_per_ang[(size_t)L * n_E + j] = h;
        }
}
