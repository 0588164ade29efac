/*
 * or_elsepa.c -- FP64 CPU restatement of tabulated (ELSEPA) elastic angles.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): loaded by tests/, never by the
 * product path.
 *
 * Follows src/pyrite/montecarlo/transport/scattering.py:
 *   :528-537  _elsepa_bracket (clamped lower node and fraction on ln E)
 *   :551-578  _elsepa_panel: the last panel k with cdf[k] <= xi, bisection
 *             confined to the guide window of xi's cell
 *   :581-610  _elsepa_invert_row: the root of the panel's quadratic CDF in the
 *             cancellation-free form t = 2r / (b + sqrt(b^2 + 2 dmu dp r))
 *   :613-630  _sample_cos_theta_elsepa: both bracketing rows inverted at one
 *             uniform, quantiles mixed linearly in ln E
 * and the device bracketing of _jit_shell_device.py:58-84.  The whole-row
 * bisection is the reference's pre-guide inversion (its test
 * test_elsepa_elastic.py:196-222), kept as the exactness check of the guide.
 */
#include <math.h>

#include "hr_oracle.h"

void or_elsepa_bracket(const double *g, int64_t row, int64_t n, double x, int64_t *lo, double *f) {
    if (x <= g[row]) {
        *lo = 0;
        *f = 0.0;
        return;
    }
    if (x >= g[row + n - 1]) {
        *lo = n - 2;
        *f = 1.0;
        return;
    }
    int64_t l = 0, h = n - 1;
    while (h - l > 1) {
        int64_t m = (l + h) / 2;
        if (g[row + m] <= x)
            l = m;
        else
            h = m;
    }
    *lo = l;
    *f = (x - g[row + l]) / (g[row + l + 1] - g[row + l]);
}

/* mu inside panel lo of `row` at xi */
static double panel_root(const or_elsepa *t, int64_t row, int lo, double xi) {
    const int64_t base = row * (int64_t)t->n_mu;
    double dmu = t->mu[lo + 1] - t->mu[lo];
    double p0 = t->pdf[base + lo], p1 = t->pdf[base + lo + 1];
    double r = xi - t->cdf[base + lo];
    double b = dmu * p0;
    double disc = b * b + 2.0 * dmu * (p1 - p0) * r;
    if (disc < 0.0) disc = 0.0;
    double den = b + sqrt(disc);
    double frac = den > 0.0 ? 2.0 * r / den : 0.0;
    if (frac < 0.0)
        frac = 0.0;
    else if (frac > 1.0)
        frac = 1.0;
    return t->mu[lo] + frac * dmu;
}

double or_elsepa_invert_row(const or_elsepa *t, int64_t row, double xi) {
    const int64_t base = row * (int64_t)t->n_mu;
    const int64_t gbase = row * (int64_t)(t->guide_cells + 1);
    double scaled = xi * (double)t->guide_cells;
    int cell;
    if (scaled >= (double)t->guide_cells)
        cell = t->guide_cells - 1;
    else if (scaled > 0.0)
        cell = (int)scaled;
    else
        cell = 0;
    int lo = t->guide[gbase + cell], hi = t->guide[gbase + cell + 1] + 1;
    while (hi - lo > 1) {
        int m = (lo + hi) / 2;
        if (t->cdf[base + m] <= xi)
            lo = m;
        else
            hi = m;
    }
    return panel_root(t, row, lo, xi);
}

double or_elsepa_invert_row_full(const or_elsepa *t, int64_t row, double xi) {
    const int64_t base = row * (int64_t)t->n_mu;
    int lo = 0, hi = t->n_mu - 1;
    while (hi - lo > 1) {
        int m = (lo + hi) / 2;
        if (t->cdf[base + m] <= xi)
            lo = m;
        else
            hi = m;
    }
    return panel_root(t, row, lo, xi);
}

void or_elsepa_energy_rows(const or_elsepa *t, int64_t el, double E_keV, int64_t *row, double *f) {
    int64_t lo;
    or_elsepa_bracket(t->logE, t->start[el], t->length[el], log(E_keV * 1.0e3), &lo, f);
    *row = t->start[el] + lo;
}

double or_elsepa_mix_mu(const or_elsepa *t, int64_t row, double f, double xi) {
    double mu0 = or_elsepa_invert_row(t, row, xi);
    double mu1 = or_elsepa_invert_row(t, row + 1, xi);
    return (1.0 - f) * mu0 + f * mu1;
}

double or_elsepa_sample_mu(const or_elsepa *t, int64_t el, double E_keV, double xi) {
    int64_t row;
    double f;
    or_elsepa_energy_rows(t, el, E_keV, &row, &f);
    return or_elsepa_mix_mu(t, row, f, xi);
}

/* Exact CDF of angular row `row` at mu = m (cone_split.py:177-197). */
double or_elsepa_cdf_at(const or_elsepa *t, int64_t row, double m) {
    int n = t->n_mu;
    if (m <= t->mu[0]) return 0.0;
    if (m >= t->mu[n - 1]) return 1.0;
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (t->mu[mid] <= m)
            lo = mid;
        else
            hi = mid;
    }
    int64_t base = row * (int64_t)n;
    double dmu = t->mu[lo + 1] - t->mu[lo];
    double x = (m - t->mu[lo]) / dmu;
    double p0 = t->pdf[base + lo], p1 = t->pdf[base + lo + 1];
    double v = t->cdf[base + lo] + dmu * (p0 * x + 0.5 * (p1 - p0) * x * x);
    return fmin(1.0, fmax(0.0, v));
}

/* Batches for the tests: mu of element slot `el` at E_keV[k], xi[k]; and row
 * inversions at (row[k], xi[k]), guided or by whole-row bisection. */
void or_elsepa_sample_mu_batch(const or_elsepa *t, int64_t el, int64_t n, const double *E_keV, const double *xi,
                               double *out) {
    for (int64_t k = 0; k < n; ++k) out[k] = or_elsepa_sample_mu(t, el, E_keV[k], xi[k]);
}

void or_elsepa_invert_batch(const or_elsepa *t, int64_t n, const int64_t *row, const double *xi, int guided,
                            double *out) {
    for (int64_t k = 0; k < n; ++k)
        out[k] = guided ? or_elsepa_invert_row(t, row[k], xi[k]) : or_elsepa_invert_row_full(t, row[k], xi[k]);
}
