// elsepa.hpp -- tabulated (ELSEPA) elastic angles for the transport step
// (the reference's elastic_model_code == 2 branch, device functions
// src/pyrite/montecarlo/transport/_jit_device.py:493-620 and the grid
// bracketing of _jit_shell_device.py:58-84; host restatement
// transport/scattering.py:540-620).
//
// Each energy row of an element holds a piecewise-linear density p(mu) on the
// shared grid mu in [0, 1] (mu = (1 - cos theta)/2) and its exact CDF.  One
// uniform xi is inverted on the two rows bracketing ln E[eV] and the two
// quantiles are mixed linearly in ln E, which keeps the sample monotone in xi
// and reproduces a node's law exactly at the node.  Inside a panel the CDF is
// quadratic in t = (mu - mu_k)/dmu; the root is taken in the cancellation-free
// form t = 2r / (b + sqrt(b^2 + 2 dmu (p_k+1 - p_k) r)), b = dmu p_k,
// r = xi - F_k.  The panel is the last k with F_k <= xi, found by bisection
// inside the window a guide table gives for xi's cell of [0, 1): the same
// comparisons as a bisection over the whole row, on about two panels.
#pragma once

#include "hr_common.hpp"

namespace hr {
namespace elsepa {

// Lower node (relative to `row`) and fraction of x on the ascending nodes
// g[row .. row + n), clamped: (0, 0) at or below the first, (n - 2, 1) at or
// above the last.
__host__ __device__ __forceinline__ void bracket(const double *__restrict__ g, int64_t row, int64_t n, double x,
                                                 int64_t &lo, double &f) {
    if (x <= g[row]) {
        lo = 0;
        f = 0.0;
        return;
    }
    if (x >= g[row + n - 1]) {
        lo = n - 2;
        f = 1.0;
        return;
    }
    int64_t l = 0, h = n - 1;
    while (h - l > 1) {
        const int64_t m = (l + h) / 2;
        if (g[row + m] <= x) l = m;
        else h = m;
    }
    lo = l;
    f = (x - g[row + l]) / (g[row + l + 1] - g[row + l]);
}

// Exact inversion of angular row `row` at xi (guided panel search).
__host__ __device__ __forceinline__ double invert_row(const hr_elsepa &t, int64_t row, double xi) {
    const int64_t base = row * (int64_t)t.n_mu;
    const int64_t gbase = row * (int64_t)(t.guide_cells + 1);
    const double scaled = xi * (double)t.guide_cells;
    int cell = 0;
    if (scaled >= (double)t.guide_cells) cell = t.guide_cells - 1;
    else if (scaled > 0.0) cell = (int)scaled;
    int lo = t.guide[gbase + cell];
    int hi = t.guide[gbase + cell + 1] + 1;
    while (hi - lo > 1) {
        const int m = (lo + hi) / 2;
        if (t.cdf[base + m] <= xi) lo = m;
        else hi = m;
    }
    const double dmu = t.mu[lo + 1] - t.mu[lo];
    const double p0 = t.pdf[base + lo], p1 = t.pdf[base + lo + 1];
    const double r = xi - t.cdf[base + lo];
    const double b = dmu * p0;
    double disc = b * b + 2.0 * dmu * (p1 - p0) * r;
    if (disc < 0.0) disc = 0.0;
    const double den = b + sqrt(disc);
    double frac = 0.0;
    if (den > 0.0) frac = 2.0 * r / den;
    if (frac < 0.0) frac = 0.0;
    if (frac > 1.0) frac = 1.0;
    return t.mu[lo] + frac * dmu;
}

// Quantile of the ln-E mixture of rows `row` and `row + 1` (weight f on the upper) at xi.
__host__ __device__ __forceinline__ double mix_mu(const hr_elsepa &t, int64_t row, double f, double xi) {
    const double mu0 = invert_row(t, row, xi);
    const double mu1 = invert_row(t, row + 1, xi);
    return (1.0 - f) * mu0 + f * mu1;
}

// Absolute first row and fraction of the element row block `el` at E_keV.
__host__ __device__ __forceinline__ void energy_rows(const hr_elsepa &t, size_t el, double E_keV, int64_t &row,
                                                     double &f) {
    const int64_t first = t.start[el];
    int64_t lo;
    bracket(t.logE, first, t.length[el], log(E_keV * 1.0e3), lo, f);
    row = first + lo;
}

// mu of element row block `el` (index into start/length) at energy E_keV and uniform xi.
__host__ __device__ __forceinline__ double sample_mu(const hr_elsepa &t, size_t el, double E_keV, double xi) {
    int64_t row;
    double f;
    energy_rows(t, el, E_keV, row, f);
    return mix_mu(t, row, f, xi);
}

// Exact CDF of angular row `row` at mu = m: the piecewise-linear density's
// quadratic inside the panel holding m, clamped to [0, 1]
// (cone_split.py:177-197).
__host__ __device__ __forceinline__ double cdf_at(const hr_elsepa &t, int64_t row, double m) {
    const int n = t.n_mu;
    if (m <= t.mu[0]) return 0.0;
    if (m >= t.mu[n - 1]) return 1.0;
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (t.mu[mid] <= m) lo = mid;
        else hi = mid;
    }
    const int64_t base = row * (int64_t)n;
    const double dmu = t.mu[lo + 1] - t.mu[lo];
    const double x = (m - t.mu[lo]) / dmu;
    const double p0 = t.pdf[base + lo], p1 = t.pdf[base + lo + 1];
    const double v = t.cdf[base + lo] + dmu * (p0 * x + 0.5 * (p1 - p0) * x * x);
    return fmin(1.0, fmax(0.0, v));
}

// Pointers and sizes the tabulated angles read; checked before any launch.
inline hr_status check(const hr_transport_args &a, const char *who) {
    const hr_elsepa &t = a.elsepa;
    HR_REQUIRE(t.on == 0 || t.on == 1, "%s: elsepa.on must be 0 or 1", who);
    if (!t.on) return HR_OK;
    HR_REQUIRE(t.start && t.length && t.logE && t.mu && t.cdf && t.pdf && t.guide,
               "%s: tabulated elastic angles need start, length, logE, mu, cdf, pdf and guide", who);
    HR_REQUIRE(t.n_mu >= 2, "%s: the angular grid needs >= 2 nodes (n_mu=%d)", who, t.n_mu);
    HR_REQUIRE(t.guide_cells >= 1 && (t.guide_cells & (t.guide_cells - 1)) == 0,
               "%s: guide_cells must be a positive power of two (%d)", who, t.guide_cells);
    return HR_OK;
}

}  // namespace elsepa
}  // namespace hr
