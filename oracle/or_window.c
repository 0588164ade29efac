/*
 * or_window.c -- support-windowed sinc^2 line reduction with the cycle-averaged
 * far tail, FP64 (TEST INFRASTRUCTURE ONLY).
 *
 * Restates the reference's per-row windowed node route:
 *   row_window_plan            spectrum/cxr/_node_window.py:225-250
 *   _sincsq_window_reduce      spectrum/cxr/_node_window.py:469-522
 *   _tail_value / correction   spectrum/cxr/_node_window.py:454-465
 *   NodeTail deposit + apply   spectrum/cxr/_node_window.py:86-180
 * for lines whose node lies in the convolution range [-n, 2n) (every line the
 * prologue's keep window admits).  The reference adds the deposited tail
 * moments by an FFT convolution; here the same sum is taken directly, which
 * differs only by the FFT's rounding.  Lines outside the convolution range
 * (rejected by the keep window upstream) and broad windows (at least a quarter
 * of the grid) are summed exactly over the whole grid, as the reference's
 * "broad" rule does.
 *
 * Pins (tests/test_window_oracle.py): with x_c large every window is broad and
 * the result equals the dense reduction or_sincsq_reduce exactly; for finite
 * x_c the windowed sum approaches the dense sum with the O(1/x_c^2) error the
 * cycle-averaged tail leaves; the tail moments reproduce sin^2/x^2's cycle
 * average 1/(2x^2) far from the line.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "hr_oracle.h"

static double sinc2(double a, double dE) {
    const double pi = 3.141592653589793;
    double x = a * dE / pi;
    if (x == 0.0) x = 1.0e-20;
    const double y = pi * x;
    const double s = sin(y) / y;
    return s * s;
}

static double tail_value(double m2, double m3, double m4, double x) {
    const double inv = 1.0 / x;
    return (m2 + (m3 + m4 * inv) * inv) * inv * inv;
}

/* Plan of one line (row_window_plan); returns 1 for a conv-mode window. */
int or_window_plan(double E_r, double a, double E0, double dE, int64_t n, double x_c, int64_t *node,
                   int64_t *lo, int64_t *hi, double *mu) {
    const double nd = rint((E_r - E0) / dE);
    const double eps = E_r - (E0 + nd * dE);
    double R = ceil((fabs(eps) + x_c / a) / dE);
    if (R > 4503599627370496.0) R = 4503599627370496.0;
    const int64_t nn = (int64_t)nd, RR = (int64_t)R;
    int64_t l = nn - RR < 0 ? 0 : nn - RR;
    int64_t h = nn + RR > n - 1 ? n - 1 : nn + RR;
    const int conv = nn + n >= 0 && nn + n < 3 * n;
    const int broad = h - l + 1 >= 0.25 * (double)n;
    *node = nn;
    const double mu2 = 0.5 / (a * a);
    mu[0] = mu2;
    mu[1] = 2.0 * eps * mu2;
    mu[2] = 3.0 * eps * eps * mu2;
    if (!conv || broad) {
        *lo = 0;
        *hi = n - 1;
        return 0;
    }
    *lo = l;
    *hi = h;
    return 1;
}

void or_window_reduce(const double *E_r, const double *a, const float *w, int64_t n_lines, double E0,
                      double dE, int64_t n, double x_c, double *out) {
    const int64_t L = 3 * n;
    double *dep = (double *)calloc((size_t)(3 * L), sizeof(double));
    double *win = (double *)calloc((size_t)n, sizeof(double));
    for (int64_t j = 0; j < n_lines; ++j) {
        const double wj = (double)w[j];
        if (wj == 0.0 || !(a[j] > 0.0)) continue;
        int64_t node, lo, hi;
        double mu[3];
        const int conv = or_window_plan(E_r[j], a[j], E0, dE, n, x_c, &node, &lo, &hi, mu);
        for (int64_t e = lo; e <= hi; ++e) {
            double v = sinc2(a[j], (E0 + (double)e * dE) - E_r[j]);
            if (conv && e != node) v -= tail_value(mu[0], mu[1], mu[2], (double)(e - node) * dE);
            win[e] += wj * v;
        }
        if (conv)
            for (int k = 0; k < 3; ++k) dep[k * L + node + n] += wj * mu[k];
    }
    for (int64_t e = 0; e < n; ++e) {
        double t = 0.0;
        for (int64_t q = 0; q < L; ++q) {
            if (dep[q] == 0.0 && dep[L + q] == 0.0 && dep[2 * L + q] == 0.0) continue;
            const int64_t off = e - (q - n);
            if (off == 0) continue;
            t += tail_value(dep[q], dep[L + q], dep[2 * L + q], (double)off * dE);
        }
        out[e] += win[e] + t;
    }
    free(dep);
    free(win);
}
