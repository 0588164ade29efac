// line_core.hpp -- per-(segment, reflection) line kinematics on gfx950, shared
// by the incoherent line prologue (score.hip) and the coherent one
// (coherent.hip).
//
// Follows the reference's batched line route: segment kinematics
// (cxr/_setup.py:981-985, cxr/_kernels.py:678-699), slab escape paths
// (segment_escape.py:78-117,190-211), the recoil pair and in-medium resonance
// root (cxr/_fused.py:335-386, cxr/_kernels.py:624-766), the table gather
// (cxr/_kernels.py:562-620) and the quantum PXR/CBS line elements
// (cxr/_recoil.py:35-76, cxr/_batched.py:789-826).  Everything is FP64: it
// runs once per line.
#pragma once

#include <math.h>

#include "hr_common.hpp"

namespace hr {
namespace line {

constexpr double kHbarc = 1973.269804;     // materials/crystal.py:48
constexpr double kAlpha = 1.0 / 137.035999; // crystal.py:49
constexpr double kMe = 510998.95;          // crystal.py:50
constexpr double kPi = 3.141592653589793;
constexpr double kRootRtol = 1e-3;         // cxr/_kernels.py:675

// numpy searchsorted(side="left") + clip(., 1, n-1) (_kernels.py:328-349),
// found from a guess instead of a bisection: the coupling grids are
// geometric, so index ~ log(x / grid[0]) / log(ratio), and a short walk from
// the guess lands on the first index with grid[i] >= x -- exactly the
// bisection's answer for any nondecreasing grid.  A walk longer than kWalk (a
// grid far from geometric) finishes by bisection.  Two or three dependent
// loads replace the bisection's dozen.
struct GridGuess {
    float g0, inv_lr;
};

__device__ __forceinline__ GridGuess grid_guess(const float *__restrict__ grid, int64_t n) {
    GridGuess q;
    q.g0 = grid[0];
    const float span = __logf(grid[n - 1] / q.g0);
    q.inv_lr = span > 0.0f ? (float)(n - 1) / span : 0.0f;
    return q;
}

__device__ __forceinline__ int64_t bracket_from(const float *__restrict__ grid, int64_t n, double x,
                                                const GridGuess &q) {
    constexpr int kWalk = 16;
    int64_t y;
    if (isnan(x)) {
        y = n;
    } else if (x > (double)grid[n - 1]) {
        y = n;
    } else {
        const float t = (float)x > q.g0 ? __logf((float)x / q.g0) * q.inv_lr : 0.0f;
        int64_t m = t < (float)(n - 1) ? (int64_t)t : n - 1;
        int steps = 0;
        if ((double)grid[m] < x) {
            // grid[n-1] >= x, so the walk up ends
            while ((double)grid[m] < x && steps < kWalk) {
                ++m;
                ++steps;
            }
            if ((double)grid[m] < x) {  // bisection over (m, n-1]
                int64_t lo = m + 1, hi = n - 1;
                while (lo < hi) {
                    const int64_t c = lo + ((hi - lo) >> 1);
                    if ((double)grid[c] < x) lo = c + 1;
                    else hi = c;
                }
                m = hi;
            }
        } else {
            while (m > 0 && (double)grid[m - 1] >= x && steps < kWalk) {
                --m;
                ++steps;
            }
            if (m > 0 && (double)grid[m - 1] >= x) {  // bisection over [0, m)
                int64_t lo = 0, hi = m - 1;
                while (lo < hi) {
                    const int64_t c = lo + ((hi - lo) >> 1);
                    if ((double)grid[c] < x) lo = c + 1;
                    else hi = c;
                }
                m = hi;
            }
        }
        y = m;
    }
    return y < 1 ? 1 : (y > n - 1 ? n - 1 : y);
}

__device__ __forceinline__ double slab_path(double z, double thickness, double nz) {
    const double inv_nz = 1.0 / fmax(fabs(nz), 1.0e-12);
    const double depth = z < 0.0 ? 0.0 : (z > thickness ? thickness : z);
    return nz < 0.0 ? depth * inv_nz : (thickness - depth) * inv_nz;
}

__device__ __forceinline__ double mean_transmission(double t0, double t1) {
    const double lo = fmin(t0, t1);
    const double d = fabs(t1 - t0);
    const double ratio = d > 1.0e-4 ? -expm1(-d) / d : 1.0 - 0.5 * d + d * d / 6.0;
    return exp(-lo) * ratio;
}

// Reflection-independent kinematics of one (segment, direction) pair.
struct SegKin {
    double vx, vy, vz, gamma, denom_row, t_L, gz, v_dot_grad, grad2, path0, path1, inv_E;
};

__device__ __forceinline__ SegKin seg_kin(double rz, double vhx, double vhy, double vhz, double L, double E,
                                          double nx, double ny, double nz, double thickness) {
    SegKin k;
    const double gg0 = 1.0 + E / 510.99895;
    const double beta = sqrt(1.0 - 1.0 / (gg0 * gg0));
    k.vx = beta * vhx;
    k.vy = beta * vhy;
    k.vz = beta * vhz;
    k.gamma = 1.0 + E / (kMe / 1e3);
    const double one_minus_beta = 1.0 / (k.gamma * k.gamma * (1.0 + beta));
    const double ddx = vhx - nx, ddy = vhy - ny, ddz = vhz - nz;
    k.denom_row = one_minus_beta + beta * (0.5 * (ddx * ddx + ddy * ddy + ddz * ddz));
    k.t_L = L / beta;
    k.gz = -1.0 / nz;  // slab escape gradient -e/(e.n), e = +z; zero outside the slab
    if (rz < 0.0 || rz > thickness) k.gz = 0.0;
    k.v_dot_grad = k.vz * k.gz;
    k.grad2 = k.gz * k.gz;
    const double half = 0.5 * L;
    k.path0 = slab_path(rz - half * vhz, thickness, nz);
    k.path1 = slab_path(rz + half * vhz, thickness, nz);
    k.inv_E = kHbarc / (k.gamma * kMe);  // _recoil.py:17-23
    return k;
}

// One (segment, direction, reflection) line: the recoil pair, the in-medium
// resonance, its width slope, mu at the resonance and both polarizations'
// complex amplitudes A = chi f_pxr + U f_cbs (before squaring).
struct LineCore {
    bool keep;
    double E_res, omega, slope, mu, vdg, r_num, r_den;
    double A_re[2], A_im[2];
};

// P carries the crystal and coupling tables: g, row, e_s, e_p, n_g, n_tab,
// E_tab, delta_tab, chi_re, chi_im, u_re, u_im, n_mu, log_mu, lo_keep, hi_keep.
template <class P>
__device__ __forceinline__ LineCore line_core(const P &p, const SegKin &k, double nx, double ny, double nz, int d,
                                              int g, const GridGuess &guess) {
    LineCore o;
    const double h = kHbarc / kMe;
    const int64_t n_tab = p.n_tab;
    const double Gx = p.g[3 * g], Gy = p.g[3 * g + 1], Gz = p.g[3 * g + 2];
    const double g2 = Gx * Gx + Gy * Gy + Gz * Gz;
    const double n_dot_g = Gx * nx + Gy * ny + Gz * nz;
    const double r_num = 0.5 * g2 * k.inv_E;  // _fused.py:335-345
    const double r_den = n_dot_g * k.inv_E;
    const double denom_vac = k.denom_row + r_den;
    const double vdg = k.vx * Gx + k.vy * Gy + k.vz * Gz;
    const double v_dot_g = vdg - r_num;
    // in-medium first-order Snell root, three fixed-point passes (_kernels.py:702-766)
    double denom = denom_vac, previous = denom, delta = 0.0;
    for (int pass = 0; pass < 3; ++pass) {
        const double Ex = kHbarc * (v_dot_g / denom);
        const int64_t i = bracket_from(p.E_tab, n_tab, Ex, guess);
        const double x0 = p.E_tab[i - 1];
        const double fr = (Ex - x0) / ((double)p.E_tab[i] - x0);
        const double f0 = p.delta_tab[i - 1];
        double y = f0 + fr * ((double)p.delta_tab[i] - f0);
        if (Ex <= (double)p.E_tab[0]) y = p.delta_tab[0];
        if (Ex >= (double)p.E_tab[n_tab - 1]) y = p.delta_tab[n_tab - 1];
        delta = y;
        previous = denom;
        denom = denom_vac - delta * k.v_dot_grad;
    }
    bool settled = fabs(denom - previous) <= kRootRtol * fabs(denom);
    settled = settled && (2.0 * fabs(delta) * k.grad2 < 1.0);
    if (!settled) denom = __builtin_nan("");
    const double n_dot_g_blk = n_dot_g + delta * (k.gz * Gz);
    const double omega = (vdg - r_num) / denom;  // _kernels.py:624-652
    const double k_dot_v = omega * ((1.0 - denom) + r_den);
    const double k_dot_g = omega * n_dot_g_blk;
    const double detuning = g2 + 2.0 * k_dot_g;
    const double radicand = k.grad2 == 0.0 ? 1.0 : 1.0 - 2.0 * delta + delta * delta * k.grad2;
    const double k_mag = omega * sqrt(radicand);
    const double E_res = kHbarc * omega;
    bool keep = (E_res > p.lo_keep) && (E_res > 10.0) && (E_res < p.hi_keep);
    const double oe = omega * k.inv_E;
    o.slope = denom * (1.0 / (1.0 - oe));
    keep = keep && (denom > 0.0) && (oe <= 1.0 - 1.0 / k.gamma);
    o.keep = keep;
    o.E_res = E_res;
    o.omega = omega;
    o.vdg = vdg;
    o.r_num = r_num;
    o.r_den = r_den;
    o.mu = 0.0;
    o.A_re[0] = o.A_re[1] = o.A_im[0] = o.A_im[1] = 0.0;
    if (!keep) return o;
    const int64_t i = bracket_from(p.E_tab, n_tab, E_res, guess);
    const double x0 = p.E_tab[i - 1], x1 = p.E_tab[i];
    const double fr = (E_res - x0) / (x1 - x0);
    const bool below = E_res <= (double)p.E_tab[0];
    const bool above = E_res >= (double)p.E_tab[n_tab - 1];
    const double bounded = fmin(fmax(E_res, x0), x1);
    const double log_frac = log1p((bounded - x0) / x0) / log1p((x1 - x0) / x0);
    const int64_t base = (int64_t)p.row[g] * n_tab;
    double tab[4];
    const float *tables[4] = {p.chi_re, p.chi_im, p.u_re, p.u_im};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const float *T = tables[t] + base;
        const double t0 = T[i - 1];
        double y = t0 + fr * ((double)T[i] - t0);
        if (below) y = T[0];
        if (above) y = T[n_tab - 1];
        tab[t] = y;
    }
    for (int el = 0; el < p.n_mu; ++el) {
        const float *T = p.log_mu + (int64_t)el * n_tab;
        const double l0 = T[i - 1];
        double lm = l0 + log_frac * ((double)T[i] - l0);
        if (below) lm = T[0];
        if (above) lm = T[n_tab - 1];
        o.mu += exp(lm);
    }
    const double gg = k.gamma * k.gamma;
    const double gamma = k.gamma;
    const double k2 = k_mag * k_mag;
    // _recoil.py:35-76, polarization-independent parts hoisted
    const double n_1 = h * g2 - 2.0 * gamma * vdg;
    const double n_2 = n_1 + 2.0 * h * k_dot_g;
    const double n_f = h * (g2 + 2.0 * k_dot_g + k2) - 2.0 * gamma * (vdg + k_dot_v);
    const double e_1 = sqrt(gg + h * n_1);
    const double e_2 = sqrt(gg + h * (h * k2 - 2.0 * gamma * k_dot_v));
    const double e_f = sqrt(gg + h * n_f);
    const double s_if = gamma + e_f;
    const double s_1 = e_1 + gamma, s_3 = e_1 + e_f, s_4 = gamma + e_2;
    const double sigma = s_if + e_1 + e_2;
    const int64_t pol_base = ((int64_t)d * p.n_g + g) * 3;
#pragma unroll
    for (int pol = 0; pol < 2; ++pol) {
        const double *e = (pol == 0 ? p.e_s : p.e_p) + pol_base;
        const double v_dot_e = k.vx * e[0] + k.vy * e[1] + k.vz * e[2];
        const double k_dot_e = omega * delta * (k.gz * e[2]);
        const double g_dot_e = Gx * e[0] + Gy * e[1] + Gz * e[2];
        const double ge_ke = g_dot_e + k_dot_e;
        const double pxr_num = (-n_f * ge_ke - k2 * (2.0 * gamma * v_dot_e - h * ge_ke)) / s_if;
        const double x = 2.0 * gamma * v_dot_e - h * k_dot_e;
        const double f_cbs = 2.0 * g_dot_e * s_1 / (s_3 * n_1) +
                             x * (n_1 * (n_f / s_if) * sigma - 2.0 * s_1 * s_4 * k_dot_g) / (s_3 * s_4 * n_1 * n_2);
        const double f_pxr = pxr_num / detuning;
        o.A_re[pol] = tab[0] * f_pxr + tab[2] * f_cbs;
        o.A_im[pol] = tab[1] * f_pxr + tab[3] * f_cbs;
    }
    return o;
}

}  // namespace line
}  // namespace hr
