// shell.hpp -- hard inelastic collisions and the soft-loss sampler of the
// shell soft/hard inelastic mode.
//
// Algorithm of PyRITE's shell mode (src/pyrite/montecarlo/transport/
// hard_inelastic.py:120-554; device twin _jit_shell_device.py:143-306): a
// layer's hard channels (one shell's distant-longitudinal, distant-transverse
// or close branch each) carry tabulated rates; a hard event picks a channel by
// its cumulative rate, inverts the channel's restricted loss CDF by safeguarded
// Newton iteration (bisection fallback), and deflects the primary by the
// channel's recoil kinematics (PENELOPE-2024 Eqs. 3.76-3.138).  Between hard
// events the soft losses follow the two-moment sampler of PENELOPE-2024
// Eqs. 4.54-4.63 on the straggling key domain.  Electrons only (branches
// 0 distant longitudinal, 1 distant transverse, 2 close Moller).
#pragma once
#include <math.h>

#include "hr_common.hpp"

namespace hr {
namespace shell {

constexpr double kMc2eV = 510998.95;
constexpr double kTruncWidth = 1.015387;  // PENELOPE-2024 Eq. 4.59
constexpr int kNewtonSteps = 200;
constexpr double kNewtonRtol = 8.881784197001252e-16;  // 2^-50
constexpr double kTwoPi = 6.283185307179586;

struct Bounds {
    double lower, upper, peak_end;
};

// One channel's hard loss interval [eV]: close losses from max(U or W, W_c) to
// (E+U)/2; bound distant losses from max(U, W_c) to min(3W - 2U clipped at E, W_max).
__host__ __device__ inline Bounds loss_bounds(double E, double U, double W, int branch, double Wc) {
    const double w_max = 0.5 * (E + U);
    Bounds b;
    if (branch == 2) {
        const double base = U > 0.0 ? U : W;
        b.lower = base > Wc ? base : Wc;
        b.upper = w_max;
        b.peak_end = 0.0;
        return b;
    }
    if (U == 0.0) {
        b.lower = b.upper = W;
        b.peak_end = 0.0;
        return b;
    }
    double pe = 3.0 * W - 2.0 * U;
    if (E <= pe) pe = E;
    b.upper = pe < w_max ? pe : w_max;
    b.lower = U > Wc ? U : Wc;
    b.peak_end = pe;
    return b;
}

// J(w) and J'(w) of a loss law: kind 0 distant triangle, 1 Moller.
__host__ __device__ inline void cdf_density(int kind, double E, double prime, double peak_end, double lower,
                                            double w, double &value, double &density) {
    if (kind == 1) {
        const double g = E / (E + kMc2eV);
        const double a = g * g;
        const double rest = prime - w;
        const double c1 = (1.0 - a) / prime;
        value = -1.0 / w + 1.0 / rest + c1 * log(rest / w) + a * w / (prime * prime);
        density = 1.0 / (w * w) + 1.0 / (rest * rest) - c1 * (1.0 / rest + 1.0 / w) + a / (prime * prime);
        return;
    }
    value = peak_end * log(w / lower) - (w - lower);
    density = peak_end / w - 1.0;
}

// Inverse of one channel's restricted hard loss CDF at u [eV].
__host__ __device__ inline double sample_transfer(double E, double U, double W, int branch, double Wc, double u) {
    const Bounds b = loss_bounds(E, U, W, branch, Wc);
    if (branch != 2 && U == 0.0) return W;
    if (!(b.upper > b.lower)) return b.lower;
    const double prime = E + U;
    const int kind = branch == 2 ? 1 : 0;
    double base, base_d, top, top_d;
    cdf_density(kind, E, prime, b.peak_end, b.lower, b.lower, base, base_d);
    cdf_density(kind, E, prime, b.peak_end, b.lower, b.upper, top, top_d);
    const double target = u * (top - base);
    if (!(target > 0.0)) return b.lower;
    double w = b.lower + target / base_d;
    if (kind != 0) {
        const double inv = 1.0 / b.lower - target;
        w = inv > 0.0 ? 1.0 / inv : b.upper;
    }
    double lo = b.lower, hi = b.upper;
    if (!(w > lo && w < hi)) w = 0.5 * (lo + hi);
    for (int k = 0; k < kNewtonSteps; ++k) {
        double v, d;
        cdf_density(kind, E, prime, b.peak_end, b.lower, w, v, d);
        const double res = v - base - target;
        if (res == 0.0) return w;
        if (res < 0.0) lo = w;
        else hi = w;
        double wn = w - res / d;
        if (!(wn > lo && wn < hi)) wn = 0.5 * (lo + hi);
        if (fabs(wn - w) <= kNewtonRtol * w || !(wn > lo && wn < hi)) return wn;
        w = wn;
    }
    return w;
}

// Minimum momentum-transfer recoil energy Q_- [eV] for a loss W at kinetic E.
__host__ __device__ inline double qmin(double E, double W) {
    const double p0 = sqrt(E * (E + 2.0 * kMc2eV));
    const double rem = E - W;
    const double p1 = sqrt(rem * (rem + 2.0 * kMc2eV));
    const double dp = W * (2.0 * (E + kMc2eV) - W) / (p0 + p1);
    return dp * dp / (sqrt(dp * dp + kMc2eV * kMc2eV) + kMc2eV);
}

// Distant longitudinal recoil Q at uniform u, with the loss it goes with (the
// modified resonance near threshold); false when the recoil interval is empty.
__host__ __device__ inline bool distant_recoil(double E, double U, double W, double u, double &Q, double &res) {
    const double two = 2.0 * kMc2eV;
    double q_up;
    if (U > 0.0) {
        const double full = 3.0 * W - 2.0 * U;
        if (E > full) {
            res = W;
            q_up = U;
        } else {
            res = (E + 2.0 * U) / 3.0;
            q_up = U * E / full;
        }
    } else {
        res = W;
        q_up = W;
    }
    if (!(res < E)) return false;
    const double q_lo = qmin(E, res);
    if (!(0.0 < q_lo && q_lo < q_up)) return false;
    const double ll = log(q_lo / (q_lo + two));
    const double lu = log(q_up / (q_up + two));
    const double lr = (1.0 - u) * ll + u * lu;
    Q = two / expm1(-lr);
    return true;
}

// Primary polar cosine after a hard collision, relative to the flight.
__host__ __device__ inline double primary_cosine(double E, double U, double W, int branch, double W_eV, double u) {
    const double two = 2.0 * kMc2eV;
    if (branch == 1) return 1.0;
    if (branch == 2) {
        const double rem = E - W_eV;
        return sqrt(rem / E * (E + two) / (rem + two));
    }
    double Q, res;
    if (!distant_recoil(E, U, W, u, Q, res)) return 1.0;
    const double p0 = E * (E + two);
    const double rem = E - res;
    const double p1 = rem * (rem + two);
    const double c = (p0 + p1 - Q * (Q + two)) / (2.0 * sqrt(p0 * p1));
    return c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
}

// Secondary polar cosine of a hard collision, relative to the flight.
__host__ __device__ inline double secondary_cosine(double E, double U, double W, int branch, double W_eV,
                                                   double u) {
    const double two = 2.0 * kMc2eV;
    if (branch == 1) return 0.5;
    if (branch == 2) return sqrt(W_eV / E * (E + two) / (W_eV + two));
    double Q, res;
    if (!distant_recoil(E, U, W, u, Q, res)) return 1.0;
    const double p0 = E * (E + two);
    const double rem = E - res;
    const double p1 = rem * (rem + two);
    const double q2 = Q * (Q + two);
    const double c = (p0 + q2 - p1) / (2.0 * sqrt(p0 * q2));
    return c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
}

// Soft energy loss [keV] with mean `mean` and variance `var`: truncated
// Gaussian (case I), uniform (case II) or a delta at zero plus a uniform
// (case III); draws from (key, counter).
__host__ __device__ inline double soft_loss(double mean, double var, uint64_t key, uint64_t &counter) {
    if (mean <= 0.0) return 0.0;
    if (var <= 0.0) return mean;
    const double sigma = sqrt(var);
    const double m2 = mean * mean;
    if (m2 > 9.0 * var) {
        const double width = kTruncWidth * sigma;
        for (;;) {
            const double u1 = stream_uniform(key, counter);
            counter += 1ULL;
            const double u2 = stream_uniform(key, counter);
            counter += 1ULL;
            const double z = sqrt(-2.0 * log(1.0 - u1)) * cos(kTwoPi * u2);
            const double off = width * z;
            if (fabs(off) < 3.0 * sigma) return mean + off;
        }
    }
    const double u = stream_uniform(key, counter);
    counter += 1ULL;
    if (m2 > 3.0 * var) {
        const double half = sqrt(3.0) * sigma;
        return mean - half + 2.0 * half * u;
    }
    const double weight = (3.0 * var - m2) / (3.0 * var + 3.0 * m2);
    const double top = (3.0 * var + 3.0 * m2) / (2.0 * mean);
    if (u < weight) return 0.0;
    return top * (u - weight) / (1.0 - weight);
}

}  // namespace shell
}  // namespace hr
