// cone.hpp -- detector-cone splitting of one elastic event on the
// screened-Rutherford law (the reference's cone_split_event,
// src/pyrite/montecarlo/transport/cone_split.py:223-335, whose analog law is
// a tabulated ELSEPA density; here the law is the transport's own
// p(mu) ~ 1/(alpha + mu)^2, mu = (1 - cos theta)/2, with the exact tail
// G(mu) = 1 - F(mu) = alpha (1 - mu)/(alpha + mu) and inverse
// mu = alpha (1 - eta)/(alpha + eta) of eta = 1 - F).
//
// The event is written once over the law (a template): the screened-
// Rutherford law above, with its exact band, or a tabulated (ELSEPA) law as
// the reference samples it -- rows bracketing ln E, their quantiles mixed in
// ln E -- whose band is bounded before the copy roulette by the two rows'
// exact CDFs at the band edges and found after it by bisection on the mixed
// quantile (cone_split.py:240-295).
#pragma once

#include "elsepa.hpp"
#include "hr_common.hpp"

namespace hr {
namespace cone {

constexpr uint64_t kSplitSalt = 0xC2B2AE3D27D4EB4FULL;  // cone_split.py:57
constexpr uint64_t kCopySalt = 0x165667B19E3779F9ULL;   // cone_split.py:58
constexpr double kConePi = 3.141592653589793;
constexpr double kConeMc2 = 510.99895;  // electron rest energy [keV]

// Key of the elastic event that closed flight `flight` (cone_split.py:126-129).
__host__ __device__ __forceinline__ uint64_t event_key(uint64_t split_key, int64_t flight) {
    return splitmix64(split_key + kGolden * ((uint64_t)flight + 1ULL));
}

struct Copy {
    bool made;
    double ox, oy, oz, w;
};

// Tail of the law above mu, computed without cancellation near F = 1.
__host__ __device__ __forceinline__ double tail(double alpha, double mu) {
    return alpha * (1.0 - mu) / (alpha + mu);
}

// Half-width of the azimuth arc about the incident direction inside the cone
// (cone_split.py:210-220).
__device__ __forceinline__ double arc_half_width(double cos_t, double sin_t, double cos_dn, double sin_dn,
                                                 double cos_c) {
    if (sin_t * sin_dn <= 1e-300) return cos_t * cos_dn >= cos_c ? kConePi : 0.0;
    const double c_star = (cos_c - cos_t * cos_dn) / (sin_t * sin_dn);
    if (c_star <= -1.0) return kConePi;
    if (c_star >= 1.0) return 0.0;
    return acos(c_star);
}

// Screened-Rutherford law of screening alpha: band probability exact from the tail.
struct SRLaw {
    double alpha, g_lo = 0.0, g_hi = 0.0, p = 0.0;
    // upper bound of the band probability (here the probability itself)
    __host__ __device__ double bound(double m_lo, double m_hi) {
        g_lo = tail(alpha, m_lo);
        g_hi = tail(alpha, m_hi);
        p = g_lo - g_hi;
        return p;
    }
    __host__ __device__ double band() { return p; }
    // mu at the band's polar uniform u: eta = 1 - xi uniform on [g_hi, g_lo]
    __host__ __device__ double mu(double u) const {
        const double eta = g_lo - u * p;
        return alpha * (1.0 - eta) / (alpha + eta);
    }
};

// Tabulated law of element row block `el` at kinetic energy E (keV).
struct TabLaw {
    static constexpr int kBisection = 64;  // cone_split.py:59
    const hr_elsepa *t;
    int64_t row = 0;
    double f = 0.0, m_lo = 0.0, m_hi = 0.0, xa = 0.0, p = 0.0;
    __host__ __device__ TabLaw(const hr_elsepa *t_, size_t el, double E) : t(t_) {
        elsepa::energy_rows(*t, el, E, row, f);
    }
    // xi >= min_k F_k(mu_lo) and xi <= max_k F_k(mu_hi) bound the mixture's band
    __host__ __device__ double bound(double lo, double hi) {
        m_lo = lo;
        m_hi = hi;
        const double fa0 = elsepa::cdf_at(*t, row, lo), fa1 = elsepa::cdf_at(*t, row + 1, lo);
        const double fb0 = elsepa::cdf_at(*t, row, hi), fb1 = elsepa::cdf_at(*t, row + 1, hi);
        return fmax(fb0, fb1) - fmin(fa0, fa1);
    }
    // the exact uniform interval [xa, xb] mapping into the band
    __host__ __device__ double band() {
        if (elsepa::mix_mu(*t, row, f, 0.0) >= m_lo) {
            xa = 0.0;
        } else {
            double lo = 0.0, hi = 1.0;
            for (int i = 0; i < kBisection; ++i) {
                const double mid = 0.5 * (lo + hi);
                if (elsepa::mix_mu(*t, row, f, mid) >= m_lo) hi = mid;
                else lo = mid;
            }
            xa = hi;
        }
        double xb;
        if (elsepa::mix_mu(*t, row, f, 1.0) <= m_hi) {
            xb = 1.0;
        } else {
            double lo = 0.0, hi = 1.0;
            for (int i = 0; i < kBisection; ++i) {
                const double mid = 0.5 * (lo + hi);
                if (elsepa::mix_mu(*t, row, f, mid) <= m_hi) lo = mid;
                else hi = mid;
            }
            xb = lo;
        }
        p = xb - xa;
        return p;
    }
    __host__ __device__ double mu(double u) const { return elsepa::mix_mu(*t, row, f, xa + u * p); }
};

// The in-cone copy of one elastic event: incident direction d, detector
// direction n, cone (cos_c, sin_c), roulette target w_t, angle roulette
// u_ref (0: off), the chosen element's law, kinetic energy E [keV], event
// key.  Counter 0 is the copy roulette, 1 the polar draw in the band, 2 the
// azimuth on the arc, 3 the angle roulette.
template <class Law>
__device__ __forceinline__ Copy event_law(double dx, double dy, double dz, double nx, double ny, double nz,
                                          double cos_c, double sin_c, double w_t, double u_ref, Law &law, double E,
                                          uint64_t key);

__device__ __forceinline__ Copy event(double dx, double dy, double dz, double nx, double ny, double nz,
                                      double cos_c, double sin_c, double w_t, double u_ref, double alpha, double E,
                                      uint64_t key) {
    SRLaw law{alpha};
    return event_law(dx, dy, dz, nx, ny, nz, cos_c, sin_c, w_t, u_ref, law, E, key);
}

template <class Law>
__device__ __forceinline__ Copy event_law(double dx, double dy, double dz, double nx, double ny, double nz,
                                          double cos_c, double sin_c, double w_t, double u_ref, Law &law, double E,
                                          uint64_t key) {
    Copy c{false, 0.0, 0.0, 0.0, 0.0};
    double cos_dn = dx * nx + dy * ny + dz * nz;
    cos_dn = fmin(1.0, fmax(-1.0, cos_dn));
    const double sin_dn = sqrt(fmax(0.0, 1.0 - cos_dn * cos_dn));
    const double theta_dn = acos(cos_dn);
    const double theta_c = asin(sin_c);
    const double theta_lo = fmax(0.0, theta_dn - theta_c);
    const double theta_hi = fmin(kConePi, theta_dn + theta_c);
    // mu = sin^2(theta/2) (= (1 - cos theta)/2 without its cancellation)
    const double s_lo = sin(0.5 * theta_lo), s_hi = sin(0.5 * theta_hi);
    const double band_bound = law.bound(s_lo * s_lo, s_hi * s_hi);
    const double arc_bound = sin_dn <= sin_c ? kConePi : asin(sin_c / sin_dn);
    const double bound = band_bound * arc_bound / kConePi;
    if (!(bound > 0.0)) return c;
    const double keep = fmin(1.0, bound / w_t);
    if (stream_uniform(key, 0ULL) >= keep) return c;
    const double p_band = law.band();  // the analog law's exact band probability
    if (!(p_band > 0.0)) return c;
    const double mu_t = law.mu(stream_uniform(key, 1ULL));
    const double cos_t = 1.0 - 2.0 * mu_t;
    const double sin_t = sqrt(fmax(0.0, 1.0 - cos_t * cos_t));
    const double half = arc_half_width(cos_t, sin_t, cos_dn, sin_dn, cos_c);
    if (!(half > 0.0)) return c;
    const double psi = (2.0 * stream_uniform(key, 2ULL) - 1.0) * half;
    // frame about d with e1 toward n
    double e1x, e1y, e1z;
    if (sin_dn > 1e-12) {
        e1x = (nx - cos_dn * dx) / sin_dn;
        e1y = (ny - cos_dn * dy) / sin_dn;
        e1z = (nz - cos_dn * dz) / sin_dn;
    } else {
        const bool use_x = fabs(dx) < 0.9;
        const double rx = use_x ? 1.0 : 0.0, ry = use_x ? 0.0 : 1.0, rz = 0.0;
        e1x = dy * rz - dz * ry;
        e1y = dz * rx - dx * rz;
        e1z = dx * ry - dy * rx;
        const double nr = sqrt(e1x * e1x + e1y * e1y + e1z * e1z);
        e1x /= nr;
        e1y /= nr;
        e1z /= nr;
    }
    const double e2x = dy * e1z - dz * e1y, e2y = dz * e1x - dx * e1z, e2z = dx * e1y - dy * e1x;
    const double ca = sin_t * cos(psi), cb = sin_t * sin(psi);
    double ox = cos_t * dx + ca * e1x + cb * e2x;
    double oy = cos_t * dy + ca * e1y + cb * e2y;
    double oz = cos_t * dz + ca * e1z + cb * e2z;
    const double on = sqrt(ox * ox + oy * oy + oz * oz);
    ox /= on;
    oy /= on;
    oz /= on;
    double w = p_band * half / kConePi / keep;
    if (u_ref > 0.0) {
        // angle roulette: the copy's direction is event state, not its future
        const double cos_on = fmin(1.0, fmax(-1.0, ox * nx + oy * ny + oz * nz));
        const double u = (1.0 + E / kConeMc2) * acos(cos_on) / u_ref;
        const double keep_u = fmin(1.0, w / (w_t * fmax(1.0, u * u)));
        if (stream_uniform(key, 3ULL) >= keep_u) return c;
        w /= keep_u;
    }
    c.made = true;
    c.ox = ox;
    c.oy = oy;
    c.oz = oz;
    c.w = w;
    return c;
}

}  // namespace cone
}  // namespace hr
