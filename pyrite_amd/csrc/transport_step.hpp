// transport_step.hpp -- one step of PyRITE's per-electron LUT core on gfx950.
//
// Physics and draw order follow src/pyrite/montecarlo/transport/cores.py:365-1250
// and its CUDA twin _jit_lut_kernel.py:64-574: screened-Rutherford elastic
// flights from a uniform-ln(E) LUT, spliced Joy-Luo/Berger-Seltzer stopping
// with the midpoint rule, Urban straggling on a per-(flight, substep) key
// domain, slab or prism geometry, and the cap/slot row contract.  Shared by the
// analog history kernel (transport.hip) and the round-based kernel of the
// learned variance reduction (vr.hip).
#pragma once
#include <math.h>

#include "cone.hpp"
#include "elsepa.hpp"
#include "hr_common.hpp"
#include "radiative.hpp"
#include "shell.hpp"

namespace hr {
namespace {

constexpr double kMc2 = 510.99895;
constexpr double kBs = 1.535e-6;
constexpr double kLn2 = 0.6931471805599453;
constexpr double kE0 = 1.0e-2;      // Urban E_0 [keV]          straggling.py:117
constexpr double kE2PerZ2 = 1.0e-2; // Urban E_2 / Z^2 [keV]    straggling.py:118
constexpr double kRate = 0.55;      // Urban r                  straggling.py:119
constexpr double kChunk = 64.0;     // Poisson chunk mean       straggling.py:125
constexpr double kPi = 3.141592653589793;

// Row event codes of montecarlo/transport/events.py:36-43 (3 is the groove
// surface, which this path never emits).
enum : int8_t { EV_ELASTIC = 0, EV_SUBSTEP = 1, EV_LAYER = 2, EV_EXIT_TOP = 4, EV_EXIT_BOTTOM = 5,
                EV_EXIT_SIDE = 6, EV_CUTOFF = 7, EV_HARD_INELASTIC = 8, EV_HARD_RADIATIVE = 9 };
// Salt of the hard-collision key: splitmix64(stream_key ^ salt) (hard_inelastic.py:40,77-80).
constexpr uint64_t kHardSalt = 0x3C6EF372FE94F82BULL;
enum : int8_t { EXIT_CUTOFF = 0, EXIT_BACK = 1, EXIT_TRANS = 2, EXIT_SIDE = 3, EXIT_STEP_LIMITED = 4,
                EXIT_NOT_ENTERED = 5, EXIT_SUSPENDED = 6 };

struct LutPos {
    int i;
    double f;
};

// lut.py:94-110: upper clamp tested first so a non-finite coordinate lands low.
__device__ __forceinline__ LutPos lut_pos(double E, double log_E_min, double inv_dlogE, int n_energy) {
    const double x = (log(E) - log_E_min) * inv_dlogE;
    const int last = n_energy - 1;
    LutPos p;
    if (x >= (double)last) {
        p.i = last - 1;
        p.f = 1.0;
    } else if (x > 0.0) {
        p.i = (int)x;
        p.f = x - (double)p.i;
    } else {
        p.i = 0;
        p.f = 0.0;
    }
    return p;
}

__device__ __forceinline__ double lerp(const double *__restrict__ t, LutPos p) {
    const double lo = t[p.i];
    return lo + p.f * (t[p.i + 1] - lo);
}

// Mixed precision (transport/_jit_precision.py:1-40): the table query ln E and
// the interpolation run in FP32 on FP32 table copies; the result is widened.
__device__ __forceinline__ LutPos lut_pos_f(double E, double log_E_min, double inv_dlogE, int n_energy) {
    const float x = (logf((float)E) - (float)log_E_min) * (float)inv_dlogE;
    const int last = n_energy - 1;
    LutPos p;
    if (x >= (float)last) {
        p.i = last - 1;
        p.f = 1.0;
    } else if (x > 0.0f) {
        p.i = (int)x;
        p.f = (double)(x - (float)p.i);
    } else {
        p.i = 0;
        p.f = 0.0;
    }
    return p;
}

__device__ __forceinline__ double lerp_f(const float *__restrict__ t, LutPos p) {
    const float lo = t[p.i];
    return (double)(lo + (float)p.f * (t[p.i + 1] - lo));
}

// Table access of one precision: kMixed selects the FP32 copies and FP32 queries.
template <bool kMixed>
struct Lut {
    const hr_transport_args &a;
    __device__ __forceinline__ LutPos pos(double E) const {
        return kMixed ? lut_pos_f(E, a.lut_log_E_min, a.lut_inv_dlogE, a.lut_n_energy)
                      : lut_pos(E, a.lut_log_E_min, a.lut_inv_dlogE, a.lut_n_energy);
    }
    __device__ __forceinline__ double rate(size_t off, LutPos p) const {
        return kMixed ? lerp_f(a.lut_total_rate_f + off, p) : lerp(a.lut_total_rate + off, p);
    }
    __device__ __forceinline__ double dEds(size_t off, LutPos p) const {
        return kMixed ? lerp_f(a.lut_dEds_f + off, p) : lerp(a.lut_dEds + off, p);
    }
    __device__ __forceinline__ double inv_beta(LutPos p) const {
        return kMixed ? lerp_f(a.lut_inv_beta_f, p) : lerp(a.lut_inv_beta, p);
    }
    __device__ __forceinline__ double cdf(size_t off, LutPos p) const {
        return kMixed ? lerp_f(a.lut_cdf_f + off, p) : lerp(a.lut_cdf + off, p);
    }
    __device__ __forceinline__ double alpha(size_t off, LutPos p) const {
        return kMixed ? lerp_f(a.lut_alpha_f + off, p) : lerp(a.lut_alpha + off, p);
    }
};

// straggling.py:305-329, one element of the spliced stopping, delta = 0.
__device__ double element_dEds(double J, double k, double coeff, double E_cross, double E) {
    const double tau = E / kMc2;
    const double gamma = 1.0 + tau;
    const double beta_sq = 1.0 - 1.0 / (gamma * gamma);
    if (E < E_cross) return -7.85e-4 / E * coeff * log(1.166 * (E + k * J) / J);
    const double f_minus = 1.0 - beta_sq + (tau * tau / 8.0 - (2.0 * tau + 1.0) * kLn2) / (gamma * gamma);
    const double I_rel = J / kMc2;
    return -kBs / beta_sq * coeff * (log(tau * tau * (tau + 2.0) / (2.0 * I_rel * I_rel)) + f_minus - 0.0);
}

// straggling.py:229-263: inverse-CDF Poisson in bounded chunks, one draw each.
__device__ int64_t urban_poisson(double lam, uint64_t key, uint64_t &counter) {
    // Same as lam <= 0 for every finite mean; also turns a NaN or absurd mean
    // (only reachable from corrupted state) into no draw instead of an
    // unbounded chunk loop.
    if (!(lam > 0.0) || !(lam <= 1.0e12)) return 0;
    const int64_t chunks = (int64_t)ceil(lam / kChunk);
    const double chunk_lam = lam / (double)chunks;
    const double p0 = exp(-chunk_lam);
    int64_t total = 0;
    for (int64_t c = 0; c < chunks; ++c) {
        const double u = stream_uniform(key, counter);
        counter += 1ULL;
        double p = p0, cdf = p0;
        int64_t k = 0;
        while (u >= cdf) {
            k += 1;
            p = p * chunk_lam / (double)k;
            const double next = cdf + p;
            if (next <= cdf) break;
            cdf = next;
        }
        total += k;
    }
    return total;
}

// straggling.py:131-201 and 278-363: Bragg-additive Urban loss of one layer row.
__device__ double urban_compound(const double *__restrict__ Zr, const double *__restrict__ Jr,
                                 const double *__restrict__ kr, const double *__restrict__ cr,
                                 const double *__restrict__ xr, int n_el, double E, double s,
                                 uint64_t key) {
    uint64_t counter = 0;
    double loss = 0.0;
    const double tau = E / kMc2;
    const double gamma = 1.0 + tau;
    const double beta_sq = 1.0 - 1.0 / (gamma * gamma);
    const double two = 2.0 * kMc2 * tau * (tau + 2.0);
    const double T_up = 0.5 * E;
    for (int el = 0; el < n_el; ++el) {
        const double Z = Zr[el], J = Jr[el];
        const double C = -element_dEds(J, kr[el], cr[el], xr[el], E) * 1.0;
        bool valid = T_up > kE0 && C > 0.0;
        double L_I = 0.0;
        if (valid) {
            L_I = log(two / J) - beta_sq;
            valid = L_I > 0.0;
        }
        double dE = 0.0;
        if (!valid || !(s > 0.0)) {
            dE = C * s;
        } else {
            const double e2 = kE2PerZ2 * Z * Z;
            double f2 = Z > 2.0 ? 2.0 / Z : 1.0;
            double f1 = 1.0, e1 = J;
            bool resolved = false;
            if (Z > 2.0 && e2 < T_up && log(two / e2) - beta_sq > 0.0) {
                const double f1c = 1.0 - f2;
                const double e1c = exp((log(J) - f2 * log(e2)) / f1c);
                if (log(two / e1c) - beta_sq > 0.0) {
                    f1 = f1c;
                    e1 = e1c;
                    resolved = true;
                }
            }
            if (!resolved) {
                f1 = 1.0;
                e1 = J;
                f2 = 0.0;
            }
            const double soft = C * (1.0 - kRate) / L_I;
            const double s1 = soft * (f1 / e1) * (log(two / e1) - beta_sq);
            const double s2 = f2 > 0.0 ? soft * (f2 / e2) * (log(two / e2) - beta_sq) : 0.0;
            const double s3 = C * kRate * (T_up - kE0) / (kE0 * T_up * log(T_up / kE0));
            const int64_t n1 = urban_poisson(s1 * s, key, counter);
            const int64_t n2 = urban_poisson(s2 * s, key, counter);
            const int64_t n3 = urban_poisson(s3 * s, key, counter);
            dE = (double)n1 * e1 + (double)n2 * e2;
            for (int64_t q = 0; q < n3; ++q) {
                const double u = stream_uniform(key, counter);
                counter += 1ULL;
                dE += kE0 / (1.0 - u * (T_up - kE0) / T_up);
            }
        }
        loss += dE;
    }
    return loss;
}

// e indexes the state arrays and the slot rows; id is the history's global
// electron id written to seg_id (equal to e except under continuation, where e
// is a lane of the active population).
struct Lane {
    double px, py, pz, dx, dy, dz, E, clock, tau, E_cut, stragg;
    uint64_t key, draws;
    int64_t e, id, nseg, flight, substep;
    int32_t steps;
    // hard-event modes: open optical depths (< 0: draw at the next step), keys, counters
    double tau_h, tau_r;
    uint64_t hkey, rkey, hdraws, rdraws;
    // detector-cone splitting: the electron's split-decision key
    uint64_t skey;
};

__device__ __forceinline__ void load_lane(const hr_transport_args &a, int64_t e, Lane &s) {
    s.e = e;
    s.id = e;
    s.px = a.pos[3 * e];
    s.py = a.pos[3 * e + 1];
    s.pz = a.pos[3 * e + 2];
    s.dx = a.dirs[3 * e];
    s.dy = a.dirs[3 * e + 1];
    s.dz = a.dirs[3 * e + 2];
    s.E = a.E_keV[e];
    s.clock = a.clock[e];
    s.E_cut = a.E_cut[e];
    s.key = a.stream_key[e];
    s.tau = -1.0;
    s.stragg = 0.0;
    s.draws = 0;
    s.nseg = 0;
    s.flight = 0;
    s.substep = 0;
    s.steps = 0;
    s.tau_h = -1.0;
    s.tau_r = -1.0;
    s.hkey = a.hard.shell_on ? splitmix64(s.key ^ kHardSalt) : 0ULL;
    s.rkey = a.hard.rad_on ? a.hard.rad_key[e] : 0ULL;
    s.hdraws = 0;
    s.rdraws = 0;
    s.skey = a.cone.on ? a.cone.split_key[e] : 0ULL;
}

__device__ __forceinline__ rad::Tables rad_tables(const hr_hard_events &h, int max_el) {
    rad::Tables t;
    t.incident = h.rad_incident;
    t.nominal = h.rad_nominal;
    t.top = h.rad_top;
    t.chi = h.rad_chi;
    t.z = h.rad_z;
    t.ncm3 = h.rad_ncm3;
    t.nt = h.rad_nt;
    t.nel = nullptr;
    t.plasma = h.rad_plasma;
    t.max_el = max_el;
    t.max_t = h.rad_max_t;
    t.n_red = h.rad_n_red;
    t.cutoff_eV = h.rad_cutoff_eV;
    return t;
}

// core_geometry.py:16-36: rotate the lane's direction by polar cosine cos_t
// and azimuth phi about itself (kMixed: FP32 sqrt(1 - cos^2) and sin/cos).
template <bool kMixed>
__device__ __forceinline__ void rotate_lane(Lane &s, double cos_t, double phi) {
    const double dx = s.dx, dy = s.dy, dz = s.dz;
    const double s2 = 1.0 - cos_t * cos_t;
    double sin_t, sin_phi, cos_phi;
    if (kMixed) {
        const float s2f = (float)s2;
        sin_t = (double)sqrtf(s2f > 0.0f ? s2f : 0.0f);
        float sf, cf;
        sincosf((float)phi, &sf, &cf);
        sin_phi = (double)sf;
        cos_phi = (double)cf;
    } else {
        sin_t = sqrt(s2 > 0.0 ? s2 : 0.0);
        sin_phi = sin(phi);
        cos_phi = cos(phi);
    }
    double refx, refy;
    if (fabs(dx) < 0.9) { refx = 1.0; refy = 0.0; } else { refx = 0.0; refy = 1.0; }
    double ux = -dz * refy, uy = dz * refx, uz = dx * refy - dy * refx;
    const double u_mag = sqrt(ux * ux + uy * uy + uz * uz);
    ux = ux / u_mag;
    uy = uy / u_mag;
    uz = uz / u_mag;
    const double wx = dy * uz - dz * uy, wy = dz * ux - dx * uz, wz = dx * uy - dy * ux;
    const double ra = sin_t * cos_phi, rb = sin_t * sin_phi;
    const double ox = cos_t * dx + ra * ux + rb * wx;
    const double oy = cos_t * dy + ra * uy + rb * wy;
    const double oz = cos_t * dz + ra * uz + rb * wz;
    const double mag = sqrt(ox * ox + oy * oy + oz * oz);
    s.dx = ox / mag;
    s.dy = oy / mag;
    s.dz = oz / mag;
}

__device__ __forceinline__ void store_lane(const hr_transport_args &a, const Lane &s, int8_t code) {
    const int64_t e = s.e, i = e - a.e_start;
    a.pos[3 * e] = s.px;
    a.pos[3 * e + 1] = s.py;
    a.pos[3 * e + 2] = s.pz;
    a.dirs[3 * e] = s.dx;
    a.dirs[3 * e + 1] = s.dy;
    a.dirs[3 * e + 2] = s.dz;
    a.E_keV[e] = s.E;
    a.clock[e] = s.clock;
    if (a.straggle_on) a.stragg_dE[e] = s.stragg;
    a.seg_count[i] = s.nseg;
    a.exit_code[i] = code;
}

// One loop iteration of cores.py:408-1239 (per_electron, lut).  Returns the
// exit code when the history ends, or -1 when it continues.  kRows = false
// never writes rows (tally-only callers), whatever a.cap says.  kMixed is the
// reference's "mixed" transport precision: FP32 table lookups, query
// logarithms, sqrt(1 - cos^2) and azimuth sin/cos; every accumulated quantity,
// free-path draw and the straggling stay FP64.
// kCone: 0 no splitting; 1 copies on the screened-Rutherford law; 2 copies on
// the tabulated (ELSEPA) law, a variant of its own so the bisection's
// registers stay out of the other kernels.
template <bool kRows = true, bool kMixed = false, bool kHard = false, int kCone = 0>
__device__ int step_lane(const hr_transport_args &a, Lane &s) {
    const Lut<kMixed> lut{a};
    const int nE = a.lut_n_energy;
    const int max_el = a.max_el;
    int L = 0;
    if (a.n_layers != 1) {
        int lo = 0, hi = a.n_layers - 1;  // searchsorted(side="right")
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (a.internal_bounds[mid] <= s.pz) lo = mid + 1;
            else hi = mid;
        }
        L = lo;
    }
    const int n_el = a.L_nel[L];
    const double z_top = a.L_top[L], z_bot = a.L_bot[L];
    const double E = s.E;
    const LutPos p = lut.pos(E);
    const size_t row_off = (size_t)L * nE;
    const double total_rate = lut.rate(row_off, p);
    const double lam = 1e8 / total_rate;
    if (s.tau < 0.0) {
        s.tau = -log(stream_uniform(s.key, s.draws));
        s.draws += 1ULL;
    }
    double step = s.tau * lam;
    // Hard-event modes: optical depths on their own keys against the row-start
    // rates; a radiative proposal that wins clears a hard-collision one.
    bool hard_j = false, rad_j = false;
    double mu_hard = 0.0, mu_rad = 0.0, rad_dEds = 0.0;
    if constexpr (kHard) {
        const hr_hard_events &h = a.hard;
        if (h.shell_on) {
            mu_hard = lerp(h.lut_sh_rate + row_off, p);
            if (s.tau_h < 0.0) {
                s.tau_h = -log(stream_uniform(s.hkey, s.hdraws));
                s.hdraws += 1ULL;
            }
            if (mu_hard > 0.0 && s.tau_h / mu_hard < step) {
                step = s.tau_h / mu_hard;
                hard_j = true;
            }
        }
        if (h.rad_on) {
            rad_dEds = lerp(h.lut_rad_soft + row_off, p);
            mu_rad = lerp(h.lut_rad_rate + row_off, p);
            if (s.tau_r < 0.0) {
                s.tau_r = -log(stream_uniform(s.rkey, s.rdraws));
                s.rdraws += 1ULL;
            }
            if (mu_rad > 0.0 && s.tau_r / mu_rad < step) {
                step = s.tau_r / mu_rad;
                rad_j = true;
                hard_j = false;
            }
        }
    }
    const double dx = s.dx, dy = s.dy, dz = s.dz, px = s.px, py = s.py, pz = s.pz;
    bool up = false, dn = false, side = false;
    if (a.finite_footprint) {
        double best = INFINITY;
        int face = -1;
        const double hw = 0.5 * a.width_ang, hh = 0.5 * a.height_ang;
        double t;
        if (dx < 0.0) { t = (-hw - px) / dx; if (t > 0.0 && t < best) { best = t; face = 0; } }
        if (dx > 0.0) { t = (hw - px) / dx; if (t > 0.0 && t < best) { best = t; face = 1; } }
        if (dy < 0.0) { t = (-hh - py) / dy; if (t > 0.0 && t < best) { best = t; face = 2; } }
        if (dy > 0.0) { t = (hh - py) / dy; if (t > 0.0 && t < best) { best = t; face = 3; } }
        if (dz < 0.0) { t = (z_top - pz) / dz; if (t > 0.0 && t < best) { best = t; face = 4; } }
        if (dz > 0.0) { t = (z_bot - pz) / dz; if (t > 0.0 && t < best) { best = t; face = 5; } }
        if (step > best) {
            step = best;
            up = face == 4;
            dn = face == 5;
            side = face >= 0 && face <= 3;
        }
    } else if (dz < 0.0) {
        const double sb = (pz - z_top) / (-dz);
        if (step > sb) { step = sb; up = true; }
    } else if (dz > 0.0) {
        const double sb = (z_bot - pz) / dz;
        if (step > sb) { step = sb; dn = true; }
    }
    bool exit_top = up && z_top <= 0.0;
    bool exit_bot = dn && z_bot >= a.z_total;
    // soft radiative stopping joins the collision stopping as a mean loss
    const double dEds = kHard ? lut.dEds(row_off, p) - rad_dEds : lut.dEds(row_off, p);
    const double inv_beta_j = lut.inv_beta(p);
    bool cutoff = false, limited = false;
    bool geometry_event = up || dn || side;
    double E_end, t_end;
    const bool midpoint = a.energy_model == 1;
    if (a.straggle_on) {
        if (a.max_dE_frac > 0.0) {
            const double se = a.max_dE_frac * E / (-dEds);
            if (se < step) {
                step = se;
                limited = true;
                up = dn = exit_top = exit_bot = side = false;
                geometry_event = false;
            }
        }
        const uint64_t ukey = splitmix64(s.key ^ kUrbanSalt);
        const uint64_t combined = ((uint64_t)s.flight << 32) + (uint64_t)s.substep;
        const uint64_t fkey = splitmix64(ukey + kGolden * combined);
        const size_t row = (size_t)L * max_el;
        double loss;
        if (kHard && a.hard.shell_on) {
            // soft collision losses only: the two-moment sampler replaces Urban's
            const double omega2 = lerp(a.hard.lut_sh_omega2 + row_off, p);
            uint64_t counter = 0;
            loss = shell::soft_loss(-(dEds + rad_dEds) * step, omega2 * step, fkey, counter);
        } else {
            loss = urban_compound(a.L_Z + row, a.L_J + row, a.L_k + row, a.L_coeff + row, a.L_E_cross + row,
                                  n_el, E, step, fkey);
        }
        if (kHard) loss += rad_dEds * step;
        s.stragg += loss;
        const double delta_cut = E - s.E_cut;
        if (loss > delta_cut || (loss == delta_cut && !geometry_event)) {
            step = loss > 0.0 ? step * (delta_cut / loss) : 0.0;
            cutoff = true;
            limited = false;
            up = dn = exit_top = exit_bot = side = false;
            E_end = s.E_cut;
        } else {
            E_end = E - loss;
        }
        if (midpoint) {
            const LutPos c = lut.pos(0.5 * (E + E_end));
            t_end = s.clock + step * lut.inv_beta(c);
        } else {
            t_end = s.clock + step * inv_beta_j;
        }
    } else {
        double cut_dist;
        if (midpoint) {
            const LutPos c = lut.pos(0.5 * (E + s.E_cut));
            double rate_c = lut.dEds(row_off, c);
            if (kHard && a.hard.rad_on) rate_c -= lerp(a.hard.lut_rad_soft + row_off, c);
            cut_dist = (s.E_cut - E) / rate_c;
        } else {
            cut_dist = (s.E_cut - E) / dEds;
        }
        if (cut_dist < step || (cut_dist == step && !geometry_event)) {
            step = cut_dist;
            cutoff = true;
            up = dn = exit_top = exit_bot = side = false;
        }
        if (a.max_dE_frac > 0.0 && !cutoff) {
            const double se = a.max_dE_frac * E / (-dEds);
            if (se < step) {
                step = se;
                limited = true;
                up = dn = exit_top = exit_bot = side = false;
            }
        }
        if (midpoint) {
            if (cutoff) {
                E_end = s.E_cut;
            } else {
                const double E_pred = E + dEds * step;
                const LutPos m = lut.pos(0.5 * (E + E_pred));
                double rate_m = lut.dEds(row_off, m);
                if (kHard && a.hard.rad_on) rate_m -= lerp(a.hard.lut_rad_soft + row_off, m);
                E_end = E + step * rate_m;
            }
            const LutPos c = lut.pos(0.5 * (E + E_end));
            t_end = s.clock + step * lut.inv_beta(c);
        } else {
            E_end = cutoff ? s.E_cut : E + dEds * step;
            t_end = s.clock + step * inv_beta_j;
        }
    }
    int8_t event = EV_ELASTIC;
    if (limited) event = EV_SUBSTEP;
    else if (cutoff) event = EV_CUTOFF;
    else if (exit_top) event = EV_EXIT_TOP;
    else if (exit_bot) event = EV_EXIT_BOTTOM;
    else if (side) event = EV_EXIT_SIDE;
    else if (up || dn) event = EV_LAYER;
    else if (kHard && hard_j) event = EV_HARD_INELASTIC;
    else if (kHard && rad_j) event = EV_HARD_RADIATIVE;

    // Hard events are sampled before the row is written (its columns carry them).
    double hard_W = 0.0, hard_cos = 1.0, hard_phi = 0.0, rad_k = 0.0;
    int16_t hard_code = -1, rad_Z = 0;
    if constexpr (kHard) {
        const hr_hard_events &h = a.hard;
        if (event == EV_HARD_INELASTIC) {
            // channel, transfer, recoil, azimuth: four draws at the row-start energy
            const double u_ch = stream_uniform(s.hkey, s.hdraws);
            const double u_w = stream_uniform(s.hkey, s.hdraws + 1ULL);
            const double u_q = stream_uniform(s.hkey, s.hdraws + 2ULL);
            const double u_phi = stream_uniform(s.hkey, s.hdraws + 3ULL);
            s.hdraws += 4ULL;
            const double target = u_ch * mu_hard;
            const int nch = h.sh_nch[L];
            int ch = nch - 1;
            double cum = 0.0;
            for (int k = 0; k < nch; ++k) {
                cum += lerp(h.lut_sh_ch_rate + ((size_t)L * h.sh_max_ch + k) * nE, p);
                if (cum > target) {
                    ch = k;
                    break;
                }
            }
            const int c = L * h.sh_max_ch + ch;
            const double E_eV = E * 1.0e3;
            const double W_eV = shell::sample_transfer(E_eV, h.sh_U[c], h.sh_W[c], h.sh_branch[c], h.sh_cutoff_eV, u_w);
            hard_W = W_eV * 1.0e-3;
            hard_code = h.sh_code[c];
            if (E_end - hard_W <= s.E_cut) {
                event = EV_CUTOFF;  // absorbed at the collision point: a terminal row
                cutoff = true;
            } else {
                hard_cos = shell::primary_cosine(E_eV, h.sh_U[c], h.sh_W[c], h.sh_branch[c], W_eV, u_q);
                hard_phi = 2.0 * kPi * u_phi;
            }
        } else if (event == EV_HARD_RADIATIVE) {
            // element and photon from the exact moments at the row-end energy
            const rad::Tables rt = rad_tables(h, max_el);
            const double E_end_eV = E_end * 1.0e3;
            const double plasma = h.rad_plasma[L];
            const double u_el = stream_uniform(s.rkey, s.rdraws);
            const double u_k = stream_uniform(s.rkey, s.rdraws + 1ULL);
            s.rdraws += 2ULL;
            double rates[8];
            double mu_event = 0.0;
            const int n_rad = n_el < 8 ? n_el : 8;
            for (int k = 0; k < n_rad; ++k) {
                const int slot = L * max_el + k;
                rates[k] = rt.ncm3[slot] * rad::kCmPerAng *
                           rad::moment(rad::HARD_RATE, rt, slot, E_end_eV, rt.cutoff_eV, plasma);
                mu_event += rates[k];
            }
            const double target = u_el * mu_event;
            int sel = n_rad - 1;
            double cum = 0.0;
            for (int k = 0; k < n_rad; ++k) {
                cum += rates[k];
                if (cum > target) {
                    sel = k;
                    break;
                }
            }
            const int slot = L * max_el + sel;
            rad_k = rad::sample_photon(rt, slot, E_end_eV, u_k, plasma);
            rad_Z = (int16_t)rt.z[slot];
            if (E_end - rad_k * 1.0e-3 <= s.E_cut) {
                event = EV_CUTOFF;
                cutoff = true;
            }
        }
    }

    if (kRows && s.nseg < a.cap) {
        const int64_t slot = (s.e - a.e_start) * a.cap + s.nseg;
        a.seg_dir[3 * slot] = dx;
        a.seg_dir[3 * slot + 1] = dy;
        a.seg_dir[3 * slot + 2] = dz;
        a.seg_mid[3 * slot] = px + 0.5 * step * dx;
        a.seg_mid[3 * slot + 1] = py + 0.5 * step * dy;
        a.seg_mid[3 * slot + 2] = pz + 0.5 * step * dz;
        a.seg_len[slot] = step;
        a.seg_E[slot] = E;
        a.seg_t0[slot] = s.clock;
        a.seg_id[slot] = s.id;
        a.seg_lay[slot] = (int16_t)L;
        if (midpoint) {
            a.seg_E_end[slot] = E_end;
            a.seg_t_end[slot] = t_end;
            a.seg_flight[slot] = s.flight;
            a.seg_substep[slot] = s.substep;
            a.seg_event[slot] = event;
            if constexpr (kHard) {
                if (a.hard.seg_hard_W) {
                    a.hard.seg_hard_W[slot] = hard_W;
                    a.hard.seg_hard_ch[slot] = hard_code;
                }
                if (a.hard.seg_rad_k) {
                    a.hard.seg_rad_k[slot] = rad_k;
                    a.hard.seg_rad_Z[slot] = rad_Z;
                }
            }
        }
    }
    s.nseg += 1;
    s.px = px + step * dx;
    s.py = py + step * dy;
    s.pz = pz + step * dz;
    s.E = E_end;
    s.clock = t_end;
    s.tau -= step / lam;
    if (s.tau < 0.0) s.tau = 0.0;
    if constexpr (kHard) {
        if (a.hard.shell_on) {
            s.tau_h -= step * mu_hard;
            if (s.tau_h < 0.0) s.tau_h = 0.0;
        }
        if (a.hard.rad_on) {
            s.tau_r -= step * mu_rad;
            if (s.tau_r < 0.0) s.tau_r = 0.0;
        }
    }
    if (limited) {
        s.substep += 1;
        s.steps += 1;
        return s.steps < a.max_steps ? -1 : EXIT_STEP_LIMITED;
    }
    if (exit_top || exit_bot || side || cutoff)
        return exit_top ? EXIT_BACK : exit_bot ? EXIT_TRANS : side ? EXIT_SIDE : EXIT_CUTOFF;
    s.flight += 1;
    s.substep = 0;
    s.tau = -1.0;
    if constexpr (kHard) {
        s.tau_h = -1.0;
        s.tau_r = -1.0;
    }
    if (up || dn) {
        s.pz += (s.dz > 0.0 ? 1.0 : -1.0) * 1e-6;
        s.steps += 1;
        return s.steps < a.max_steps ? -1 : EXIT_STEP_LIMITED;
    }
    if constexpr (kHard) {
        if (event == EV_HARD_RADIATIVE) {  // debit the photon; no deflection, no elastic draw
            s.E -= rad_k * 1.0e-3;
            s.steps += 1;
            return s.steps < a.max_steps ? -1 : EXIT_STEP_LIMITED;
        }
        if (event == EV_HARD_INELASTIC) {  // the loss and recoil deflection replace the elastic draw
            s.E -= hard_W;
            rotate_lane<kMixed>(s, hard_cos, hard_phi);
            s.steps += 1;
            return s.steps < a.max_steps ? -1 : EXIT_STEP_LIMITED;
        }
    }
    // Elastic collision: element (cores.py:1094-1105), screened-Rutherford angle.
    const size_t el_row = (size_t)L * max_el;
    int i_el = 0;
    if (n_el != 1) {
        const double u = stream_uniform(s.key, s.draws);
        s.draws += 1ULL;
        i_el = n_el - 1;
        for (int k = 0; k < n_el; ++k) {
            if (lut.cdf((el_row + k) * nE, p) > u) {
                i_el = k;
                break;
            }
        }
    }
    const double cos_u = stream_uniform(s.key, s.draws);
    const double phi_u = stream_uniform(s.key, s.draws + 1ULL);
    s.draws += 2ULL;
    double alpha = 0.0, cos_t;
    if (a.elsepa.on) {  // tabulated angles at the same uniform (_jit_kernel.py:1149-1160)
        cos_t = 1.0 - 2.0 * elsepa::sample_mu(a.elsepa, el_row + i_el, s.E, cos_u);
    } else {
        alpha = lut.alpha((el_row + i_el) * nE, lut.pos(s.E));
        cos_t = 1.0 - 2.0 * alpha * cos_u / (1.0 + alpha - cos_u);
    }
    bool split_here = false;
    if constexpr (kCone != 0) {
        // the in-cone copy of this event, from the incident direction and the
        // chosen element, on the (electron, flight) event key (cores.py:1189-1211)
        const hr_cone_split &c = a.cone;
        split_here = s.E >= c.min_energy_keV;
        if (split_here) {
            const int64_t fl = s.flight - 1;
            cone::Copy cp;
            if constexpr (kCone == 2) {  // the reference's law: tabulated angles of the chosen element
                cone::TabLaw law(&a.elsepa, el_row + i_el, s.E);
                cp = cone::event_law(s.dx, s.dy, s.dz, c.nx, c.ny, c.nz, c.cos_c, c.sin_c, c.target_weight,
                                     c.angle_scale, law, s.E, cone::event_key(s.skey, fl));
            } else {
                cp = cone::event(s.dx, s.dy, s.dz, c.nx, c.ny, c.nz, c.cos_c, c.sin_c, c.target_weight,
                                 c.angle_scale, alpha, s.E, cone::event_key(s.skey, fl));
            }
            if (cp.made) {
                const int64_t i = s.e - a.e_start;
                const int64_t k = c.count[i];
                if (k < c.cap) {
                    const int64_t r = i * c.cap + k;
                    c.pos[3 * r] = s.px;
                    c.pos[3 * r + 1] = s.py;
                    c.pos[3 * r + 2] = s.pz;
                    c.dir[3 * r] = cp.ox;
                    c.dir[3 * r + 1] = cp.oy;
                    c.dir[3 * r + 2] = cp.oz;
                    c.E_keV[r] = s.E;
                    c.clock[r] = s.clock;
                    c.weight[r] = cp.w;
                    c.flight[r] = fl;
                }
                c.count[i] = k + 1;
            }
        }
    }
    rotate_lane<kMixed>(s, cos_t, 2.0 * kPi * phi_u);
    if constexpr (kCone != 0) {
        // the analog outcome lies in the cone, which the copies represent: the
        // primary ends after this event (cores.py:1215-1226)
        if (split_here && s.dx * a.cone.nx + s.dy * a.cone.ny + s.dz * a.cone.nz >= a.cone.cos_c) {
            a.cone.killed[s.e - a.e_start] = 1;
            return HR_EXIT_CONE_SPLIT;
        }
    }
    s.steps += 1;
    return s.steps < a.max_steps ? -1 : EXIT_STEP_LIMITED;
}

}  // namespace
}  // namespace hr
