// radiative.hpp -- coupled soft/hard bremsstrahlung of the transport step.
//
// Algorithm of PyRITE's coupled radiative mode, device form
// (src/pyrite/montecarlo/transport/_jit_radiative_device.py:40-481, host twin
// _jit_radiative.py:26-288, dielectric cells dielectric_radiative.py:19-80):
// a per-element scaled SDCS chi(T, kappa) [mb], linear in the reduced photon
// energy kappa = k/T on each cell and linear in ln T between incident rows, is
// integrated analytically cell by cell below the radiative cutoff k_c (soft
// first moment, a continuous loss) and above it (hard zeroth moment, the rate
// of explicit photons).  With a plasma energy the cells carry the
// Ter-Mikaelian factor k^2/(k^2 + (gamma hbar w_p)^2), integrated in closed
// form (a six-term series below k/(gamma hbar w_p) = 0.1).  A hard photon's
// energy inverts the zeroth-moment CDF: the cell by cumulative mass, then 52
// bisection steps on the cell's partial integral.
//
// Table layout, element slot s = layer*max_el + i: incident energies [keV]
// and row tips at s*max_t, nominal reduced energies at s*n_red, row r of chi
// at (s*max_t + r)*n_red.
#pragma once
#include <math.h>

#include "hr_common.hpp"

namespace hr {
namespace rad {

constexpr double kMbCm2 = 1.0e-27;
constexpr double kMc2eV = 510998.95;
constexpr double kCmPerAng = 1.0e-8;
constexpr int kBisection = 52;
constexpr int kTaylorTerms = 6;
enum : int { SOFT = 0, HARD_RATE = 1 };

struct Tables {
    const double *incident;  // [slots*max_t] keV
    const double *nominal;   // [slots*n_red]
    const double *top;       // [slots*max_t]
    const double *chi;       // [slots*max_t*n_red] mb
    const double *z;         // [slots]
    const double *ncm3;      // [slots] atoms/cm^3
    const int32_t *nt;       // [slots] incident rows in use
    const int32_t *nel;      // [layers]
    const double *plasma;    // [layers] hbar w_p [eV]
    int32_t max_el, max_t, n_red;
    double cutoff_eV;
};

// Lower incident row of energy_eV (clamped to the first/last interval).
__host__ __device__ inline int row_of(const double *inc, int n_t, double energy_eV) {
    const double le = log(energy_eV * 1.0e-3);
    if (le <= log(inc[0])) return 0;
    if (le >= log(inc[n_t - 1])) return n_t - 2;
    int lo = 0, hi = n_t - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (log(inc[mid]) <= le) lo = mid;
        else hi = mid;
    }
    return lo;
}

// ln-energy fraction above row `row`.
__host__ __device__ inline double fraction_of(const double *inc, int n_t, double energy_eV, int row) {
    const double le = log(energy_eV * 1.0e-3);
    if (le <= log(inc[0])) return 0.0;
    if (le >= log(inc[n_t - 1])) return 1.0;
    const double lower = log(inc[row]);
    return (le - lower) / (log(inc[row + 1]) - lower);
}

// One incident row's piecewise-linear chi at reduced energy x; the last cell
// ends at the row's tip.
__host__ __device__ inline double row_chi(double x, const double *nom, int n_red, double tip, const double *chi_row) {
    if (x >= tip) return chi_row[n_red - 1];
    int lo = 0, hi = n_red - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (x >= nom[mid]) lo = mid;
        else hi = mid;
    }
    const double x0 = nom[lo];
    const double x1 = lo == n_red - 2 ? tip : nom[lo + 1];
    const double t = (x - x0) / (x1 - x0);
    return chi_row[lo] + t * (chi_row[lo + 1] - chi_row[lo]);
}

// Smallest node of the union {nominal, both row tips, kc/T, 1} strictly above x, else +inf.
__host__ __device__ inline double next_node(double x, const double *nom, int n_red, double tip_a, double tip_b,
                                            double cut_red) {
    int lo = 0, hi = n_red - 1;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (nom[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    double best = INFINITY;
    if (lo < n_red - 1) best = nom[lo];
    if (tip_a > x && tip_a < best) best = tip_a;
    if (tip_b > x && tip_b < best) best = tip_b;
    if (cut_red > x && cut_red < best) best = cut_red;
    if (1.0 > x && 1.0 < best) best = 1.0;
    return best;
}

// Zeroth moment of a linear chi cell, int chi/k dk.
__host__ __device__ inline double cell_rate(double left, double right, double cl, double cr) {
    if (right <= left) return 0.0;
    const double slope = (cr - cl) / (right - left);
    const double icpt = cl - slope * left;
    return icpt * log(right / left) + slope * (right - left);
}

// Suppressed cell: first (SOFT) or zeroth (HARD_RATE) moment of chi*F, F = k^2/(k^2+s^2).
__host__ __device__ inline double dielectric_cell(int which, double left, double right, double cl, double cr,
                                                  double s) {
    if (right <= left) return 0.0;
    const double slope = (cr - cl) / (right - left);
    const double icpt = cl - slope * left;
    const double u = left / s, v = right / s;
    if (v < 0.1) {
        double acc = 0.0, sign = 1.0;
        for (int n = 0; n < kTaylorTerms; ++n) {
            const int p = 2 * n + 2;
            if (which == SOFT) {
                acc += sign * s *
                       (icpt * (pow(v, p + 1) - pow(u, p + 1)) / (double)(p + 1) +
                        slope * s * (pow(v, p + 2) - pow(u, p + 2)) / (double)(p + 2));
            } else {
                acc += sign * (icpt * (pow(v, p) - pow(u, p)) / (double)p +
                               slope * s * (pow(v, p + 1) - pow(u, p + 1)) / (double)(p + 1));
            }
            sign = -sign;
        }
        return acc;
    }
    const double dc = 0.5 * (log1p(v * v) - log1p(u * u));
    const double da = (v - u) - (atan(v) - atan(u));
    if (which == HARD_RATE) return icpt * dc + slope * s * da;
    const double db = 0.5 * (v * v - u * u) - dc;
    return s * (icpt * da + slope * s * db);
}

struct Row {
    int row;
    double frac, tip_a, tip_b, cut_red, scale, sup;
    const double *inc, *nom, *chi;
    int n_red;
};

__host__ __device__ inline Row open_row(const Tables &t, int slot, double energy_eV, double cutoff_eV,
                                        double plasma_eV) {
    Row r;
    r.inc = t.incident + (size_t)slot * t.max_t;
    const int n_t = t.nt[slot];
    r.row = row_of(r.inc, n_t, energy_eV);
    r.frac = fraction_of(r.inc, n_t, energy_eV, r.row);
    const double *tip = t.top + (size_t)slot * t.max_t;
    r.tip_a = tip[r.row];
    r.tip_b = tip[r.row + 1];
    r.cut_red = cutoff_eV / energy_eV;
    const double zz = t.z[slot];
    r.scale = kMbCm2 * zz * zz;
    r.sup = (1.0 + energy_eV / kMc2eV) * plasma_eV;
    r.nom = t.nominal + (size_t)slot * t.n_red;
    r.chi = t.chi + (size_t)slot * t.max_t * t.n_red;
    r.n_red = t.n_red;
    return r;
}

// Energy-interpolated chi at reduced energy x.
__host__ __device__ inline double chi_at(const Row &r, double x) {
    const double v0 = row_chi(x, r.nom, r.n_red, r.tip_a, r.chi + (size_t)r.row * r.n_red);
    const double v1 = row_chi(x, r.nom, r.n_red, r.tip_b, r.chi + (size_t)(r.row + 1) * r.n_red);
    return v0 + r.frac * (v1 - v0);
}

// Photon energy of a reduced node; the cutoff node is kept exact.
__host__ __device__ inline double photon_of(double red, double energy_eV, double cutoff_eV, double cut_red) {
    return red == cut_red ? cutoff_eV : red * energy_eV;
}

// Soft first moment [eV cm^2] or hard rate [cm^2] per atom of element slot `slot`.
__host__ __device__ inline double moment(int which, const Tables &t, int slot, double energy_eV, double cutoff_eV,
                                         double plasma_eV) {
    const Row r = open_row(t, slot, energy_eV, cutoff_eV, plasma_eV);
    double acc = 0.0;
    double lred = which == SOFT ? next_node(-1.0, r.nom, r.n_red, r.tip_a, r.tip_b, r.cut_red) : r.cut_red;
    double left = photon_of(lred, energy_eV, cutoff_eV, r.cut_red);
    double cl = chi_at(r, left / energy_eV);
    double rred = next_node(lred, r.nom, r.n_red, r.tip_a, r.tip_b, r.cut_red);
    while (rred < INFINITY) {
        const double right = photon_of(rred, energy_eV, cutoff_eV, r.cut_red);
        if (which == SOFT && right > cutoff_eV) break;
        const double cr = chi_at(r, right / energy_eV);
        if (r.sup == 0.0) {
            if (left == 0.0) {
                acc += r.scale * 0.5 * (cl + cr) * right;
            } else if (right <= cutoff_eV) {
                if (right > left) acc += r.scale * (0.5 * (cl + cr) * (right - left));
            } else {
                acc += r.scale * cell_rate(left, right, cl, cr);
            }
        } else if (right <= cutoff_eV) {
            acc += r.scale * dielectric_cell(SOFT, left, right, cl, cr, r.sup);
        } else {
            acc += r.scale * dielectric_cell(HARD_RATE, left, right, cl, cr, r.sup);
        }
        left = right;
        cl = cr;
        lred = rred;
        rred = next_node(lred, r.nom, r.n_red, r.tip_a, r.tip_b, r.cut_red);
    }
    return acc;
}

// Layer soft loss [eV/A] (SOFT) or hard rate [1/A] (HARD_RATE).
__host__ __device__ inline double layer_moment(int which, const Tables &t, int layer, double energy_eV) {
    double total = 0.0;
    for (int i = 0; i < t.nel[layer]; ++i) {
        const int slot = layer * t.max_el + i;
        total += t.ncm3[slot] * kCmPerAng * moment(which, t, slot, energy_eV, t.cutoff_eV, t.plasma[layer]);
    }
    return total;
}

// Hard photon energy [eV] of element slot `slot` at uniform u.
__host__ __device__ inline double sample_photon(const Tables &t, int slot, double energy_eV, double u,
                                                double plasma_eV) {
    const double cutoff_eV = t.cutoff_eV;
    const double rate = moment(HARD_RATE, t, slot, energy_eV, cutoff_eV, plasma_eV);
    if (rate <= 0.0) return energy_eV;
    const Row r = open_row(t, slot, energy_eV, cutoff_eV, plasma_eV);
    const double target = u * rate;
    double cum = 0.0;
    double lred = next_node(-1.0, r.nom, r.n_red, r.tip_a, r.tip_b, r.cut_red);
    double left = photon_of(lred, energy_eV, cutoff_eV, r.cut_red);
    double cl = chi_at(r, left / energy_eV);
    double rred = next_node(lred, r.nom, r.n_red, r.tip_a, r.tip_b, r.cut_red);
    while (rred < INFINITY) {
        const double right = photon_of(rred, energy_eV, cutoff_eV, r.cut_red);
        const double cr = chi_at(r, right / energy_eV);
        const double nred = next_node(rred, r.nom, r.n_red, r.tip_a, r.tip_b, r.cut_red);
        if (right > cutoff_eV) {
            const double mass = r.scale * (r.sup == 0.0 ? cell_rate(left, right, cl, cr)
                                                        : dielectric_cell(HARD_RATE, left, right, cl, cr, r.sup));
            if (cum + mass >= target || nred == INFINITY) {
                const double local = target - cum;
                double lo = left, hi = right;
                for (int k = 0; k < kBisection; ++k) {
                    const double mid = 0.5 * (lo + hi);
                    const double cm = cl + (cr - cl) * (mid - left) / (right - left);
                    const double part = r.sup == 0.0 ? cell_rate(left, mid, cl, cm)
                                                     : dielectric_cell(HARD_RATE, left, mid, cl, cm, r.sup);
                    if (r.scale * part < local) lo = mid;
                    else hi = mid;
                }
                return 0.5 * (lo + hi);
            }
            cum += mass;
        }
        left = right;
        cl = cr;
        rred = nred;
    }
    return energy_eV;
}

}  // namespace rad
}  // namespace hr
