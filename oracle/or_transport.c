/*
 * or_transport.c -- FP64 CPU restatement of PyRITE's per-electron LUT transport.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): loaded by tests/, the smoke check and
 * bench.py's cpu_baseline leg, never by the product path.
 *
 * Follows, in order of authority:
 *   src/pyrite/montecarlo/transport/cores.py:365-1250   (CPU per-electron LUT body)
 *   src/pyrite/montecarlo/transport/_jit_lut_kernel.py:64-574 (device twin)
 *   src/pyrite/montecarlo/transport/kinematics.py:198-252 (SplitMix64 streams)
 *   src/pyrite/montecarlo/transport/straggling.py:36-363 (Urban sampler)
 *   src/pyrite/montecarlo/transport/stopping.py:305-411 (spliced stopping)
 *   src/pyrite/montecarlo/transport/lut.py:94-124      (LUT clamp order)
 *   src/pyrite/montecarlo/transport/core_geometry.py:16-70 (rotation, prism exit)
 */
#include "hr_oracle.h"

#include <math.h>
#include <stddef.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* kinematics.py:181-193 */
#define SM64_GOLDEN 0x9E3779B97F4A7C15ULL
#define SM64_MIX1 0xBF58476D1CE4E5B9ULL
#define SM64_MIX2 0x94D049BB133111EBULL
#define U53_SCALE (1.0 / 9007199254740992.0)
/* straggling.py:33 */
#define URBAN_SALT 0xD6E8FEB86659FD93ULL
/* straggling.py:117-127 */
#define URBAN_E0 1.0e-2
#define URBAN_E2_PER_Z2 1.0e-2
#define URBAN_RATE 0.55
#define URBAN_CHUNK_MAX 64.0
/* stopping.py:195-197 */
#define BS_PREFACTOR 1.535e-6
#define MC2_KEV 510.99895
#define LN2 0.6931471805599453

/* events.py:36-131 */
/* events.py:36-43: ELASTIC 0, SUBSTEP 1, LAYER_BOUNDARY 2, GROOVE_SURFACE 3 (never
   emitted on this path), EXIT_TOP 4, EXIT_BOTTOM 5, EXIT_SIDE 6, CUTOFF 7,
   HARD_INELASTIC 8, HARD_RADIATIVE 9 */
enum { EV_ELASTIC = 0, EV_SUBSTEP = 1, EV_LAYER = 2, EV_GROOVE = 3, EV_EXIT_TOP = 4, EV_EXIT_BOTTOM = 5,
       EV_EXIT_SIDE = 6, EV_CUTOFF = 7, EV_HARD_INELASTIC = 8, EV_HARD_RADIATIVE = 9 };
/* hard_inelastic.py:40 */
#define HARD_SALT 0x3C6EF372FE94F82BULL
/* cores.py:90-95 */
enum { EXIT_CUTOFF = 0, EXIT_BACK = 1, EXIT_TRANS = 2, EXIT_SIDE = 3, EXIT_STEP_LIMITED = 4,
       EXIT_NOT_ENTERED = 5 };

/* kinematics.py:198-204 */
uint64_t or_splitmix64(uint64_t x) {
    x = (x ^ (x >> 30)) * SM64_MIX1;
    x = (x ^ (x >> 27)) * SM64_MIX2;
    return x ^ (x >> 31);
}

/* kinematics.py:212-216 */
uint64_t or_stream_key(uint64_t seed, uint64_t e) {
    return or_splitmix64(seed + SM64_GOLDEN * (e + 1ULL));
}

/* kinematics.py:219-223 */
double or_stream_uniform(uint64_t key, uint64_t counter) {
    uint64_t z = or_splitmix64(key + SM64_GOLDEN * (counter + 1ULL));
    return (double)(z >> 11) * U53_SCALE;
}

/* kinematics.py:226-245 */
void or_stream_keys(uint64_t seed, int64_t start, int64_t n, uint64_t *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = or_stream_key(seed, (uint64_t)(start + i));
}

/* straggling.py:36-39 */
uint64_t or_urban_stream_key(uint64_t stream_key) { return or_splitmix64(stream_key ^ URBAN_SALT); }

/* straggling.py:42-54 */
uint64_t or_urban_flight_key(uint64_t urban_key, int64_t flight, int64_t substep) {
    uint64_t combined = ((uint64_t)flight << 32) + (uint64_t)substep;
    return or_splitmix64(urban_key + SM64_GOLDEN * combined);
}

/* straggling.py:305-329 (element), with delta = 0 as every call site passes */
double or_dEds_spliced_element(double J, double k, double coeff, double E_cross, double E) {
    double tau = E / MC2_KEV;
    double gamma = 1.0 + tau;
    double beta_sq = 1.0 - 1.0 / (gamma * gamma);
    if (E < E_cross) return -7.85e-4 / E * coeff * log(1.166 * (E + k * J) / J);
    double f_minus = 1.0 - beta_sq + (tau * tau / 8.0 - (2.0 * tau + 1.0) * LN2) / (gamma * gamma);
    double I_rel = J / MC2_KEV;
    return -BS_PREFACTOR / beta_sq * coeff *
           (log(tau * tau * (tau + 2.0) / (2.0 * I_rel * I_rel)) + f_minus - 0.0);
}

/* stopping.py:384-411 */
double or_dEds_spliced_packed(const double *J, const double *k, const double *coeff,
                              const double *E_cross, int n_el, double E) {
    double tau = E / MC2_KEV;
    double gamma = 1.0 + tau;
    double beta_sq = 1.0 - 1.0 / (gamma * gamma);
    double f_minus = 1.0 - beta_sq + (tau * tau / 8.0 - (2.0 * tau + 1.0) * LN2) / (gamma * gamma);
    double jl = 0.0, bs = 0.0;
    for (int i = 0; i < n_el; ++i) {
        double Ji = J[i];
        double ci = coeff[i];
        if (E < E_cross[i]) {
            jl += ci * log(1.166 * (E + k[i] * Ji) / Ji);
        } else {
            double I_rel = Ji / MC2_KEV;
            bs += ci * (log(tau * tau * (tau + 2.0) / (2.0 * I_rel * I_rel)) + f_minus - 0.0);
        }
    }
    return -7.85e-4 / E * jl - BS_PREFACTOR / beta_sq * bs;
}

/* straggling.py:229-263 */
int64_t or_urban_poisson(double lam, uint64_t key, uint64_t *counter) {
    if (lam <= 0.0) return 0;
    int64_t n_chunks = (int64_t)ceil(lam / URBAN_CHUNK_MAX);
    double chunk_lam = lam / (double)n_chunks;
    int64_t total = 0;
    for (int64_t c = 0; c < n_chunks; ++c) {
        double u = or_stream_uniform(key, *counter);
        *counter += 1ULL;
        double p = exp(-chunk_lam);
        double cdf = p;
        int64_t kk = 0;
        while (u >= cdf) {
            kk += 1;
            p = p * chunk_lam / (double)kk;
            double next_cdf = cdf + p;
            if (next_cdf <= cdf) break;
            cdf = next_cdf;
        }
        total += kk;
    }
    return total;
}

/* straggling.py:131-201: returns valid flag, fills the channel rates */
static int urban_channels(double Z, double J, double C, double E, double *s1, double *E1,
                          double *s2, double *E2, double *s3) {
    double tau = E / MC2_KEV;
    double gamma = 1.0 + tau;
    double beta_sq = 1.0 - 1.0 / (gamma * gamma);
    double two = 2.0 * MC2_KEV * tau * (tau + 2.0);
    double T_up = 0.5 * E;
    if (T_up <= URBAN_E0 || C <= 0.0) return 0;
    double L_I = log(two / J) - beta_sq;
    if (L_I <= 0.0) return 0;
    /* _urban_levels_scalar (straggling.py:131-159) */
    double e2 = URBAN_E2_PER_Z2 * Z * Z;
    double f2 = Z > 2.0 ? 2.0 / Z : 1.0;
    double f1 = 1.0, e1 = J;
    int resolved = 0;
    if (Z > 2.0 && e2 < T_up && log(two / e2) - beta_sq > 0.0) {
        double f1c = 1.0 - f2;
        double e1c = exp((log(J) - f2 * log(e2)) / f1c);
        if (log(two / e1c) - beta_sq > 0.0) {
            f1 = f1c;
            e1 = e1c;
            resolved = 1;
        }
    }
    if (!resolved) {
        f1 = 1.0;
        e1 = J;
        f2 = 0.0;
    }
    double soft = C * (1.0 - URBAN_RATE) / L_I;
    *s1 = soft * (f1 / e1) * (log(two / e1) - beta_sq);
    *s2 = 0.0;
    if (f2 > 0.0) *s2 = soft * (f2 / e2) * (log(two / e2) - beta_sq);
    *s3 = C * URBAN_RATE * (T_up - URBAN_E0) / (URBAN_E0 * T_up * log(T_up / URBAN_E0));
    *E1 = e1;
    *E2 = e2;
    return 1;
}

/* straggling.py:204-226 */
void or_urban_moments_element(double Z, double J, double C, double E, double s, double *mean,
                              double *var) {
    double s1, E1, s2, E2, s3;
    if (!urban_channels(Z, J, C, E, &s1, &E1, &s2, &E2, &s3)) {
        *mean = C * s;
        *var = 0.0;
        return;
    }
    double T_up = 0.5 * E;
    double mean3 = URBAN_E0 * T_up * log(T_up / URBAN_E0) / (T_up - URBAN_E0);
    *mean = s * (s1 * E1 + s2 * E2 + s3 * mean3);
    *var = s * (s1 * E1 * E1 + s2 * E2 * E2 + s3 * URBAN_E0 * T_up);
}

/* straggling.py:278-302 */
static double urban_sample_element(double Z, double J, double C, double E, double s, uint64_t key,
                                   uint64_t *counter) {
    double s1, E1, s2, E2, s3;
    if (!urban_channels(Z, J, C, E, &s1, &E1, &s2, &E2, &s3) || s <= 0.0) return C * s;
    int64_t n1 = or_urban_poisson(s1 * s, key, counter);
    int64_t n2 = or_urban_poisson(s2 * s, key, counter);
    int64_t n3 = or_urban_poisson(s3 * s, key, counter);
    double dE = (double)n1 * E1 + (double)n2 * E2;
    double T_up = 0.5 * E;
    for (int64_t q = 0; q < n3; ++q) {
        double u = or_stream_uniform(key, *counter);
        *counter += 1ULL;
        dE += URBAN_E0 / (1.0 - u * (T_up - URBAN_E0) / T_up); /* straggling.py:266-275 */
    }
    return dE;
}

/* straggling.py:332-363 */
double or_urban_sample_compound(const double *Z, const double *J, const double *k,
                                const double *coeff, const double *E_cross, int n_el, double E,
                                double s, double stopping_scale, uint64_t key, uint64_t *counter) {
    double dE = 0.0;
    for (int i = 0; i < n_el; ++i) {
        double C = -or_dEds_spliced_element(J[i], k[i], coeff[i], E_cross[i], E);
        C *= stopping_scale;
        dE += urban_sample_element(Z[i], J[i], C, E, s, key, counter);
    }
    return dE;
}

/* lut.py:94-110; branch order: upper clamp first */
void or_lut_index_frac(double E, double log_E_min, double inv_dlogE, int n_energy, int *i,
                       double *f) {
    double x = (log(E) - log_E_min) * inv_dlogE;
    int last = n_energy - 1;
    if (x >= (double)last) {
        *i = last - 1;
        *f = 1.0;
        return;
    }
    if (x > 0.0) {
        int ii = (int)x;
        *i = ii;
        *f = x - (double)ii;
        return;
    }
    *i = 0;
    *f = 0.0;
}

static inline double lerp_row(const double *t, int i, double f) {
    return t[i] + f * (t[i + 1] - t[i]); /* lut.py:113-124 */
}

/* core_geometry.py:16-36 */
static void rotate_direction(double *dx, double *dy, double *dz, double cos_t, double phi) {
    double x = *dx, y = *dy, z = *dz;
    double s2 = 1.0 - cos_t * cos_t;
    double sin_t = sqrt(s2 > 0.0 ? s2 : 0.0);
    double cos_phi = cos(phi), sin_phi = sin(phi);
    double refx, refy;
    if (fabs(x) < 0.9) {
        refx = 1.0;
        refy = 0.0;
    } else {
        refx = 0.0;
        refy = 1.0;
    }
    double ux = -z * refy, uy = z * refx, uz = x * refy - y * refx;
    double u_mag = sqrt(ux * ux + uy * uy + uz * uz);
    ux = ux / u_mag;
    uy = uy / u_mag;
    uz = uz / u_mag;
    double wx = y * uz - z * uy, wy = z * ux - x * uz, wz = x * uy - y * ux;
    double a = sin_t * cos_phi, b = sin_t * sin_phi;
    double ox = cos_t * x + a * ux + b * wx;
    double oy = cos_t * y + a * uy + b * wy;
    double oz = cos_t * z + a * uz + b * wz;
    double mag = sqrt(ox * ox + oy * oy + oz * oz);
    *dx = ox / mag;
    *dy = oy / mag;
    *dz = oz / mag;
}

/* core_geometry.py:47-70, faces X_MIN..Z_MAX = 0..5 (montecarlo/geometry.py) */
static double first_prism_exit(double px, double py, double pz, double dx, double dy, double dz,
                               double z_min, double z_max, double width, double height,
                               int *face) {
    double best_t = INFINITY;
    int best = -1;
    double hw = 0.5 * width, hh = 0.5 * height, t;
    if (dx < 0.0) {
        t = (-hw - px) / dx;
        if (0.0 < t && t < best_t) { best_t = t; best = 0; }
    }
    if (dx > 0.0) {
        t = (hw - px) / dx;
        if (0.0 < t && t < best_t) { best_t = t; best = 1; }
    }
    if (dy < 0.0) {
        t = (-hh - py) / dy;
        if (0.0 < t && t < best_t) { best_t = t; best = 2; }
    }
    if (dy > 0.0) {
        t = (hh - py) / dy;
        if (0.0 < t && t < best_t) { best_t = t; best = 3; }
    }
    if (dz < 0.0) {
        t = (z_min - pz) / dz;
        if (0.0 < t && t < best_t) { best_t = t; best = 4; }
    }
    if (dz > 0.0) {
        t = (z_max - pz) / dz;
        if (0.0 < t && t < best_t) { best_t = t; best = 5; }
    }
    *face = best;
    return best_t;
}

/* core_geometry.py searchsorted(side="right") */
static int searchsorted_right(const double *b, double x, int n) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (b[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

/* One electron: cores.py:408-1239 with per_electron=True, lut=True; the shell
 * (inelastic) and coupled radiative modes by a->hard, detector-cone splitting
 * by a->cone (or_cone.c), every other optional mode (groove, mixed) off.  The radiative mode reads its rates
 * from LUT columns on the transport grid (the reference's table moment mode,
 * cores.py:467-470). */
static void transport_one(const or_transport_args *a, int64_t e) {
    const int64_t i = e - a->e_start;
    const int nE = a->lut_n_energy;
    const int max_el = a->max_el;
    a->seg_count[i] = 0;
    a->exit_code[i] = EXIT_NOT_ENTERED;
    if (a->straggle_on) a->stragg_dE[e] = 0.0;
    const or_cone_split *cone = &a->cone;
    if (cone->on) {
        cone->count[i] = 0;
        cone->killed[i] = 0;
    }
    if (!a->alive[e]) return;
    a->exit_code[i] = EXIT_STEP_LIMITED;

    const uint64_t key = a->stream_key[e];
    uint64_t draws = 0;
    int64_t nseg = 0, flight = 0, substep = 0;
    double tau_left = -1.0;
    const double E_cut = a->E_cut[e];
    const int energy_controlled = a->max_dE_frac > 0.0;
    int32_t steps = 0;
    double *pos = a->pos + 3 * e;
    double *dir = a->dirs + 3 * e;
    const or_hard_modes *h = &a->hard;
    const uint64_t hard_key = h->shell_on ? or_splitmix64(key ^ HARD_SALT) : 0;
    const uint64_t rad_key = h->rad_on ? h->rad_key[e] : 0;
    uint64_t hard_draws = 0, rad_draws = 0;
    double tau_h = -1.0, tau_rad = -1.0;

    while (steps < a->max_steps) {
        int L = a->n_layers == 1 ? 0 : searchsorted_right(a->internal_bounds, pos[2], a->n_layers - 1);
        int n_el = a->L_nel[L];
        double z_top = a->L_top[L], z_bot = a->L_bot[L];
        double E = a->E_keV[e];
        int li;
        double lf;
        or_lut_index_frac(E, a->lut_log_E_min, a->lut_inv_dlogE, nE, &li, &lf);
        double total_rate = lerp_row(a->lut_total_rate + (size_t)L * nE, li, lf);
        double lam = 1e8 / total_rate;
        if (tau_left < 0.0) {
            tau_left = -log(or_stream_uniform(key, draws));
            draws += 1ULL;
        }
        double step = tau_left * lam;
        int hard_j = 0, rad_j = 0;
        double mu_hard = 0.0, mu_rad = 0.0, rad_soft = 0.0;
        if (h->shell_on) {
            mu_hard = lerp_row(h->lut_sh_rate + (size_t)L * nE, li, lf);
            if (tau_h < 0.0) {
                tau_h = -log(or_stream_uniform(hard_key, hard_draws));
                hard_draws += 1ULL;
            }
            if (mu_hard > 0.0 && tau_h / mu_hard < step) {
                step = tau_h / mu_hard;
                hard_j = 1;
            }
        }
        if (h->rad_on) {
            rad_soft = lerp_row(h->lut_rad_soft + (size_t)L * nE, li, lf);
            mu_rad = lerp_row(h->lut_rad_rate + (size_t)L * nE, li, lf);
            if (tau_rad < 0.0) {
                tau_rad = -log(or_stream_uniform(rad_key, rad_draws));
                rad_draws += 1ULL;
            }
            if (mu_rad > 0.0 && tau_rad / mu_rad < step) {
                step = tau_rad / mu_rad;
                rad_j = 1;
                hard_j = 0;
            }
        }

        double dx = dir[0], dy = dir[1], dz = dir[2];
        double px = pos[0], py = pos[1], pz = pos[2];
        int cross_up = 0, cross_dn = 0, exit_side = 0;
        if (a->finite_footprint) {
            int face;
            double exit_distance = first_prism_exit(px, py, pz, dx, dy, dz, z_top, z_bot,
                                                    a->width_ang, a->height_ang, &face);
            if (step > exit_distance) {
                step = exit_distance;
                cross_up = face == 4;
                cross_dn = face == 5;
                exit_side = face >= 0 && face <= 3;
            }
        } else if (dz < 0.0) {
            double sb = (pz - z_top) / (-dz);
            if (step > sb) {
                step = sb;
                cross_up = 1;
            }
        } else if (dz > 0.0) {
            double sb = (z_bot - pz) / dz;
            if (step > sb) {
                step = sb;
                cross_dn = 1;
            }
        }
        int exit_top = cross_up && z_top <= 0.0;
        int exit_bot = cross_dn && z_bot >= a->z_total;

        double dEds = lerp_row(a->lut_dEds + (size_t)L * nE, li, lf);
        double inv_beta_j = lerp_row(a->lut_inv_beta, li, lf);
        /* soft radiative stopping joins as a mean loss (cores.py:605-611) */
        double rad_dEds = 0.0;
        if (h->rad_on) {
            rad_dEds = rad_soft;
            dEds -= rad_dEds;
        }
        int cutoff = 0, limited = 0;
        int geometry_event = cross_up || cross_dn || exit_side;
        double E_end, t_end;
        const double clock0 = a->clock[e];
        if (a->straggle_on) {
            if (energy_controlled) {
                double step_energy = a->max_dE_frac * E / (-dEds);
                if (step_energy < step) {
                    step = step_energy;
                    limited = 1;
                    cross_up = cross_dn = exit_top = exit_bot = exit_side = 0;
                    geometry_event = 0;
                }
            }
            uint64_t urban_key = or_urban_stream_key(key);
            uint64_t flight_key = or_urban_flight_key(urban_key, flight, substep);
            uint64_t counter = 0;
            const size_t row = (size_t)L * max_el;
            double loss;
            if (h->shell_on) { /* cores.py:616-640: the two-moment soft sampler */
                double omega2 = lerp_row(h->lut_sh_omega2 + (size_t)L * nE, li, lf);
                loss = or_shell_soft_loss(-(dEds + rad_dEds) * step, omega2 * step, flight_key, &counter);
            } else {
                loss = or_urban_sample_compound(a->L_Z + row, a->L_J + row, a->L_k + row, a->L_coeff + row,
                                                a->L_E_cross + row, n_el, E, step, 1.0, flight_key, &counter);
            }
            loss += rad_dEds * step;
            a->stragg_dE[e] += loss;
            double delta_cut = E - E_cut;
            if (loss > delta_cut || (loss == delta_cut && !geometry_event)) {
                step = loss > 0.0 ? step * (delta_cut / loss) : 0.0;
                cutoff = 1;
                limited = 0;
                cross_up = cross_dn = exit_top = exit_bot = exit_side = 0;
                E_end = E_cut;
            } else {
                E_end = E - loss;
            }
            if (a->energy_model == 1) {
                int ci;
                double cf;
                or_lut_index_frac(0.5 * (E + E_end), a->lut_log_E_min, a->lut_inv_dlogE, nE, &ci, &cf);
                t_end = clock0 + step * lerp_row(a->lut_inv_beta, ci, cf);
            } else {
                t_end = clock0 + step * inv_beta_j;
            }
        } else {
            double cutoff_distance;
            if (a->energy_model == 1) {
                double E_mid = 0.5 * (E + E_cut);
                int ci;
                double cf;
                or_lut_index_frac(E_mid, a->lut_log_E_min, a->lut_inv_dlogE, nE, &ci, &cf);
                double cutoff_rate = lerp_row(a->lut_dEds + (size_t)L * nE, ci, cf);
                if (h->rad_on) cutoff_rate -= lerp_row(h->lut_rad_soft + (size_t)L * nE, ci, cf);
                cutoff_distance = (E_cut - E) / cutoff_rate;
            } else {
                cutoff_distance = (E_cut - E) / dEds;
            }
            if (cutoff_distance < step || (cutoff_distance == step && !geometry_event)) {
                step = cutoff_distance;
                cutoff = 1;
                cross_up = cross_dn = exit_top = exit_bot = exit_side = 0;
            }
            if (energy_controlled && !cutoff) {
                double step_energy = a->max_dE_frac * E / (-dEds);
                if (step_energy < step) {
                    step = step_energy;
                    limited = 1;
                    cross_up = cross_dn = exit_top = exit_bot = exit_side = 0;
                }
            }
            if (a->energy_model == 1) {
                if (cutoff) {
                    E_end = E_cut;
                } else {
                    double E_pred = E + dEds * step;
                    double E_mid = 0.5 * (E + E_pred);
                    int mi;
                    double mf;
                    or_lut_index_frac(E_mid, a->lut_log_E_min, a->lut_inv_dlogE, nE, &mi, &mf);
                    double mid_rate = lerp_row(a->lut_dEds + (size_t)L * nE, mi, mf);
                    if (h->rad_on) mid_rate -= lerp_row(h->lut_rad_soft + (size_t)L * nE, mi, mf);
                    E_end = E + step * mid_rate;
                }
                int ci;
                double cf;
                or_lut_index_frac(0.5 * (E + E_end), a->lut_log_E_min, a->lut_inv_dlogE, nE, &ci, &cf);
                t_end = clock0 + step * lerp_row(a->lut_inv_beta, ci, cf);
            } else {
                E_end = cutoff ? E_cut : E + dEds * step;
                t_end = clock0 + step * inv_beta_j;
            }
        }

        int8_t event;
        if (limited) event = EV_SUBSTEP;
        else if (cutoff) event = EV_CUTOFF;
        else if (exit_top) event = EV_EXIT_TOP;
        else if (exit_bot) event = EV_EXIT_BOTTOM;
        else if (exit_side) event = EV_EXIT_SIDE;
        else if (cross_up || cross_dn) event = EV_LAYER;
        else if (h->shell_on && hard_j) event = EV_HARD_INELASTIC;
        else if (h->rad_on && rad_j) event = EV_HARD_RADIATIVE;
        else event = EV_ELASTIC;

        double hard_W_keV = 0.0, hard_cos = 1.0, hard_phi = 0.0, rad_k_eV = 0.0;
        int16_t hard_code = -1, rad_Z = 0;
        if (event == EV_HARD_INELASTIC) { /* cores.py:850-880, _sample_hard_collision */
            double u_ch = or_stream_uniform(hard_key, hard_draws);
            double u_w = or_stream_uniform(hard_key, hard_draws + 1ULL);
            double u_q = or_stream_uniform(hard_key, hard_draws + 2ULL);
            double u_phi = or_stream_uniform(hard_key, hard_draws + 3ULL);
            hard_draws += 4ULL;
            double target = u_ch * mu_hard;
            int nch = h->sh_nch[L], ch = nch - 1;
            double cum = 0.0;
            for (int k = 0; k < nch; ++k) {
                cum += lerp_row(h->lut_sh_ch_rate + ((size_t)L * h->sh_max_ch + k) * nE, li, lf);
                if (cum > target) { ch = k; break; }
            }
            size_t c = (size_t)L * h->sh_max_ch + ch;
            double E_eV = E * 1.0e3;
            double W_eV = or_shell_sample_transfer(E_eV, h->sh_U[c], h->sh_W[c], h->sh_branch[c], h->sh_cutoff_eV, u_w);
            hard_W_keV = W_eV * 1.0e-3;
            hard_code = h->sh_code[c];
            if (E_end - hard_W_keV <= E_cut) {
                event = EV_CUTOFF;
                cutoff = 1;
            } else {
                hard_cos = or_shell_primary_cosine(E_eV, h->sh_U[c], h->sh_W[c], h->sh_branch[c], W_eV, u_q);
                hard_phi = 2.0 * M_PI * u_phi;
            }
        } else if (event == EV_HARD_RADIATIVE) { /* cores.py:880-906, _sample_hard_photon */
            double rates[16], soft_e, mu_event;
            or_rad_layer_moments(&h->rad, L, n_el, max_el, E_end * 1e3, &soft_e, &mu_event, rates);
            double u_el = or_stream_uniform(rad_key, rad_draws);
            double u_k = or_stream_uniform(rad_key, rad_draws + 1ULL);
            rad_draws += 2ULL;
            double target = u_el * mu_event, cum = 0.0;
            int sel = n_el - 1;
            for (int k = 0; k < n_el; ++k) {
                cum += rates[k];
                if (cum > target) { sel = k; break; }
            }
            int slot = L * max_el + sel;
            rad_k_eV = or_rad_sample_photon(&h->rad, slot, E_end * 1e3, h->rad.cutoff_eV, u_k, h->rad.plasma[L]);
            rad_Z = (int16_t)h->rad.z[slot];
            if (E_end - rad_k_eV * 1e-3 <= E_cut) {
                event = EV_CUTOFF;
                cutoff = 1;
            }
        }

        if (nseg < a->cap) {
            const int64_t slot = i * a->cap + nseg;
            a->seg_dir[3 * slot + 0] = dx;
            a->seg_dir[3 * slot + 1] = dy;
            a->seg_dir[3 * slot + 2] = dz;
            a->seg_mid[3 * slot + 0] = px + 0.5 * step * dx;
            a->seg_mid[3 * slot + 1] = py + 0.5 * step * dy;
            a->seg_mid[3 * slot + 2] = pz + 0.5 * step * dz;
            a->seg_len[slot] = step;
            a->seg_E[slot] = E;
            a->seg_t0[slot] = clock0;
            a->seg_id[slot] = e;
            a->seg_lay[slot] = (int16_t)L;
            if (a->energy_model == 1) {
                a->seg_E_end[slot] = E_end;
                a->seg_t_end[slot] = t_end;
                a->seg_flight[slot] = flight;
                a->seg_substep[slot] = substep;
                a->seg_event[slot] = event;
                if (h->seg_hard_W) {
                    h->seg_hard_W[slot] = hard_W_keV;
                    h->seg_hard_ch[slot] = hard_code;
                }
                if (h->seg_rad_k) {
                    h->seg_rad_k[slot] = rad_k_eV;
                    h->seg_rad_Z[slot] = rad_Z;
                }
            }
        }
        nseg += 1;
        pos[0] = px + step * dx;
        pos[1] = py + step * dy;
        pos[2] = pz + step * dz;
        a->E_keV[e] = E_end;
        a->clock[e] = t_end;
        tau_left -= step / lam;
        if (tau_left < 0.0) tau_left = 0.0;
        if (h->shell_on) {
            tau_h -= step * mu_hard;
            if (tau_h < 0.0) tau_h = 0.0;
        }
        if (h->rad_on) {
            tau_rad -= step * mu_rad;
            if (tau_rad < 0.0) tau_rad = 0.0;
        }

        if (limited) {
            substep += 1;
            steps += 1;
            continue;
        }
        if (exit_top || exit_bot || exit_side || cutoff) {
            a->exit_code[i] = exit_top ? EXIT_BACK : exit_bot ? EXIT_TRANS : exit_side ? EXIT_SIDE : EXIT_CUTOFF;
            break;
        }
        flight += 1;
        substep = 0;
        tau_left = -1.0;
        tau_h = -1.0;
        tau_rad = -1.0;
        if (cross_up || cross_dn) {
            pos[2] += (dir[2] > 0.0 ? 1.0 : -1.0) * 1e-6;
            steps += 1;
            continue;
        }
        if (event == EV_HARD_RADIATIVE) { /* a photon: debit, no deflection */
            a->E_keV[e] -= rad_k_eV * 1e-3;
            steps += 1;
            continue;
        }
        if (event == EV_HARD_INELASTIC) { /* the loss and recoil deflection replace the elastic draw */
            a->E_keV[e] -= hard_W_keV;
            rotate_direction(&dir[0], &dir[1], &dir[2], hard_cos, hard_phi);
            steps += 1;
            continue;
        }
        /* elastic: element choice (cores.py:1094-1105), then SR angle */
        int i_el;
        const size_t el_row = (size_t)L * max_el;
        if (n_el == 1) {
            i_el = 0;
        } else {
            double u = or_stream_uniform(key, draws);
            draws += 1ULL;
            i_el = n_el - 1;
            for (int kel = 0; kel < n_el; ++kel) {
                if (lerp_row(a->lut_cdf + (el_row + kel) * nE, li, lf) > u) {
                    i_el = kel;
                    break;
                }
            }
        }
        int ai;
        double af;
        or_lut_index_frac(a->E_keV[e], a->lut_log_E_min, a->lut_inv_dlogE, nE, &ai, &af);
        double alpha = lerp_row(a->lut_alpha + (el_row + i_el) * nE, ai, af);
        double cos_u = or_stream_uniform(key, draws);
        draws += 1ULL;
        double phi_u = or_stream_uniform(key, draws);
        draws += 1ULL;
        double cos_t;
        if (a->elsepa.on) /* tabulated angles at the same uniform (_jit_kernel.py:1149-1160) */
            cos_t = 1.0 - 2.0 * or_elsepa_sample_mu(&a->elsepa, (int64_t)(el_row + i_el), a->E_keV[e], cos_u);
        else
            cos_t = 1.0 - 2.0 * alpha * cos_u / (1.0 + alpha - cos_u); /* scattering.py:234 */
        /* cone split (cores.py:1189-1226): the in-cone copy from the incident
         * direction and chosen element, on the key of the flight just closed */
        int split_here = cone->on && a->E_keV[e] >= cone->min_energy_keV;
        if (split_here) {
            double cp[4];
            uint64_t ek = or_cone_event_key(cone->split_key[e], flight - 1);
            int made = a->elsepa.on ? or_cone_event_tab(dir[0], dir[1], dir[2], cone, &a->elsepa,
                                                        (int64_t)(el_row + i_el), a->E_keV[e], ek, cp)
                                    : or_cone_event(dir[0], dir[1], dir[2], cone, alpha, a->E_keV[e], ek, cp);
            if (made) {
                int64_t k = cone->count[i];
                if (k < cone->cap) {
                    int64_t r = i * cone->cap + k;
                    for (int d = 0; d < 3; ++d) {
                        cone->pos[3 * r + d] = pos[d];
                        cone->dir[3 * r + d] = cp[d];
                    }
                    cone->E_keV[r] = a->E_keV[e];
                    cone->clock[r] = a->clock[e];
                    cone->weight[r] = cp[3];
                    cone->flight[r] = flight - 1;
                }
                cone->count[i] = k + 1;
            }
        }
        rotate_direction(&dir[0], &dir[1], &dir[2], cos_t, 2.0 * M_PI * phi_u);
        if (split_here && dir[0] * cone->nx + dir[1] * cone->ny + dir[2] * cone->nz >= cone->cos_c) {
            cone->killed[i] = 1;
            a->exit_code[i] = 6; /* EXIT_CONE_SPLIT, cone_split.py:55 */
            break;
        }
        steps += 1;
    }
    a->seg_count[i] = nseg;
}

void or_rotate_direction(double *d, double cos_t, double phi) {
    rotate_direction(&d[0], &d[1], &d[2], cos_t, phi);
}

void or_transport_lut(const or_transport_args *a, int threads) {
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
#endif
    for (int64_t e = a->e_start; e < a->e_start + a->e_count; ++e) transport_one(a, e);
    (void)threads;
}
