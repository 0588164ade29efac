/*
 * or_cone.c -- FP64 CPU restatement of detector-cone splitting of elastic
 * events, on the screened-Rutherford law of this transport.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): loaded by tests/, never by the
 * product path.
 *
 * Follows src/pyrite/montecarlo/transport/cone_split.py:
 *   :120-130  split keys and the per-event key
 *   :140-165  power-of-two weight ladder (counter 4 of the event key)
 *   :168-179  copy stream keys
 *   :210-220  azimuth arc half-width
 *   :223-335  cone_split_event: band bound, copy roulette (counter 0), polar
 *             draw in the band (counter 1), azimuth on the arc (counter 2),
 *             frame about the incident direction with e1 toward n, weight
 *             p_band * dphi / pi / keep, angle roulette (counter 3)
 *   :412-444  collect: copies in (electron, event) order
 * and cores.py:1189-1226 (call site: copy from the incident direction and
 * chosen element before the analog rotation; a primary whose analog outcome
 * lands in the cone ends with exit code 6).
 *
 * The reference's analog law is a tabulated ELSEPA density whose band edges
 * it finds by bisection (or_cone_event_tab restates that); or_cone_event runs
 * the transport's own screened-Rutherford law, p(mu) ~ 1/(alpha + mu)^2 on
 * mu = (1 - cos theta)/2 with CDF F(mu) = (1 + alpha) mu / (alpha + mu), so
 * the band's uniform interval [F(mu_lo), F(mu_hi)] is exact.  The interval
 * is handled through the tail 1 - F = alpha (1 - mu)/(alpha + mu), which
 * keeps its width exact when F is within 1e-6 of 1 (MeV electrons).
 */
#include "hr_oracle.h"

#include <math.h>
#include <stddef.h>

#define CONE_SPLIT_SALT 0xC2B2AE3D27D4EB4FULL
#define CONE_COPY_SALT 0x165667B19E3779F9ULL
#define CONE_GOLDEN 0x9E3779B97F4A7C15ULL
#define CONE_MC2 510.99895

uint64_t or_cone_split_key(uint64_t stream_key) { return or_splitmix64(stream_key ^ CONE_SPLIT_SALT); }

uint64_t or_cone_event_key(uint64_t split_key, int64_t flight) {
    return or_splitmix64(split_key + CONE_GOLDEN * ((uint64_t)flight + 1ULL));
}

uint64_t or_cone_copy_key(uint64_t split_key, int64_t flight) {
    return or_splitmix64(or_cone_event_key(split_key, flight) ^ CONE_COPY_SALT);
}

static double sr_tail(double alpha, double mu) { return alpha * (1.0 - mu) / (alpha + mu); }

static double arc_half(double cos_t, double sin_t, double cos_dn, double sin_dn, double cos_c) {
    if (sin_t * sin_dn <= 1e-300) return cos_t * cos_dn >= cos_c ? M_PI : 0.0;
    double c_star = (cos_c - cos_t * cos_dn) / (sin_t * sin_dn);
    if (c_star <= -1.0) return M_PI;
    if (c_star >= 1.0) return 0.0;
    return acos(c_star);
}

/* The analog law of one event: screened Rutherford (tab == NULL) with its
 * exact band, or a tabulated row pair mixed in ln E whose band is bounded by
 * the two rows' exact CDFs before the copy roulette and found by 64 bisection
 * steps on the mixed quantile after it (cone_split.py:240-295). */
typedef struct cone_law {
    const or_elsepa *tab;
    double alpha;                        /* screened Rutherford */
    int64_t row;                         /* tabulated: first bracketing row */
    double f, m_lo, m_hi, xa, p, tail_a; /* state between the steps */
} cone_law;

static double law_bound(cone_law *L, double mu_lo, double mu_hi) {
    if (!L->tab) {
        L->tail_a = sr_tail(L->alpha, mu_lo);
        L->p = L->tail_a - sr_tail(L->alpha, mu_hi); /* F(mu_hi) - F(mu_lo) */
        return L->p;
    }
    L->m_lo = mu_lo;
    L->m_hi = mu_hi;
    double fa0 = or_elsepa_cdf_at(L->tab, L->row, mu_lo), fa1 = or_elsepa_cdf_at(L->tab, L->row + 1, mu_lo);
    double fb0 = or_elsepa_cdf_at(L->tab, L->row, mu_hi), fb1 = or_elsepa_cdf_at(L->tab, L->row + 1, mu_hi);
    return fmax(fb0, fb1) - fmin(fa0, fa1);
}

static double law_band(cone_law *L) {
    if (!L->tab) return L->p;
    double lo, hi;
    if (or_elsepa_mix_mu(L->tab, L->row, L->f, 0.0) >= L->m_lo) {
        L->xa = 0.0;
    } else {
        lo = 0.0;
        hi = 1.0;
        for (int i = 0; i < 64; ++i) {
            double mid = 0.5 * (lo + hi);
            if (or_elsepa_mix_mu(L->tab, L->row, L->f, mid) >= L->m_lo)
                hi = mid;
            else
                lo = mid;
        }
        L->xa = hi;
    }
    double xb;
    if (or_elsepa_mix_mu(L->tab, L->row, L->f, 1.0) <= L->m_hi) {
        xb = 1.0;
    } else {
        lo = 0.0;
        hi = 1.0;
        for (int i = 0; i < 64; ++i) {
            double mid = 0.5 * (lo + hi);
            if (or_elsepa_mix_mu(L->tab, L->row, L->f, mid) <= L->m_hi)
                lo = mid;
            else
                hi = mid;
        }
        xb = lo;
    }
    L->p = xb - L->xa;
    return L->p;
}

static double law_mu(const cone_law *L, double u) {
    if (!L->tab) {
        /* xi = F(mu_lo) + u p_band, i.e. 1 - xi = tail_a - u p_band */
        double one_minus_xi = L->tail_a - u * L->p;
        return L->alpha * (1.0 - one_minus_xi) / (L->alpha + one_minus_xi);
    }
    return or_elsepa_mix_mu(L->tab, L->row, L->f, L->xa + u * L->p);
}

static int cone_event_law(double dx, double dy, double dz, const or_cone_split *c, cone_law *law, double E_keV,
                          uint64_t key, double *out) {
    out[0] = out[1] = out[2] = out[3] = 0.0;
    double cos_dn = dx * c->nx + dy * c->ny + dz * c->nz;
    if (cos_dn > 1.0) cos_dn = 1.0;
    if (cos_dn < -1.0) cos_dn = -1.0;
    double sin_dn = sqrt(fmax(0.0, 1.0 - cos_dn * cos_dn));
    double theta_dn = acos(cos_dn);
    double theta_c = asin(c->sin_c);
    double theta_lo = fmax(0.0, theta_dn - theta_c);
    double theta_hi = fmin(M_PI, theta_dn + theta_c);
    double h_lo = sin(0.5 * theta_lo), h_hi = sin(0.5 * theta_hi);
    double mu_lo = h_lo * h_lo, mu_hi = h_hi * h_hi; /* (1 - cos theta)/2 */
    double band_bound = law_bound(law, mu_lo, mu_hi);
    double arc_bound = sin_dn <= c->sin_c ? M_PI : asin(c->sin_c / sin_dn);
    double bound = band_bound * arc_bound / M_PI;
    if (!(bound > 0.0)) return 0;
    double keep = bound / c->target_weight;
    if (keep > 1.0) keep = 1.0;
    if (or_stream_uniform(key, 0) >= keep) return 0;
    double p_band = law_band(law);
    if (!(p_band > 0.0)) return 0;
    double mu_t = law_mu(law, or_stream_uniform(key, 1));
    double cos_t = 1.0 - 2.0 * mu_t;
    double sin_t = sqrt(fmax(0.0, 1.0 - cos_t * cos_t));
    double half = arc_half(cos_t, sin_t, cos_dn, sin_dn, c->cos_c);
    if (!(half > 0.0)) return 0;
    double psi = (2.0 * or_stream_uniform(key, 2) - 1.0) * half;
    double e1[3];
    if (sin_dn > 1e-12) {
        e1[0] = (c->nx - cos_dn * dx) / sin_dn;
        e1[1] = (c->ny - cos_dn * dy) / sin_dn;
        e1[2] = (c->nz - cos_dn * dz) / sin_dn;
    } else {
        double r[3] = {0.0, 0.0, 0.0};
        if (fabs(dx) < 0.9) r[0] = 1.0;
        else r[1] = 1.0;
        e1[0] = dy * r[2] - dz * r[1];
        e1[1] = dz * r[0] - dx * r[2];
        e1[2] = dx * r[1] - dy * r[0];
        double nr = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        e1[0] /= nr;
        e1[1] /= nr;
        e1[2] /= nr;
    }
    double e2[3] = {dy * e1[2] - dz * e1[1], dz * e1[0] - dx * e1[2], dx * e1[1] - dy * e1[0]};
    double ca = sin_t * cos(psi), cb = sin_t * sin(psi);
    double o[3] = {cos_t * dx + ca * e1[0] + cb * e2[0], cos_t * dy + ca * e1[1] + cb * e2[1],
                   cos_t * dz + ca * e1[2] + cb * e2[2]};
    double on = sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
    o[0] /= on;
    o[1] /= on;
    o[2] /= on;
    double w = p_band * half / M_PI / keep;
    if (c->angle_scale > 0.0) {
        double cos_on = o[0] * c->nx + o[1] * c->ny + o[2] * c->nz;
        if (cos_on > 1.0) cos_on = 1.0;
        if (cos_on < -1.0) cos_on = -1.0;
        double u = (1.0 + E_keV / CONE_MC2) * acos(cos_on) / c->angle_scale;
        double target = c->target_weight * (u * u > 1.0 ? u * u : 1.0);
        double keep_u = w / target;
        if (keep_u > 1.0) keep_u = 1.0;
        if (or_stream_uniform(key, 3) >= keep_u) return 0;
        w /= keep_u;
    }
    out[0] = o[0];
    out[1] = o[1];
    out[2] = o[2];
    out[3] = w;
    return 1;
}

int or_cone_event(double dx, double dy, double dz, const or_cone_split *c, double alpha, double E_keV,
                  uint64_t key, double *out) {
    cone_law law = {0};
    law.alpha = alpha;
    return cone_event_law(dx, dy, dz, c, &law, E_keV, key, out);
}

int or_cone_event_tab(double dx, double dy, double dz, const or_cone_split *c, const or_elsepa *t, int64_t el,
                      double E_keV, uint64_t key, double *out) {
    cone_law law = {0};
    law.tab = t;
    or_elsepa_energy_rows(t, el, E_keV, &law.row, &law.f);
    return cone_event_law(dx, dy, dz, c, &law, E_keV, key, out);
}

void or_cone_event_tab_batch(const or_cone_split *c, const or_elsepa *t, int64_t n, const double *d,
                             const int64_t *el, const double *E_keV, const uint64_t *key, int32_t *made,
                             double *out) {
    for (int64_t i = 0; i < n; ++i)
        made[i] = or_cone_event_tab(d[3 * i], d[3 * i + 1], d[3 * i + 2], c, t, el[i], E_keV[i], key[i], out + 4 * i);
}

/* Power-of-two ladder weight: w when w is a power of two, else the next one up. */
double or_cone_ladder_weight(double w) {
    int e = 0;
    double m = frexp(w, &e);
    return m == 0.5 ? w : ldexp(1.0, e);
}

int64_t or_cone_collect(const or_cone_split *c, int64_t e_start, int64_t e_count, int ladder, int64_t out_cap,
                        int64_t *parent, int64_t *flight, double *pos, double *dir, double *E_keV, double *clock,
                        double *weight, uint64_t *stream_key, int64_t *max_count) {
    int64_t o = 0, biggest = 0;
    for (int64_t i = 0; i < e_count; ++i) {
        int64_t cnt = c->count[i];
        if (cnt > biggest) biggest = cnt;
        int64_t n = cnt < c->cap ? cnt : c->cap;
        uint64_t sk = c->split_key[e_start + i];
        for (int64_t k = 0; k < n; ++k) {
            int64_t r = i * c->cap + k;
            double w = c->weight[r], W = w;
            if (ladder) {
                W = or_cone_ladder_weight(w);
                double u = or_stream_uniform(or_cone_event_key(sk, c->flight[r]), 4);
                if (!(u < w / W)) continue;
            }
            if (o < out_cap) {
                parent[o] = i;
                flight[o] = c->flight[r];
                for (int d = 0; d < 3; ++d) {
                    pos[3 * o + d] = c->pos[3 * r + d];
                    dir[3 * o + d] = c->dir[3 * r + d];
                }
                E_keV[o] = c->E_keV[r];
                clock[o] = c->clock[r];
                weight[o] = W;
                stream_key[o] = or_cone_copy_key(sk, c->flight[r]);
            }
            ++o;
        }
    }
    *max_count = biggest;
    return o;
}
