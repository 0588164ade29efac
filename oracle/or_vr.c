/*
 * or_vr.c -- CPU restatement of the learned variance reduction's weight window,
 * child keys and line-mass labels (TEST INFRASTRUCTURE ONLY).
 *
 * The variance reduction has no counterpart in the reference (its variance
 * reduction is cone splitting, src/pyrite/_cone_splitting.py:57, and line
 * roulette, src/pyrite/_line_roulette.py:29); these functions define the
 * decisions the device must reproduce bit for bit (counts, keys, roulette
 * survivors) or to one rounding (weights).  Pins: weight conservation (every split conserves w
 * exactly, every roulette conserves it in expectation), the SplitMix64 stream
 * (pinned by the reference's known-answer keys), and the line mass against a
 * brute-force fine-grid sum of sinc^2.
 */
#include <math.h>
#include <stdint.h>

#include "hr_oracle.h"

#define OR_DECISION_COUNTER (1ULL << 40)
#define OR_SPAWN_SALT 0x6A09E667F3BCC909ULL
#define OR_GOLDEN 0x9E3779B97F4A7C15ULL

uint64_t or_vr_child_key(uint64_t key, int64_t j) {
    return or_splitmix64(or_splitmix64(key ^ OR_SPAWN_SALT) + OR_GOLDEN * (uint64_t)(j + 1));
}

void or_vr_window(int64_t n, const int8_t *exit_code, const float *log_importance, const double *weight,
                  const uint64_t *key, double log_I_ref, double w_min, double w_max, double ratio_lo,
                  double ratio_hi, int32_t n_max, int32_t *count, double *w_child, const double *weight2,
                  double *w2_child) {
    for (int64_t i = 0; i < n; ++i) {
        int32_t c = 0;
        double wc = 0.0, w2c = 0.0;
        const double w2 = weight2 ? weight2[i] : 0.0;
        if (exit_code[i] == 6) {
            const double w = weight[i];
            double wt = exp(log_I_ref - (double)log_importance[i]);
            wt = fmin(fmax(wt, w_min), w_max);
            const double r = w / wt;
            const double u = or_stream_uniform(key[i], OR_DECISION_COUNTER);
            if (r > ratio_hi) {
                double fl = floor(r);
                if (u < r - fl) fl += 1.0;
                c = (int32_t)(fl < (double)n_max ? fl : (double)n_max);
                wc = w / (double)c;
                w2c = w2;
            } else if (r >= ratio_lo) {
                c = 1;
                wc = w;
                w2c = w2;
            } else if (u < r) {
                c = 1;
                wc = wt;
                w2c = w2 / r;
            }
        }
        count[i] = c;
        w_child[i] = wc;
        if (w2_child) w2_child[i] = w2c;
    }
}

void or_line_roulette(float *w, int64_t n_seg, int32_t n_dir, int32_t n_g, const double *mass, double m_thr,
                      uint64_t key, int64_t row0) {
    for (int64_t j = 0; j < n_seg; ++j) {
        const double m = mass[j];
        if (m >= m_thr) continue;
        const double p = m > 0.0 ? m / m_thr : 0.0;
        const double u = or_stream_uniform(key, (uint64_t)(row0 + j));
        const float scale = u < p ? (float)(1.0 / p) : 0.0f;
        for (int d = 0; d < n_dir; ++d)
            for (int g = 0; g < n_g; ++g) {
                const int64_t k = ((int64_t)d * n_seg + j) * n_g + g;
                w[k] = w[k] * scale;
            }
    }
}

void or_line_mass(const double *E_r, const double *a, const float *w, int64_t n_seg, int32_t n_dir,
                  int32_t n_g, double lo, double hi, double dE, double *mass) {
    const double pi = 3.141592653589793;
    for (int64_t j = 0; j < n_seg; ++j) {
        double m = 0.0;
        for (int d = 0; d < n_dir; ++d)
            for (int g = 0; g < n_g; ++g) {
                const int64_t k = ((int64_t)d * n_seg + j) * n_g + g;
                const double wk = (double)w[k];
                if (wk != 0.0 && a[k] > 0.0 && E_r[k] >= lo && E_r[k] < hi) m += wk * pi * (1.0 / dE) / a[k];
            }
        mass[j] = m;
    }
}
