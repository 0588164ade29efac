/*
 * or_mlp.c -- FP64 CPU oracle of the importance MLP (6 -> 64 -> 64 -> 1, SiLU).
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  The reference has no learned
 * component (its closest variance reduction is detector-cone splitting,
 * src/pyrite/_cone_splitting.py); this network and its training follow the
 * definitions below, which the device kernels implement with BF16 operands on
 * MFMA and FP32 accumulation:
 *
 *   parameters (FP32 master, 4673 floats, this order):
 *     W1[64][6], b1[64], W2[64][64], b2[64], w3[64], b3
 *   forward, per sample (the device rounds x, W1, W2 and h1 to bfloat16):
 *     z1 = W1 x + b1, h1 = silu(z1), z2 = W2 h1 + b2, h2 = silu(z2),
 *     y  = w3 . h2 + b3,            silu(z) = z / (1 + exp(-z))
 *   loss: L = (1/N) sum_s (y_s - t_s)^2, gradient by the chain rule with the
 *   bfloat16 roundings treated as identity (straight-through).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "hr_oracle.h"

#define H 64
#define NIN 6

/* 1: round x, W1, W2 and h1 to bfloat16 as the device does; 0: exact FP64
 * (lets the tests pin the chain rule by finite differences). */
static int g_round = 1;
void or_mlp_set_rounding(int on) { g_round = on; }

static double rnd(double v);

/* round-to-nearest-even float -> bfloat16 -> float */
double or_bf16(double v) {
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return f; /* inf / nan */
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return f;
}

static double rnd(double v) { return g_round ? or_bf16(v) : v; }

static double silu(double z) { return z / (1.0 + exp(-z)); }
static double dsilu(double z) {
    double s = 1.0 / (1.0 + exp(-z));
    return s * (1.0 + z * (1.0 - s));
}

typedef struct {
    const float *W1, *b1, *W2, *b2, *w3, *b3;
} params_t;

static params_t split(const float *p) {
    params_t q;
    q.W1 = p;
    q.b1 = q.W1 + H * NIN;
    q.W2 = q.b1 + H;
    q.b2 = q.W2 + H * H;
    q.w3 = q.b2 + H;
    q.b3 = q.w3 + H;
    return q;
}

static void forward_one(const params_t *q, const float *x, double *z1, double *h1, double *z2,
                        double *h2, double *y) {
    double xb[NIN];
    for (int k = 0; k < NIN; ++k) xb[k] = rnd(x[k]);
    for (int n = 0; n < H; ++n) {
        double s = 0.0;
        for (int k = 0; k < NIN; ++k) s += rnd(q->W1[n * NIN + k]) * xb[k];
        z1[n] = s + q->b1[n];
        h1[n] = rnd(silu(z1[n]));
    }
    for (int n = 0; n < H; ++n) {
        double s = 0.0;
        for (int k = 0; k < H; ++k) s += rnd(q->W2[n * H + k]) * h1[k];
        z2[n] = s + q->b2[n];
        h2[n] = silu(z2[n]);
    }
    double s = 0.0;
    for (int n = 0; n < H; ++n) s += (double)q->w3[n] * h2[n];
    *y = s + q->b3[0];
}

void or_mlp_forward(const float *params, const float *x, int64_t n, double *y) {
    params_t q = split(params);
    double z1[H], h1[H], z2[H], h2[H];
    for (int64_t s = 0; s < n; ++s) forward_one(&q, x + NIN * s, z1, h1, z2, h2, &y[s]);
}

double or_mlp_loss_grad(const float *params, const float *x, const float *t, int64_t n,
                        double *grad) {
    params_t q = split(params);
    double *gW1 = grad, *gb1 = gW1 + H * NIN, *gW2 = gb1 + H, *gb2 = gW2 + H * H;
    double *gw3 = gb2 + H, *gb3 = gw3 + H;
    memset(grad, 0, sizeof(double) * (H * NIN + H + H * H + H + H + 1));
    double loss = 0.0;
    double z1[H], h1[H], z2[H], h2[H], y, dz2[H], dz1[H];
    for (int64_t s = 0; s < n; ++s) {
        const float *xs = x + NIN * s;
        forward_one(&q, xs, z1, h1, z2, h2, &y);
        double r = y - t[s];
        loss += r * r;
        double dy = 2.0 * r / (double)n;
        *gb3 += dy;
        for (int j = 0; j < H; ++j) {
            gw3[j] += dy * h2[j];
            dz2[j] = dy * q.w3[j] * dsilu(z2[j]);
            gb2[j] += dz2[j];
        }
        for (int j = 0; j < H; ++j)
            for (int k = 0; k < H; ++k) gW2[j * H + k] += dz2[j] * h1[k];
        for (int k = 0; k < H; ++k) {
            double dh = 0.0;
            for (int j = 0; j < H; ++j) dh += dz2[j] * rnd(q.W2[j * H + k]);
            dz1[k] = dh * dsilu(z1[k]);
            gb1[k] += dz1[k];
            for (int m = 0; m < NIN; ++m) gW1[k * NIN + m] += dz1[k] * rnd(xs[m]);
        }
    }
    return loss / (double)n;
}
