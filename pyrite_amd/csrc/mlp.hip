// mlp.hip -- the importance network (6 -> 64 -> 64 -> 1, SiLU) on gfx950 MFMA.
//
// Parameters: FP32 master vector (4673 floats: W1[64][6], b1[64], W2[64][64],
// b2[64], w3[64], b3) plus a BF16 shadow of W1 (K padded 6 -> 8) and W2 that
// the MFMA operands read.  Activations entering a GEMM (x, h1, dz2) are rounded
// to BF16; every product is accumulated in FP32 by v_mfma_f32_32x32x8bf16_1k.
// A wave owns 32 samples; a 256-lane block owns 128.
//
// MFMA 32x32x8 (bf16) lane layout, wave64, lane l:
//   A (32 x 8, rows x K):  A[l % 32][4*(l/32) + i], i = 0..3
//   B (8 x 32, K x cols):  B[4*(l/32) + i][l % 32]
//   D (32 x 32):           D[8*(i/4) + 4*(l/32) + (i%4)][l % 32], i = 0..15
#include <math.h>

#include "hr_common.hpp"

namespace hr {
namespace {

constexpr int H = 64;
constexpr int NIN = 6;
constexpr int KIN = 8;
constexpr int OW1 = 0, OB1 = OW1 + H * NIN, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H;
constexpr int NPARAM = OB3 + 1;  // 4673
static_assert(NPARAM == 4673, "parameter count");
constexpr int WAVE_SAMPLES = 32;
constexpr int BLOCK = 256;
constexpr int WAVES = BLOCK / 64;
constexpr int BLOCK_SAMPLES = WAVES * WAVE_SAMPLES;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ f32x16 mfma(bf16x4 a, bf16x4 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x4 pack4(const unsigned short *p) {
    bf16x4 v;
    v[0] = (short)p[0];
    v[1] = (short)p[1];
    v[2] = (short)p[2];
    v[3] = (short)p[3];
    return v;
}
__device__ __forceinline__ bf16x4 pack4s(const unsigned short *p, int stride) {
    bf16x4 v;
    v[0] = (short)p[0];
    v[1] = (short)p[stride];
    v[2] = (short)p[2 * stride];
    v[3] = (short)p[3 * stride];
    return v;
}

// LDS layout: weights plus per-wave scratch.  Inference blocks hold 4 waves and
// keep no pre-activations; training blocks hold one wave and keep z1, z2, dz2.
template <int NW, bool TRAIN>
struct Shared {
    unsigned short W1[H * KIN];  // [feature][k]
    unsigned short W2[H * H];    // [out][in]
    float b1[H], b2[H], w3[H];
    float b3;
    unsigned short x[NW][WAVE_SAMPLES * KIN];  // [sample][k]
    unsigned short h1[NW][WAVE_SAMPLES * H];   // [sample][feature]
    float red[NW][WAVE_SAMPLES * H];
    float z1[TRAIN ? NW : 0][TRAIN ? WAVE_SAMPLES * H : 1];
    float z2[TRAIN ? NW : 0][TRAIN ? WAVE_SAMPLES * H : 1];
    unsigned short dz2b[TRAIN ? NW : 0][TRAIN ? WAVE_SAMPLES * H : 1];
};
using FwdShared = Shared<WAVES, false>;
using GradShared = Shared<1, true>;
static_assert(sizeof(FwdShared) <= 64 * 1024, "forward LDS");
static_assert(sizeof(GradShared) <= 64 * 1024, "training LDS");

template <class S>
__device__ void load_weights(S &sh, const unsigned short *__restrict__ w1b,
                             const unsigned short *__restrict__ w2b, const float *__restrict__ p) {
    for (int i = threadIdx.x; i < H * KIN; i += blockDim.x) sh.W1[i] = w1b[i];
    for (int i = threadIdx.x; i < H * H; i += blockDim.x) sh.W2[i] = w2b[i];
    for (int i = threadIdx.x; i < H; i += blockDim.x) {
        sh.b1[i] = p[OB1 + i];
        sh.b2[i] = p[OB2 + i];
        sh.w3[i] = p[OW3 + i];
    }
    if (threadIdx.x == 0) sh.b3 = p[OB3];
}

// Forward of one wave's 32 samples.  Leaves z1, h1, z2 (pre-activation) and
// red = w3 * silu(z2) in LDS; returns y for lanes 0..31.
template <class S, bool TRAIN>
__device__ float forward_wave(S &sh, int wv, int lane, const float *__restrict__ x, int64_t s0, int64_t n) {
    // x -> bf16 [sample][k], k = 6, 7 zero
    if (lane < WAVE_SAMPLES) {
        const int64_t s = s0 + lane;
        for (int k = 0; k < KIN; ++k) {
            const float v = (s < n && k < NIN) ? x[s * NIN + k] : 0.0f;
            sh.x[wv][lane * KIN + k] = f2bf(v);
        }
    }
    __syncthreads();
    const int m = lane & 31, hi = lane >> 5;
    // layer 1
    for (int fb = 0; fb < 2; ++fb) {
        f32x16 acc = {0};
        const bf16x4 a = pack4(&sh.x[wv][m * KIN + 4 * hi]);
        const bf16x4 b = pack4(&sh.W1[(fb * 32 + m) * KIN + 4 * hi]);
        acc = mfma(a, b, acc);
        for (int i = 0; i < 16; ++i) {
            const int row = 8 * (i >> 2) + 4 * hi + (i & 3);
            const int col = fb * 32 + m;
            const float z = acc[i] + sh.b1[col];
            if constexpr (TRAIN) sh.z1[wv][row * H + col] = z;
            sh.h1[wv][row * H + col] = f2bf(silu(z));
        }
    }
    __syncthreads();
    // layer 2
    for (int fb = 0; fb < 2; ++fb) {
        f32x16 acc = {0};
        for (int kk = 0; kk < H / 8; ++kk) {
            const bf16x4 a = pack4(&sh.h1[wv][m * H + kk * 8 + 4 * hi]);
            const bf16x4 b = pack4(&sh.W2[(fb * 32 + m) * H + kk * 8 + 4 * hi]);
            acc = mfma(a, b, acc);
        }
        for (int i = 0; i < 16; ++i) {
            const int row = 8 * (i >> 2) + 4 * hi + (i & 3);
            const int col = fb * 32 + m;
            const float z = acc[i] + sh.b2[col];
            if constexpr (TRAIN) sh.z2[wv][row * H + col] = z;
            sh.red[wv][row * H + col] = sh.w3[col] * silu(z);
        }
    }
    __syncthreads();
    float y = 0.0f;
    if (lane < WAVE_SAMPLES) {
        float s = 0.0f;
        for (int j = 0; j < H; ++j) s += sh.red[wv][lane * H + j];
        y = s + sh.b3;
    }
    return y;
}

__global__ void __launch_bounds__(BLOCK) mlp_forward_kernel(const unsigned short *__restrict__ w1b,
                                                            const unsigned short *__restrict__ w2b,
                                                            const float *__restrict__ params,
                                                            const float *__restrict__ x, int64_t n,
                                                            float *__restrict__ y) {
    __shared__ FwdShared sh;
    load_weights(sh, w1b, w2b, params);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t s0 = (int64_t)blockIdx.x * BLOCK_SAMPLES + wv * WAVE_SAMPLES;
    __syncthreads();
    const float out = forward_wave<FwdShared, false>(sh, wv, lane, x, s0, n);
    if (lane < WAVE_SAMPLES && s0 + lane < n) y[s0 + lane] = out;
}

// Forward + backward of MSE loss; each wave writes its partial gradient
// (NPARAM floats) and its partial loss to partial[wave_global][0..NPARAM].
__global__ void __launch_bounds__(64) mlp_grad_kernel(const unsigned short *__restrict__ w1b,
                                                         const unsigned short *__restrict__ w2b,
                                                         const float *__restrict__ params,
                                                         const float *__restrict__ x,
                                                         const float *__restrict__ t, int64_t n,
                                                         float *__restrict__ partial) {
    __shared__ GradShared sh;
    load_weights(sh, w1b, w2b, params);
    const int wv = 0, lane = threadIdx.x & 63;
    const int64_t s0 = (int64_t)blockIdx.x * WAVE_SAMPLES;
    __syncthreads();
    const float y = forward_wave<GradShared, true>(sh, wv, lane, x, s0, n);
    float *g = partial + (int64_t)blockIdx.x * (NPARAM + 1);
    // dy, loss; red[s][0] <- dy (only lanes < 32 hold samples)
    float r2 = 0.0f;
    if (lane < WAVE_SAMPLES) {
        const int64_t s = s0 + lane;
        float dy = 0.0f;
        if (s < n) {
            const float r = y - t[s];
            r2 = r * r;
            dy = 2.0f * r / (float)n;
        }
        sh.red[wv][lane * H] = dy;
    }
    __syncthreads();
    // dz2 = dy * w3 * silu'(z2) (FP32 and BF16 copies); gw3 partial via red
    for (int idx = lane; idx < WAVE_SAMPLES * H; idx += 64) {
        const int s = idx / H, j = idx % H;
        const float dy = sh.red[wv][s * H];
        const float z = sh.z2[wv][idx];
        const float d = dy * sh.w3[j] * dsilu(z);
        sh.dz2b[wv][idx] = f2bf(d);
    }
    __syncthreads();
    // per-feature sums over the wave's samples: gw3, gb2; gb3 and loss
    {
        const int j = lane;  // 64 lanes = 64 features
        float sw3 = 0.0f, sb2 = 0.0f;
        for (int s = 0; s < WAVE_SAMPLES; ++s) {
            const float dy = sh.red[wv][s * H];
            const float z = sh.z2[wv][s * H + j];
            sw3 += dy * silu(z);
            sb2 += dy * sh.w3[j] * dsilu(z);
        }
        g[OW3 + j] = sw3;
        g[OB2 + j] = sb2;
        if (lane == 0) {
            float sb3 = 0.0f;
            for (int s = 0; s < WAVE_SAMPLES; ++s) sb3 += sh.red[wv][s * H];
            g[OB3] = sb3;
        }
    }
    // loss partial (sum of r^2 over the wave's lanes 0..31)
    for (int off = 32; off > 0; off >>= 1) r2 += __shfl_down(r2, off);
    if (lane == 0) g[NPARAM] = r2;
    __syncthreads();  // every read of the dy slots precedes the dz1 writes below
    const int m = lane & 31, hi = lane >> 5;
    // gW2[j][k] = sum_s dz2[s][j] h1[s][k]; M = j, N = k, K = samples
    for (int jb = 0; jb < 2; ++jb) {
        for (int kb = 0; kb < 2; ++kb) {
            f32x16 acc = {0};
            for (int ks = 0; ks < WAVE_SAMPLES / 8; ++ks) {
                const int sbase = ks * 8 + 4 * hi;
                const bf16x4 a = pack4s(&sh.dz2b[wv][sbase * H + jb * 32 + m], H);
                const bf16x4 b = pack4s(&sh.h1[wv][sbase * H + kb * 32 + m], H);
                acc = mfma(a, b, acc);
            }
            for (int i = 0; i < 16; ++i) {
                const int row = 8 * (i >> 2) + 4 * hi + (i & 3);
                g[OW2 + (jb * 32 + row) * H + kb * 32 + m] = acc[i];
            }
        }
    }
    // dh1[s][k] = sum_j dz2[s][j] W2[j][k]; M = s, N = k, K = j  -> dz1 in z1 slot
    for (int kb = 0; kb < 2; ++kb) {
        f32x16 acc = {0};
        for (int kk = 0; kk < H / 8; ++kk) {
            const bf16x4 a = pack4(&sh.dz2b[wv][m * H + kk * 8 + 4 * hi]);
            const bf16x4 b = pack4s(&sh.W2[(kk * 8 + 4 * hi) * H + kb * 32 + m], H);
            acc = mfma(a, b, acc);
        }
        for (int i = 0; i < 16; ++i) {
            const int row = 8 * (i >> 2) + 4 * hi + (i & 3);
            const int col = kb * 32 + m;
            const float z = sh.z1[wv][row * H + col];
            sh.red[wv][row * H + col] = acc[i] * dsilu(z);  // dz1 (dy slots already consumed)
        }
    }
    __syncthreads();
    // gW1[n][k] = sum_s dz1[s][n] x[s][k], gb1[n] = sum_s dz1[s][n]
    {
        const int nf = lane;
        float acc[NIN] = {0, 0, 0, 0, 0, 0};
        float sb1 = 0.0f;
        for (int s = 0; s < WAVE_SAMPLES; ++s) {
            const float d = (s0 + s < n) ? sh.red[wv][s * H + nf] : 0.0f;
            sb1 += d;
            for (int k = 0; k < NIN; ++k) acc[k] += d * bf2f(sh.x[wv][s * KIN + k]);
        }
        for (int k = 0; k < NIN; ++k) g[OW1 + nf * NIN + k] = acc[k];
        g[OB1 + nf] = sb1;
    }
}

// grad[p] = sum_w partial[w][p], fixed order; loss = sum_w partial[w][NPARAM] / n
__global__ void __launch_bounds__(256) mlp_reduce_kernel(const float *__restrict__ partial, int64_t n_waves,
                                                         int64_t n, float *__restrict__ grad, float *__restrict__ loss) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > NPARAM) return;
    double s = 0.0;
    for (int64_t w = 0; w < n_waves; ++w) s += partial[w * (NPARAM + 1) + p];
    if (p < NPARAM) grad[p] = (float)s;
    else if (loss) *loss = (float)(s / (double)n);
}

// Adam (FP32 master), then refresh the BF16 shadows.
__global__ void __launch_bounds__(256) mlp_adam_kernel(float *__restrict__ params, const float *__restrict__ grad,
                                                       float *__restrict__ m1, float *__restrict__ m2, float lr,
                                                       float beta1, float beta2, float eps, int step,
                                                       unsigned short *__restrict__ w1b, unsigned short *__restrict__ w2b) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    const float gr = grad[p];
    const float a = beta1 * m1[p] + (1.0f - beta1) * gr;
    const float b = beta2 * m2[p] + (1.0f - beta2) * gr * gr;
    m1[p] = a;
    m2[p] = b;
    const float ah = a / (1.0f - powf(beta1, (float)step));
    const float bh = b / (1.0f - powf(beta2, (float)step));
    const float v = params[p] - lr * ah / (sqrtf(bh) + eps);
    params[p] = v;
    if (p < OB1) {
        const int f = p / NIN, k = p % NIN;
        w1b[f * KIN + k] = f2bf(v);
    } else if (p >= OW2 && p < OB2) {
        w2b[p - OW2] = f2bf(v);
    }
}

// BF16 shadows from the master vector (K padding of W1 zeroed).
__global__ void mlp_shadow_kernel(const float *__restrict__ params, unsigned short *__restrict__ w1b,
                                  unsigned short *__restrict__ w2b) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < H * KIN) {
        const int f = p / KIN, k = p % KIN;
        w1b[p] = k < NIN ? f2bf(params[OW1 + f * NIN + k]) : (unsigned short)0;
    }
    if (p < H * H) w2b[p] = f2bf(params[OW2 + p]);
}

}  // namespace
}  // namespace hr

extern "C" {

int32_t hr_mlp_param_count(void) { return hr::NPARAM; }

size_t hr_mlp_workspace_bytes(int64_t n) {
    const int64_t waves = (n + hr::WAVE_SAMPLES - 1) / hr::WAVE_SAMPLES;
    return (size_t)waves * (hr::NPARAM + 1) * sizeof(float);
}

hr_status hr_mlp_shadow(const float *params_d, uint16_t *w1b_d, uint16_t *w2b_d, void *stream) {
    HR_TRACE("hr_mlp_shadow");
    HR_REQUIRE(params_d && w1b_d && w2b_d, "hr_mlp_shadow: null pointer");
    hipLaunchKernelGGL(hr::mlp_shadow_kernel, dim3((hr::H * hr::H + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, params_d, (unsigned short *)w1b_d, (unsigned short *)w2b_d);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

hr_status hr_mlp_forward(const hr_mlp *net, const float *x_d, int64_t n, float *y_d, void *stream) {
    HR_TRACE("hr_mlp_forward");
    using namespace hr;
    HR_REQUIRE(net && net->params && net->w1b && net->w2b, "hr_mlp_forward: null network");
    HR_REQUIRE(n >= 0, "hr_mlp_forward: negative n");
    if (n == 0) return HR_OK;
    HR_REQUIRE(x_d && y_d, "hr_mlp_forward: null x or y");
    const int64_t blocks = (n + BLOCK_SAMPLES - 1) / BLOCK_SAMPLES;
    HR_REQUIRE(blocks < (1LL << 31), "hr_mlp_forward: batch too large");
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)blocks), dim3(BLOCK), 0,
                       (hipStream_t)stream, (const unsigned short *)net->w1b, (const unsigned short *)net->w2b,
                       net->params, x_d, n, y_d);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

hr_status hr_mlp_grad(const hr_mlp *net, const float *x_d, const float *t_d, int64_t n, float *workspace_d,
                      float *grad_d, float *loss_d, void *stream) {
    HR_TRACE("hr_mlp_grad");
    using namespace hr;
    HR_REQUIRE(net && net->params && net->w1b && net->w2b, "hr_mlp_grad: null network");
    HR_REQUIRE(n > 0 && x_d && t_d && workspace_d && grad_d, "hr_mlp_grad: bad arguments");
    const int64_t blocks = (n + WAVE_SAMPLES - 1) / WAVE_SAMPLES;
    HR_REQUIRE(blocks < (1LL << 31), "hr_mlp_grad: batch too large");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(mlp_grad_kernel, dim3((unsigned)blocks), dim3(64), 0, s,
                       (const unsigned short *)net->w1b, (const unsigned short *)net->w2b, net->params, x_d, t_d,
                       n, workspace_d);
    HR_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, s, workspace_d,
                       blocks, n, grad_d, loss_d);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

hr_status hr_mlp_adam(hr_mlp *net, const float *grad_d, float *m1_d, float *m2_d, float lr, int32_t step,
                      void *stream) {
    HR_TRACE("hr_mlp_adam");
    using namespace hr;
    HR_REQUIRE(net && net->params && grad_d && m1_d && m2_d && step >= 1, "hr_mlp_adam: bad arguments");
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       net->params, grad_d, m1_d, m2_d, lr, 0.9f, 0.999f, 1e-8f, (int)step,
                       (unsigned short *)net->w1b, (unsigned short *)net->w2b);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}
}
