// Learned initialiser: a small convolutional network that maps an observed
// image to a diagonal Gaussian over the unconstrained lens parameters.  It
// has no counterpart in the reference; it seeds the multi-start MAP of
// ModellingSequence.MAP (src/gigalens/tf/inference.py:17-41), whose prior
// draws stay the fallback.
//
// Architecture: 4 x (3x3 conv, pad 1, bias, tanh-GELU), strides 1,2,1,2,
// global average pool, Linear(C4 -> hidden) + GELU, Linear(hidden -> 2 d)
// giving (mean, log-std).  Activations are NHWC bfloat16.  The convolution
// passes (cnn_conv.hip) run on BF16 MFMA (v_mfma_f32_16x16x32_bf16, FP32
// accumulation); this file holds the bf16 helpers, the input transform, the
// head, the Gaussian NLL, the reductions and AdamW (FP32 on the vector ALUs).
//
// Rounding points (the CPU oracle applies the same ones): the input
// asinh(img / scale), the weights, Z = conv + bias and Y = gelu(Z) are
// rounded to bfloat16; dZ = dY * gelu'(Z) is rounded to bfloat16; dY, the
// weight gradients and the head are FP32.
#pragma once
#include <hip/hip_runtime.h>

namespace gl {
namespace cnn {

typedef unsigned short u16;
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned short v8u __attribute__((ext_vector_type(8)));

__host__ __device__ inline u16 f2bf(float f) {
    unsigned int u = __builtin_bit_cast(unsigned int, f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return 0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);                                           // round to nearest even
    return (u16)(u >> 16);
}
__host__ __device__ inline float bf2f(u16 h) { return __builtin_bit_cast(float, (unsigned int)h << 16); }

constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2/pi)
constexpr float kGeluA = 0.044715f;
__device__ inline float gelu(float x) {
    const float u = kGeluC * (x + kGeluA * x * x * x);
    return 0.5f * x * (1.0f + tanhf(u));
}
__device__ inline float gelu_grad(float x) {
    const float u = kGeluC * (x + kGeluA * x * x * x);
    const float t = tanhf(u);
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * kGeluC * (1.0f + 3.0f * kGeluA * x * x);
}

// ---------------------------------------------------------------------------
// Small kernels
// ---------------------------------------------------------------------------
// X0 = bf16(asinh(img / scale)), NHWC with one channel
__global__ void input_kernel(const float* img, u16* x0, long n, float inv_scale) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x0[i] = f2bf(asinhf(img[i] * inv_scale));
}

// split-K reduction of the weight gradient: part[S][cout][9 cin + 1] -> grad
// block 256 = kWrOut outputs x kWrLanes split lanes: lane l sums splits l, l + 32, ...
// of its output in order, then the lanes are added in a fixed tree (deterministic)
constexpr int kWrOut = 8, kWrLanes = 32;
__global__ __launch_bounds__(256) void wgrad_reduce(const float* part, int splits, int cout, int ncol, float* gw,
                                                    float* gb) {
    __shared__ float red[kWrLanes][kWrOut];
    const int n = cout * ncol;
    const int o = threadIdx.x % kWrOut, l = threadIdx.x / kWrOut;
    const int i = blockIdx.x * kWrOut + o;
    float s = 0.0f;
    if (i < n)
        for (int q = l; q < splits; q += kWrLanes) s += part[(size_t)q * n + i];
    red[l][o] = s;
    __syncthreads();
    for (int w = kWrLanes / 2; w > 0; w >>= 1) {
        if (l < w) red[l][o] += red[l + w][o];
        __syncthreads();
    }
    if (l == 0 && i < n) {
        const float t = red[0][o];
        const int co = i / ncol, kk = i - co * ncol;
        if (kk < ncol - 1) gw[co * (ncol - 1) + kk] = t;
        else gb[co] = t;
    }
}

// bf16 GEMM operands from the FP32 master weights W[co][9 cin]:
// wt[co][kp] (zero padded to kp) and w2t[ci][(kh*3+kw)*cout + co]
__global__ void pack_weights(const float* w, int cin, int cout, int kp, u16* wt, u16* w2t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cout * kp) return;
    const int co = i / kp, kk = i - co * kp;
    const int kv = 9 * cin;
    const float v = kk < kv ? w[co * kv + kk] : 0.0f;
    wt[i] = f2bf(v);
    if (kk < kv && w2t) {
        const int khw = kk / cin, ci = kk - khw * cin;
        w2t[(size_t)ci * 9 * cout + khw * cout + co] = f2bf(v);
    }
}

// Head, one workgroup (256 threads) per image: global average pool over P
// pixels of C channels, Linear(C -> Hd) + GELU, Linear(Hd -> 2 d).  With a
// target: Gaussian NLL per image (log-std clamped to [-7, 5]), its gradient
// (scaled by inv_bs), the head's per-image back-propagation and dZ of the
// last convolution.
struct HeadArgs {
    const u16* y4;  // [bs][P][C]
    const u16* z4;  // [bs][P][C]
    int P, C, Hd, d;
    const float* w1;  // [Hd][C]
    const float* b1;
    const float* w2;  // [2d][Hd]
    const float* b2;
    float* h0;    // [bs][C]
    float* h1;    // [bs][Hd]
    float* da1;   // [bs][Hd]
    float* out;   // [bs][2d]
    float* dout;  // [bs][2d]
    float* lossb; // [bs]
    u16* dz4;     // [bs][P][C]
    const float* target;  // [bs][d] or null
    float inv_bs;
};

constexpr int kHeadMaxC = 256, kHeadMaxHd = 512, kHeadMaxOut = 256;

__global__ __launch_bounds__(256) void head_kernel(const HeadArgs A) {
    __shared__ float s_h0[kHeadMaxC];
    __shared__ float s_part[4][kHeadMaxC];
    __shared__ float s_a1[kHeadMaxHd];
    __shared__ float s_h1[kHeadMaxHd];
    __shared__ float s_out[kHeadMaxOut];
    __shared__ float s_dout[kHeadMaxOut];
    __shared__ float s_da1[kHeadMaxHd];
    __shared__ float s_dh0[kHeadMaxC];
    __shared__ float s_red[256];
    const int b = blockIdx.x, t = threadIdx.x;
    const int P = A.P, C = A.C, Hd = A.Hd, d = A.d;
    const u16* y = A.y4 + (size_t)b * P * C;
    // global average pool: 4 pixel groups x 64 channel lanes
    for (int c0 = 0; c0 < C; c0 += 64) {
        const int c = c0 + (t & 63), g = t >> 6;
        float s = 0.0f;
        if (c < C)
            for (int p = g; p < P; p += 4) s += bf2f(y[(size_t)p * C + c]);
        if (c < C) s_part[g][c] = s;
    }
    __syncthreads();
    for (int c = t; c < C; c += 256) {
        const float h = (s_part[0][c] + s_part[1][c] + s_part[2][c] + s_part[3][c]) / (float)P;
        s_h0[c] = h;
        A.h0[(size_t)b * C + c] = h;
    }
    __syncthreads();
    for (int j = t; j < Hd; j += 256) {
        float a = A.b1[j];
        for (int c = 0; c < C; ++c) a = fmaf(A.w1[j * C + c], s_h0[c], a);
        s_a1[j] = a;
        const float h = gelu(a);
        s_h1[j] = h;
        A.h1[(size_t)b * Hd + j] = h;
    }
    __syncthreads();
    for (int o = t; o < 2 * d; o += 256) {
        float a = A.b2[o];
        for (int j = 0; j < Hd; ++j) a = fmaf(A.w2[o * Hd + j], s_h1[j], a);
        s_out[o] = a;
        A.out[(size_t)b * 2 * d + o] = a;
    }
    __syncthreads();
    if (!A.target) return;
    // Gaussian NLL: sum_o [ls + 0.5 (t - mu)^2 exp(-2 ls)], ls = clamp(raw, -7, 5)
    float l = 0.0f;
    for (int o = t; o < d; o += 256) {
        const float mu = s_out[o], raw = s_out[d + o];
        const float ls = fminf(fmaxf(raw, -7.0f), 5.0f);
        const float r = A.target[(size_t)b * d + o] - mu;
        const float iv = expf(-2.0f * ls);
        l += ls + 0.5f * r * r * iv;
        s_dout[o] = -r * iv * A.inv_bs;
        const float inside = (raw > -7.0f && raw < 5.0f) ? 1.0f : 0.0f;
        s_dout[d + o] = (1.0f - r * r * iv) * inside * A.inv_bs;
    }
    s_red[t] = l;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) s_red[t] += s_red[t + w];
        __syncthreads();
    }
    if (t == 0) A.lossb[b] = s_red[0];
    for (int o = t; o < 2 * d; o += 256) A.dout[(size_t)b * 2 * d + o] = s_dout[o];
    for (int j = t; j < Hd; j += 256) {
        float g = 0.0f;
        for (int o = 0; o < 2 * d; ++o) g = fmaf(A.w2[o * Hd + j], s_dout[o], g);
        const float da = g * gelu_grad(s_a1[j]);
        s_da1[j] = da;
        A.da1[(size_t)b * Hd + j] = da;
    }
    __syncthreads();
    for (int c = t; c < C; c += 256) {
        float g = 0.0f;
        for (int j = 0; j < Hd; ++j) g = fmaf(A.w1[j * C + c], s_da1[j], g);
        s_dh0[c] = g / (float)P;
    }
    __syncthreads();
    const u16* z = A.z4 + (size_t)b * P * C;
    u16* dz = A.dz4 + (size_t)b * P * C;
    for (int i = t; i < P * C; i += 256) {
        const int c = i % C;
        dz[i] = f2bf(s_dh0[c] * gelu_grad(bf2f(z[i])));
    }
}

// Head weight gradients: sums over the batch of the per-image outer products,
// in two passes.  Pass 1 (grid x: weights, y: kHeadChunks image ranges) writes
// each range's sum to part[chunk][weight]; pass 2 adds the ranges in order, so
// the result does not depend on scheduling.  Weight index i runs over
// [gw1 (Hd x C) | gb1 (Hd) | gw2 (2d x Hd) | gb2 (2d)].
constexpr int kHeadChunks = 16;
__global__ void head_wgrad(const HeadArgs A, int bs, float* part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int C = A.C, Hd = A.Hd, d2 = 2 * A.d;
    const int n1 = Hd * C, n2 = Hd, n3 = d2 * Hd, n4 = d2;
    const int n = n1 + n2 + n3 + n4;
    if (i >= n) return;
    const int per = (bs + kHeadChunks - 1) / kHeadChunks;
    const int b0 = blockIdx.y * per, b1 = min(bs, b0 + per);
    float s = 0.0f;
    if (i < n1) {
        const int j = i / C, c = i - j * C;
        for (int b = b0; b < b1; ++b) s = fmaf(A.da1[(size_t)b * Hd + j], A.h0[(size_t)b * C + c], s);
    } else if (i < n1 + n2) {
        const int j = i - n1;
        for (int b = b0; b < b1; ++b) s += A.da1[(size_t)b * Hd + j];
    } else if (i < n1 + n2 + n3) {
        const int q = i - n1 - n2;
        const int o = q / Hd, j = q - o * Hd;
        for (int b = b0; b < b1; ++b) s = fmaf(A.dout[(size_t)b * d2 + o], A.h1[(size_t)b * Hd + j], s);
    } else {
        const int o = i - n1 - n2 - n3;
        for (int b = b0; b < b1; ++b) s += A.dout[(size_t)b * d2 + o];
    }
    part[(size_t)blockIdx.y * n + i] = s;
}

__global__ void head_wgrad_reduce(const HeadArgs A, const float* part, float* gw1, float* gb1, float* gw2,
                                  float* gb2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int C = A.C, Hd = A.Hd, d2 = 2 * A.d;
    const int n1 = Hd * C, n2 = Hd, n3 = d2 * Hd, n4 = d2;
    const int n = n1 + n2 + n3 + n4;
    if (i >= n) return;
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < kHeadChunks; ++c) s += part[(size_t)c * n + i];
    if (i < n1) gw1[i] = s;
    else if (i < n1 + n2) gb1[i - n1] = s;
    else if (i < n1 + n2 + n3) gw2[i - n1 - n2] = s;
    else gb2[i - n1 - n2 - n3] = s;
}

// mean of the per-image losses (one workgroup)
__global__ void loss_mean(const float* lossb, int bs, float* loss) {
    __shared__ float r[256];
    float s = 0.0f;
    for (int b = threadIdx.x; b < bs; b += 256) s += lossb[b];
    r[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) r[threadIdx.x] += r[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = r[0] / (float)bs;
}

// AdamW with torch.optim.AdamW semantics: p *= 1 - lr wd; m, v moments;
// p -= lr / (1 - b1^t) * m / (sqrt(v) / sqrt(1 - b2^t) + eps)
__global__ void adamw(float* p, float* m, float* v, const float* g, int n, float lr, float b1, float b2, float eps,
                      float wd, float bc1, float bc2_sqrt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    float pi = p[i] * (1.0f - lr * wd);
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    pi -= (lr / bc1) * mi / (sqrtf(vi) / bc2_sqrt + eps);
    p[i] = pi;
}

}  // namespace cnn
}  // namespace gl
