// Convolution passes of the learned initialiser's network as band-tiled direct
// convolutions.  A workgroup takes one image and a band of rows, copies the band of
// its input (with the 3x3 halo and zero padding) from the NHWC tensor into LDS with
// contiguous 16-byte loads, and reads the MFMA operands straight from LDS at the
// shifted pixel of each tap, so every input element is fetched from HBM about once.
// LDS rows hold C + 8 bf16 per pixel (16-byte aligned, and 16 consecutive pixels of
// a b128 read fall on distinct banks).
//
//   conv_fwd<S>    Z = conv(X, W) + b, Y = gelu(Z)          rows: output pixels, cols: cout
//   conv_dgrad<S>  dZprev = conv^T(dZ, W) * gelu'(Zprev)      rows: input pixels, cols: cin
//   conv_wgrad<S>  dW = sum_pixels dZ (x) X~, db = sum dZ     rows: cout, cols: (tap, cin)
//   conv1_fwd / conv1_wgrad: the single-channel first layer on the vector ALUs.
//
// v_mfma_f32_16x16x32_bf16 lane layout: lane l supplies row (l % 16) of A and column
// (l % 16) of B with the 8 k-values 8 (l / 16) .. + 7; the result element
// (4 (l / 16) + r, l % 16) sits in lane l, register r.
#pragma once
#include <hip/hip_runtime.h>

#include "cnn_kernels.hip"

namespace gl {
namespace cnn {

__device__ __forceinline__ v8bf ld8(const u16* p) { return __builtin_bit_cast(v8bf, *reinterpret_cast<const v8u*>(p)); }
__device__ __forceinline__ v4f mfma(v8bf a, v8bf b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
__host__ __device__ inline int floordiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

constexpr int kWaves = 8;           // waves of a band workgroup (512 threads)
constexpr int kGPW = 2;             // pixel groups of 16 per wave and pass
constexpr int kGroupsPerPass = kWaves * kGPW;
constexpr int kStageF = 40;         // bf16 row stride of the forward output staging (32 channels + 8)
constexpr int kStageD = 36;         // fp32 row stride of the data-gradient staging (32 channels + 4)

// a / d for 0 <= a < 2^22 through a float reciprocal and one correction step
struct Div {
    int d;
    float inv;
    __device__ explicit Div(int d_) : d(d_), inv(1.0f / (float)d_) {}
    __device__ int q(int a) const {
        int t = (int)((float)a * inv);
        t -= t * d > a;
        t += (t + 1) * d <= a;
        return t;
    }
};

// GELU (tanh form) and its derivative with tanh(u) = 1 - 2 / (exp(2u) + 1) on the
// hardware exp and reciprocal: the results are rounded to bfloat16 right after, so
// the last-bit differences from tanhf change a rounded value only at a tie.
__device__ __forceinline__ float tanh_fast(float u) { return 1.0f - 2.0f * __frcp_rn(__expf(2.0f * u) + 1.0f); }
__device__ __forceinline__ float gelu_f(float x) {
    return 0.5f * x * (1.0f + tanh_fast(kGeluC * (x + kGeluA * x * x * x)));
}
__device__ __forceinline__ float gelu_grad_f(float x) {
    const float t = tanh_fast(kGeluC * (x + kGeluA * x * x * x));
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * kGeluC * (1.0f + 3.0f * kGeluA * x * x);
}

// 16-byte copies into LDS with eight loads in flight per thread: for i in [0, n),
// m(i, &off) returns the source vector (or null for zeros) and sets its LDS offset.
template <class Map>
__device__ __forceinline__ void copy16(int n, u16* dst, const Map& m) {
    constexpr int U = 8;
    for (int i0 = threadIdx.x; i0 < n; i0 += U * blockDim.x) {
        v8u val[U];
        int off[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * blockDim.x;
            val[u] = v8u{0, 0, 0, 0, 0, 0, 0, 0};
            off[u] = -1;
            if (i < n) {
                const u16* src = m(i, off[u]);
                if (src) val[u] = *reinterpret_cast<const v8u*>(src);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (off[u] >= 0) *reinterpret_cast<v8u*>(dst + off[u]) = val[u];
    }
}

// Rows [r_lo, r_lo + nr) x columns [-1, W] of image b of an NHWC tensor (C channels,
// C % 8 == 0) into LDS as [nr][W + 2][C + 8], zeros outside the image.
struct BandMap {
    const u16* src;
    int b, H, W, C, r_lo;
    Div dv, dc;
    __device__ const u16* operator()(int i, int& off) const {
        const int pc = dv.q(i), v = i - pc * dv.d;
        const int r = dc.q(pc), c = pc - r * dc.d;
        off = (r * dc.d + c) * (C + 8) + v * 8;
        const int h = r_lo + r, w = c - 1;
        return (h >= 0 && h < H && w >= 0 && w < W) ? src + (((size_t)b * H + h) * W + w) * C + v * 8 : nullptr;
    }
};
__device__ __forceinline__ void load_band(const u16* src, int b, int H, int W, int C, int r_lo, int nr, u16* dst) {
    const BandMap m{src, b, H, W, C, r_lo, Div(C >> 3), Div(W + 2)};
    copy16(nr * (W + 2) * (C >> 3), dst, m);
}

// Register-staged band copy: issue() starts the (at most kPre per thread) 16-byte
// loads of a band, commit() writes them to LDS, so the loads of the next band can
// be in flight while the current one is computed.
constexpr int kPre = 8;
struct BandPrefetch {
    v8u val[kPre];
    int off[kPre];
    __device__ void issue(const u16* src, int b, int H, int W, int C, int r_lo, int nr) {
        const int vec = C >> 3, n = nr * (W + 2) * vec;
        const BandMap m{src, b, H, W, C, r_lo, Div(vec), Div(W + 2)};
#pragma unroll
        for (int u = 0; u < kPre; ++u) {
            const int i = threadIdx.x + u * blockDim.x;
            val[u] = v8u{0, 0, 0, 0, 0, 0, 0, 0};
            off[u] = -1;
            if (i < n) {
                const u16* p = m(i, off[u]);
                if (p) val[u] = *reinterpret_cast<const v8u*>(p);
            }
        }
    }
    __device__ void commit(u16* dst) const {
#pragma unroll
        for (int u = 0; u < kPre; ++u)
            if (off[u] >= 0) *reinterpret_cast<v8u*>(dst + off[u]) = val[u];
    }
};

// 32-row weight tile [9][32][C + 8] of rows j0 .. j0 + 31 of w[rows][9][C]
struct WeightMap {
    const u16* w;
    int C, j0;
    Div dv;
    __device__ const u16* operator()(int i, int& off) const {
        const int tj = dv.q(i), v = i - tj * dv.d;
        const int tap = tj >> 5, j = tj & 31;
        off = (tap * 32 + j) * (C + 8) + v * 8;
        return w + ((size_t)(j0 + j) * 9 + tap) * C + v * 8;
    }
};
__device__ __forceinline__ void load_weights(const u16* w, int C, int j0, u16* dst) {
    const WeightMap m{w, C, j0, Div(C >> 3)};
    copy16(9 * 32 * (C >> 3), dst, m);
}

// ---------------------------------------------------------------------------
// Forward: persistent workgroups of 512 threads over the (image, band) items, one
// 32-channel tile of cout at a time (weight tile loaded once per tile).  w:
// [cout][9][cin] bf16.  LDS: input band, weight tile, output staging.  The next
// item's band is loaded into registers while the current one is computed; each pass
// of up to 16 pixel groups leaves Z in the staging, then Z and Y = gelu(Z) go out
// as 16-byte stores of 8 channels.
// ---------------------------------------------------------------------------
struct FwdArgs {
    const u16* x;
    const u16* w;
    const float* bias;
    u16* Z;
    u16* Y;
    int B, H, W, cin, cout, Ho, Wo, R;
};

template <int S>
__global__ __launch_bounds__(512) void conv_fwd(const FwdArgs A) {
    extern __shared__ __attribute__((aligned(16))) u16 lds[];
    const int CP = A.cin + 8, NC = A.W + 2;
    const int NRmax = S * (A.R - 1) + 3;
    u16* band = lds;
    u16* wl = band + (size_t)NRmax * NC * CP;  // [9][32][CP]
    u16* stage = wl + (size_t)9 * 32 * CP;     // [<= 256][kStageF]
    const int nb = (A.Ho + A.R - 1) / A.R, items = A.B * nb;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nk = A.cin >> 5;
    const Div dwo(A.Wo);
    BandPrefetch pf;
    for (int ct = 0; ct < A.cout / 32; ++ct) {
        __syncthreads();
        load_weights(A.w, A.cin, ct * 32, wl);
        int item = blockIdx.x;
        if (item < items) {
            const int b = item / nb, ho0 = (item - b * nb) * A.R, R = min(A.R, A.Ho - ho0);
            pf.issue(A.x, b, A.H, A.W, A.cin, S * ho0 - 1, S * (R - 1) + 3);
        }
        while (item < items) {
            const int b = item / nb, ho0 = (item - b * nb) * A.R, R = min(A.R, A.Ho - ho0);
            __syncthreads();
            pf.commit(band);
            __syncthreads();
            const int next = item + gridDim.x;
            if (next < items) {
                const int b2 = next / nb, h2 = (next - b2 * nb) * A.R, R2 = min(A.R, A.Ho - h2);
                pf.issue(A.x, b2, A.H, A.W, A.cin, S * h2 - 1, S * (R2 - 1) + 3);
            }
            const int P = R * A.Wo, G = (P + 15) / 16;
            for (int g0 = 0; g0 < G; g0 += kGroupsPerPass) {
                int boff[kGPW];
                int ng = 0;
#pragma unroll
                for (int gi = 0; gi < kGPW; ++gi) {
                    const int g = g0 + wave + kWaves * gi;
                    boff[gi] = 0;
                    if (g < G && g < g0 + kGroupsPerPass) {
                        const int q = min(g * 16 + (lane & 15), P - 1);
                        const int hr = dwo.q(q), wo = q - hr * A.Wo;
                        boff[gi] = ((S * hr) * NC + S * wo) * CP + 8 * (lane >> 4);
                        ng = gi + 1;
                    }
                }
                v4f acc[kGPW][2];
#pragma unroll
                for (int gi = 0; gi < kGPW; ++gi) acc[gi][0] = acc[gi][1] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
                for (int tap = 0; tap < 9; ++tap) {
                    const int kh = tap / 3, kw = tap - 3 * kh;
                    const int toff = (kh * NC + kw) * CP;
                    for (int kc = 0; kc < nk; ++kc) {
                        const v8bf b0 = ld8(wl + ((size_t)tap * 32 + (lane & 15)) * CP + kc * 32 + 8 * (lane >> 4));
                        const v8bf b1 =
                            ld8(wl + ((size_t)tap * 32 + 16 + (lane & 15)) * CP + kc * 32 + 8 * (lane >> 4));
#pragma unroll
                        for (int gi = 0; gi < kGPW; ++gi) {
                            if (gi < ng) {
                                const v8bf a = ld8(band + boff[gi] + toff + kc * 32);
                                acc[gi][0] = mfma(a, b0, acc[gi][0]);
                                acc[gi][1] = mfma(a, b1, acc[gi][1]);
                            }
                        }
                    }
                }
                // Z = bf16(acc + bias) into the staging, row = pixel within the pass
#pragma unroll
                for (int gi = 0; gi < kGPW; ++gi) {
                    if (gi >= ng) continue;
                    const int g = g0 + wave + kWaves * gi;
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        const int cl = 16 * j + (lane & 15);
                        const float bias = A.bias[ct * 32 + cl];
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            stage[((g - g0) * 16 + 4 * (lane >> 4) + r) * kStageF + cl] = f2bf(acc[gi][j][r] + bias);
                    }
                }
                __syncthreads();
                const int npx = min(P - g0 * 16, kGroupsPerPass * 16);
                for (int i = threadIdx.x; i < npx * 4; i += blockDim.x) {
                    const int slot = i >> 2, v = i & 3;
                    const int q = g0 * 16 + slot;
                    const int hr = dwo.q(q), wo = q - hr * A.Wo;
                    const size_t o = (((size_t)b * A.Ho + ho0 + hr) * A.Wo + wo) * A.cout + ct * 32 + 8 * v;
                    const v8u z = *reinterpret_cast<const v8u*>(stage + slot * kStageF + 8 * v);
                    v8u y;
#pragma unroll
                    for (int e = 0; e < 8; ++e) y[e] = f2bf(gelu_f(bf2f(z[e])));
                    *reinterpret_cast<v8u*>(A.Z + o) = z;
                    *reinterpret_cast<v8u*>(A.Y + o) = y;
                }
                if (g0 + kGroupsPerPass < G) __syncthreads();
            }
            item = next;
        }
    }
}

// ---------------------------------------------------------------------------
// Data gradient: persistent workgroups of 512 threads over the (image, band of input
// rows) items, one 32-channel tile of cin at a time.  w2t: [cin][9][cout] bf16.
// Input pixels (h, w) of the band are taken class by class of (h mod S, w mod S), so
// that every pixel group of 16 uses the same taps: with stride 2 only
// kh = (h + 1) mod 2 (+ 2) reach an output row.  The next item's dZ band is loaded
// into registers during the current item, and each pass's Zprev vectors before its
// MFMAs; the pass leaves its FP32 sums in the staging, then dZprev = bf16(sum *
// gelu'(Zprev)) goes out 8 channels at a time with 16-byte stores.
// ---------------------------------------------------------------------------
struct DgradArgs {
    const u16* dz;   // [B][Ho][Wo][cout]
    const u16* w2t;  // [cin][9][cout]
    const u16* zp;   // [B][H][W][cin]
    u16* dzp;        // [B][H][W][cin]
    int B, H, W, cin, cout, Ho, Wo, R;
};

template <int S>
struct PixelClasses {  // pixel classes c = (ph, pw) of a band of rows [h0, h0 + R)
    int fr[S], nc[S], cnt[S * S], gstart[S * S + 1];  // constant indices only (registers)
    __device__ PixelClasses(int h0, int R, int W) {
        int nr[S];
#pragma unroll
        for (int t = 0; t < S; ++t) {
            fr[t] = ((t - h0) % S + S) % S;
            nr[t] = fr[t] < R ? (R - fr[t] + S - 1) / S : 0;
            nc[t] = (W - t + S - 1) / S;
        }
        gstart[0] = 0;
#pragma unroll
        for (int c = 0; c < S * S; ++c) {
            cnt[c] = nr[c / S] * nc[c % S];
            gstart[c + 1] = gstart[c] + (cnt[c] + 15) / 16;
        }
    }
    __device__ int groups() const { return gstart[S * S]; }
    __device__ int cls(int g) const {
        int c = 0;
#pragma unroll
        for (int t = 1; t < S * S; ++t) c += g >= gstart[t];
        return c;
    }
    __device__ int first_group(int c) const {
        int v = 0;
#pragma unroll
        for (int t = 0; t < S * S; ++t) v = t == c ? gstart[t] : v;
        return v;
    }
    __device__ int count(int c) const {
        int v = 0;
#pragma unroll
        for (int t = 0; t < S * S; ++t) v = t == c ? cnt[t] : v;
        return v;
    }
    // row offset within the band and column of the k-th pixel of class c
    __device__ void pixel(int c, int k, int& hr, int& w) const {
        const int ph = c / S, pw = c % S;
        int f = 0, n = 1;
#pragma unroll
        for (int t = 0; t < S; ++t) {
            f = t == ph ? fr[t] : f;
            n = t == pw ? nc[t] : n;
        }
        const int i = k / n;
        hr = f + S * i;
        w = pw + S * (k - i * n);
    }
};

template <int S>
__global__ __launch_bounds__(512) void conv_dgrad(const DgradArgs A) {
    extern __shared__ __attribute__((aligned(16))) u16 lds[];
    const int CP = A.cout + 8, NC = A.Wo + 2;
    const int NRmax = (A.R + S) / S + 1;  // most dZ rows any band of A.R rows needs
    u16* band = lds;
    u16* wl = band + (size_t)NRmax * NC * CP;                           // [9][32][CP]
    float* stage = reinterpret_cast<float*>(wl + (size_t)9 * 32 * CP);  // [<= 256][kStageD]
    const int nb = (A.H + A.R - 1) / A.R, items = A.B * nb;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nk = A.cout >> 5;
    constexpr int U = 2;  // Zprev vectors per thread and pass: 256 pixels x 4 vectors / 512 threads
    BandPrefetch pf;
    for (int ct = 0; ct < A.cin / 32; ++ct) {
        __syncthreads();
        load_weights(A.w2t, A.cout, ct * 32, wl);
        int item = blockIdx.x;
        if (item < items) {
            const int b = item / nb, h0 = (item - b * nb) * A.R, R = min(A.R, A.H - h0);
            const int lo = floordiv(h0 - 1, S);
            pf.issue(A.dz, b, A.Ho, A.Wo, A.cout, lo, floordiv(h0 + R, S) - lo + 1);
        }
        while (item < items) {
            const int b = item / nb, h0 = (item - b * nb) * A.R, R = min(A.R, A.H - h0);
            const int ho_lo = floordiv(h0 - 1, S);
            __syncthreads();
            pf.commit(band);
            __syncthreads();
            const int next = item + gridDim.x;
            if (next < items) {
                const int b2 = next / nb, h2 = (next - b2 * nb) * A.R, R2 = min(A.R, A.H - h2);
                const int lo2 = floordiv(h2 - 1, S);
                pf.issue(A.dz, b2, A.Ho, A.Wo, A.cout, lo2, floordiv(h2 + R2, S) - lo2 + 1);
            }
            const PixelClasses<S> pc(h0, R, A.W);
            const int G = pc.groups();
            for (int g0 = 0; g0 < G; g0 += kGroupsPerPass) {
                const int nslot = min(G - g0, kGroupsPerPass) * 16;
                // this pass's Zprev vectors (in flight during the MFMAs)
                v8u zv[U];
                size_t zo[U];
                bool zok[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int i = threadIdx.x + u * blockDim.x;
                    const int slot = i >> 2, v = i & 3;
                    zok[u] = false;
                    zo[u] = 0;
                    zv[u] = v8u{0, 0, 0, 0, 0, 0, 0, 0};
                    if (i < nslot * 4) {
                        const int g = g0 + (slot >> 4);
                        const int c = pc.cls(g);
                        const int k = (g - pc.first_group(c)) * 16 + (slot & 15);
                        if (k < pc.count(c)) {
                            int hr, w;
                            pc.pixel(c, k, hr, w);
                            zo[u] = (((size_t)b * A.H + h0 + hr) * A.W + w) * A.cin + ct * 32 + 8 * v;
                            zv[u] = *reinterpret_cast<const v8u*>(A.zp + zo[u]);
                            zok[u] = true;
                        }
                    }
                }
                int boff[kGPW], cl[kGPW];
                int ng = 0;
#pragma unroll
                for (int gi = 0; gi < kGPW; ++gi) {
                    const int g = g0 + wave + kWaves * gi;
                    boff[gi] = 0;
                    cl[gi] = 0;
                    if (g < G && g < g0 + kGroupsPerPass) {
                        const int c = pc.cls(g);
                        const int ph = c / S, pw = c % S;
                        const int k = min((g - pc.first_group(c)) * 16 + (lane & 15), pc.count(c) - 1);
                        int hr, w;
                        pc.pixel(c, k, hr, w);
                        // dZ row of tap kh: (h + 1 - kh) / S = (h - ph) / S + (ph + 1 - kh) / S
                        boff[gi] = (((h0 + hr - ph) / S - ho_lo) * NC + (w - pw) / S + 1) * CP + 8 * (lane >> 4);
                        cl[gi] = c;
                        ng = gi + 1;
                    }
                }
                v4f acc[kGPW][2];
#pragma unroll
                for (int gi = 0; gi < kGPW; ++gi) acc[gi][0] = acc[gi][1] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
                for (int tap = 0; tap < 9; ++tap) {
                    const int kh = tap / 3, kw = tap - 3 * kh;
                    for (int kc = 0; kc < nk; ++kc) {
                        const v8bf b0 = ld8(wl + ((size_t)tap * 32 + (lane & 15)) * CP + kc * 32 + 8 * (lane >> 4));
                        const v8bf b1 =
                            ld8(wl + ((size_t)tap * 32 + 16 + (lane & 15)) * CP + kc * 32 + 8 * (lane >> 4));
#pragma unroll
                        for (int gi = 0; gi < kGPW; ++gi) {
                            if (gi >= ng) continue;
                            const int ph = cl[gi] / S, pw = cl[gi] % S;
                            if ((ph + 1 - kh) % S || (pw + 1 - kw) % S) continue;  // tap misses every output pixel
                            const int dh = floordiv(ph + 1 - kh, S), dw = floordiv(pw + 1 - kw, S);
                            const v8bf a = ld8(band + boff[gi] + (dh * NC + dw) * CP + kc * 32);
                            acc[gi][0] = mfma(a, b0, acc[gi][0]);
                            acc[gi][1] = mfma(a, b1, acc[gi][1]);
                        }
                    }
                }
#pragma unroll
                for (int gi = 0; gi < kGPW; ++gi) {
                    if (gi >= ng) continue;
                    const int g = g0 + wave + kWaves * gi;
#pragma unroll
                    for (int j = 0; j < 2; ++j)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            stage[((g - g0) * 16 + 4 * (lane >> 4) + r) * kStageD + 16 * j + (lane & 15)] =
                                acc[gi][j][r];
                }
                __syncthreads();
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (!zok[u]) continue;
                    const int i = threadIdx.x + u * blockDim.x;
                    const float* sv = stage + (i >> 2) * kStageD + 8 * (i & 3);
                    v8u o;
#pragma unroll
                    for (int e = 0; e < 8; ++e) o[e] = f2bf(sv[e] * gelu_grad_f(bf2f(zv[u][e])));
                    *reinterpret_cast<v8u*>(A.dzp + zo[u]) = o;
                }
                if (g0 + kGroupsPerPass < G) __syncthreads();
            }
            item = next;
        }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: grid (blocks over images, tile chunks), 512 threads.  Output
// tiles are (co group of 16) x (tap, ci group of 16) plus one bias tile per co
// group (B = ones); a chunk covers co groups [cg0, cg0 + ncg) (ncg <= 4) and pairs
// [p0, p0 + np) of the 9 cin / 16 + 1 (tap, ci group) pairs, wave w taking pairs
// p0 + w, p0 + w + kWaves, ... (<= kMaxPairs).  Pixels are the reduction dimension:
// k-steps of 32 pixels of the band; the operands gather 8 pixels per lane from LDS.
// part[blockIdx.x][cout][9 cin + 1]: per-block sums, added by wgrad_reduce.
// ---------------------------------------------------------------------------
constexpr int kMaxPairs = 3;

struct WgradArgs {
    const u16* dz;  // [B][Ho][Wo][cout]
    const u16* x;   // [B][H][W][cin]
    float* part;
    int B, H, W, cin, cout, Ho, Wo, R;
    int ncg, npairs;  // per chunk; gridDim.y = co chunks x pair chunks
};

struct RowsMap {  // P consecutive pixels of C channels into LDS rows of C + 8
    const u16* src;
    int C;
    Div dv;
    __device__ const u16* operator()(int i, int& off) const {
        const int q = dv.q(i), v = i - q * dv.d;
        off = q * (C + 8) + v * 8;
        return src + (size_t)q * C + v * 8;
    }
};

template <int S>
__global__ __launch_bounds__(512) void conv_wgrad(const WgradArgs A) {
    extern __shared__ __attribute__((aligned(16))) u16 lds[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int CPo = A.cout + 8, CPi = A.cin + 8, NC = A.W + 2;
    const int n_pairs_total = 9 * (A.cin >> 4) + 1;
    const int n_pchunks = (n_pairs_total + A.npairs - 1) / A.npairs;
    const int cchunk = blockIdx.y / n_pchunks, pchunk = blockIdx.y % n_pchunks;
    const int cg0 = cchunk * A.ncg;
    const int ncg = min(A.ncg, (A.cout >> 4) - cg0);
    const int p0 = pchunk * A.npairs;
    const int np = min(A.npairs, n_pairs_total - p0);
    u16* dzb = lds;                            // [R Wo][CPo]
    u16* xb = lds + (size_t)A.R * A.Wo * CPo;  // [S (R - 1) + 3][NC][CPi]
    // my pairs: (tap, ci group) or the bias pair (index 9 cin / 16)
    int ptap[kMaxPairs], pci[kMaxPairs];
    int mp = 0;
#pragma unroll
    for (int i = 0; i < kMaxPairs; ++i) {
        const int p = p0 + wave + kWaves * i;
        ptap[i] = -1;
        pci[i] = 0;
        if (wave + kWaves * i < np) {
            if (p < n_pairs_total - 1) {
                ptap[i] = p / (A.cin >> 4);
                pci[i] = (p % (A.cin >> 4)) * 16;
            } else {
                ptap[i] = 9;  // bias
            }
            mp = i + 1;
        }
    }
    v4f acc[kMaxPairs][4];
#pragma unroll
    for (int i = 0; i < kMaxPairs; ++i)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[i][c] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    const v8u ones = {0x3f80, 0x3f80, 0x3f80, 0x3f80, 0x3f80, 0x3f80, 0x3f80, 0x3f80};
    const int nbands = (A.Ho + A.R - 1) / A.R;
    const Div dwo(A.Wo);
    for (int b = blockIdx.x; b < A.B; b += gridDim.x) {
        for (int band = 0; band < nbands; ++band) {
            const int ho0 = band * A.R, R = min(A.R, A.Ho - ho0);
            const int P = R * A.Wo;
            __syncthreads();
            {  // dZ rows [ho0, ho0 + R): contiguous pixels
                const RowsMap m{A.dz + (((size_t)b * A.Ho + ho0) * A.Wo) * A.cout, A.cout, Div(A.cout >> 3)};
                copy16(P * (A.cout >> 3), dzb, m);
            }
            load_band(A.x, b, A.H, A.W, A.cin, S * ho0 - 1, S * (R - 1) + 3, xb);
            __syncthreads();
            for (int k0 = 0; k0 < P; k0 += 32) {
                // the 8 pixels of this lane
                int dofs[8], xofs[8];
                bool ok[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int q = k0 + 8 * (lane >> 4) + j;
                    ok[j] = q < P;
                    const int qq = ok[j] ? q : 0;
                    const int hr = dwo.q(qq), wo = qq - hr * A.Wo;
                    dofs[j] = qq * CPo;
                    xofs[j] = ((S * hr) * NC + S * wo) * CPi;
                }
                v8bf av[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    v8u t = {0, 0, 0, 0, 0, 0, 0, 0};
                    if (c < ncg) {
                        const int co = (cg0 + c) * 16 + (lane & 15);
#pragma unroll
                        for (int j = 0; j < 8; ++j) t[j] = ok[j] ? dzb[dofs[j] + co] : (u16)0;
                    }
                    av[c] = __builtin_bit_cast(v8bf, t);
                }
#pragma unroll
                for (int i = 0; i < kMaxPairs; ++i) {
                    if (i >= mp) continue;
                    v8u t = ones;
                    if (ptap[i] < 9) {
                        const int kh = ptap[i] / 3, kw = ptap[i] - 3 * kh;
                        const int toff = (kh * NC + kw) * CPi + pci[i] + (lane & 15);
#pragma unroll
                        for (int j = 0; j < 8; ++j) t[j] = xb[xofs[j] + toff];
                    }
                    const v8bf bv = __builtin_bit_cast(v8bf, t);
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        if (c < ncg) acc[i][c] = mfma(av[c], bv, acc[i][c]);
                }
            }
        }
    }
    const int ncol = 9 * A.cin + 1;
    float* part = A.part + (size_t)blockIdx.x * A.cout * ncol;
#pragma unroll
    for (int i = 0; i < kMaxPairs; ++i) {
        if (i >= mp) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (c >= ncg) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int co = (cg0 + c) * 16 + 4 * (lane >> 4) + r;
                if (ptap[i] < 9)
                    part[(size_t)co * ncol + ptap[i] * A.cin + pci[i] + (lane & 15)] = acc[i][c][r];
                else if ((lane & 15) == 0)
                    part[(size_t)co * ncol + 9 * A.cin] = acc[i][c][r];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// First layer (one input channel) on the vector ALUs.
// conv1_fwd: grid (bands of R rows, images), one thread per (pixel, 8 output
// channels) so that a wave's 16-byte stores are contiguous; the weights (FP32 copies
// of the bf16 values) and the input band with its halo in LDS.  wt: [cout][kp] bf16,
// taps first.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void conv1_fwd(const u16* x0, const u16* wt, int kp, const float* bias, int H,
                                                 int W, int cout, int R, u16* Z, u16* Y) {
    extern __shared__ float sm[];
    float* ws = sm;               // [cout][9], then bias[cout]
    float* xs = sm + cout * 10;   // [R + 2][W + 2]
    const int b = blockIdx.y, h0 = blockIdx.x * R;
    const int Rr = min(R, H - h0), NC = W + 2, ng = cout >> 3;
    for (int i = threadIdx.x; i < cout * 9; i += blockDim.x) ws[i] = bf2f(wt[(size_t)(i / 9) * kp + i % 9]);
    for (int i = threadIdx.x; i < cout; i += blockDim.x) ws[cout * 9 + i] = bias[i];
    for (int i = threadIdx.x; i < (Rr + 2) * NC; i += blockDim.x) {
        const int r = i / NC, c = i - r * NC;
        const int h = h0 - 1 + r, w = c - 1;
        xs[i] = (h >= 0 && h < H && w >= 0 && w < W) ? bf2f(x0[((size_t)b * H + h) * W + w]) : 0.0f;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < Rr * W * ng; t += blockDim.x) {
        const int p = t / ng, g = t - p * ng;
        const int hr = p / W, w = p - hr * W;
        float xv[9];
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) xv[tap] = xs[(hr + tap / 3) * NC + w + tap % 3];
        const size_t o = (((size_t)b * H + h0 + hr) * W + w) * cout + 8 * g;
        v8u zo, yo;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int co = 8 * g + e;
            float a = 0.0f;
#pragma unroll
            for (int tap = 0; tap < 9; ++tap) a = fmaf(ws[co * 9 + tap], xv[tap], a);
            const u16 zb = f2bf(a + ws[cout * 9 + co]);
            zo[e] = zb;
            yo[e] = f2bf(gelu_f(bf2f(zb)));
        }
        *reinterpret_cast<v8u*>(Z + o) = zo;
        *reinterpret_cast<v8u*>(Y + o) = yo;
    }
}

// conv1_wgrad: one workgroup per image, the image (with its zero halo) in LDS as
// FP32; wave w takes output-channel groups of 8 (w, w + 4, ...), its lanes stride
// over the pixels with eight dZ loads in flight; 8 x 10 sums (9 taps + bias) per
// lane, reduced across the wave.  part[b][cout][10].
__global__ __launch_bounds__(256) void conv1_wgrad(const u16* x0, const u16* dz, int H, int W, int cout,
                                                   float* part) {
    extern __shared__ float xs[];  // [H + 2][W + 2]
    const int b = blockIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int P = H * W, NC = W + 2;
    for (int i = threadIdx.x; i < (H + 2) * NC; i += blockDim.x) {
        const int r = i / NC, c = i - r * NC;
        const int h = r - 1, w = c - 1;
        xs[i] = (h >= 0 && h < H && w >= 0 && w < W) ? bf2f(x0[(size_t)b * P + h * W + w]) : 0.0f;
    }
    __syncthreads();
    constexpr int U = 8;
    for (int g = wave; g < (cout >> 3); g += 4) {
        float s[8][10];
#pragma unroll
        for (int e = 0; e < 8; ++e)
#pragma unroll
            for (int k = 0; k < 10; ++k) s[e][k] = 0.0f;
        for (int p0 = lane; p0 < P; p0 += 64 * U) {
            v8u d[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int p = p0 + 64 * u;
                d[u] = p < P ? *reinterpret_cast<const v8u*>(dz + ((size_t)b * P + p) * cout + 8 * g)
                             : v8u{0, 0, 0, 0, 0, 0, 0, 0};
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int p = min(p0 + 64 * u, P - 1);  // beyond P: d = 0
                const int h = p / W, w = p - h * W;
                float xv[10];
#pragma unroll
                for (int tap = 0; tap < 9; ++tap) xv[tap] = xs[(h + tap / 3) * NC + w + tap % 3];
                xv[9] = 1.0f;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float de = bf2f(d[u][e]);
#pragma unroll
                    for (int k = 0; k < 10; ++k) s[e][k] = fmaf(de, xv[k], s[e][k]);
                }
            }
        }
#pragma unroll
        for (int e = 0; e < 8; ++e)
#pragma unroll
            for (int k = 0; k < 10; ++k) {
                float v = s[e][k];
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
                if (lane == 0) part[((size_t)b * cout + 8 * g + e) * 10 + k] = v;
            }
    }
}

}  // namespace cnn
}  // namespace gl
