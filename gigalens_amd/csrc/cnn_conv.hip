// Convolution passes of the learned initialiser's network on the MFMA
// (v_mfma_f32_16x16x32_bf16), NHWC bfloat16 activations.
//
//   conv_wave<0, S>  Z = conv(X, W) + b, Y = gelu(Z)          D[cout][pixel], operands from global memory
//   conv_wave<1, S>  dZprev = conv^T(dZ, W) * gelu'(Zprev)      D[cin][pixel], operands from global memory
//   conv_wgrad<S>    dW = sum_pixels dZ (x) X~, db = sum dZ     rows: cout, cols: (tap, cin), band in LDS
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
// hardware exp and reciprocal (v_rcp_f32, 1 ulp; exp overflow gives rcp(inf) = 0 and
// tanh = 1): the results are rounded to bfloat16 right after, so the last-bit
// differences from tanhf change a rounded value only at a tie.
__device__ __forceinline__ float tanh_fast(float u) {
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(__expf(2.0f * u) + 1.0f);
}
__device__ __forceinline__ float gelu_f(float x) {
    return 0.5f * x * (1.0f + tanh_fast(kGeluC * (x + kGeluA * x * x * x)));
}
__device__ __forceinline__ float gelu_grad_f(float x) {
    const float t = tanh_fast(kGeluC * (x + kGeluA * x * x * x));
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * kGeluC * (1.0f + 3.0f * kGeluA * x * x);
}

// Pixels of a band [h0, h0 + R) of an image W pixels wide taken class by class of
// (h mod S, w mod S); 16-pixel groups numbered class after class.
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

// ---------------------------------------------------------------------------
// Wave-streaming forward and data-gradient passes.  The MFMA runs in the transposed
// form D[channel][pixel] = sum_k A[channel][k] B[k][pixel]: A is a 16-channel tile of
// the weights, held in LDS for the whole launch and stored lane-major so that a wave's
// operand is one contiguous 1 KB read; B holds the operand activations of 16 output
// pixels at one tap, each lane loading 8 channels (16 bytes) of one pixel straight
// from global memory; the result leaves lane l with 4 consecutive channels of pixel
// l % 16, the 8-byte vectors the epilogue loads (Zprev) and stores.  The waves of a
// workgroup take the 16-pixel groups of its (image, band of rows) items on their own:
// after the weight load there is no barrier, so every wave keeps its loads in flight,
// and the rows that neighbouring groups share come from the L1 and L2 of one CU.  A
// wave issues the loads of its next group before the epilogue of the current one:
// vector stores and loads retire in order, so a load issued after a group's stores
// would wait for them.
//   KIND 0 (forward):       output pixel (ho, wo), tap (kh, kw) -> operand pixel
//                           (S ho + kh - 1, S wo + kw - 1);  Z = acc + b, Y = gelu(Z).
//   KIND 1 (data gradient): output pixel (h, w), tap (kh, kw) -> dZ pixel
//                           ((h + 1 - kh) / S, (w + 1 - kw) / S) where S divides;
//                           out = acc * gelu'(Zprev).  With S = 2 the pixels are taken
//                           class by class of (ph, pw) = (h mod 2, w mod 2), so a group
//                           uses 1, 2 or 4 taps: slot (i, j), i <= ph, j <= pw, is tap
//                           (kh, kw) = (ph ? 2 i : 1, pw ? 2 j : 1), dZ pixel
//                           ((h - ph) / 2 + (ph && i == 0), (w - pw) / 2 + (pw && j == 0)).
// Zero padding: a tap whose operand pixel lies outside the image loads from a zeroed
// tail the host allocates after the operand tensor (zoff), so all loads of a group
// issue together and need no select.  Lanes past a class's last pixel repeat its last
// pixel and store nothing.  Offsets are 32-bit element indices (the host keeps every
// activation tensor below 2^31 elements).
// ---------------------------------------------------------------------------
constexpr int kWvWaves = 8;  // waves of a wave-streaming workgroup (512 threads)
typedef unsigned short v4u __attribute__((ext_vector_type(4)));

struct WaveArgs {
    const u16* x;       // operand activations [B][Hx][Wx][K] (forward: Y of the previous layer; data gradient: dZ)
    const u16* w;       // weights [Nout][9][K] (forward: wt; data gradient: w2t)
    const float* bias;  // forward: [Nout]
    const u16* zp;      // data gradient: Zprev [B][Ho][Wo][Nout]
    u16* out0;          // forward: Z; data gradient: dZprev; [B][Ho][Wo][Nout]
    u16* out1;          // forward: Y
    int B, Ho, Wo, Hx, Wx, K, Nout;
    int R;          // output rows per item
    unsigned zoff;  // element offset of K zeros after the operand tensor
};

// bfloat16 rounding of 4 values on v_cvt_pk_bf16_f32: round to nearest even, as f2bf
typedef __bf16 v4bf __attribute__((ext_vector_type(4)));
__device__ __forceinline__ v4u bf4(v4f x) { return __builtin_bit_cast(v4u, __builtin_convertvector(x, v4bf)); }

// LDS bytes of the weight tiles: 9 taps x K / 32 chunks x NMT 16-channel tiles of 1 KB
__host__ __device__ constexpr int wave_lds(int K, int nmt) { return 9 * (K >> 5) * nmt * 1024; }

// one 16-pixel group of a wave: where its outputs go and where its operands are
template <int KIND, int S>
struct WaveGroup {
    static constexpr bool kClasses = KIND == 1 && S == 2;
    static constexpr int T = kClasses ? 4 : 9;  // operand slots
    unsigned obase;  // the lane's output vector (channel c0 + 4 (lane / 16))
    int ocen;        // the lane's operand vector at dy = dx = 0
    unsigned vmask;  // bit t: slot t's operand pixel lies in the image
    bool ok;         // the lane holds a pixel of the group
    int ph, pw;      // pixel class (uniform)
    __device__ bool used(int t) const { return !kClasses || ((t >> 1) <= ph && (t & 1) <= pw); }
    __device__ int tap(int t) const {
        if (!kClasses) return t;
        const int kh = ph ? 2 * (t >> 1) : 1, kw = pw ? 2 * (t & 1) : 1;
        return 3 * kh + kw;
    }
    __device__ int dy(int t) const { return kClasses ? (ph && (t >> 1) == 0) : KIND == 0 ? t / 3 : -(t / 3); }
    __device__ int dx(int t) const { return kClasses ? (pw && (t & 1) == 0) : KIND == 0 ? t % 3 : -(t % 3); }
};

template <int KIND, int S, int SC>
__device__ __forceinline__ void wave_group_setup(const WaveArgs& A, const PixelClasses<SC>& pc, int b, int h0, int g,
                                                 int lane, int c0, WaveGroup<KIND, S>& u) {
    const int c = pc.cls(g);
    const int cnt = pc.count(c);
    const int k = (g - pc.first_group(c)) * 16 + (lane & 15);
    int hr, w;
    pc.pixel(c, min(k, cnt - 1), hr, w);
    const int h = h0 + hr;
    u.ok = k < cnt;
    u.ph = SC == 2 ? c / 2 : 0;
    u.pw = SC == 2 ? c % 2 : 0;
    u.obase = (unsigned)(((b * A.Ho + h) * A.Wo + w) * A.Nout + c0 + 4 * (lane >> 4));
    int yc, xc;  // operand pixel at dy = dx = 0
    if (KIND == 0) {
        yc = S * h - 1;
        xc = S * w - 1;
    } else if (S == 1) {
        yc = h + 1;
        xc = w + 1;
    } else {
        yc = (h - u.ph) / 2;
        xc = (w - u.pw) / 2;
    }
    u.ocen = ((b * A.Hx + yc) * A.Wx + xc) * A.K + 8 * (lane >> 4);
    unsigned m = 0;
#pragma unroll
    for (int t = 0; t < WaveGroup<KIND, S>::T; ++t) {
        const bool v = (unsigned)(yc + u.dy(t)) < (unsigned)A.Hx && (unsigned)(xc + u.dx(t)) < (unsigned)A.Wx;
        m |= (unsigned)v << t;
    }
    u.vmask = m;
}

template <int KIND, int S, int NK>
__device__ __forceinline__ void wave_group_load(const WaveArgs& A, const WaveGroup<KIND, S>& u, int lane, int kq,
                                                v8u (&bv)[WaveGroup<KIND, S>::T][NK]) {
    const unsigned zlane = A.zoff + 8 * (lane >> 4);
#pragma unroll
    for (int t = 0; t < WaveGroup<KIND, S>::T; ++t) {
        const unsigned off = (u.vmask >> t) & 1u ? (unsigned)(u.ocen + (u.dy(t) * A.Wx + u.dx(t)) * A.K) : zlane;
#pragma unroll
        for (int kc = 0; kc < NK; ++kc) {
            bv[t][kc] = v8u{0, 0, 0, 0, 0, 0, 0, 0};
            if (!u.used(t)) continue;  // uniform
            bv[t][kc] = *reinterpret_cast<const v8u*>(A.x + off + 32 * (kq + kc));
        }
    }
}

template <int KIND, int S, int NMT, int NK>
__device__ __forceinline__ void wave_group_mfma(const WaveGroup<KIND, S>& u, const v8u* wl, int lane, int nkc, int kq,
                                                const v8u (&bv)[WaveGroup<KIND, S>::T][NK], v4f (&acc)[NMT]) {
#pragma unroll
    for (int t = 0; t < WaveGroup<KIND, S>::T; ++t) {
        if (!u.used(t)) continue;  // uniform
        const int tap = u.tap(t);
#pragma unroll
        for (int kc = 0; kc < NK; ++kc) {
            const v8bf bop = __builtin_bit_cast(v8bf, bv[t][kc]);
#pragma unroll
            for (int mt = 0; mt < NMT; ++mt) {
                const v8bf aop = __builtin_bit_cast(v8bf, wl[((tap * nkc + kq + kc) * NMT + mt) * 64 + lane]);
                acc[mt] = mfma(aop, bop, acc[mt]);
            }
        }
    }
}

template <int KIND, int S, int NMT>
__device__ __forceinline__ void wave_group_store(const WaveArgs& A, const WaveGroup<KIND, S>& u, const v4f (&acc)[NMT],
                                                 const v4u (&zv)[NMT], const v4f (&bias)[NMT]) {
    if (!u.ok) return;
#pragma unroll
    for (int mt = 0; mt < NMT; ++mt) {
        if (KIND == 0) {
            const v4u zo = bf4(acc[mt] + bias[mt]);
            v4f y;
#pragma unroll
            for (int r = 0; r < 4; ++r) y[r] = gelu_f(bf2f(zo[r]));
            *reinterpret_cast<v4u*>(A.out0 + u.obase + 16 * mt) = zo;
            *reinterpret_cast<v4u*>(A.out1 + u.obase + 16 * mt) = bf4(y);
        } else {
            v4f o;
#pragma unroll
            for (int r = 0; r < 4; ++r) o[r] = acc[mt][r] * gelu_grad_f(bf2f(zv[mt][r]));
            *reinterpret_cast<v4u*>(A.out0 + u.obase + 16 * mt) = bf4(o);
        }
    }
}

// Persistent workgroups of 512 threads over the (image, band of R output rows) items;
// the output channels in chunks of 16 NMT (one chunk when they fit, weights reloaded
// per chunk otherwise); the K operand channels in steps of 32 NK.  Needs K % (32 NK)
// == 0, Nout % (16 NMT) == 0 and wave_lds(K, NMT) bytes of dynamic LDS.
template <int KIND, int S, int NMT, int NK>
__global__ __launch_bounds__(512) void conv_wave(const WaveArgs A) {
    extern __shared__ __attribute__((aligned(16))) v8u wl[];
    constexpr int SC = KIND == 1 ? S : 1;
    constexpr int T = WaveGroup<KIND, S>::T;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nkc = A.K >> 5;
    const int nvec = 9 * nkc * NMT * 64;
    const int nb = (A.Ho + A.R - 1) / A.R, items = A.B * nb;
    for (int c0 = 0; c0 < A.Nout; c0 += 16 * NMT) {
        __syncthreads();
        for (int i = threadIdx.x; i < nvec; i += blockDim.x) {
            const int t = i >> 6, l = i & 63;
            const int mt = t % NMT, kk = (t / NMT) % nkc, tap = t / (NMT * nkc);
            wl[i] = *reinterpret_cast<const v8u*>(A.w + ((size_t)(c0 + 16 * mt + (l & 15)) * 9 + tap) * A.K + 32 * kk +
                                                  8 * (l >> 4));
        }
        __syncthreads();
        v4f bias[NMT];
#pragma unroll
        for (int mt = 0; mt < NMT; ++mt)
            bias[mt] = KIND == 0 ? *reinterpret_cast<const v4f*>(A.bias + c0 + 16 * mt + 4 * (lane >> 4))
                                 : v4f{0.0f, 0.0f, 0.0f, 0.0f};
        for (int item = blockIdx.x; item < items; item += gridDim.x) {
            const int b = item / nb, h0 = (item - b * nb) * A.R, R = min(A.R, A.Ho - h0);
            const PixelClasses<SC> pc(h0, R, A.Wo);
            const int G = pc.groups();
            // the first group rotates with the item so that no wave always takes the remainder
            int g = (wave + item) & (kWvWaves - 1);
            if (g >= G) continue;
            WaveGroup<KIND, S> u;
            v8u bv[T][NK];
            v4u zv[NMT];
            wave_group_setup<KIND, S, SC>(A, pc, b, h0, g, lane, c0, u);
            wave_group_load<KIND, S, NK>(A, u, lane, 0, bv);
            if (KIND == 1) {
#pragma unroll
                for (int mt = 0; mt < NMT; ++mt) zv[mt] = *reinterpret_cast<const v4u*>(A.zp + u.obase + 16 * mt);
            }
            for (;;) {
                v4f acc[NMT];
#pragma unroll
                for (int mt = 0; mt < NMT; ++mt) acc[mt] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
                // the weight tiles are read from LDS per group: an opaque lane index keeps the
                // compiler from hoisting all of them into registers (up to 144 VGPRs)
                int wlane = lane;
                asm volatile("" : "+v"(wlane));
                wave_group_mfma<KIND, S, NMT, NK>(u, wl, wlane, nkc, 0, bv, acc);
                for (int kq = NK; kq < nkc; kq += NK) {
                    wave_group_load<KIND, S, NK>(A, u, lane, kq, bv);
                    wave_group_mfma<KIND, S, NMT, NK>(u, wl, wlane, nkc, kq, bv, acc);
                }
                // the next group's loads go out before this group's stores
                const int gn = g + kWvWaves;
                const bool more = gn < G;
                WaveGroup<KIND, S> un;
                v4u zn[NMT];
                if (more) {
                    wave_group_setup<KIND, S, SC>(A, pc, b, h0, gn, lane, c0, un);
                    wave_group_load<KIND, S, NK>(A, un, lane, 0, bv);
                    if (KIND == 1) {
#pragma unroll
                        for (int mt = 0; mt < NMT; ++mt)
                            zn[mt] = *reinterpret_cast<const v4u*>(A.zp + un.obase + 16 * mt);
                    }
                }
                wave_group_store<KIND, S, NMT>(A, u, acc, zv, bias);
                if (!more) break;
                u = un;
#pragma unroll
                for (int mt = 0; mt < NMT; ++mt) zv[mt] = zn[mt];
                g = gn;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: grid (blocks over images, tile chunks), 512 threads.  Output
// tiles are (co group of 16) x (tap, ci group of 16) plus one bias tile per co
// group (B = ones); a chunk covers co groups [cg0, cg0 + ncg) (ncg <= 4) and pairs
// [p0, p0 + np) of the 9 cin / 16 + 1 (tap, ci group) pairs, wave w taking pairs
// p0 + w, p0 + w + kWaves, ... (<= kMaxPairs) with every co group of the chunk.
// Pixels are the reduction dimension, so each MFMA operand lane needs 8 consecutive
// pixels of one channel: a band of R output rows is stored in LDS transposed,
//   dZT[co][hr][wo]        wo < Wp (Wo rounded up to 8; zeros past Wo), row stride DS
//   XT[kw][ci][r][cc] = X[S ho0 - 1 + r][S cc + kw - 1]   (zeros outside the image)
// one plane per tap column kw, so that the operand of tap (kh, kw) at output pixel
// (hr, wo) is XT[kw][ci][S hr + kh][wo .. wo + 7], and every operand is one aligned
// 16-byte LDS read whatever the stride and tap.  The fill transposes in registers: a
// thread loads the 16-byte vectors (8 channels) of 8 or 9 pixels of one row, turns
// each 8 x 8 block into 8 channel rows of 8 pixels with byte permutes, and writes
// them as aligned 16-byte LDS stores.  9 pixels one step apart feed two planes (S = 1:
// columns cc0 - 1 .. cc0 + 7 for kw = 0, 1; S = 2: the odd columns for kw = 0, 2),
// 8 pixels the third (S = 1: cc0 + 1 .. cc0 + 8 for kw = 2; S = 2: the even columns
// for kw = 1).  k-steps of 32 pixels = 4 segments of 8
// (one per quarter-wave).  part[blockIdx.x][cout][9 cin + 1]: per-block sums, added
// by wgrad_reduce.
// ---------------------------------------------------------------------------
constexpr int kMaxPairs = 3;

struct WgradArgs {
    const u16* dz;  // [B][Ho][Wo][cout]
    const u16* x;   // [B][H][W][cin]
    float* part;
    int B, H, W, cin, cout, Ho, Wo, R;
    int ncg, npairs;  // per chunk; gridDim.y = co chunks x pair chunks
    int Wp;           // Wo rounded up to a multiple of 8
};

// LDS elements of the transposed band (host and device agree on the layout)
__host__ __device__ inline int wgrad_ds(int R, int Wp) { return R * Wp + 8; }
__host__ __device__ inline int wgrad_xs(int S, int R, int Wp) { return (S * (R - 1) + 3) * Wp + 8; }
__host__ __device__ inline long wgrad_lds_elems(int S, int R, int Wp, int cin, int cout) {
    return (long)cout * wgrad_ds(R, Wp) + 3L * cin * wgrad_xs(S, R, Wp);
}

// 8 pixels x 8 channels (p[i] = the channels of pixel i) -> 8 channels x 8 pixels,
// written as out + c * stride for channel c (16-byte aligned)
__device__ __forceinline__ void transpose_store8(const v8u* p, u16* out, int stride) {
    typedef unsigned int v4w __attribute__((ext_vector_type(4)));
    v4w d[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) d[i] = __builtin_bit_cast(v4w, p[i]);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        v4w o;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const unsigned a = d[2 * m][c >> 1], b = d[2 * m + 1][c >> 1];
            o[m] = (c & 1) ? __builtin_amdgcn_perm(b, a, 0x07060302u) : __builtin_amdgcn_perm(b, a, 0x05040100u);
        }
        *reinterpret_cast<v4w*>(out + c * stride) = o;
    }
}

template <int S>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void conv_wgrad(const WgradArgs A) {
    extern __shared__ __attribute__((aligned(16))) u16 lds[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int Wp = A.Wp, spr = Wp >> 3;
    const int DS = wgrad_ds(A.R, Wp), XS = wgrad_xs(S, A.R, Wp);
    const int PL = A.cin * XS;  // elements of one XT plane
    u16* dzt = lds;
    u16* xt = lds + A.cout * DS;
    const int n_pairs_total = 9 * (A.cin >> 4) + 1;
    const int n_pchunks = (n_pairs_total + A.npairs - 1) / A.npairs;
    const int cchunk = blockIdx.y / n_pchunks, pchunk = blockIdx.y % n_pchunks;
    const int cg0 = cchunk * A.ncg;
    const int ncg = min(A.ncg, (A.cout >> 4) - cg0);
    const int p0 = pchunk * A.npairs;
    const int np = min(A.npairs, n_pairs_total - p0);
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
    const v8u zero8 = {0, 0, 0, 0, 0, 0, 0, 0};
    const int nbands = (A.Ho + A.R - 1) / A.R;
    const int vo = A.cout >> 3, vi = A.cin >> 3, nblk = Wp >> 3;
    constexpr int NP = 9;  // pixels a fill thread loads (one row)
    for (int b = blockIdx.x; b < A.B; b += gridDim.x) {
        for (int band = 0; band < nbands; ++band) {
            const int ho0 = band * A.R, R = min(A.R, A.Ho - ho0);
            const int nrx = S * (R - 1) + 3, xr0 = S * ho0 - 1;
            __syncthreads();
            // fill units: X (kind e, channel block pair, row, column block, channel block) then dZ (channel
            // block pair, row, column block, channel block)
            const int nxu = 2 * nrx * nblk * vi, ndu = R * nblk * vo;
            for (int u = threadIdx.x; u < nxu + ndu; u += blockDim.x) {
                v8u px[NP];
                if (u < nxu) {
                    // two channel blocks fastest (32 contiguous bytes of a pixel per lane pair),
                    // then column block, then row: the channel-block rows lie an odd multiple
                    // of 128 bytes apart, so 16 consecutive lanes store 16 distinct 16-byte bank
                    // slots (no LDS bank conflicts)
                    int q = u;
                    const int v_lo = q & 1;
                    q >>= 1;
                    const int cb = q % nblk;
                    q /= nblk;
                    const int r = q % nrx;
                    q /= nrx;
                    const int v = v_lo + 2 * (q % (vi >> 1)), e = q / (vi >> 1);
                    const int xr = xr0 + r, cc0 = cb * 8;
                    // pixel i at column x0 + S i.  e = 1: 9 pixels for the planes kw = 0 and
                    // (S = 1) 1 or (S = 2) 2, one pixel apart; e = 0: 8 pixels for the third plane
                    const int x0 = S == 1 ? (e ? cc0 - 1 : cc0 + 1) : (e ? 2 * cc0 - 1 : 2 * cc0);
                    const int np = e ? 9 : 8;
                    const bool rin = xr >= 0 && xr < A.H;
                    const u16* row = A.x + (((size_t)b * A.H + (rin ? xr : 0)) * A.W) * A.cin + 8 * v;
#pragma unroll
                    for (int i = 0; i < NP; ++i) {
                        const int x = x0 + S * i;
                        px[i] = v8u{0, 0, 0, 0, 0, 0, 0, 0};
                        if (i < np && rin && x >= 0 && x < A.W) px[i] = *reinterpret_cast<const v8u*>(row + (size_t)x * A.cin);
                    }
                    u16* dst = xt + 8 * v * XS + r * Wp + cc0;
                    if (e) {
                        transpose_store8(px, dst, XS);                       // kw = 0
                        transpose_store8(px + 1, dst + (S == 1 ? 1 : 2) * PL, XS);  // kw = 1 (S = 1), kw = 2 (S = 2)
                    } else {
                        transpose_store8(px, dst + (S == 1 ? 2 : 1) * PL, XS);  // kw = 2 (S = 1), kw = 1 (S = 2)
                    }
                } else {
                    int q = u - nxu;
                    const int v_lo = q & 1;
                    q >>= 1;
                    const int wb = q % nblk;
                    q /= nblk;
                    const int hr = q % R, v = v_lo + 2 * (q / R);
                    const int wo0 = wb * 8;
                    const u16* row = A.dz + (((size_t)b * A.Ho + ho0 + hr) * A.Wo) * A.cout + 8 * v;
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        px[i] = v8u{0, 0, 0, 0, 0, 0, 0, 0};
                        if (wo0 + i < A.Wo) px[i] = *reinterpret_cast<const v8u*>(row + (size_t)(wo0 + i) * A.cout);
                    }
                    transpose_store8(px, dzt + 8 * v * DS + hr * Wp + wo0, DS);
                }
            }
            __syncthreads();
            const int nseg = R * spr;
            for (int s0 = 0; s0 < nseg; s0 += 4) {
                int seg = s0 + (lane >> 4);
                const bool sv = seg < nseg;
                seg = sv ? seg : nseg - 1;
                const int hr = seg / spr, wo0 = (seg - hr * spr) * 8;
                v8bf av[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    v8u t = zero8;
                    if (c < ncg) {
                        const v8u d = *reinterpret_cast<const v8u*>(dzt + ((cg0 + c) * 16 + (lane & 15)) * DS +
                                                                     hr * Wp + wo0);
                        t = sv ? d : zero8;
                    }
                    av[c] = __builtin_bit_cast(v8bf, t);
                }
#pragma unroll
                for (int i = 0; i < kMaxPairs; ++i) {
                    if (i >= mp) continue;
                    v8u t = ones;
                    if (ptap[i] < 9) {
                        const int kh = ptap[i] / 3, kw = ptap[i] - 3 * kh;
                        t = *reinterpret_cast<const v8u*>(xt + kw * PL + (pci[i] + (lane & 15)) * XS +
                                                          (S * hr + kh) * Wp + wo0);
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
    // thread -> (channel group g, first pixel): g stays fixed, so its 72 weights and 8
    // biases sit in registers for the whole band
    const int pstep = blockDim.x / ng, g = threadIdx.x % ng;
    if (threadIdx.x >= pstep * ng) return;
    float wr[8][9], br[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) wr[e][tap] = ws[(8 * g + e) * 9 + tap];
        br[e] = ws[cout * 9 + 8 * g + e];
    }
    for (int p = threadIdx.x / ng; p < Rr * W; p += pstep) {
        const int hr = p / W, w = p - hr * W;
        float xv[9];
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) xv[tap] = xs[(hr + tap / 3) * NC + w + tap % 3];
        const size_t o = (((size_t)b * H + h0 + hr) * W + w) * cout + 8 * g;
        v4f z0, z1;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            float a = 0.0f;
#pragma unroll
            for (int tap = 0; tap < 9; ++tap) a = fmaf(wr[e][tap], xv[tap], a);
            if (e < 4) z0[e] = a + br[e];
            else z1[e - 4] = a + br[e];
        }
        const v4u zo0 = bf4(z0), zo1 = bf4(z1);
        v4f y0, y1;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            y0[e] = gelu_f(bf2f(zo0[e]));
            y1[e] = gelu_f(bf2f(zo1[e]));
        }
        const v4u yo0 = bf4(y0), yo1 = bf4(y1);
        *reinterpret_cast<v8u*>(Z + o) = v8u{zo0[0], zo0[1], zo0[2], zo0[3], zo1[0], zo1[1], zo1[2], zo1[3]};
        *reinterpret_cast<v8u*>(Y + o) = v8u{yo0[0], yo0[1], yo0[2], yo0[3], yo1[0], yo1[1], yo1[2], yo1[3]};
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
