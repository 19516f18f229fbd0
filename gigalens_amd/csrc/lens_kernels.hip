// Fused per-particle lens-model kernels for gfx950 (MI355X).
//
// One workgroup owns one parameter particle and keeps the particle's whole
// supersampled image resident in LDS:
//   A  prologue   z -> x (bijectors), log prior, per-component constants,
//                 EPL series coefficient tables (LDS)
//   B  ray-trace  every sub-pixel: lens light(theta) + source(beta), NaN -> 0,
//                 written into a zero-haloed LDS image
//   C  conv+pool  stride-ss correlation with the PSF pre-folded with the ss x ss
//                 box (== conv2d SAME + avg_pool2d, tf/simulator.py:83-96),
//                 x conversion factor, then the Gaussian log-likelihood terms
//   D  adjoint    (optional) transposed strided correlation of dL/dimage, in
//                 polyphase order so the tap weights are wave-uniform
//   E  recompute  (optional) ray-trace again per sub-pixel and accumulate the
//                 reverse-mode statistics of every component
//   F  reduce     wavefront shuffles + LDS, per-particle chain rule, Adam
//
// Reference semantics: src/gigalens/tf/simulator.py:57-97 (simulate),
// src/gigalens/tf/model.py:53-86 (ForwardProbModel.log_prob),
// src/gigalens/tf/inference.py:28-35 (MAP train_step).
#pragma once
#include <hip/hip_runtime.h>
#include "lens_math.h"
#include "lens_plan.h"

namespace gl {

enum Mode { M_SIM = 0, M_LOGPROB = 1, M_VJP = 2, M_MAP = 3 };

struct KArgs {
    const float* in;      // x (SIM, VJP) or z (LOGPROB, MAP): bs x d
    float* z_io;          // MAP: z updated in place
    float* img;           // SIM: bs x N x N
    const float* g_img;   // VJP: upstream image gradient
    float* lp;
    float* chi2;
    const float* g_lp;
    const float* g_chi2;
    float* grad;          // LOGPROB: grad_z, VJP: grad_x
    float* m;
    float* v;
    const int* lens;      // optional per-particle observed-image index (bs)
    const float* weight;  // MAP: optional per-particle loss weight (bs); default 1/bs
    const int* order;     // optional: particle of workgroup b is order[b] (launch order, see order_kernel)
    float lr_t, beta1, beta2, eps, gscale;
    int mode;
    int bs;
    int limit_k;
    const int* limit;     // optional device scalar: the launch does nothing when limit_k >= *limit (a step
                          // of a graph-recorded loop whose length is decided on the device)
};

template <int K> struct LensAcc { static constexpr int n = 0; };
template <> struct LensAcc<glm::K_EPL> { static constexpr int n = 8; };
template <> struct LensAcc<glm::K_SHEAR> { static constexpr int n = 2; };
template <> struct LensAcc<glm::K_SIS> { static constexpr int n = 3; };
template <> struct LensAcc<glm::K_SIE> { static constexpr int n = 5; };
template <> struct LensAcc<glm::K_TNFW> { static constexpr int n = 5; };
template <int K> struct LightAcc {
    static constexpr int n = (K == glm::K_SERSIC || K == glm::K_SERSIC_ELLIPSE) ? 8 : K == glm::K_CORE_SERSIC ? 11 : 0;
};

template <int L0, int L1, int L2, int A0, int A1, int S0, int S1>
struct ModelT {
    static constexpr int lens[3] = {L0, L1, L2};
    static constexpr int light[4] = {A0, A1, S0, S1};  // lens light then source
    static constexpr int NL = (L0 != 0) + (L1 != 0) + (L2 != 0);
    static constexpr int NLL = (A0 != 0) + (A1 != 0);
    static constexpr int NS = (S0 != 0) + (S1 != 0);
    static constexpr int lens_off(int i) {
        return i == 0 ? 0 : i == 1 ? LensAcc<L0>::n : LensAcc<L0>::n + LensAcc<L1>::n;
    }
    static constexpr int NLENSACC = LensAcc<L0>::n + LensAcc<L1>::n + LensAcc<L2>::n;
    // light accumulator slot j: lens lights first (j < NLL), then sources
    static constexpr int light_kind_acc(int j) { return j < NLL ? light[j] : light[2 + j - NLL]; }
    static constexpr int light_n(int k) {
        return (k == glm::K_SERSIC || k == glm::K_SERSIC_ELLIPSE) ? 8 : k == glm::K_CORE_SERSIC ? 11 : 0;
    }
    static constexpr int light_off(int j) {
        int o = NLENSACC;
        for (int t = 0; t < j; ++t) o += light_n(light_kind_acc(t));
        return o;
    }
    static constexpr int NACC = NLENSACC + LightAcc<A0>::n + LightAcc<A1>::n + LightAcc<S0>::n + LightAcc<S1>::n;
    static constexpr int NRED = NACC + 2;  // + log-likelihood, chi^2
};

// Component index (in the plan's component table) of light slot i of the model.
template <class M>
__device__ __forceinline__ int light_comp_index(int i) {
    // lens-light slots map to comp[NL + i], source slots to comp[NL + NLL + (i-2)]
    return i < 2 ? M::NL + i : M::NL + M::NLL + (i - 2);
}

// ---------------------------------------------------------------------------
// Conv + pool (forward) for one strip of RB outputs in row I.
// ---------------------------------------------------------------------------
template <int SS, int KF, int RB>
__device__ __forceinline__ void conv_strip_static(const float* __restrict__ sub, const float* __restrict__ wg,
                                                  int PW, int I, int J0, float* acc) {
    // sub: LDS image, row (SS*I + p) of the strip starts 16-byte aligned (PW % 4 == 0,
    // SS*J0 % 4 == 0); wg: folded weights in global memory (wave-uniform -> scalar loads)
    constexpr int SEG = SS * (RB - 1) + KF;
    constexpr int SEG4 = (SEG + 3) / 4;
#pragma unroll
    for (int r = 0; r < RB; ++r) acc[r] = 0.0f;
    const float* base = sub + (SS * I) * PW + SS * J0;
    for (int p = 0; p < KF; ++p) {
        const float4* row4 = reinterpret_cast<const float4*>(base + p * PW);
        float seg[SEG4 * 4];
#pragma unroll
        for (int c = 0; c < SEG4; ++c) {
            float4 t = row4[c];
            seg[4 * c] = t.x; seg[4 * c + 1] = t.y; seg[4 * c + 2] = t.z; seg[4 * c + 3] = t.w;
        }
        const float* wp = wg + p * KF;
#pragma unroll
        for (int q = 0; q < KF; ++q) {
            const float wq = wp[q];
#pragma unroll
            for (int r = 0; r < RB; ++r) acc[r] = fmaf(wq, seg[SS * r + q], acc[r]);
        }
    }
}

template <int RB>
__device__ __forceinline__ void conv_strip_dyn(const float* __restrict__ sub, const float* __restrict__ w, int PW,
                                               int ss, int kf, int I, int J0, float* acc) {
#pragma unroll
    for (int r = 0; r < RB; ++r) acc[r] = 0.0f;
    const float* base = sub + (ss * I) * PW + ss * J0;
    for (int p = 0; p < kf; ++p) {
        const float* row = base + p * PW;
        const float* wp = w + p * kf;
        for (int q = 0; q < kf; ++q) {
            float wq = wp[q];
#pragma unroll
            for (int r = 0; r < RB; ++r) acc[r] = fmaf(wq, row[ss * r + q], acc[r]);
        }
    }
}

// ---------------------------------------------------------------------------
// Adjoint conv for one strip of RB sub-pixels of phase (a, b):
//   gs[r] = sum_{m,n} W[ss m + ra][ss n + rb] G[Y + da - m][X0 + r + db - n]
// G: padded adjoint image, idx(I,J) = (I+gp0)*GW + (J+gp0).
// ---------------------------------------------------------------------------
template <int SS, int KF, int RB>
__device__ __forceinline__ void adj_strip_static(const float* __restrict__ G, const float* __restrict__ wg, int GW,
                                                 int gp0, bool aligned, int a, int b, int Y, int X0, float* gs) {
    constexpr int PB = (KF - SS) / 2;  // (k-1)/2 with k = KF - SS + 1
    const int ra = (a + PB) % SS, da = (a + PB) / SS;
    const int rb = (b + PB) % SS, db = (b + PB) / SS;
    const int Ma = (KF - ra + SS - 1) / SS;
    constexpr int MMAX = (KF + SS - 1) / SS;
    constexpr int SEGA = RB + MMAX - 1;
    constexpr int SEGA4 = (SEGA + 3) / 4;
    const int Mb = (KF - rb + SS - 1) / SS;
#pragma unroll
    for (int r = 0; r < RB; ++r) gs[r] = 0.0f;
    for (int m = 0; m < Ma; ++m) {
        const float* row = G + (Y + da - m + gp0) * GW + (X0 + db - (Mb - 1) + gp0);
        const float* wr = wg + (SS * m + ra) * KF + rb;
        float seg[SEGA4 * 4];
        if (aligned) {
            const float4* row4 = reinterpret_cast<const float4*>(row);
#pragma unroll
            for (int c = 0; c < SEGA4; ++c) {
                float4 t = row4[c];
                seg[4 * c] = t.x; seg[4 * c + 1] = t.y; seg[4 * c + 2] = t.z; seg[4 * c + 3] = t.w;
            }
        } else {
#pragma unroll
            for (int j = 0; j < SEGA4 * 4; ++j) seg[j] = (j < RB + Mb - 1) ? row[j] : 0.0f;
        }
#pragma unroll
        for (int n = 0; n < MMAX; ++n) {
            if (n < Mb) {
                const float wq = wr[SS * n];
#pragma unroll
                for (int r = 0; r < RB; ++r) gs[r] = fmaf(wq, seg[r + Mb - 1 - n], gs[r]);
            }
        }
    }
}

template <int RB>
__device__ __forceinline__ void adj_strip_dyn(const float* __restrict__ G, const float* __restrict__ w, int GW,
                                              int gp0, int ss, int kf, int pb, int a, int b, int Y, int X0,
                                              float* gs) {
    const int ra = (a + pb) % ss, da = (a + pb) / ss;
    const int rb = (b + pb) % ss, db = (b + pb) / ss;
    const int Ma = (kf - ra + ss - 1) / ss;
    const int Mb = (kf - rb + ss - 1) / ss;
#pragma unroll
    for (int r = 0; r < RB; ++r) gs[r] = 0.0f;
    for (int m = 0; m < Ma; ++m) {
        const float* row = G + (Y + da - m + gp0) * GW + (X0 + db + gp0);
        const float* wr = w + (ss * m + ra) * kf + rb;
        for (int n = 0; n < Mb; ++n) {
            float wq = wr[ss * n];
#pragma unroll
            for (int r = 0; r < RB; ++r) gs[r] = fmaf(wq, row[r - n], gs[r]);
        }
    }
}

// Light slot i (0,1 lens light; 2,3 source) -> accumulator slot index.
// Lens-light slots occupy 0..NLL-1, source slots NLL..NLL+NS-1.
template <class M>
__host__ __device__ constexpr int light_acc_slot(int i) {
    return i < 2 ? i : M::NLL + (i - 2);
}


// ---------------------------------------------------------------------------
// Separable (low-rank) PSF path: W = sum_r a_r b_r^T with R <= kMaxRank.
//   forward  T_r[y][J]  = sum_q b_r[q] sub[y][ss J + q - pb]          (horizontal)
//            out[I][J]  = sum_r sum_p a_r[p] T_r[ss I + p - pb][J]       (vertical)
//   adjoint  U_r[I][x]  = sum_J b_r[x + pb - ss J] G[I][J]              (transposed horizontal)
//            gsub[y][x] = sum_r sum_I a_r[y + pb - ss I] U_r[I][x]       (transposed vertical)
// ---------------------------------------------------------------------------
constexpr int kMaxRank = 2;
constexpr int kTB = 4;     // J per horizontal-pass task
constexpr int kMaxTT = 5;  // horizontal-pass tasks per thread
constexpr int kRV = 4;     // output rows per vertical-pass task

template <int SS, int KF>
__device__ __forceinline__ void sep_hpass(const float* __restrict__ sub, const float* __restrict__ sb, int R, int PW,
                                          int yy, int J0, float (&t)[kMaxRank][kTB]) {
    constexpr int SEG = SS * (kTB - 1) + KF;
    constexpr int SEG4 = (SEG + 3) / 4;
    const float4* row4 = reinterpret_cast<const float4*>(sub + yy * PW + SS * J0);
    float seg[SEG4 * 4];
#pragma unroll
    for (int c = 0; c < SEG4; ++c) {
        float4 v = row4[c];
        seg[4 * c] = v.x; seg[4 * c + 1] = v.y; seg[4 * c + 2] = v.z; seg[4 * c + 3] = v.w;
    }
#pragma unroll
    for (int r = 0; r < kMaxRank; ++r) {
        if (r < R) {
#pragma unroll
            for (int j = 0; j < kTB; ++j) t[r][j] = 0.0f;
#pragma unroll
            for (int q = 0; q < KF; ++q) {
                const float w = sb[r * KF + q];
#pragma unroll
                for (int j = 0; j < kTB; ++j) t[r][j] = fmaf(w, seg[SS * j + q], t[r][j]);
            }
        }
    }
}

template <int SS, int KF>
__device__ __forceinline__ void sep_vpass(const float* __restrict__ T, const float* __restrict__ sa, int R, int TR,
                                          int TW, int I0, int J, float (&o)[kRV]) {
#pragma unroll
    for (int i = 0; i < kRV; ++i) o[i] = 0.0f;
#pragma unroll
    for (int r = 0; r < kMaxRank; ++r) {
        if (r < R) {
            // rows at or beyond TR only reach outputs I0 + i >= N, which are discarded; the
            // plan keeps SS (kRV - 1) rows of LDS after the last rank for them
            const float* col = T + r * TR * TW + SS * I0 * TW + J;
#pragma unroll
            for (int c = 0; c < SS * (kRV - 1) + KF; ++c) {
                const float v = col[c * TW];
#pragma unroll
                for (int i = 0; i < kRV; ++i) {
                    const int pp = c - SS * i;
                    if (pp >= 0 && pp < KF) o[i] = fmaf(sa[r * KF + pp], v, o[i]);
                }
            }
        }
    }
}

// Geometry of the transposed horizontal pass for phase b (sub-pixel column x = SS X + b):
// taps b_r[SS n + rb], n < Mb, read G columns X + db - n; sep_off(b) = db - (Mb - 1).
template <int SS, int KF>
struct SepGeom {
    static constexpr int PB = (KF - SS) / 2;
    static constexpr int MMAX = (KF + SS - 1) / SS;
    static constexpr int rb(int b) { return (b + PB) % SS; }
    static constexpr int db(int b) { return (b + PB) / SS; }
    static constexpr int Mb(int b) { return (KF - rb(b) + SS - 1) / SS; }
    static constexpr int off(int b) { return db(b) - (Mb(b) - 1); }
    static constexpr int off_min() {
        int m = off(0);
        for (int b = 1; b < SS; ++b) m = off(b) < m ? off(b) : m;
        return m;
    }
    static constexpr int off_max() {
        int m = off(0);
        for (int b = 1; b < SS; ++b) m = off(b) > m ? off(b) : m;
        return m;
    }
    static constexpr int SEGA = kTB + MMAX - 1 + off_max() - off_min();
    static constexpr int SEGA4 = (SEGA + 3) / 4;
};

// One task: kTB adjoint columns X0.. of output row I, every phase b, every rank;
// writes U_r[SS X + b][I] for X < N (U column-major, column stride US; the caller
// passes U at row I of column 0).  The phase loop is unrolled, so every tap index
// is a compile-time constant and the factors arrive by scalar loads.
template <int SS, int KF>
__device__ __forceinline__ void sep_upass(const float* __restrict__ G, float* __restrict__ U,
                                          const float* __restrict__ sb, int R, int GW, int gp0, int ustride,
                                          int US, bool aligned, int I, int X0, int N) {
    using Gm = SepGeom<SS, KF>;
    constexpr int OMIN = Gm::off_min();
    const float* row = G + (I + gp0) * GW + (X0 + OMIN + gp0);
    float seg[Gm::SEGA4 * 4];
    if (aligned) {
        const float4* row4 = reinterpret_cast<const float4*>(row);
#pragma unroll
        for (int c = 0; c < Gm::SEGA4; ++c) {
            float4 v = row4[c];
            seg[4 * c] = v.x; seg[4 * c + 1] = v.y; seg[4 * c + 2] = v.z; seg[4 * c + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < Gm::SEGA4 * 4; ++j) seg[j] = j < Gm::SEGA ? row[j] : 0.0f;
    }
#pragma unroll
    for (int r = 0; r < kMaxRank; ++r) {
        if (r < R) {
            float o[SS * kTB];
#pragma unroll
            for (int b = 0; b < SS; ++b) {
                const int sh = Gm::off(b) - OMIN;
#pragma unroll
                for (int j = 0; j < kTB; ++j) o[SS * j + b] = 0.0f;
#pragma unroll
                for (int n = 0; n < Gm::MMAX; ++n) {
                    if (n < Gm::Mb(b)) {
                        const float w = sb[r * KF + SS * n + Gm::rb(b)];
#pragma unroll
                        for (int j = 0; j < kTB; ++j)
                            o[SS * j + b] = fmaf(w, seg[sh + j + Gm::Mb(b) - 1 - n], o[SS * j + b]);
                    }
                }
            }
            float* dst = U + r * ustride + SS * X0 * US;
#pragma unroll
            for (int j = 0; j < kTB; ++j)
                if (X0 + j < N)
#pragma unroll
                    for (int b = 0; b < SS; ++b) dst[(SS * j + b) * US] = o[SS * j + b];
        }
    }
}

template <int SS, int KF>
__device__ __forceinline__ float sep_gsub(const float* __restrict__ U, const float* __restrict__ sa, int R,
                                          int ustride, int US, int up0, int a, int Yc, int x) {
    constexpr int PB = (KF - SS) / 2;
    constexpr int MMAX = (KF + SS - 1) / SS;
    const int ra = (a + PB) % SS, da = (a + PB) / SS;
    const int Ma = (KF - ra + SS - 1) / SS;
    float g = 0.0f;
#pragma unroll
    for (int r = 0; r < kMaxRank; ++r) {
        if (r < R) {
            // column x of U is contiguous: the taps are constant offsets from one address
            const float* col = U + r * ustride + x * US + (Yc + da + up0);
#pragma unroll
            for (int m = 0; m < MMAX; ++m)
                if (m < Ma) g = fmaf(sa[r * KF + SS * m + ra], col[-m], g);
        }
    }
    return g;
}

// ---------------------------------------------------------------------------
// Per-sub-pixel model evaluation with the slot structure of M.
// ---------------------------------------------------------------------------
template <class M>
struct Consts {
    glm::EplC epl[3];
    glm::ShearC shr[3];
    glm::SisC sis[3];
    glm::SieC sie[3];
    glm::TnfwC tnfw[3];
    glm::SerC ser[4];
    glm::CoreSerC core[4];
    const float* coefA[3];
    const float4* coefB[3];
};

// Read a constant block from LDS into wave-uniform (scalar) registers.
template <class T>
__device__ __forceinline__ T uniform_load(const float* src) {
    static_assert(sizeof(T) % 4 == 0, "constant blocks are 32-bit words");
    T out;
    int* o = reinterpret_cast<int*>(&out);
    const int* w = reinterpret_cast<const int*>(src);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) o[i] = __builtin_amdgcn_readfirstlane(w[i]);
    return out;
}

template <class M>
__device__ __forceinline__ void load_consts(const float* s_c, const float* s_coef, const DPlan& P, Consts<M>& C) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int k = M::lens[i];
        if (k == glm::K_EPL) {
            C.epl[i] = uniform_load<glm::EplC>(s_c + i * kConstStride);
            C.coefA[i] = s_coef + P.coef_off[i];
            C.coefB[i] = reinterpret_cast<const float4*>(s_coef + P.coef_off[i] + glm::epl_tab_a_floats(P.max_n));
        } else if (k == glm::K_SHEAR) {
            C.shr[i] = uniform_load<glm::ShearC>(s_c + i * kConstStride);
        } else if (k == glm::K_SIS) {
            C.sis[i] = uniform_load<glm::SisC>(s_c + i * kConstStride);
        } else if (k == glm::K_SIE) {
            C.sie[i] = uniform_load<glm::SieC>(s_c + i * kConstStride);
        } else if (k == glm::K_TNFW) {
            C.tnfw[i] = uniform_load<glm::TnfwC>(s_c + i * kConstStride);
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (M::light[i] == glm::K_CORE_SERSIC) {
            C.core[i] = uniform_load<glm::CoreSerC>(s_c + light_comp_index<M>(i) * kConstStride);
        } else if (M::light[i] != 0) {
            C.ser[i] = uniform_load<glm::SerC>(s_c + light_comp_index<M>(i) * kConstStride);
        }
    }
}

template <class M>
__device__ __forceinline__ float eval_forward(const Consts<M>& C, float X, float Y) {
    float val = 0.0f;
    glm::SerState st;
    glm::CoreSerState cst;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        if (M::light[i] == glm::K_CORE_SERSIC) val += glm::core_sersic_forward(C.core[i], X, Y, cst);
        else if (M::light[i] != 0) val += glm::sersic_forward(C.ser[i], X, Y, st);
    }
    float bx = X, by = Y;
    glm::EplState es;
    float sis_st[3];
    glm::SieState sie_st;
    glm::TnfwState tn_st;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float ax = 0.0f, ay = 0.0f;
        if (M::lens[i] == glm::K_EPL) glm::epl_forward<false>(C.epl[i], C.coefA[i], nullptr, X, Y, ax, ay, es);
        else if (M::lens[i] == glm::K_SHEAR) glm::shear_forward(C.shr[i], X, Y, ax, ay);
        else if (M::lens[i] == glm::K_SIS) glm::sis_forward(C.sis[i], X, Y, ax, ay, sis_st);
        else if (M::lens[i] == glm::K_SIE) glm::sie_forward<false>(C.sie[i], X, Y, ax, ay, sie_st);
        else if (M::lens[i] == glm::K_TNFW) glm::tnfw_forward<false>(C.tnfw[i], X, Y, ax, ay, tn_st);
        bx -= ax;
        by -= ay;
    }
#pragma unroll
    for (int i = 2; i < 4; ++i) {
        if (M::light[i] == glm::K_CORE_SERSIC) val += glm::core_sersic_forward(C.core[i], bx, by, cst);
        else if (M::light[i] != 0) val += glm::sersic_forward(C.ser[i], bx, by, st);
    }
    return val;
}

template <class M>
__device__ __forceinline__ void eval_adjoint(const Consts<M>& C, float X, float Y, float g, float* acc) {
    float val = 0.0f;
    glm::SerState lst[4];
    glm::CoreSerState cst[4];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        if (M::light[i] == glm::K_CORE_SERSIC) val += glm::core_sersic_forward(C.core[i], X, Y, cst[i]);
        else if (M::light[i] != 0) val += glm::sersic_forward<true>(C.ser[i], X, Y, lst[i]);
    }
    float bx = X, by = Y;
    glm::EplState es[3];
    float sis_st[3][3];
    glm::SieState sie_st[3];
    glm::TnfwState tn_st[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float ax = 0.0f, ay = 0.0f;
        if (M::lens[i] == glm::K_EPL) glm::epl_forward<true>(C.epl[i], C.coefA[i], C.coefB[i], X, Y, ax, ay, es[i]);
        else if (M::lens[i] == glm::K_SHEAR) glm::shear_forward(C.shr[i], X, Y, ax, ay);
        else if (M::lens[i] == glm::K_SIS) glm::sis_forward(C.sis[i], X, Y, ax, ay, sis_st[i]);
        else if (M::lens[i] == glm::K_SIE) glm::sie_forward<true>(C.sie[i], X, Y, ax, ay, sie_st[i]);
        else if (M::lens[i] == glm::K_TNFW) glm::tnfw_forward<true>(C.tnfw[i], X, Y, ax, ay, tn_st[i]);
        bx -= ax;
        by -= ay;
    }
#pragma unroll
    for (int i = 2; i < 4; ++i) {
        if (M::light[i] == glm::K_CORE_SERSIC) val += glm::core_sersic_forward(C.core[i], bx, by, cst[i]);
        else if (M::light[i] != 0) val += glm::sersic_forward<true>(C.ser[i], bx, by, lst[i]);
    }
    if (isnan(val)) return;  // zeroed in the forward pass: no contribution
    float gbx = 0.0f, gby = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (M::light[i] != 0) {
            float gdx, gdy;
            float* la = acc + M::light_off(light_acc_slot<M>(i));
            if (M::light[i] == glm::K_CORE_SERSIC) glm::core_sersic_adjoint(C.core[i], cst[i], g, la, gdx, gdy);
            else glm::sersic_adjoint(C.ser[i], lst[i], g, la, gdx, gdy);
            if (i >= 2) { gbx += gdx; gby += gdy; }
        }
    }
    const float Hx = -gbx, Hy = -gby;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (M::lens[i] == glm::K_EPL) glm::epl_adjoint(C.epl[i], es[i], Hx, Hy, acc + M::lens_off(i));
        else if (M::lens[i] == glm::K_SHEAR) glm::shear_adjoint(X, Y, Hx, Hy, acc + M::lens_off(i));
        else if (M::lens[i] == glm::K_SIS) glm::sis_adjoint(C.sis[i], sis_st[i], Hx, Hy, acc + M::lens_off(i));
        else if (M::lens[i] == glm::K_SIE) glm::sie_adjoint(C.sie[i], sie_st[i], Hx, Hy, acc + M::lens_off(i));
        else if (M::lens[i] == glm::K_TNFW) glm::tnfw_adjoint(C.tnfw[i], tn_st[i], Hx, Hy, acc + M::lens_off(i));
    }
}

// ---------------------------------------------------------------------------
// The kernel
// ---------------------------------------------------------------------------
constexpr int kRB = 4;    // outputs per thread strip (forward conv)
constexpr int kRB2 = 4;   // sub-pixels per thread strip (adjoint conv)
constexpr int kMaxT = 4;  // forward strips per thread held in registers

// Sum over the 64 lanes of a wave, returned wave-uniform.  Every lane must be
// active.  DPP moves within each row of 16 lanes (quad_perm [1,0,3,2],
// quad_perm [2,3,0,1], row_ror 4, row_ror 8) leave the row sum in all 16
// lanes; the four row sums are then read from lanes 0, 16, 32 and 48.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float wave_sum(float v) {
    v += dpp_f<0xB1>(v);
    v += dpp_f<0x4E>(v);
    v += dpp_f<0x124>(v);
    v += dpp_f<0x128>(v);
    const int b = __builtin_bit_cast(int, v);
    return (__builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 0)) +
            __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 16))) +
           (__builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 32)) +
            __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 48)));
}

// EPL coefficient n = lane (0 <= n <= NN < 64) of glm::epl_coef, computed by one
// wave: the serial recurrence (a, da) -> (a r_k, da r_k + a dr_k) is associative in
// the pairs (r_k, dr_k), so an inclusive scan over the lanes (r_k, dr_k scaled by -f
// for k >= 2, identity elsewhere) gives pw a_n and pw da_n in log2(64) steps.
// Every lane of the wave must be active.
__device__ __forceinline__ void epl_coef_scan(int n, int NN, double f, double t, float& c, float& dt, float& df) {
    double r = 1.0, dr = 0.0;
    if (n >= 1 && n <= NN) {
        const double inv = 1.0 / (2.0 * n + 2.0 - t);
        r = (2.0 * n - 2.0 + t) * inv;
        dr = 4.0 * n * inv * inv;
        if (n > 1) { r *= -f; dr *= -f; }
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double pr = __shfl_up(r, o, 64), pdr = __shfl_up(dr, o, 64);
        if (n >= o) { dr = pdr * r + pr * dr; r = pr * r; }
    }
    if (n == 0) { c = 1.0f; dt = 0.0f; df = 0.0f; return; }
    c = (float)(-f * r);
    dt = (float)(-f * dr);
    df = (float)(-(double)n * r);
}

// ---- Per-particle steps of the banded kernel (band_kernels.hip) -----------------------
// The resident-image particle_kernel below performs the same prologue (A) and reductions
// (F) written out inline: calling these helpers from it costs the headline instantiation
// 68 bytes of scratch spills per lane (its register allocation sits at the limit).
//
// A: z -> x through the bijectors, log prior + fldj (thread tid < d handles parameter tid;
// zin is its z, or its x when !use_z).  s_x = [x | dxdz | dlpdx | dfldjdz | gx | lpj].
__device__ __forceinline__ void prologue_prior(const DPlan& P, bool use_z, int tid, float zin, float* s_x) {
    const int d = P.d;
    if (tid < d) {
        glm::PriorOut o;
        if (use_z) {
            glm::prior_eval(P.prior_kind[tid], P.prior_args + 8 * tid, zin, P.fldj_at_x, o);
        } else {
            o.x = zin; o.dxdz = 1.0f; o.lp = 0.0f; o.dlpdx = 0.0f; o.fldj = 0.0f; o.dfldjdz = 0.0f;
        }
        s_x[tid] = o.x;
        s_x[d + tid] = o.dxdz;
        s_x[2 * d + tid] = o.dlpdx;
        s_x[3 * d + tid] = o.dfldjdz;
        s_x[4 * d + tid] = 0.0f;
        s_x[5 * d + tid] = o.lp + o.fldj;
    }
}

// Component constants: component ci is prepared by lane 0 of wave ci (the components run
// side by side) into s_c + ci * kConstStride.
__device__ __forceinline__ void prologue_components(const DPlan& P, int tid, const float* s_x, float* s_c) {
    if ((tid & 63) == 0 && (tid >> 6) < P.ncomp) {
        const int ci = tid >> 6;
        const DComp& cd = P.comp[ci];
        float pv[GL_MAX_COMPONENT_PARAMS];
#pragma unroll
        for (int j = 0; j < GL_MAX_COMPONENT_PARAMS; ++j) pv[j] = cd.col[j] >= 0 ? s_x[cd.col[j]] : cd.cst[j];
        float* dst = s_c + ci * kConstStride;
        if (cd.kind == glm::K_EPL) {
            glm::EplC c;
            int nov = P.batch_n ? P.batch_n[ci] : -1;
            glm::epl_prepare(pv[0], pv[1], pv[2], pv[3], pv[4], pv[5], cd.niter, nov, c);
            if (c.N > P.max_n) c.N = P.max_n;
            *reinterpret_cast<glm::EplC*>(dst) = c;
        } else if (cd.kind == glm::K_SHEAR) {
            glm::ShearC c{pv[0], pv[1]};
            *reinterpret_cast<glm::ShearC*>(dst) = c;
        } else if (cd.kind == glm::K_SIS) {
            glm::SisC c{pv[0], pv[1], pv[2]};
            *reinterpret_cast<glm::SisC*>(dst) = c;
        } else if (cd.kind == glm::K_SIE) {
            glm::SieC c;
            glm::sie_prepare(pv[0], pv[1], pv[2], pv[3], pv[4], c);
            *reinterpret_cast<glm::SieC*>(dst) = c;
        } else if (cd.kind == glm::K_TNFW) {
            glm::TnfwC c;
            glm::tnfw_prepare(pv[0], pv[1], pv[2], pv[3], pv[4], c);
            *reinterpret_cast<glm::TnfwC*>(dst) = c;
        } else if (cd.kind == glm::K_SERSIC_ELLIPSE) {
            glm::SerC c;
            glm::sersic_prepare(pv[0], pv[1], pv[2], pv[3], pv[4], pv[5], cd.lstsq ? 1.0f : pv[6], 1, c);
            *reinterpret_cast<glm::SerC*>(dst) = c;
        } else if (cd.kind == glm::K_SERSIC) {
            glm::SerC c;
            glm::sersic_prepare(pv[0], pv[1], 0.0f, 0.0f, pv[2], pv[3], cd.lstsq ? 1.0f : pv[4], 0, c);
            *reinterpret_cast<glm::SerC*>(dst) = c;
        } else if (cd.kind == glm::K_CORE_SERSIC) {
            glm::CoreSerC c;
            glm::core_sersic_prepare(pv[0], pv[1], pv[2], pv[3], pv[4], pv[5], pv[6], pv[7], pv[8],
                                     cd.lstsq ? 1.0f : pv[9], c);
            *reinterpret_cast<glm::CoreSerC*>(dst) = c;
        }
    }
}

// EPL series coefficient tables: the table of EPL slot i is built by wave i at
// s_coef + P.coef_off[i] (c_n, then (c_n, dc_n/dt, dc_n/df, 0)), zero padded above N.
template <class M>
__device__ __forceinline__ void prologue_tables(const DPlan& P, int tid, const float* s_c, float* s_coef) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (M::lens[i] == glm::K_EPL && (tid >> 6) == i) {
            const glm::EplC* c = reinterpret_cast<const glm::EplC*>(s_c + i * kConstStride);
            float* ca = s_coef + P.coef_off[i];
            float4* cb = reinterpret_cast<float4*>(ca + glm::epl_tab_a_floats(P.max_n));
            const int NN = c->N;
            const int n = tid & 63;
            if (NN < 64) {
                float cv, dt, df;
                epl_coef_scan(n, NN, (double)c->f, (double)c->t, cv, dt, df);
                if (n <= NN) {
                    ca[n] = cv;
                    cb[n] = make_float4(cv, dt, df, 0.0f);
                }
            } else {
                for (int k = n; k <= NN; k += 64) {
                    float cv, dt, df;
                    glm::epl_coef(k, (double)c->f, (double)c->t, cv, dt, df);
                    ca[k] = cv;
                    cb[k] = make_float4(cv, dt, df, 0.0f);
                }
            }
            for (int k = NN + 1 + n; k <= glm::epl_pad(NN); k += 64) {  // zero padding above N
                ca[k] = 0.0f;
                cb[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
        }
    }
}

// F: reductions, per-particle chain rule, outputs.  Every thread of the workgroup calls it (after a barrier) with its partial sums of
// the log-likelihood, chi^2 and the reverse-mode statistics acc[].  Writes lp / chi2
// (LOGPROB, MAP), grad (LOGPROB: d/dz, VJP: d/dx) or the Keras-Adam update of z, m, v
// (MAP; zin, m_in, v_in: thread tid < d's entries).  glp: the weight of lp in the
// gradient (MAP: minus the loss weight).
template <int BLOCK, class M, bool GRAD>
__device__ __forceinline__ void finish_particle(const DPlan& P, const Consts<M>& C, const KArgs& A, int tid, int pid,
                                                float ll_acc, float chi_acc, const float* acc, float glp,
                                                bool bad_lens, float zin, float m_in, float v_in, float* s_x,
                                                float* s_red) {
    const int d = P.d;
    const int N = P.N;
    const float inv_np = 1.0f / (float)(N * N);
    const int lane = tid & 63, wave = tid >> 6;
    constexpr int NW = BLOCK / 64;
    float part = wave_sum(ll_acc);
    if (lane == 0) s_red[wave * M::NRED + 0] = part;
    part = wave_sum(chi_acc);
    if (lane == 0) s_red[wave * M::NRED + 1] = part;
    if (GRAD) {
#pragma unroll
        for (int e = 0; e < M::NACC; ++e) {
            part = wave_sum(acc[e]);
            if (lane == 0) s_red[wave * M::NRED + 2 + e] = part;
        }
    }
    __syncthreads();
    float* s_tot = s_red + NW * M::NRED;
    if (tid < M::NRED) {
        float t = 0.0f;
#pragma unroll
        for (int w = 0; w < NW; ++w) t += s_red[w * M::NRED + tid];
        s_tot[tid] = t;
    }
    __syncthreads();

    if (tid == 0) {
        float lp_like = s_tot[0] - (float)(N * N) * glm::kHalfLn2Pi;
        float chi2 = s_tot[1] * inv_np;
        float lp_prior = 0.0f;
        for (int j = 0; j < d; ++j) lp_prior += s_x[5 * d + j];
        if (A.mode == M_LOGPROB || A.mode == M_MAP) {
            if (A.lp) A.lp[pid] = bad_lens ? __builtin_nanf("") : lp_like + lp_prior;
            if (A.chi2) A.chi2[pid] = bad_lens ? __builtin_nanf("") : chi2;
        }
    }
    if (GRAD && A.mode != M_SIM) {
        // per-component chain rule: lens slot i on lane 0 of wave 1 + i, light slot i on
        // lane 0 of wave 4 + i; contributions meet in gx through LDS atomics
        const float* tot = s_tot + 2;
        float* gx = s_x + 4 * d;
        const int w = tid >> 6;
        if ((tid & 63) == 0) {
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const int k = M::lens[i];
                if (k == 0 || w != 1 + i) continue;
                const DComp& cd = P.comp[i];
                float g[8];
                int np = 0;
                if (k == glm::K_EPL) { glm::epl_finish(C.epl[i], tot + M::lens_off(i), g); np = 6; }
                else if (k == glm::K_SHEAR) { g[0] = tot[M::lens_off(i)]; g[1] = tot[M::lens_off(i) + 1]; np = 2; }
                else if (k == glm::K_SIS) { glm::sis_finish(tot + M::lens_off(i), g); np = 3; }
                else if (k == glm::K_SIE) { glm::sie_finish(C.sie[i], tot + M::lens_off(i), g); np = 5; }
                else if (k == glm::K_TNFW) { glm::tnfw_finish(C.tnfw[i], tot + M::lens_off(i), g); np = 5; }
                for (int p = 0; p < np; ++p)
                    if (cd.col[p] >= 0) atomicAdd(&gx[cd.col[p]], g[p]);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (M::light[i] == 0 || w != 4 + i) continue;
                const DComp& cd = P.comp[light_comp_index<M>(i)];
                float g[10];
                if (M::light[i] == glm::K_CORE_SERSIC) {
                    glm::core_sersic_finish(C.core[i], tot + M::light_off(light_acc_slot<M>(i)), g);
                    for (int p = 0; p < (cd.lstsq ? 9 : 10); ++p)
                        if (cd.col[p] >= 0) atomicAdd(&gx[cd.col[p]], g[p]);
                    continue;
                }
                glm::sersic_finish(C.ser[i], tot + M::light_off(light_acc_slot<M>(i)), g);
                if (M::light[i] == glm::K_SERSIC_ELLIPSE) {
                    for (int p = 0; p < (cd.lstsq ? 6 : 7); ++p)
                        if (cd.col[p] >= 0) atomicAdd(&gx[cd.col[p]], g[p]);
                } else {
                    const int map[5] = {0, 1, 4, 5, 6};
                    for (int p = 0; p < (cd.lstsq ? 4 : 5); ++p)
                        if (cd.col[p] >= 0) atomicAdd(&gx[cd.col[p]], g[map[p]]);
                }
            }
        }
        __syncthreads();
        if (tid < d) {
            const float gxj = s_x[4 * d + tid];
            if (A.mode == M_VJP) {
                A.grad[(size_t)pid * d + tid] = gxj;
            } else {
                const float gz = (gxj + glp * s_x[2 * d + tid]) * s_x[d + tid] + glp * s_x[3 * d + tid];
                if (A.mode == M_LOGPROB) {
                    A.grad[(size_t)pid * d + tid] = gz;
                } else {  // M_MAP: Keras Adam
                    const size_t o = (size_t)pid * d + tid;
                    float mm = A.beta1 * m_in + (1.0f - A.beta1) * gz;
                    float vv = A.beta2 * v_in + (1.0f - A.beta2) * gz * gz;
                    A.m[o] = mm;
                    A.v[o] = vv;
                    A.z_io[o] = zin - A.lr_t * mm / (sqrtf(vv) + A.eps);
                }
            }
        }
    }
}

template <int BLOCK, class M, bool GRAD>
__global__ __launch_bounds__(BLOCK) void particle_kernel(const DPlan P, const KArgs A) {
    if (A.limit && A.limit_k >= *A.limit) return;  // uniform over the launch
    extern __shared__ float4 smem4[];
    float* smem = reinterpret_cast<float*>(smem4);
    const int pid = A.order ? A.order[blockIdx.x] : (int)blockIdx.x;
    const int tid = threadIdx.x;
    const int d = P.d;
    float* s_sub = smem;
    float* s_x = smem + P.off_x;  // [x | dxdz | dlpdx | dfldjdz | gx | lpj]
    float* s_c = smem + P.off_c;
    float* s_coef = smem + P.off_coef;
    float* s_red = smem + P.off_red;
    const bool use_z = (A.mode == M_LOGPROB || A.mode == M_MAP);

    // ---- A: prologue -----------------------------------------------------
    // z, and for MAP the Adam moments, are loaded once here and kept in registers
    float zin = 0.0f, m_in = 0.0f, v_in = 0.0f;
    if (tid < d) {
        const size_t o = (size_t)pid * d + tid;
        zin = A.in[o];
        if (A.mode == M_MAP) { m_in = A.m[o]; v_in = A.v[o]; }
    }
    for (int i = tid; i < (P.sub_size + 3) / 4; i += BLOCK) smem4[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (tid < d) {
        glm::PriorOut o;
        if (use_z) {
            glm::prior_eval(P.prior_kind[tid], P.prior_args + 8 * tid, zin, P.fldj_at_x, o);
        } else {
            o.x = zin; o.dxdz = 1.0f; o.lp = 0.0f; o.dlpdx = 0.0f; o.fldj = 0.0f; o.dfldjdz = 0.0f;
        }
        s_x[tid] = o.x;
        s_x[d + tid] = o.dxdz;
        s_x[2 * d + tid] = o.dlpdx;
        s_x[3 * d + tid] = o.dfldjdz;
        s_x[4 * d + tid] = 0.0f;
        s_x[5 * d + tid] = o.lp + o.fldj;
    }
    __syncthreads();
    // component ci is prepared by lane 0 of wave ci (the components run side by side)
    if ((tid & 63) == 0 && (tid >> 6) < P.ncomp) {
        const int ci = tid >> 6;
        const DComp& cd = P.comp[ci];
        float pv[GL_MAX_COMPONENT_PARAMS];
#pragma unroll
        for (int j = 0; j < GL_MAX_COMPONENT_PARAMS; ++j) pv[j] = cd.col[j] >= 0 ? s_x[cd.col[j]] : cd.cst[j];
        float* dst = s_c + ci * kConstStride;
        if (cd.kind == glm::K_EPL) {
            glm::EplC c;
            int nov = P.batch_n ? P.batch_n[ci] : -1;
            glm::epl_prepare(pv[0], pv[1], pv[2], pv[3], pv[4], pv[5], cd.niter, nov, c);
            if (c.N > P.max_n) c.N = P.max_n;
            *reinterpret_cast<glm::EplC*>(dst) = c;
        } else if (cd.kind == glm::K_SHEAR) {
            glm::ShearC c{pv[0], pv[1]};
            *reinterpret_cast<glm::ShearC*>(dst) = c;
        } else if (cd.kind == glm::K_SIS) {
            glm::SisC c{pv[0], pv[1], pv[2]};
            *reinterpret_cast<glm::SisC*>(dst) = c;
        } else if (cd.kind == glm::K_SIE) {
            glm::SieC c;
            glm::sie_prepare(pv[0], pv[1], pv[2], pv[3], pv[4], c);
            *reinterpret_cast<glm::SieC*>(dst) = c;
        } else if (cd.kind == glm::K_TNFW) {
            glm::TnfwC c;
            glm::tnfw_prepare(pv[0], pv[1], pv[2], pv[3], pv[4], c);
            *reinterpret_cast<glm::TnfwC*>(dst) = c;
        } else if (cd.kind == glm::K_SERSIC_ELLIPSE) {
            glm::SerC c;
            glm::sersic_prepare(pv[0], pv[1], pv[2], pv[3], pv[4], pv[5], cd.lstsq ? 1.0f : pv[6], 1, c);
            *reinterpret_cast<glm::SerC*>(dst) = c;
        } else if (cd.kind == glm::K_SERSIC) {
            glm::SerC c;
            glm::sersic_prepare(pv[0], pv[1], 0.0f, 0.0f, pv[2], pv[3], cd.lstsq ? 1.0f : pv[4], 0, c);
            *reinterpret_cast<glm::SerC*>(dst) = c;
        } else if (cd.kind == glm::K_CORE_SERSIC) {
            glm::CoreSerC c;
            glm::core_sersic_prepare(pv[0], pv[1], pv[2], pv[3], pv[4], pv[5], pv[6], pv[7], pv[8],
                                     cd.lstsq ? 1.0f : pv[9], c);
            *reinterpret_cast<glm::CoreSerC*>(dst) = c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        // the table of EPL slot i is built by wave i
        if (M::lens[i] == glm::K_EPL && (tid >> 6) == i) {
            const glm::EplC* c = reinterpret_cast<const glm::EplC*>(s_c + i * kConstStride);
            float* ca = s_coef + P.coef_off[i];
            float4* cb = reinterpret_cast<float4*>(ca + glm::epl_tab_a_floats(P.max_n));
            const int NN = c->N;
            const int n = tid & 63;
            if (NN < 64) {
                float cv, dt, df;
                epl_coef_scan(n, NN, (double)c->f, (double)c->t, cv, dt, df);
                if (n <= NN) {
                    ca[n] = cv;
                    cb[n] = make_float4(cv, dt, df, 0.0f);
                }
            } else {
                for (int k = n; k <= NN; k += 64) {
                    float cv, dt, df;
                    glm::epl_coef(k, (double)c->f, (double)c->t, cv, dt, df);
                    ca[k] = cv;
                    cb[k] = make_float4(cv, dt, df, 0.0f);
                }
            }
            for (int k = NN + 1 + n; k <= glm::epl_pad(NN); k += 64) {  // zero padding above N
                ca[k] = 0.0f;
                cb[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
        }
    }
    __syncthreads();

    Consts<M> C;
    load_consts<M>(s_c, s_coef, P, C);

    const int N = P.N, ss = P.ss, nsub = P.nsub, PW = P.PW;
    const float* sub0 = s_sub + P.sub_base;  // sub0[(y+pb)*PW + x]

    // ---- B: ray-trace into LDS -----------------------------------------------
    {
        int i = tid / nsub, j = tid - (tid / nsub) * nsub;
        const int nsub2 = nsub * nsub;
        for (int s = tid; s < nsub2; s += BLOCK) {
            float X = fmaf(P.M00, (float)j, fmaf(P.M01, (float)i, P.ox));
            float Y = fmaf(P.M10, (float)j, fmaf(P.M11, (float)i, P.oy));
            float val = eval_forward<M>(C, X, Y);
            if (isnan(val)) val = 0.0f;
            s_sub[P.sub_base + (i + P.pb) * PW + j] = val;
            j += BLOCK;
            while (j >= nsub) { j -= nsub; ++i; }
        }
    }
    __syncthreads();

    // ---- C: conv + pool + likelihood ------------------------------------------
    float ll_acc = 0.0f, chi_acc = 0.0f;
    const float cf = P.cf;
    const float inv_np = 1.0f / (float)(N * N);
    // observed image of this particle's lens; an index out of range gives NaN lp / chi2
    int lens_i = A.lens ? A.lens[pid] : 0;
    const bool bad_lens = lens_i < 0 || lens_i >= P.n_obs;
    if (bad_lens) lens_i = 0;
    const size_t obs_off = (size_t)lens_i * N * N;
    const float* obs_p = P.obs + obs_off;
    const float* ivar_p = P.inv_var_obs + obs_off;
    const float* lsig_p = P.log_sig_obs + obs_off;
    const float glp = A.mode == M_MAP ? (A.weight ? A.gscale * A.weight[pid] : A.gscale)
                                      : (A.g_lp ? A.g_lp[pid] : 1.0f);
    const float gch = (A.mode == M_LOGPROB && A.g_chi2) ? A.g_chi2[pid] : 0.0f;
    const int ssk = ss * 1000 + P.kf;
    const int sepk = P.sep_rank > 0 ? ssk : 0;  // separable path specialisation (0: direct)
    const int R = P.sep_rank;

    // per-output pixel: write the image or the likelihood terms; returns cf * dL/dimage
    auto out_pixel = [&](int I, int J, float conv) -> float {
        const float im = cf * conv;
        const int pix = I * N + J;
        if (A.mode == M_SIM) {
            A.img[(size_t)pid * N * N + pix] = im;
            return 0.0f;
        }
        if (A.mode == M_VJP) return cf * A.g_img[(size_t)pid * N * N + pix];
        float ll, r2, dll, dr2;
        if (P.lik == GL_LIK_FORWARD)
            glm::lik_forward(im, obs_p[pix], P.bkg2, P.inv_texp, ll, r2, dll, dr2);
        else
            glm::lik_backward(im, obs_p[pix], ivar_p[pix], lsig_p[pix], ll, r2, dll, dr2);
        ll_acc += ll;
        chi_acc += r2;
        return cf * (glp * dll + gch * inv_np * dr2);
    };

    // forward outputs of this thread; gv[tt][r] belongs to output
    //   direct:    (I, J0 + r) with task = I*NSTR + J0/kRB
    //   separable: (I0 + r, J) with task = (I0/kRV)*N + J
    float gv[kMaxT][kRB];
    const int NSTR = (N + kRB - 1) / kRB;
    const int ntask = N * NSTR;
    static_assert(kRB == kRV, "register layout shared by both paths");
#pragma unroll
    for (int tt = 0; tt < kMaxT; ++tt)
#pragma unroll
        for (int r = 0; r < kRB; ++r) gv[tt][r] = 0.0f;

    if (sepk) {
        // C1: horizontal pass over all padded sub rows -> registers
        const int TR = P.nsub + P.k - 1;
        const int TW = N;
        const int htask = TR * NSTR;
        float tv[kMaxTT][kMaxRank][kTB];
#pragma unroll
        for (int tt = 0; tt < kMaxTT; ++tt) {
            const int task = tid + tt * BLOCK;
            if (task < htask) {
                const int yy = task / NSTR;
                const int J0 = (task - yy * NSTR) * kTB;
                if (sepk == 2022) sep_hpass<2, 22>(s_sub, P.sb, R, PW, yy, J0, tv[tt]);
                else if (sepk == 1021) sep_hpass<1, 21>(s_sub, P.sb, R, PW, yy, J0, tv[tt]);
                else sep_hpass<3, 23>(s_sub, P.sb, R, PW, yy, J0, tv[tt]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int tt = 0; tt < kMaxTT; ++tt) {
            const int task = tid + tt * BLOCK;
            if (task < htask) {
                const int yy = task / NSTR;
                const int J0 = (task - yy * NSTR) * kTB;
#pragma unroll
                for (int r = 0; r < kMaxRank; ++r)
                    if (r < R)
#pragma unroll
                        for (int j = 0; j < kTB; ++j)
                            if (J0 + j < N) s_sub[r * TR * TW + yy * TW + J0 + j] = tv[tt][r][j];
            }
        }
        __syncthreads();
        // C2: vertical pass, lanes along J
#pragma unroll
        for (int tt = 0; tt < kMaxT; ++tt) {
            const int task = tid + tt * BLOCK;
            if (task < ntask) {
                const int I0 = (task / N) * kRV;
                const int J = task - (task / N) * N;
                float o[kRV];
                if (sepk == 2022) sep_vpass<2, 22>(s_sub, P.sa, R, TR, TW, I0, J, o);
                else if (sepk == 1021) sep_vpass<1, 21>(s_sub, P.sa, R, TR, TW, I0, J, o);
                else sep_vpass<3, 23>(s_sub, P.sa, R, TR, TW, I0, J, o);
#pragma unroll
                for (int i = 0; i < kRV; ++i)
                    if (I0 + i < N) gv[tt][i] = out_pixel(I0 + i, J, o[i]);
            }
        }
    } else {
#pragma unroll
        for (int tt = 0; tt < kMaxT; ++tt) {
            const int task = tid + tt * BLOCK;
            if (task < ntask) {
                const int I = task / NSTR;
                const int J0 = (task - I * NSTR) * kRB;
                float acc[kRB];
                const float* base = s_sub + P.sub_base - P.pb;  // column offset -pb folded in
                if (ssk == 2022) conv_strip_static<2, 22, kRB>(base, P.w, PW, I, J0, acc);
                else if (ssk == 2026) conv_strip_static<2, 26, kRB>(base, P.w, PW, I, J0, acc);
                else if (ssk == 1021) conv_strip_static<1, 21, kRB>(base, P.w, PW, I, J0, acc);
                else conv_strip_dyn<kRB>(base, P.w, PW, ss, P.kf, I, J0, acc);
#pragma unroll
                for (int r = 0; r < kRB; ++r)
                    if (J0 + r < N) gv[tt][r] = out_pixel(I, J0 + r, acc[r]);
            }
        }
    }

    if (A.mode == M_SIM) return;

    float acc[M::NACC > 0 ? M::NACC : 1];
#pragma unroll
    for (int e = 0; e < M::NACC; ++e) acc[e] = 0.0f;

    if (GRAD) {
        __syncthreads();
        for (int i = tid; i < P.g_size / 4; i += BLOCK) smem4[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        __syncthreads();
        const int GW = P.GW, gp0 = P.gp0;
#pragma unroll
        for (int tt = 0; tt < kMaxT; ++tt) {
            const int task = tid + tt * BLOCK;
            if (task < ntask) {
                if (sepk) {
                    const int I0 = (task / N) * kRV;
                    const int J = task - (task / N) * N;
#pragma unroll
                    for (int i = 0; i < kRV; ++i)
                        if (I0 + i < N) s_sub[(I0 + i + gp0) * GW + (J + gp0)] = gv[tt][i];
                } else {
                    const int I = task / NSTR;
                    const int J0 = (task - I * NSTR) * kRB;
#pragma unroll
                    for (int r = 0; r < kRB; ++r)
                        if (J0 + r < N) s_sub[(I + gp0) * GW + (J0 + r + gp0)] = gv[tt][r];
                }
            }
        }
        __syncthreads();

        if (sepk) {
            // ---- D1: transposed horizontal pass G -> U, lanes along X ---------------
            // U lives after G in LDS (no overlap), so tasks store straight to it.
            const int XS = (N + kTB - 1) / kTB;
            const int utask = N * XS;
            // U_r is column-major: nsub columns of P.urows rows, column stride US (odd,
            // so lanes on consecutive columns hit distinct banks)
            const int UR = P.urows, US = P.ucol, up0 = gp0;
            const int ustride = nsub * US;
            float* s_u = s_sub + P.u_off;
            const int zr = UR - N;  // rows of a column outside [up0, up0 + N): zero
            for (int i = tid; i < R * nsub * zr; i += BLOCK) {
                const int rc = i / zr, k = i - rc * zr;  // rc = r * nsub + column
                s_u[rc * US + (k < up0 ? k : k + N)] = 0.0f;
            }
            for (int task = tid; task < utask; task += BLOCK) {
                const int I = task / XS;
                const int X0 = (task - I * XS) * kTB;
                float* urow = s_u + (I + up0);
                if (sepk == 2022)
                    sep_upass<2, 22>(s_sub, urow, P.sb, R, GW, gp0, ustride, US, P.sep_aligned, I, X0, N);
                else if (sepk == 1021)
                    sep_upass<1, 21>(s_sub, urow, P.sb, R, GW, gp0, ustride, US, P.sep_aligned, I, X0, N);
                else
                    sep_upass<3, 23>(s_sub, urow, P.sb, R, GW, gp0, ustride, US, P.sep_aligned, I, X0, N);
            }
            __syncthreads();
            // ---- D2 + E: per sub-pixel transposed vertical pass + recompute, one row per wave ----
            const int lane_ = tid & 63, wave_ = __builtin_amdgcn_readfirstlane(tid >> 6);
            const int H = (nsub + 63) / 64;
            for (int wt = wave_; wt < nsub * H; wt += BLOCK / 64) {
                const int y = wt / H;
                const int x = (wt - y * H) * 64 + lane_;
                const int a = y % ss, Yc = y / ss;
                if (x < nsub) {
                    float g;
                    if (sepk == 2022) g = sep_gsub<2, 22>(s_u, P.sa, R, ustride, US, up0, a, Yc, x);
                    else if (sepk == 1021) g = sep_gsub<1, 21>(s_u, P.sa, R, ustride, US, up0, a, Yc, x);
                    else g = sep_gsub<3, 23>(s_u, P.sa, R, ustride, US, up0, a, Yc, x);
                    float X = fmaf(P.M00, (float)x, fmaf(P.M01, (float)y, P.ox));
                    float Y = fmaf(P.M10, (float)x, fmaf(P.M11, (float)y, P.oy));
                    eval_adjoint<M>(C, X, Y, g, acc);
                }
            }
        } else {
            // ---- D + E: adjoint conv (polyphase) and recompute ---------------------
            const int XSTR = (N + kRB2 - 1) / kRB2;
            const int ptask = N * XSTR;
            for (int a = 0; a < ss; ++a) {
                for (int b = 0; b < ss; ++b) {
                    for (int task = tid; task < ptask; task += BLOCK) {
                        const int Yc = task / XSTR;
                        const int X0 = (task - Yc * XSTR) * kRB2;
                        float gs[kRB2];
                        if (ssk == 2022) adj_strip_static<2, 22, kRB2>(s_sub, P.w, GW, gp0, P.adj_aligned, a, b, Yc, X0, gs);
                        else if (ssk == 2026) adj_strip_static<2, 26, kRB2>(s_sub, P.w, GW, gp0, P.adj_aligned, a, b, Yc, X0, gs);
                        else if (ssk == 1021) adj_strip_static<1, 21, kRB2>(s_sub, P.w, GW, gp0, P.adj_aligned, a, b, Yc, X0, gs);
                        else adj_strip_dyn<kRB2>(s_sub, P.w, GW, gp0, ss, P.kf, P.pb, a, b, Yc, X0, gs);
                        const int yi = ss * Yc + a;
#pragma unroll
                        for (int r = 0; r < kRB2; ++r) {
                            const int xj = ss * (X0 + r) + b;
                            if (X0 + r < N) {
                                float X = fmaf(P.M00, (float)xj, fmaf(P.M01, (float)yi, P.ox));
                                float Y = fmaf(P.M10, (float)xj, fmaf(P.M11, (float)yi, P.oy));
                                eval_adjoint<M>(C, X, Y, gs[r], acc);
                            }
                        }
                    }
                }
            }
        }
    }

    // ---- F: reduction ------------------------------------------------------
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6;
    constexpr int NW = BLOCK / 64;
    float part = wave_sum(ll_acc);
    if (lane == 0) s_red[wave * M::NRED + 0] = part;
    part = wave_sum(chi_acc);
    if (lane == 0) s_red[wave * M::NRED + 1] = part;
    if (GRAD) {
#pragma unroll
        for (int e = 0; e < M::NACC; ++e) {
            part = wave_sum(acc[e]);
            if (lane == 0) s_red[wave * M::NRED + 2 + e] = part;
        }
    }
    __syncthreads();
    float* s_tot = s_red + NW * M::NRED;
    if (tid < M::NRED) {
        float t = 0.0f;
#pragma unroll
        for (int w = 0; w < NW; ++w) t += s_red[w * M::NRED + tid];
        s_tot[tid] = t;
    }
    __syncthreads();

    if (tid == 0) {
        float lp_like = s_tot[0] - (float)(N * N) * glm::kHalfLn2Pi;
        float chi2 = s_tot[1] * inv_np;
        float lp_prior = 0.0f;
        for (int j = 0; j < d; ++j) lp_prior += s_x[5 * d + j];
        if (A.mode == M_LOGPROB || A.mode == M_MAP) {
            if (A.lp) A.lp[pid] = bad_lens ? __builtin_nanf("") : lp_like + lp_prior;
            if (A.chi2) A.chi2[pid] = bad_lens ? __builtin_nanf("") : chi2;
        }
    }
    if (GRAD && A.mode != M_SIM) {
        // per-component chain rule: lens slot i on lane 0 of wave 1 + i, light slot i on
        // lane 0 of wave 4 + i; contributions meet in gx through LDS atomics
        const float* tot = s_tot + 2;
        float* gx = s_x + 4 * d;
        const int w = tid >> 6;
        if ((tid & 63) == 0) {
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const int k = M::lens[i];
                if (k == 0 || w != 1 + i) continue;
                const DComp& cd = P.comp[i];
                float g[8];
                int np = 0;
                if (k == glm::K_EPL) { glm::epl_finish(C.epl[i], tot + M::lens_off(i), g); np = 6; }
                else if (k == glm::K_SHEAR) { g[0] = tot[M::lens_off(i)]; g[1] = tot[M::lens_off(i) + 1]; np = 2; }
                else if (k == glm::K_SIS) { glm::sis_finish(tot + M::lens_off(i), g); np = 3; }
                else if (k == glm::K_SIE) { glm::sie_finish(C.sie[i], tot + M::lens_off(i), g); np = 5; }
                else if (k == glm::K_TNFW) { glm::tnfw_finish(C.tnfw[i], tot + M::lens_off(i), g); np = 5; }
                for (int p = 0; p < np; ++p)
                    if (cd.col[p] >= 0) atomicAdd(&gx[cd.col[p]], g[p]);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (M::light[i] == 0 || w != 4 + i) continue;
                const DComp& cd = P.comp[light_comp_index<M>(i)];
                float g[10];
                if (M::light[i] == glm::K_CORE_SERSIC) {
                    glm::core_sersic_finish(C.core[i], tot + M::light_off(light_acc_slot<M>(i)), g);
                    for (int p = 0; p < (cd.lstsq ? 9 : 10); ++p)
                        if (cd.col[p] >= 0) atomicAdd(&gx[cd.col[p]], g[p]);
                    continue;
                }
                glm::sersic_finish(C.ser[i], tot + M::light_off(light_acc_slot<M>(i)), g);
                if (M::light[i] == glm::K_SERSIC_ELLIPSE) {
                    for (int p = 0; p < (cd.lstsq ? 6 : 7); ++p)
                        if (cd.col[p] >= 0) atomicAdd(&gx[cd.col[p]], g[p]);
                } else {
                    const int map[5] = {0, 1, 4, 5, 6};
                    for (int p = 0; p < (cd.lstsq ? 4 : 5); ++p)
                        if (cd.col[p] >= 0) atomicAdd(&gx[cd.col[p]], g[map[p]]);
                }
            }
        }
        __syncthreads();
        if (tid < d) {
            const float gxj = s_x[4 * d + tid];
            if (A.mode == M_VJP) {
                A.grad[(size_t)pid * d + tid] = gxj;
            } else {
                const float gz = (gxj + glp * s_x[2 * d + tid]) * s_x[d + tid] + glp * s_x[3 * d + tid];
                if (A.mode == M_LOGPROB) {
                    A.grad[(size_t)pid * d + tid] = gz;
                } else {  // M_MAP: Keras Adam
                    const size_t o = (size_t)pid * d + tid;
                    float mm = A.beta1 * m_in + (1.0f - A.beta1) * gz;
                    float vv = A.beta2 * v_in + (1.0f - A.beta2) * gz * gz;
                    A.m[o] = mm;
                    A.v[o] = vv;
                    A.z_io[o] = zin - A.lr_t * mm / (sqrtf(vv) + A.eps);
                }
            }
        }
    }
}

// Launch order of a fused launch: particles by descending EPL series length
// (summed over the EPL slots), which sets a workgroup's run time, so the longest
// workgroups start first and the tail of the launch holds the shortest.  The
// result of every particle is independent of the order.  One workgroup: keys
// in LDS, counting sort over the series lengths; the length is the epl.py:41
// rule evaluated in FP32 (an ordering estimate only).
constexpr int kOrderCap = 16384;
constexpr int kOrderBins = 3 * GL_MAX_EPL_ITER + 1;
#ifdef GL_FUSED_AUX_KERNELS  // defined by the one translation unit that launches them
__global__ __launch_bounds__(1024) void order_kernel(const DPlan P, const float* in, int bs, int use_z, int* order) {
    constexpr int kScanWaves = (kOrderBins + 63) / 64;
    __shared__ unsigned short key[kOrderCap];
    __shared__ int cnt[kOrderBins];
    __shared__ int wtot[kScanWaves];
    for (int i = threadIdx.x; i < kOrderBins; i += blockDim.x) cnt[i] = 0;
    // bin of every particle; no atomics in this loop, so the loads of several
    // particles per thread can be in flight together
#pragma unroll 4
    for (int p = threadIdx.x; p < bs; p += blockDim.x) {
        int n = 0;
#pragma unroll
        for (int ci = 0; ci < kMaxLens; ++ci) {
            if (ci >= P.nlens) break;
            const DComp& cd = P.comp[ci];
            if (cd.kind != glm::K_EPL) continue;
            float e[2];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int col = cd.col[2 + j];
                float v = col >= 0 ? in[(size_t)p * P.d + col] : cd.cst[2 + j];
                if (col >= 0 && use_z) {
                    glm::PriorOut o;
                    glm::prior_eval(P.prior_kind[col], P.prior_args + 8 * col, v, 0, o);
                    v = o.x;
                }
                e[j] = v;
            }
            const float c = fminf(fmaxf(sqrtf(e[0] * e[0] + e[1] * e[1]), 0.0f), 1.0f);
            const float q = (1.0f - c) / (1.0f + c);
            const float f = (1.0f - q) / (1.0f + q);
            int m = cd.niter;
            if (f > 0.0f && f < 1.0f) m = (int)fminf(ceilf(-27.631021f / logf(f) + 2.0f) - 1.0f, (float)cd.niter);
            if (!(f > 0.0f)) m = 1;
            n += m < P.max_n ? m : P.max_n;
        }
        key[p] = (unsigned short)(kOrderBins - 1 - min(max(n, 0), kOrderBins - 1));
    }
    __syncthreads();
    for (int p = threadIdx.x; p < bs; p += blockDim.x) atomicAdd(&cnt[key[p]], 1);
    __syncthreads();
    // exclusive prefix sum of the bins: an inclusive scan per wave of 64 bins (every
    // lane of waves 0 .. kScanWaves-1 active), then the totals of the waves before
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int c = 0, incl = 0;
    if (w < kScanWaves) {
        c = threadIdx.x < kOrderBins ? cnt[threadIdx.x] : 0;
        incl = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        if (lane == 63) wtot[w] = incl;
    }
    __syncthreads();
    if (threadIdx.x < kOrderBins) {
        int base = 0;
        for (int k = 0; k < w; ++k) base += wtot[k];
        cnt[threadIdx.x] = base + incl - c;
    }
    __syncthreads();
    for (int p = threadIdx.x; p < bs; p += blockDim.x) order[atomicAdd(&cnt[key[p]], 1)] = p;
}

// Batch-coupled EPL series length (epl.py:41): one block reduces max f over
// the batch for every EPL component and stores the loop count.
__global__ void epl_batch_series_kernel(const DPlan P, const float* in, int bs, int use_z, int* out) {
    __shared__ float red[256];
    for (int ci = 0; ci < P.nlens; ++ci) {
        const DComp& cd = P.comp[ci];
        if (cd.kind != glm::K_EPL) {
            if (threadIdx.x == 0) out[ci] = -1;
            continue;
        }
        float fmax = -1.0f;
        bool any_nan = false;
        for (int p = threadIdx.x; p < bs; p += blockDim.x) {
            float pv[4];
            for (int j = 2; j < 4; ++j) {
                int col = cd.col[j];
                float v = col >= 0 ? in[(size_t)p * P.d + col] : cd.cst[j];
                if (col >= 0 && use_z) {
                    glm::PriorOut o;
                    glm::prior_eval(P.prior_kind[col], P.prior_args + 8 * col, v, 0, o);
                    v = o.x;
                }
                pv[j] = v;
            }
            float c = fminf(fmaxf(sqrtf(pv[2] * pv[2] + pv[3] * pv[3]), 0.0f), 1.0f);
            float q = (1.0f - c) / (1.0f + c);
            float f = (1.0f - q) / (1.0f + q);
            if (isnan(f)) any_nan = true;
            fmax = fmaxf(fmax, f);
        }
        if (any_nan) fmax = NAN;
        red[threadIdx.x] = fmax;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) {
                float a = red[threadIdx.x], b = red[threadIdx.x + s];
                red[threadIdx.x] = (isnan(a) || isnan(b)) ? NAN : fmaxf(a, b);
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) out[ci] = glm::epl_series_len((double)red[0], cd.niter);
        __syncthreads();
    }
}
#endif  // GL_FUSED_AUX_KERNELS

}  // namespace gl
