// HBM-staged lens-model pipeline for gfx950.  It serves three callers:
//   * LensSimulator.lstsq_simulate (src/gigalens/tf/simulator.py:99-165) and
//     BackwardProbModel.log_prob (src/gigalens/tf/model.py:134-163): linear
//     light amplitudes solved per particle by least squares;
//   * LensSimulator.simulate / ForwardProbModel.log_prob (tf/simulator.py:65-97,
//     tf/model.py:53-86) for images too large for the LDS-resident particle
//     kernel or profile combinations it is not compiled for: fixed amplitudes
//     (the light parameters include Ie), one channel;
//   * batches of lens systems with different image sizes, PSFs and particle
//     counts: every stage is one launch over all lenses.
//
// Work is addressed through a lens table (McTable): per lens a DPlan, a
// McPlan (its own workspace) and prefix sums of each stage's work units over
// the lenses; a unit finds its lens by binary search and its particle inside
// the lens by division.  A single plan is a table of one lens passed by value.
//
// Stages, one launch each, stream-ordered:
//   mc_prologue   per particle: z -> x (bijectors), prior terms, component
//                 constants, EPL coefficient tables -> workspace
//   mc_basis      per sub-pixel: the unit-amplitude channels (lens light on
//                 theta, sources -- Sersic or the shapelet basis -- on beta) or,
//                 with fixed amplitudes, their sum; NaN -> 0; staged in HBM
//                 with zero column margins
//   mc_conv       per (particle, channel, 4 output rows, 64 output columns):
//                 stride-ss correlation with the PSF folded with the pooling box
//                 (depthwise conv2d SAME + avg_pool2d); one lane per output
//                 column streams the contributing sub-rows through L1
//   mc_sep_h      (separable PSF, in place of mc_conv) horizontal pass of every
//                 sub-row and rank into H, 256 flattened outputs per wave
//   mc_sep_v      (separable PSF) polyphase vertical pass: a workgroup stages the
//                 H rows of 64 output rows in LDS per (rank, phase)
//   mc_gram      least squares: per particle and pixel range [X | obs]^T W^2 [X | obs]
//                 on FP32 MFMA (16x16x4) tiles, flushed into FP64 every 128 pixels
//   mc_solve      least squares: per particle FP64 parallel cyclic Jacobi, then the
//                 pseudo-inverse with the reference's cutoff (1e-6, tf/simulator.py:161)
//   mc_image      per particle: image = sum_c coef_c X_c, likelihood (forward: sigma
//                 from the model image; backward: from the observed image), chi^2,
//                 prior -> lp; dL/dimage
//   mc_adj_h      (separable PSF) per image row segment: transposed horizontal pass
//   mc_adj_v      (separable PSF) per sub-row block: transposed vertical pass
//   mc_adjoint    per sub-pixel row segment: transposed correlation of dL/dimage
//   mc_grad       per sub-pixel: recompute and accumulate reverse-mode statistics
//                 (least squares: amplitudes held at their optimum, envelope theorem);
//                 instances with the lens kinds (mc_grad_lean) or every component
//                 kind (mc_grad_shp4, mc_grad_m) fixed at compile time keep fewer
//                 statistics in registers
//   mc_finish     per particle: chain rule -> dlp/dz, optionally the Keras-Adam
//                 update of a MAP step (tf/inference.py:28-35)
#pragma once
#include <hip/hip_runtime.h>
#include "lens_math.h"
#include "lens_plan.h"

namespace gl {

// Channels plus the observed-image column are tiled by 16 for the Gram MFMAs
// (at most 5 tiles); the solver keeps two FP64 (D+1)^2 matrices in LDS.
constexpr int kMcMaxChan = 79;
constexpr int kMcGramMaxNB = 5;
constexpr int kMcGramMaxSplit = 8;  // pixel-range splits per particle (x 4 waves = partials)
constexpr int kStatStride = 16;  // statistics floats per component (core-Sersic: 11)
constexpr int kMcStat = kStatStride * kMaxComp;  // statistics floats per particle
constexpr int kMcMaxRank = 8;  // separable channel convolution: at most this rank
constexpr int kMcIB = 4;     // default output rows per conv lane (McPlan.conv_ib: 2, 4 or 8)

// Read-only factor tables addressed in the constant space: uniform reads become
// scalar loads even in kernels that store to global memory in the same loop.
typedef const float __attribute__((address_space(4))) CFloat;

struct McPlan {
    int lstsq;          // 1: linear amplitudes solved by least squares; 0: fixed amplitudes
    int famp;           // fixed amplitudes per channel (a shapelet light with sampled amplitudes): the
                        // unit-amplitude channels of least squares, weighted by cf x the parameters
                        // (Sersic Ie, shapelet amplitudes); 0 with lstsq = 0: one channel (the sum)
    int conv_ib;        // output rows per lane of mc_conv (a tuning choice: 2, 4 or 8)
    int sep_rank;       // > 0: separable channel convolution with this rank (W ~= sum_r sa_r sb_r^T)
    const float* sa;    // sep_rank x kf vertical factors
    const float* sap;   // sep_rank x ss x kfm polyphase vertical factors, sap[(r ss + a) kfm + m] =
                        // sa_r[ss m + a] (zero where ss m + a >= kf), kfm = (4 kf4 + ss - 1) / ss
    const float* sbp;   // sep_rank x wpad_w horizontal factors, zero padded
    float* H;           // bs x D x sep_rank x nsub x N horizontal-pass intermediate
    int adj_rank;       // > 0: separable adjoint conv with this rank (mc_adj_h, mc_adj_v)
    const float* asa;   // adj_rank x kf vertical factors of the folded PSF
    const float* asap;  // adj_rank x ss x kfm polyphase vertical factors, asap[(r ss + a) kfm + m] =
                        // asa_r[ss m + (a + pb) % ss] for m < Ma(a), else 0; kfm = (4 kf4 + ss - 1) / ss
    const float* asb;   // adj_rank x kf horizontal factors
    const float* asbq;  // adj_rank x ss x (kfm + 1) horizontal factors at the window positions of mc_adj_h:
                        // asbq[(r ss + b)(kfm + 1) + j] = asb_r[ss nn + rb] with nn = kfm - 1 + db(b) - db(0) - j
    float* U;           // bs x adj_rank x N x ss x N transposed horizontal pass, phase-major rows
    int D;              // channels
    int BW;             // staged row width (floats): pb + nsub + right margin, multiple of 4
    int pstride;        // floats of per-particle constants
    int coef_base;      // offset of EPL coefficient tables in a particle's constants
    int shp_comp;       // component index of the shapelet source (-1: none)
    int shp_nmax, shp_interp, shp_chan0;
    int rcond_mode;     // 0: reference cutoff
    float rcond;
    int lik_cf;         // 1: multiply by the conversion factor (never for lstsq)
    const float* shp_table;  // (n_max+1) x 6000
    const float* shp_pref;   // n_max+2
    const float* wpad;       // kf x (4*KF4max) folded weights, zero padded columns
    int wpad_w;              // row width of wpad
    float* consts;           // bs x pstride
    float* B;                // bs x D x nsub x BW
    float* X;                // bs x D x N x N
    const float* sw;         // N x N: 1 / err
    const float* swobs;      // N x N: obs / err
    int gram_nb;             // 16-column tiles of [X | obs]
    double* gpart;           // bs x parts x (nb(nb+1)/2) x 256 FP64 Gram partials
    float* coef;             // bs x D
    float* gimg;             // bs x (N + 2 gp) x (N + 2 gp) padded dL/dimage
    int gp;                  // padding of gimg
    // the sub-grid adjoint image dL/dsub reuses B (dead after mc_conv): channel 0's
    // interior, row stride BW, column offset pb -- the zero margins stay untouched
    float* stats;            // bs x kMcStat
    float* gamp;             // famp: bs x D gradient w.r.t. each channel's amplitude parameter
    float* lp_like;          // bs (log-likelihood + chi^2 scratch)
};

struct McArgs {
    const float* in;  // z (log_prob) or x (simulate)
    int use_z;
    int bs;
    float* lp;
    float* chi2;
    float* img;        // optional image out (bs x N x N)
    float* coef_out;   // optional (bs x D)
    float* stacked;    // optional (bs x N x N x D)
    const float* g_lp;
    const float* g_chi2;
    float* grad;       // optional dlp/dz
    const float* g_img;  // VJP of simulate: upstream image gradient (bs x N x N); no likelihood
    int map;           // 1: MAP step -- the loss is -weight * lp / N^2 and z, m, v get a Keras-Adam update
    float* z_io;
    float* m;
    float* v;
    float lr_t, beta1, beta2, eps;
};

// One lens of a launch: its plan, its workspace, and its particles
// [first, first + count) of the launch; weight = MAP loss weight per particle.
struct McLens {
    DPlan P;
    McPlan M;
    int first, count;
    float weight;
    int goff;  // a launch's particle gp of this lens is particle gp + goff of the caller's arrays
};

// Lens table of a launch.  Single-lens launches use `one` (a kernel argument);
// multi-lens launches use lens[0..n) and the n + 1 prefix sums pre_* of each
// stage's units over the lenses (device memory).
struct McTable {
    McLens one;
    const McLens* lens;
    int n;
    const int* pre_part;   // particles
    const int* pre_basis;  // 256-thread blocks of sub-pixels (mc_basis, mc_grad)
    const int* pre_conv;   // conv waves
    const int* pre_adj;    // adjoint waves (a multiple of 4 per particle)
    const int* pre_grad;   // mc_grad blocks (256 x kMcGradK sub-pixels each)
    const int* pre_adjh;   // separable adjoint horizontal-pass waves (a multiple of 4 per particle; 0 without)
    const int* pre_adjv;   // separable adjoint vertical-pass waves (a multiple of 4 per particle; 0 without)
    const int* pre_seph;   // separable channel conv: horizontal-pass waves (0 for direct-conv lenses)
    const int* pre_sepv;   // separable channel conv: vertical-pass workgroups
    int tot_part, tot_basis, tot_conv, tot_adj, tot_grad, tot_adjh, tot_adjv, tot_seph, tot_sepv;
    int lds_consts;        // bytes of LDS for one particle's constant block (max over lenses)
    int lds_w;             // bytes of LDS for the folded weights kf x kf (max over lenses)
};

// lens holding unit x of a stage, and x's offset inside that lens.  MULTI =
// false: the single lens passed by value (its fields stay kernel arguments).
template <bool MULTI>
__device__ __forceinline__ const McLens& mc_locate(const McTable& T, const int* pre, int x, int& local) {
    if (!MULTI) {
        local = x;
        return T.one;
    }
    int lo = 0, hi = T.n - 1;
    while (lo < hi) {  // largest L with pre[L] <= x
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= x) lo = mid; else hi = mid - 1;
    }
    lo = __builtin_amdgcn_readfirstlane(lo);  // x is uniform over the caller's wave
    local = x - pre[lo];
    // the table is read-only for the launch: address it in the constant space so
    // the fields come in through scalar loads
    typedef const McLens __attribute__((address_space(4))) ConstLens;
    const ConstLens* tab = (const ConstLens*)T.lens;
    return *(const McLens*)(tab + lo);
}

__device__ __forceinline__ int mc_bpp(const DPlan& P) { return (P.nsub * P.nsub + 255) / 256; }
constexpr int kMcGradK = 8;  // sub-pixels per thread in mc_grad
__device__ __forceinline__ int mc_gpp(const DPlan& P) { return (P.nsub * P.nsub + 256 * kMcGradK - 1) / (256 * kMcGradK); }

// Copy particle p's constant block (pstride floats: prior terms, component
// constants, EPL coefficient tables) into LDS; every thread of the block calls
// it.  The per-sub-pixel kernels then read constants and series coefficients
// from LDS (wave-uniform addresses, broadcast) instead of global memory.
__device__ __forceinline__ const float* mc_stage_consts(const McPlan& M, int p, float* lds) {
    const float4* src = reinterpret_cast<const float4*>(M.consts + (size_t)p * M.pstride);
    float4* dst = reinterpret_cast<float4*>(lds);
    for (int i = threadIdx.x; i < M.pstride / 4; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
    return lds;
}
__device__ __forceinline__ int mc_adj_wpp4(const DPlan& P) {
    return ((P.nsub * P.ss * ((P.N + 63) / 64)) + 3) & ~3;
}
__device__ __forceinline__ int mc_adjh_wpp4(const DPlan& P) {
    return ((P.N * ((P.N + 63) / 64)) + 3) & ~3;
}
constexpr int kAdjKY = 8;  // sub-rows of one phase per lane in the separable vertical pass
__device__ __forceinline__ int mc_adjv_wpp4(const DPlan& P) {
    return ((P.ss * ((P.N + kAdjKY - 1) / kAdjKY) * P.ss * ((P.N + 63) / 64)) + 3) & ~3;
}

// gradient weight of particle p's log posterior: MAP -> -weight / N^2, else g_lp (default 1)
__device__ __forceinline__ float mc_glp(const McArgs& A, const McLens& E, int p) {
    if (A.map) return -E.weight / (float)(E.P.N * E.P.N);
    return A.g_lp ? A.g_lp[p] : 1.0f;
}

// ---------------------------------------------------------------------------
// constants of particle p
// ---------------------------------------------------------------------------
__device__ __forceinline__ float* pconst(const McPlan& M, int p) { return M.consts + (size_t)p * M.pstride; }

// Position of Ie among a light profile's parameters (the canonical order of the C ABI).
__host__ __device__ constexpr int mc_ie_index(int kind) {
    return kind == glm::K_SERSIC ? 4 : kind == glm::K_CORE_SERSIC ? 9 : 6;
}

// Value of a Sersic-family light (circular, ellipse, core) with constants cst at (x, y).
__device__ __forceinline__ float mc_sersic_light(int kind, const float* cst, float x, float y) {
    if (kind == glm::K_CORE_SERSIC) {
        glm::CoreSerState st;
        return glm::core_sersic_forward(*reinterpret_cast<const glm::CoreSerC*>(cst), x, y, st);
    }
    glm::SerState st;
    return glm::sersic_forward(*reinterpret_cast<const glm::SerC*>(cst), x, y, st);
}

template <bool MULTI>
__global__ void mc_prologue(const McTable T, const McArgs A) {
    const int p = blockIdx.x;  // particle of the launch (inputs, outputs)
    int q;                     // particle inside its lens (workspace)
    const McLens& E = mc_locate<MULTI>(T, T.pre_part, p, q);
    const DPlan& P = E.P;
    const McPlan& M = E.M;
    const int tid = threadIdx.x;
    const int d = P.d;
    float* pc = pconst(M, q);
    float* s_x = pc;  // [x | dxdz | dlpdx | dfldjdz | gx | lpj]
    float* s_c = pc + ((6 * d + 3) & ~3);
    if (tid < d) {
        const float zin = A.in[(size_t)(p + E.goff) * d + tid];
        glm::PriorOut o;
        if (A.use_z) {
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
    __threadfence_block();
    __syncthreads();
    // component constants, one thread per component (compile-time component index)
#pragma unroll
    for (int ci = 0; ci < kMaxComp; ++ci) {
        if (ci != tid || ci >= P.ncomp) continue;
        const DComp& cd = P.comp[ci];
        float pv[GL_MAX_COMPONENT_PARAMS];
#pragma unroll
        for (int j = 0; j < GL_MAX_COMPONENT_PARAMS; ++j) pv[j] = cd.col[j] >= 0 ? s_x[cd.col[j]] : cd.cst[j];
        float* dst = s_c + ci * kConstStride;
        if (cd.kind == glm::K_EPL) {
            glm::EplC c;
            glm::epl_prepare(pv[0], pv[1], pv[2], pv[3], pv[4], pv[5], cd.niter, -1, c);
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
            glm::sersic_prepare(pv[0], pv[1], pv[2], pv[3], pv[4], pv[5], (cd.lstsq || M.famp) ? 1.0f : pv[6], 1, c);
            *reinterpret_cast<glm::SerC*>(dst) = c;
        } else if (cd.kind == glm::K_SERSIC) {
            glm::SerC c;
            glm::sersic_prepare(pv[0], pv[1], 0.0f, 0.0f, pv[2], pv[3], (cd.lstsq || M.famp) ? 1.0f : pv[4], 0, c);
            *reinterpret_cast<glm::SerC*>(dst) = c;
        } else if (cd.kind == glm::K_CORE_SERSIC) {
            glm::CoreSerC c;
            glm::core_sersic_prepare(pv[0], pv[1], pv[2], pv[3], pv[4], pv[5], pv[6], pv[7], pv[8],
                                     (cd.lstsq || M.famp) ? 1.0f : pv[9], c);
            *reinterpret_cast<glm::CoreSerC*>(dst) = c;
        } else if (cd.kind == glm::K_SHAPELETS) {
            dst[0] = pv[0];  // beta
            dst[1] = pv[1];  // center_x
            dst[2] = pv[2];  // center_y
        }
    }
    __threadfence_block();
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kMaxLens; ++i) {
        if (i >= P.nlens || P.comp[i].kind != glm::K_EPL) continue;
        const glm::EplC* c = reinterpret_cast<const glm::EplC*>(s_c + i * kConstStride);
        float* ca = pc + M.coef_base + P.coef_off[i];
        float4* cb = reinterpret_cast<float4*>(ca + glm::epl_tab_a_floats(P.max_n));
        const int NN = c->N;
        for (int n = tid; n <= glm::epl_pad(NN); n += blockDim.x) {
            float cv = 0.0f, dt = 0.0f, df = 0.0f;  // zero padding above N
            if (n <= NN) glm::epl_coef(n, (double)c->f, (double)c->t, cv, dt, df);
            ca[n] = cv;
            cb[n] = make_float4(cv, dt, df, 0.0f);
        }
    }
    if (tid < kMcStat) M.stats[(size_t)q * kMcStat + tid] = 0.0f;
    if (!M.lstsq && !M.famp && tid == 0) M.coef[(size_t)q * M.D] = P.cf;  // fixed amplitudes: image = cf * conv
    if (M.famp && tid == 0) {
        // per-channel amplitudes: image = sum_c cf a_c X_c (the sampled Ie / amp parameters)
        int ch = 0;
        for (int i = P.nlens; i < P.ncomp; ++i) {
            const DComp& cd = P.comp[i];
            if (cd.kind == glm::K_SHAPELETS) {
                const int nl = (cd.n_max + 1) * (cd.n_max + 2) / 2;
                for (int k = 0; k < nl; ++k) M.coef[(size_t)q * M.D + ch + k] = P.cf * s_x[cd.amp_col0 + k];
                ch += nl;
            } else {
                const int ie = mc_ie_index(cd.kind);
                M.coef[(size_t)q * M.D + ch++] = P.cf * (cd.col[ie] >= 0 ? s_x[cd.col[ie]] : cd.cst[ie]);
            }
        }
    }
}

// Per-lens states kept by the recompute for the deflection adjoint.
struct LensStates {
    glm::EplState es[kMaxLens];
    float sis[kMaxLens][3];
    glm::SieState sie[kMaxLens];
    glm::TnfwState tn[kMaxLens];
};

// Lens kinds a kernel instance handles, per lens slot: byte i of the mask has bit k
// set when slot i may hold glm kind k.  A narrower mask compiles the other kinds'
// branches and per-slot states out (fewer registers).
constexpr int lens_bit(int kind) { return 1 << kind; }
constexpr int kLensAny = lens_bit(glm::K_EPL) | lens_bit(glm::K_SHEAR) | lens_bit(glm::K_SIS) |
                         lens_bit(glm::K_SIE) | lens_bit(glm::K_TNFW);
constexpr int kLensAll = kLensAny | (kLensAny << 8) | (kLensAny << 16);
// EPL in slot 0, shear in slot 1, no third lens (the reference's default lens model)
constexpr int kLensEplShear = lens_bit(glm::K_EPL) | (lens_bit(glm::K_SHEAR) << 8);
__host__ __device__ constexpr bool lens_in(int LM, int slot, int kind) { return (LM >> (8 * slot)) & lens_bit(kind); }

// Deflection of all lenses of particle p at (X, Y); keeps the lens states when ADJ.
// Every lens kind of the plan must be in LM.
template <bool ADJ, int LM = kLensAll>
__device__ __forceinline__ void mc_beta(const DPlan& P, const McPlan& M, const float* pc, float X, float Y, float& bx,
                                        float& by, LensStates* ls) {
    const float* s_c = pc + ((6 * P.d + 3) & ~3);
    bx = X;
    by = Y;
#pragma unroll
    for (int i = 0; i < kMaxLens; ++i) {
        if (i >= P.nlens) break;
        const int kind = P.comp[i].kind;
        const float* cst = s_c + i * kConstStride;
        float ax = 0.0f, ay = 0.0f;
        if (lens_in(LM, i, glm::K_EPL) && kind == glm::K_EPL) {
            const glm::EplC C = *reinterpret_cast<const glm::EplC*>(cst);
            const float* ca = pc + M.coef_base + P.coef_off[i];
            const float4* cb = reinterpret_cast<const float4*>(ca + glm::epl_tab_a_floats(P.max_n));
            glm::EplState st;
            glm::epl_forward<ADJ>(C, ca, cb, X, Y, ax, ay, st);
            if (ADJ) ls->es[i] = st;
        } else if (lens_in(LM, i, glm::K_SHEAR) && kind == glm::K_SHEAR) {
            const glm::ShearC C = *reinterpret_cast<const glm::ShearC*>(cst);
            glm::shear_forward(C, X, Y, ax, ay);
        } else if (lens_in(LM, i, glm::K_SIS) && kind == glm::K_SIS) {
            const glm::SisC C = *reinterpret_cast<const glm::SisC*>(cst);
            float st[3];
            glm::sis_forward(C, X, Y, ax, ay, st);
            if (ADJ) { ls->sis[i][0] = st[0]; ls->sis[i][1] = st[1]; ls->sis[i][2] = st[2]; }
        } else if (lens_in(LM, i, glm::K_SIE) && kind == glm::K_SIE) {
            const glm::SieC C = *reinterpret_cast<const glm::SieC*>(cst);
            glm::SieState st;
            glm::sie_forward<ADJ>(C, X, Y, ax, ay, st);
            if (ADJ) ls->sie[i] = st;
        } else if (lens_in(LM, i, glm::K_TNFW) && kind == glm::K_TNFW) {
            const glm::TnfwC C = *reinterpret_cast<const glm::TnfwC*>(cst);
            glm::TnfwState st;
            glm::tnfw_forward<ADJ>(C, X, Y, ax, ay, st);
            if (ADJ) ls->tn[i] = st;
        }
        bx -= ax;
        by -= ay;
    }
}

// Shapelet 1-D functions of the plan's mode.
__device__ __forceinline__ void mc_phi(const McPlan& M, float u, float* phi, float* dphi) {
    if (M.shp_interp) glm::shapelet_1d_table(M.shp_table, M.shp_nmax, u, phi, dphi);
    else glm::shapelet_1d_direct(M.shp_pref, M.shp_nmax, u, phi, dphi);
}

// Fixed per-channel amplitudes: the simulated sub-pixel sum_c coef_c channel_c (NaN when
// the forward pass zeroed the sub-pixel).
__device__ __forceinline__ float mc_famp_sum(const DPlan& P, const McPlan& M, const float* s_c, const float* cp,
                                             float X, float Y, float bx, float by) {
    float v = 0.0f;
    int ch = 0;
    for (int i = P.nlens; i < P.ncomp; ++i) {
        const DComp& cd = P.comp[i];
        const float* cst = s_c + i * kConstStride;
        const bool src = cd.role == GL_ROLE_SOURCE;
        const float ux = src ? bx : X, uy = src ? by : Y;
        if (cd.kind == glm::K_SHAPELETS) {
            const float beta = cst[0];
            float pu[glm::kMaxShapeletN + 1], pw[glm::kMaxShapeletN + 1];
            mc_phi(M, (ux - cst[1]) / beta, pu, nullptr);
            mc_phi(M, (uy - cst[2]) / beta, pw, nullptr);
#pragma unroll
            for (int so = 0; so <= glm::kMaxShapeletN; ++so) {
                if (so > M.shp_nmax) break;
#pragma unroll
                for (int j = 0; j <= so; ++j) v = fmaf(cp[ch + so * (so + 1) / 2 + j], pu[so - j] * pw[j], v);
            }
            ch += (M.shp_nmax + 1) * (M.shp_nmax + 2) / 2;
        } else {
            v = fmaf(cp[ch++], mc_sersic_light(cd.kind, cst, ux, uy), v);
        }
    }
    return v;
}

// The column margins of a staged row (pb zeros on the left, BW - pb - nsub on the
// right) that the channel convolution reads: written by the thread of column x of the
// row, for every channel, so a workspace shared by lenses of different shapes needs
// no clearing.
__device__ __forceinline__ void mc_zero_margins(const McPlan& M, const DPlan& P, float* row, int x, int nch) {
    const int nsub = P.nsub, rm = M.BW - P.pb - nsub;
    const size_t cstride = (size_t)nsub * M.BW;
    for (int c = 0; c < nch; ++c) {
        for (int m = x; m < P.pb; m += nsub) row[cstride * c + m] = 0.0f;
        for (int m = x; m < rm; m += nsub) row[cstride * c + P.pb + nsub + m] = 0.0f;
    }
}

// ---------------------------------------------------------------------------
// mc_basis: grid (bs, ceil(nsub*nsub/256)), block 256.  Bound by its stores (every
// channel of every sub-pixel): held to 6 waves per SIMD for more stores in flight, at
// the price of the shapelet tables in scratch (12.2 -> 11.2 ms for 512 shapelet
// particles, 67 channels).
// ---------------------------------------------------------------------------
template <bool MULTI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void mc_basis(const McTable T) {
    int local;
    const McLens& E = mc_locate<MULTI>(T, T.pre_basis, blockIdx.x, local);
    const DPlan& P = E.P;
    const McPlan& M = E.M;
    const int bpp = mc_bpp(P);
    const int p = local / bpp;
    const int s = (local - p * bpp) * 256 + threadIdx.x;
    const int nsub = P.nsub;
    extern __shared__ float4 mc_dyn4[];
    const float* pc = mc_stage_consts(M, p, reinterpret_cast<float*>(mc_dyn4));
    if (s >= nsub * nsub) return;
    const int y = s / nsub, x = s - (s / nsub) * nsub;
    const float* s_c = pc + ((6 * P.d + 3) & ~3);
    const float X = fmaf(P.M00, (float)x, fmaf(P.M01, (float)y, P.ox));
    const float Y = fmaf(P.M10, (float)x, fmaf(P.M11, (float)y, P.oy));
    float bx, by;
    mc_beta<false>(P, M, pc, X, Y, bx, by, nullptr);
    float* Bp = M.B + (size_t)p * M.D * nsub * M.BW + (size_t)y * M.BW + P.pb + x;
    const size_t cstride = (size_t)nsub * M.BW;
    mc_zero_margins(M, P, Bp - P.pb - x, x, M.D);
    if (!M.lstsq && !M.famp) {
        // fixed amplitudes: the sum of every light, NaN -> 0 (tf/simulator.py:74-81)
        float v = 0.0f;
        for (int i = P.nlens; i < P.ncomp; ++i) {
            const DComp& cd = P.comp[i];
            const bool src = cd.role == GL_ROLE_SOURCE;
            v += mc_sersic_light(cd.kind, s_c + i * kConstStride, src ? bx : X, src ? by : Y);
        }
        Bp[0] = isnan(v) ? 0.0f : v;
        return;
    }
    // least squares: NaN -> 0 per channel (tf/simulator.py:125); fixed per-channel
    // amplitudes: the weighted sum of the channels is the simulated sub-pixel, and a NaN
    // there zeroes every channel of it (tf/simulator.py:74-81)
    const float* cp = M.coef + (size_t)p * M.D;
    float wsum = 0.0f;
    int ch = 0;
    for (int i = P.nlens; i < P.ncomp; ++i) {
        const DComp& cd = P.comp[i];
        const float* cst = s_c + i * kConstStride;
        const bool src = cd.role == GL_ROLE_SOURCE;
        const float ux = src ? bx : X, uy = src ? by : Y;
        if (cd.kind == glm::K_SERSIC || cd.kind == glm::K_SERSIC_ELLIPSE || cd.kind == glm::K_CORE_SERSIC) {
            float v = mc_sersic_light(cd.kind, cst, ux, uy);
            if (M.famp) wsum = fmaf(cp[ch], v, wsum);
            Bp[cstride * ch++] = (isnan(v) && !M.famp) ? 0.0f : v;
        } else if (cd.kind == glm::K_SHAPELETS) {
            const float beta = cst[0];
            const float u = (ux - cst[1]) / beta, w = (uy - cst[2]) / beta;
            float pu[glm::kMaxShapeletN + 1], pw[glm::kMaxShapeletN + 1];
            mc_phi(M, u, pu, nullptr);
            mc_phi(M, w, pw, nullptr);
#pragma unroll
            for (int so = 0; so <= glm::kMaxShapeletN; ++so) {
                if (so > M.shp_nmax) break;
#pragma unroll
                for (int j = 0; j <= so; ++j) {  // (n1, n2) = (so - j, j): order of shapelets.py:27-46
                    const float v = pu[so - j] * pw[j];
                    const int c = ch + so * (so + 1) / 2 + j;
                    if (M.famp) wsum = fmaf(cp[c], v, wsum);
                    Bp[cstride * c] = (isnan(v) && !M.famp) ? 0.0f : v;
                }
            }
            ch += (M.shp_nmax + 1) * (M.shp_nmax + 2) / 2;
        }
    }
    if (M.famp && isnan(wsum))
        for (int c = 0; c < M.D; ++c) Bp[cstride * c] = 0.0f;
}

// ---------------------------------------------------------------------------
// mc_conv: one wave = (particle, channel, IB output rows, 64 output columns)
// ---------------------------------------------------------------------------
// SS = 0: the supersampling factor is read from the plan (lenses of several ss)
template <int SS, int KF4, bool MULTI, int IB>
__global__ __launch_bounds__(256) void mc_conv(const McTable T) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int gtask = blockIdx.x * 4 + wave;
    if (gtask >= T.tot_conv) return;
    int task;
    const McLens& E = mc_locate<MULTI>(T, T.pre_conv, gtask, task);
    const DPlan& P = E.P;
    const McPlan& M = E.M;
    const int N = P.N, nsub = P.nsub, kf = P.kf, pb = P.pb;
    const int ss = SS ? SS : P.ss;
    const int nib = (N + IB - 1) / IB, njb = (N + 63) / 64;
    const int jb = task % njb; task /= njb;
    const int ib = task % nib; task /= nib;
    const int c = task % M.D;
    const int p = task / M.D;
    const int J = jb * 64 + lane;
    const int I0 = ib * IB;
    const bool active = J < N;
    const float* Bc = M.B + ((size_t)p * M.D + c) * nsub * M.BW;
    float acc[IB];
#pragma unroll
    for (int i = 0; i < IB; ++i) acc[i] = 0.0f;
    const int nrows = ss * (IB - 1) + kf;
    for (int r = 0; r < nrows; ++r) {
        const int y = ss * I0 + r - pb;
        if (y < 0 || y >= nsub) continue;
        // staged row: column x sits at index pb + x; this lane needs x = ss*J + q - pb
        const float* row = Bc + (size_t)y * M.BW + ss * J;
        float v[4 * KF4];
        if (active) {
            if (SS == 4) {
                const float4* r4 = reinterpret_cast<const float4*>(row);
#pragma unroll
                for (int j = 0; j < KF4; ++j) {
                    float4 t = r4[j];
                    v[4 * j] = t.x; v[4 * j + 1] = t.y; v[4 * j + 2] = t.z; v[4 * j + 3] = t.w;
                }
            } else {
#pragma unroll
                for (int j = 0; j < 4 * KF4; ++j) v[j] = row[j];
            }
        } else {
#pragma unroll
            for (int j = 0; j < 4 * KF4; ++j) v[j] = 0.0f;
        }
#pragma unroll
        for (int i = 0; i < IB; ++i) {
            const int pp = r - ss * i;
            if (pp >= 0 && pp < kf) {
                const float* w = M.wpad + pp * M.wpad_w;
                float a = acc[i];
#pragma unroll
                for (int j = 0; j < 4 * KF4; ++j) a = fmaf(w[j], v[j], a);
                acc[i] = a;
            }
        }
    }
    if (active) {
        float* Xc = M.X + ((size_t)p * M.D + c) * N * N;
#pragma unroll
        for (int i = 0; i < IB; ++i) {
            const int I = I0 + i;
            if (I < N) {
                const float v = acc[i];
                Xc[I * N + J] = isnan(v) ? 0.0f : v;  // NaN -> 0 after pooling (tf/simulator.py:153)
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Separable channel convolution (single plans): W ~= sum_r sa_r sb_r^T, so the
// stride-ss correlation splits into a horizontal pass over every sub-row,
//   H_r[y][J] = sum_q sb_r[q] B[y][ss J + q - pb],
// and a vertical pass over the output rows,
//   X[I][J] = sum_r sum_p sa_r[p] H_r[ss I + p - pb][J].
// mc_sep_h: one wave per (particle, channel, 256 consecutive outputs of the
// flattened (sub-row y, column J) range), each lane 4 of them, so no lane idles
// when N is not a multiple of 64.  A batch launches both passes once per (ss, kf4)
// group of its separable lenses.
// ---------------------------------------------------------------------------
template <bool MULTI, int SS, int KF4>
__global__ __launch_bounds__(256) void mc_sep_h(const McTable T) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int gtask = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wave);
    if (gtask >= T.tot_seph) return;
    int task;
    const McLens& E = mc_locate<MULTI>(T, T.pre_seph, gtask, task);
    const DPlan& P = E.P;
    const McPlan& M = E.M;
    const int N = P.N, nsub = P.nsub, R = M.sep_rank;
    // a wave covers 4 x 64 consecutive outputs of the flattened (sub-row, column) range,
    // so every lane works whatever N is
    const int nout = nsub * N, nfb = (nout + 255) / 256;
    const int fb = task % nfb; task /= nfb;
    const int c = task % M.D;
    const int p = task / M.D;
    const float* Bc = M.B + ((size_t)p * M.D + c) * nsub * M.BW;
    float* Hc = M.H + ((size_t)p * M.D + c) * R * nsub * N;
    // taps in two halves: half the window in registers at a time (more waves per SIMD)
    constexpr int KH4 = (KF4 + 1) / 2;
#pragma unroll 1
    for (int i = 0; i < 4; ++i) {
        const int f = (fb * 4 + i) * 64 + lane;
        if (f >= nout) break;
        const int y = f / N, J = f - y * N;
        const float* row = Bc + (size_t)y * M.BW + SS * J;
        float h[kMcMaxRank];
#pragma unroll
        for (int r = 0; r < kMcMaxRank; ++r) h[r] = 0.0f;
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) {
            const int j0 = hf * KH4;
            const int nj = hf == 0 ? KH4 : KF4 - KH4;
            float v[4 * KH4];
            if (SS == 4) {
                const float4* r4 = reinterpret_cast<const float4*>(row) + j0;
#pragma unroll
                for (int j = 0; j < KH4; ++j) {
                    if (j < nj) {
                        const float4 t = r4[j];
                        v[4 * j] = t.x; v[4 * j + 1] = t.y; v[4 * j + 2] = t.z; v[4 * j + 3] = t.w;
                    }
                }
            } else {
#pragma unroll
                for (int j = 0; j < 4 * KH4; ++j)
                    if (j < 4 * nj) v[j] = row[4 * j0 + j];
            }
#pragma unroll
            for (int r = 0; r < kMcMaxRank; ++r) {
                if (r >= R) break;
                const CFloat* w = (const CFloat*)M.sbp + r * M.wpad_w + 4 * j0;
#pragma unroll
                for (int j = 0; j < 4 * KH4; ++j)
                    if (j < 4 * nj) h[r] = fmaf(w[j], v[j], h[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < kMcMaxRank; ++r) {
            if (r >= R) break;
            Hc[((size_t)r * nsub + y) * N + J] = h[r];
        }
    }
}

// mc_sep_v: one workgroup per (particle, channel, kSepRows output rows, 64 columns J),
// in polyphase order: tap p = SS m + a of output row I reads H row SS (I + m) + a - pb,
// so for each phase a the output rows of a column form a 1-D correlation over H rows
// SS k + a - pb.  For each (rank, phase) the workgroup stages the kSepRows + KFM - 1 H
// rows of its block in LDS once (one wave-wide 256-byte row load per row), and each wave
// takes 16 output rows from them: every H row is read from memory once per block
// rather than once per 16 rows.  The next (rank, phase)'s rows are loaded into registers
// while the current one's taps run.  With the tap loops unrolled every factor index is
// a compile-time constant (scalar loads of the zero-padded polyphase table); the padded
// taps (SS m + a >= kf) carry zero factors and reach at most SS - 1 rows past the window.
constexpr int kSepRows = 64;
template <bool MULTI, int SS, int KF4>
__global__ __launch_bounds__(256) void mc_sep_v(const McTable T) {
    constexpr int KFM = (4 * KF4 + SS - 1) / SS;
    constexpr int KR = kSepRows + KFM - 1;  // staged H rows per (rank, phase)
    constexpr int NL = (KR + 3) / 4;        // staged rows per wave
    __shared__ float hs[KR * 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // workgroups go round-robin over the 8 XCDs: consecutive logical blocks (the row
    // blocks of one channel, which share KFM - 1 H rows) are placed on one XCD and its
    // L2.  The grid is a multiple of 8, so every logical block is covered.
    const int per = gridDim.x >> 3;
    const int lblk = __builtin_amdgcn_readfirstlane((blockIdx.x & 7) * per + (blockIdx.x >> 3));
    if (lblk >= T.tot_sepv) return;  // whole workgroup
    int task;
    const McLens& E = mc_locate<MULTI>(T, T.pre_sepv, lblk, task);
    const DPlan& P = E.P;
    const McPlan& M = E.M;
    const int N = P.N, nsub = P.nsub, pb = P.pb, R = M.sep_rank;
    const int nib = (N + kSepRows - 1) / kSepRows, njb = (N + 63) / 64;
    const int jb = task % njb; task /= njb;
    const int ib = task % nib; task /= nib;
    const int c = task % M.D;
    const int p = task / M.D;
    const int J = jb * 64 + lane;
    const bool jok = J < N;
    const int I0 = ib * kSepRows;
    const float* Hc = M.H + ((size_t)p * M.D + c) * R * nsub * N;
    // A wave stages whole rows, so every row test below is wave-uniform.  The loads take
    // the nearest image row (always in bounds); the window's zero padding is applied
    // when the rows are stored to LDS.  Columns past N load column N - 1, whose values
    // reach only outputs that are not stored.  Only the last row slot can fall past KR.
    const int swave = __builtin_amdgcn_readfirstlane(wave);
    const unsigned jbyte = 4u * (unsigned)(jok ? J : N - 1);  // lane's byte offset in a row
    float pre[NL];
    auto fetch = [&](int it) {
        const int r = it / SS, a = it - (it / SS) * SS;
        const float* Hr = Hc + (size_t)r * nsub * N;
        const int y0 = SS * I0 + a - pb;
#pragma unroll
        for (int t = 0; t < NL; ++t) {
            if (t < NL - 1 || swave + 4 * t < KR) {
                const int y = y0 + SS * (swave + 4 * t);
                const char* row = (const char*)(Hr + (size_t)min(max(y, 0), nsub - 1) * N);  // uniform
                pre[t] = *(const float*)(row + jbyte);
            }
        }
    };
    auto stage = [&](int it) {
        const int a = it - (it / SS) * SS;
        const int y0 = SS * I0 + a - pb;
#pragma unroll
        for (int t = 0; t < NL; ++t) {
            if (t < NL - 1 || swave + 4 * t < KR) {
                const int y = y0 + SS * (swave + 4 * t);
                hs[(swave + 4 * t) * 64 + lane] = (y >= 0 && y < nsub) ? pre[t] : 0.0f;
            }
        }
    };
    const int i0 = wave * 16;
    const bool live = I0 + i0 < N;  // wave-uniform: this wave has output rows
    float acc[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
    const int nit = R * SS;
    fetch(0);
    stage(0);
    __syncthreads();
#pragma unroll 1
    for (int it = 0; it < nit; ++it) {
        if (it + 1 < nit) fetch(it + 1);
        if (live) {
            const int r = it / SS, a = it - (it / SS) * SS;
            const CFloat* wa = (const CFloat*)M.sap + (r * SS + a) * KFM;
            const float* hw = hs + i0 * 64 + lane;
#pragma unroll
            for (int k = 0; k < 16 + KFM - 1; ++k) {
                const float v = hw[k * 64];
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int m = k - i;
                    if (m >= 0 && m < KFM) acc[i] = fmaf(wa[m], v, acc[i]);
                }
            }
        }
        if (it + 1 < nit) {
            __syncthreads();  // every wave is done with this (rank, phase)'s rows
            stage(it + 1);
            __syncthreads();
        }
    }
    if (!live || !jok) return;
    float* Xc = M.X + ((size_t)p * M.D + c) * N * N;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int I = I0 + i0 + i;
        if (I < N) Xc[I * N + J] = isnan(acc[i]) ? 0.0f : acc[i];  // NaN -> 0 after pooling
    }
}

// ---------------------------------------------------------------------------
// mc_gram: grid (bs, ksplit), block 256 (4 waves).  Y = [X | obs] / err is
// staged 64 pixels at a time in LDS as Y^T (channel rows); each wave owns 16
// of the 64 pixels and accumulates the upper-triangle 16x16 tiles of Y^T Y with
// v_mfma_f32_16x16x4f32.  Lane l of a wave reads the float4 of pixels
// 4(l/16)..+3 of channel row l%16 of a tile; MFMA step s uses element s, so
// A and B see the same pixel order and the K sum covers the 16 pixels once.
// FP32 tiles are added into FP64 accumulators every kFlush chunks.
// ---------------------------------------------------------------------------
typedef float mc_v4f __attribute__((ext_vector_type(4)));

template <int NB, bool MULTI>
__global__ __launch_bounds__(256) void mc_gram(const McTable T, int ksplit) {
    constexpr int NT = NB * (NB + 1) / 2;
    constexpr int CP = 16 * NB;
    constexpr int KC = 64;
    constexpr int LS = KC + 4;
    constexpr int LPT = CP * KC / 256;
    constexpr int kFlush = 8;
    __shared__ __attribute__((aligned(16))) float yt[CP * LS];
    int p;
    const McLens& E = mc_locate<MULTI>(T, T.pre_part, blockIdx.x, p);
    const DPlan& P = E.P;
    const McPlan& M = E.M;
    const int ks = blockIdx.y;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int D = M.D, NP = P.N * P.N;
    const int nchunk = (NP + KC - 1) / KC;
    const int c0 = (int)((long)nchunk * ks / ksplit), c1 = (int)((long)nchunk * (ks + 1) / ksplit);
    const float* Xp = M.X + (size_t)p * D * NP;
    mc_v4f acc[NT];
    double acc64[NT][4];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        acc[t] = mc_v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int r = 0; r < 4; ++r) acc64[t][r] = 0.0;
    }
    float pre[LPT];
    auto fetch = [&](int chunk) {
#pragma unroll
        for (int t = 0; t < LPT; ++t) {
            const int e = threadIdx.x + 256 * t;
            const int c = e / KC, k = e - (e / KC) * KC;
            const int pix = chunk * KC + k;
            float v = 0.0f;
            if (pix < NP) {
                if (c < D) v = Xp[(size_t)c * NP + pix] * M.sw[pix];
                else if (c == D) v = M.swobs[pix];
            }
            pre[t] = v;
        }
    };
    if (c0 < c1) fetch(c0);
    for (int ch = c0; ch < c1; ++ch) {
        __syncthreads();
#pragma unroll
        for (int t = 0; t < LPT; ++t) {
            const int e = threadIdx.x + 256 * t;
            const int c = e / KC, k = e - (e / KC) * KC;
            yt[c * LS + k] = pre[t];
        }
        __syncthreads();
        if (ch + 1 < c1) fetch(ch + 1);
        mc_v4f a[NB];
#pragma unroll
        for (int I = 0; I < NB; ++I)
            a[I] = *reinterpret_cast<const mc_v4f*>(yt + (16 * I + (lane & 15)) * LS + wave * 16 + 4 * (lane >> 4));
        int t = 0;
#pragma unroll
        for (int I = 0; I < NB; ++I) {
#pragma unroll
            for (int J = I; J < NB; ++J) {
#pragma unroll
                for (int s4 = 0; s4 < 4; ++s4) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[I][s4], a[J][s4], acc[t], 0, 0, 0);
                ++t;
            }
        }
        if ((ch - c0) % kFlush == kFlush - 1 || ch == c1 - 1) {
#pragma unroll
            for (int u = 0; u < NT; ++u) {
#pragma unroll
                for (int r = 0; r < 4; ++r) acc64[u][r] += (double)acc[u][r];
                acc[u] = mc_v4f{0.0f, 0.0f, 0.0f, 0.0f};
            }
        }
    }
    // tile element (i, j) = (4 (lane / 16) + r, lane % 16)
    double* out = M.gpart + ((size_t)(p * ksplit + ks) * 4 + wave) * NT * 256;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) out[t * 256 + (4 * (lane >> 4) + r) * 16 + (lane & 15)] = acc64[t][r];
}

// Entry (i, j) of [X | obs]^T W^2 [X | obs] summed over the partials of particle p.
__device__ __forceinline__ double mc_gram_entry(const McPlan& M, int p, int parts, int i, int j) {
    const int a = i < j ? i : j, b = i < j ? j : i;
    const int nb = M.gram_nb, I = a >> 4, J = b >> 4;
    const int t = I * nb - I * (I - 1) / 2 + (J - I);
    const int nt = nb * (nb + 1) / 2;
    const double* g = M.gpart + (size_t)p * parts * nt * 256 + t * 256 + (a & 15) * 16 + (b & 15);
    double s = 0.0;
    for (int q = 0; q < parts; ++q) s += g[(size_t)q * nt * 256];
    return s;
}

// ---------------------------------------------------------------------------
// mc_solve: grid bs, block 256.  FP64 parallel cyclic Jacobi + pseudo-inverse.
// ---------------------------------------------------------------------------
template <bool MULTI>
__global__ __launch_bounds__(1024) void mc_solve(const McTable T, float rcond, int parts) {
    extern __shared__ double s_d[];
    int p;
    const McPlan& M = mc_locate<MULTI>(T, T.pre_part, blockIdx.x, p).M;
    const int D = M.D;
    const int n = D + (D & 1);  // even size (dummy index n-1 when D is odd)
    const int ns = n + 1;       // odd row stride: column sweeps hit distinct LDS banks
    double* A = s_d;            // n x ns
    double* V = A + n * ns;     // n x ns
    double* rot = V + n * ns;   // (n/2) x 2
    int* pr = reinterpret_cast<int*>(rot + n);  // (n/2) x 2 pairs
    double* red = reinterpret_cast<double*>(pr + n + (n & 1));
    double* bv = red + 2;       // n: X^T W^2 obs
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n, j = e - (e / n) * n;
        A[i * ns + j] = (i < D && j < D && i <= j) ? mc_gram_entry(M, p, parts, i, j) : 0.0;
        V[i * ns + j] = i == j ? 1.0 : 0.0;
    }
    for (int i = threadIdx.x; i < D; i += blockDim.x) bv[i] = mc_gram_entry(M, p, parts, i, D);
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n, j = e - (e / n) * n;
        if (i > j) A[i * ns + j] = A[j * ns + i];
    }
    __syncthreads();
    const int half = n / 2;
    for (int sweep = 0; sweep < 20; ++sweep) {
        // convergence: off-diagonal mass relative to the diagonal
        if (threadIdx.x == 0) { red[0] = 0.0; red[1] = 0.0; }
        __syncthreads();
        double off = 0.0, dia = 0.0;
        for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
            const int i = e / n, j = e - (e / n) * n;
            const double a = A[i * ns + j];
            if (i == j) dia += a * a; else off += a * a;
        }
        for (int o = 32; o >= 1; o >>= 1) {
            off += __shfl_xor(off, o, 64);
            dia += __shfl_xor(dia, o, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&red[0], off);
            atomicAdd(&red[1], dia);
        }
        __syncthreads();
        const bool done = red[0] <= 1e-30 * red[1] || red[1] == 0.0;
        __syncthreads();
        if (done) break;
        for (int round = 0; round < n - 1; ++round) {
            // round-robin pairing: index 0 fixed, the others rotate
            if (threadIdx.x < half) {
                const int k = threadIdx.x;
                int a = k == 0 ? 0 : 1 + (k - 1 + round) % (n - 1);
                int b = 1 + (n - 2 - k + round) % (n - 1);
                if (a > b) { int t = a; a = b; b = t; }
                double c = 1.0, s = 0.0;
                const double apq = A[a * ns + b];
                if (apq != 0.0) {
                    const double th = (A[b * ns + b] - A[a * ns + a]) / (2.0 * apq);
                    const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                    c = 1.0 / sqrt(t * t + 1.0);
                    s = t * c;
                }
                rot[2 * k] = c;
                rot[2 * k + 1] = s;
                pr[2 * k] = a;
                pr[2 * k + 1] = b;
            }
            __syncthreads();
            // rows: A <- J^T A
            for (int e = threadIdx.x; e < half * n; e += blockDim.x) {
                const int k = e / n, j = e - (e / n) * n;
                const int a = pr[2 * k], b = pr[2 * k + 1];
                const double c = rot[2 * k], s = rot[2 * k + 1];
                const double x = A[a * ns + j], y = A[b * ns + j];
                A[a * ns + j] = c * x - s * y;
                A[b * ns + j] = s * x + c * y;
            }
            __syncthreads();
            // columns: A <- A J, V <- V J
            for (int e = threadIdx.x; e < half * n; e += blockDim.x) {
                const int k = e / n, i = e - (e / n) * n;
                const int a = pr[2 * k], b = pr[2 * k + 1];
                const double c = rot[2 * k], s = rot[2 * k + 1];
                const double x = A[i * ns + a], y = A[i * ns + b];
                A[i * ns + a] = c * x - s * y;
                A[i * ns + b] = s * x + c * y;
                const double vx = V[i * ns + a], vy = V[i * ns + b];
                V[i * ns + a] = c * vx - s * vy;
                V[i * ns + b] = s * vx + c * vy;
            }
            __syncthreads();
        }
    }
    // pseudo-inverse: coef = sum_{|l_i| > rcond max|l|} (v_i . b / l_i) v_i
    double* w = rot;  // reuse: n doubles of projections
    if (threadIdx.x == 0) {
        double mx = 0.0;
        for (int i = 0; i < D; ++i) mx = fmax(mx, fabs(A[i * ns + i]));
        red[0] = mx * (double)rcond;
    }
    __syncthreads();
    const double cut = red[0];
    for (int i = threadIdx.x; i < D; i += blockDim.x) {
        double s = 0.0;
        for (int j = 0; j < D; ++j) s += V[j * ns + i] * bv[j];
        const double l = A[i * ns + i];
        w[i] = fabs(l) > cut ? s / l : 0.0;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < D; j += blockDim.x) {
        double s = 0.0;
        for (int i = 0; i < D; ++i) s += V[j * ns + i] * w[i];
        M.coef[(size_t)p * D + j] = (float)s;
    }
}

// ---------------------------------------------------------------------------
// mc_image: grid bs, block 256.  image, likelihood, chi^2, lp, dL/dimage.
// ---------------------------------------------------------------------------
template <bool MULTI>
__global__ __launch_bounds__(256) void mc_image(const McTable T, const McArgs A, int want_grad) {
    int p;  // particle inside its lens
    const McLens& E = mc_locate<MULTI>(T, T.pre_part, blockIdx.x, p);
    const int gp = blockIdx.x + E.goff;  // particle of the caller's arrays
    const DPlan& P = E.P;
    const McPlan& M = E.M;
    const int N = P.N, NP = N * N, D = M.D;
    const float* Xp = M.X + (size_t)p * D * NP;
    const float* cp = M.coef + (size_t)p * D;
    const float glp = mc_glp(A, E, gp);
    const float gch = (!A.map && A.g_chi2) ? A.g_chi2[gp] : 0.0f;
    const float inv_np = 1.0f / (float)NP;
    const int GW = N + 2 * M.gp;
    float ll = 0.0f, chi = 0.0f;
    for (int pix = threadIdx.x; pix < NP; pix += blockDim.x) {
        float im = 0.0f;
        for (int c = 0; c < D; ++c) im = fmaf(Xp[(size_t)c * NP + pix], cp[c], im);
        if (A.img) A.img[(size_t)gp * NP + pix] = im;
        if (A.stacked)
            for (int c = 0; c < D; ++c) A.stacked[((size_t)gp * NP + pix) * D + c] = Xp[(size_t)c * NP + pix];
        if (A.g_img) {
            if (want_grad) {
                const int I = pix / N, J = pix - (pix / N) * N;
                M.gimg[(size_t)p * GW * GW + (I + M.gp) * GW + (J + M.gp)] = A.g_img[(size_t)gp * NP + pix];
            }
            continue;
        }
        float l, r2, dll, dr2;
        if (P.lik == GL_LIK_FORWARD)
            glm::lik_forward(im, P.obs[pix], P.bkg2, P.inv_texp, l, r2, dll, dr2);
        else
            glm::lik_backward(im, P.obs[pix], P.inv_var_obs[pix], P.log_sig_obs[pix], l, r2, dll, dr2);
        ll += l;
        chi += r2;
        if (want_grad) {
            const int I = pix / N, J = pix - (pix / N) * N;
            M.gimg[(size_t)p * GW * GW + (I + M.gp) * GW + (J + M.gp)] = glp * dll + gch * inv_np * dr2;
        }
    }
    if (want_grad) {  // the zero margins of the padded dL/dimage (a shared workspace is not cleared)
        float* Gp = M.gimg + (size_t)p * GW * GW;
        for (int i = threadIdx.x; i < GW * GW; i += blockDim.x) {
            const int r = i / GW, c = i - (i / GW) * GW;
            if (r < M.gp || r >= M.gp + N || c < M.gp || c >= M.gp + N) Gp[i] = 0.0f;
        }
    }
    for (int o = 32; o >= 1; o >>= 1) {
        ll += __shfl_xor(ll, o, 64);
        chi += __shfl_xor(chi, o, 64);
    }
    __shared__ float red[2][8];
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = ll;
        red[1][threadIdx.x >> 6] = chi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float L = 0.0f, Cq = 0.0f;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { L += red[0][w]; Cq += red[1][w]; }
        L -= (float)NP * glm::kHalfLn2Pi;
        const float* px = pconst(M, p);
        float lpri = 0.0f;
        for (int j = 0; j < P.d; ++j) lpri += px[5 * P.d + j];
        if (A.lp) A.lp[gp] = L + lpri;
        if (A.chi2) A.chi2[gp] = Cq * inv_np;
    }
    if (A.coef_out)
        for (int c = threadIdx.x; c < D; c += blockDim.x) A.coef_out[(size_t)gp * D + c] = cp[c];
}

// ---------------------------------------------------------------------------
// mc_adjoint: transposed correlation of dL/dimage onto the sub-grid.
// block 256 = 4 waves; one wave per (particle, sub-row y, column phase b, 64
// columns Xc); lane = sub-pixel x = ss Xc + b.  For row phase a = y % ss the
// contributing image rows are Yc + da - m (m < Ma) and, in each, the columns
// Xc + db - nn (nn < Mb); the weights W[ss m + ra][ss nn + rb] are wave-uniform
// and read from an LDS copy of the folded PSF.  Each particle owns a multiple
// of 4 waves, so a block never spans two particles (or lenses).
// ---------------------------------------------------------------------------
template <bool MULTI>
__global__ __launch_bounds__(256) void mc_adjoint(const McTable T) {
    int local;
    const McLens& E = mc_locate<MULTI>(T, T.pre_adj, blockIdx.x * 4, local);
    const DPlan& P = E.P;
    const McPlan& M = E.M;
    const int N = P.N, nsub = P.nsub, ss = P.ss, kf = P.kf, pb = P.pb;
    extern __shared__ float4 mc_dyn4[];
    float* s_w = reinterpret_cast<float*>(mc_dyn4);
    for (int i = threadIdx.x; i < kf * kf; i += 256) s_w[i] = P.w[i];
    __syncthreads();
    local += threadIdx.x >> 6;
    const int njb = (N + 63) / 64, wpp4 = mc_adj_wpp4(P);
    const int p = local / wpp4;
    int t = local - p * wpp4;
    if (t >= nsub * ss * njb) return;
    const int jb = t % njb; t /= njb;
    const int b = t % ss; t /= ss;
    const int y = t;
    const int Xc = jb * 64 + (threadIdx.x & 63);
    if (Xc >= N) return;
    const int a = y % ss, Yc = y / ss;
    const int ra = (a + pb) % ss, da = (a + pb) / ss, Ma = (kf - ra + ss - 1) / ss;
    const int rb = (b + pb) % ss, db = (b + pb) / ss, Mb = (kf - rb + ss - 1) / ss;
    const int GW = N + 2 * M.gp;
    const float* G = M.gimg + (size_t)p * GW * GW;
    float g = 0.0f;
    for (int m = 0; m < Ma; ++m) {
        const float* row = G + (size_t)(Yc + da - m + M.gp) * GW + (Xc + db + M.gp);
        const float* wr = s_w + (ss * m + ra) * kf + rb;
        for (int nn = 0; nn < Mb; ++nn) g = fmaf(wr[ss * nn], row[-nn], g);
    }
    M.B[((size_t)p * M.D * nsub + y) * M.BW + pb + ss * Xc + b] = g;
}

// mc_adj_v: separable adjoint, transposed vertical pass.  block 256 = 4 waves;
// one wave per (particle, row phase a, kAdjKY sub-rows y = ss (Yc0 + k) + a,
// column phase b, 64 columns Xc); lane = sub-pixel column x = ss Xc + b:
//   dL/dsub[y][x] = sum_r sum_{m < Ma} asa_r[ss m + ra] U_r[Yc + da - m][b][Xc].
// The lane walks the KFM + kAdjKY - 1 rows of U its sub-rows need once and adds
// each into every sub-row it reaches.  SS and KF4 are the launch group's: with
// the loops unrolled, the tap of every (row, sub-row) pair is a compile-time
// index into the zero-padded polyphase table (scalar loads, no per-tap test);
// the padded taps (m >= Ma) carry zero factors.
template <bool MULTI, int SS, int KF4>
__global__ __launch_bounds__(256) void mc_adj_v(const McTable T) {
    constexpr int KFM = (4 * KF4 + SS - 1) / SS;
    int local;
    const McLens& E = mc_locate<MULTI>(T, T.pre_adjv, blockIdx.x * 4, local);
    const DPlan& P = E.P;
    const McPlan& M = E.M;
    constexpr int ss = SS;
    const int N = P.N, nsub = P.nsub, pb = P.pb, R = M.adj_rank;
    local += __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int njb = (N + 63) / 64, nyb = (N + kAdjKY - 1) / kAdjKY, wpp4 = mc_adjv_wpp4(P);
    const int p = local / wpp4;
    int t = local - p * wpp4;
    if (t >= ss * nyb * ss * njb) return;
    const int jb = t % njb; t /= njb;
    const int b = t % ss; t /= ss;
    const int yb = t % nyb; t /= nyb;
    const int a = t;
    const int Xc = jb * 64 + (threadIdx.x & 63);
    if (Xc >= N) return;
    const int da = (a + pb) / ss;
    const int Yc0 = yb * kAdjKY;
    float g[kAdjKY];
#pragma unroll
    for (int k = 0; k < kAdjKY; ++k) g[k] = 0.0f;
    // uniform base; lanes index by Xc
    const float* Ub = M.U + (size_t)p * R * N * ss * N + (size_t)b * N;
    // rows I = I0 + j, j < KFM + kAdjKY - 1; row j reaches sub-row k with tap m = KFM - 1 - j + k
    const int I0 = Yc0 + da - (KFM - 1);
    const long rstride = (long)ss * N;  // floats from one U row to the next
    for (int r = 0; r < R; ++r) {
        const CFloat* wa = (const CFloat*)M.asap + (r * ss + a) * KFM;
        const float* row = Ub + (size_t)r * N * ss * N + I0 * rstride;  // row I0 (may lie outside the image)
#pragma unroll
        for (int j = 0; j < KFM + kAdjKY - 1; ++j, row += rstride) {
            const int I = I0 + j;
            const float u = (I >= 0 && I < N) ? row[Xc] : 0.0f;
#pragma unroll
            for (int k = 0; k < kAdjKY; ++k) {
                const int m = KFM - 1 - j + k;
                if (m >= 0 && m < KFM) g[k] = fmaf(wa[m], u, g[k]);
            }
        }
    }
    float* out = M.B + (size_t)p * M.D * nsub * M.BW + pb + ss * Xc + b;
#pragma unroll
    for (int k = 0; k < kAdjKY; ++k) {
        const int Yc = Yc0 + k;
        if (Yc < N) out[(size_t)(ss * Yc + a) * M.BW] = g[k];
    }
}

// mc_adj_h: separable adjoint, transposed horizontal pass.  block 256 = 4 waves;
// one wave per (particle, image row I, 64 columns Xc); each lane computes every
// column phase b and rank r:
//   U_r[I][b][Xc] = sum_{nn < Mb} asb_r[ss nn + rb] dL/dimage[I][Xc + db - nn].
// The lane loads the KFM + 1 gradient values Xc + db(0) - (KFM - 1) .. Xc + db(0) + 1
// once (db(b) - db(0) is 0 or 1); the host's table asbq places each phase's factors
// at their window positions (zero elsewhere), so with SS and KF4 fixed per launch
// group every tap index is a compile-time constant.  The padded gradient image
// (margin gp >= KFM) supplies the zero columns outside the image.
template <bool MULTI, int SS, int KF4>
__global__ __launch_bounds__(256) void mc_adj_h(const McTable T) {
    constexpr int KFM = (4 * KF4 + SS - 1) / SS;
    constexpr int WN = KFM + 1;
    int local;
    const McLens& E = mc_locate<MULTI>(T, T.pre_adjh, blockIdx.x * 4, local);
    const DPlan& P = E.P;
    const McPlan& M = E.M;
    const int N = P.N, pb = P.pb, R = M.adj_rank;
    local += __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int njb = (N + 63) / 64, wpp4 = mc_adjh_wpp4(P);
    const int p = local / wpp4;
    int t = local - p * wpp4;
    if (t >= N * njb) return;
    const int jb = t % njb;
    const int I = t / njb;
    const int Xc = jb * 64 + (threadIdx.x & 63);
    if (Xc >= N) return;
    const int db0 = pb / SS;
    const int GW = N + 2 * M.gp;
    const float* row = M.gimg + (size_t)p * GW * GW + (size_t)(I + M.gp) * GW + (Xc + db0 - (KFM - 1) + M.gp);
    float w[WN];
#pragma unroll
    for (int j = 0; j < WN; ++j) w[j] = row[j];
    float* U = M.U + (size_t)p * R * N * SS * N + (size_t)I * SS * N + Xc;
    for (int r = 0; r < R; ++r) {
#pragma unroll
        for (int b = 0; b < SS; ++b) {
            const CFloat* q = (const CFloat*)M.asbq + (r * SS + b) * WN;
            float u = 0.0f;
#pragma unroll
            for (int j = 0; j < WN; ++j) u = fmaf(q[j], w[j], u);
            U[(size_t)r * N * SS * N + (size_t)b * N] = u;
        }
    }
}

// ---------------------------------------------------------------------------
// mc_grad: block 256, each block 256 x kMcGradK sub-pixels of one particle.
// Reverse-mode statistics with the solved amplitudes; block-reduced, then
// atomically added per particle.
// Statistics layout: per lens slot 8 (EPL 8 / shear 2 / SIS 3), per light slot 8
// (Sersic) or 3 (shapelets: sum g L_u, sum g L_v, sum g (u L_u + v L_v)).
// ---------------------------------------------------------------------------
// SHP: the plan has a shapelet source (otherwise that branch is compiled out).
// Each thread accumulates kMcGradK sub-pixels (256 apart) before the block
// reduction, so the 8-per-component statistics are reduced once per 2,048
// sub-pixels.  A sub-pixel whose forward value was zeroed for NaN contributes
// nothing: per channel with least squares (the channel was zeroed), the whole
// sub-pixel with fixed amplitudes (their sum was zeroed, tf/simulator.py:81).
// LM: the lens kinds of every lens of the launch; NC: an upper bound on its
// components (statistics kept in registers: up to 11 per component).  KS: the kind of
// each component slot, 8 bits per slot (0: read from the plan), so that an instance for
// one model keeps only the statistics its kinds update.
__host__ __device__ constexpr int slot_kinds(int k0, int k1, int k2, int k3) {
    return k0 | (k1 << 8) | (k2 << 16) | (k3 << 24);
}
template <int KS>
__device__ __forceinline__ int slot_kind(int i, int plan_kind) {
    const int k = (KS >> (8 * i)) & 255;
    return k ? k : plan_kind;
}
template <bool MULTI, bool SHP, int LM, int NC, int KS = 0>
__device__ __forceinline__ void mc_grad_body(const McTable& T) {
    static_assert(kStatStride * NC <= kMcStat, "statistics per particle");
    int local;
    const McLens& E = mc_locate<MULTI>(T, T.pre_grad, blockIdx.x, local);
    const DPlan& P = E.P;
    const McPlan& M = E.M;
    const int gpp = mc_gpp(P);
    const int p = local / gpp;
    const int blk = local - p * gpp;
    const int nsub = P.nsub;
    float st[kStatStride * NC];
#pragma unroll
    for (int e = 0; e < kStatStride * NC; ++e) st[e] = 0.0f;
    extern __shared__ float4 mc_dyn4[];
    const float* pc = mc_stage_consts(M, p, reinterpret_cast<float*>(mc_dyn4));
    const float* s_c = pc + ((6 * P.d + 3) & ~3);
    const float* cp = M.coef + (size_t)p * M.D;
    for (int it = 0; it < kMcGradK; ++it) {
        const int s = (blk * kMcGradK + it) * 256 + threadIdx.x;
        if (s >= nsub * nsub) break;
        const int y = s / nsub, x = s - (s / nsub) * nsub;
        const float g = M.B[((size_t)p * M.D * nsub + y) * M.BW + P.pb + x];
        const float X = fmaf(P.M00, (float)x, fmaf(P.M01, (float)y, P.ox));
        const float Y = fmaf(P.M10, (float)x, fmaf(P.M11, (float)y, P.oy));
        LensStates ls;
        float bx, by;
        mc_beta<true, LM>(P, M, pc, X, Y, bx, by, &ls);
        if (!M.lstsq && !M.famp) {
            // the sum of the lights: a NaN zeroed the whole sub-pixel in the forward pass
            float v = 0.0f;
            for (int i = P.nlens; i < P.ncomp; ++i) {
                const bool src = P.comp[i].role == GL_ROLE_SOURCE;
                v += mc_sersic_light(P.comp[i].kind, s_c + i * kConstStride, src ? bx : X, src ? by : Y);
            }
            if (isnan(v)) continue;
        }
        if (SHP && M.famp && isnan(mc_famp_sum(P, M, s_c, cp, X, Y, bx, by))) continue;
        int ch = 0;
        float gbx = 0.0f, gby = 0.0f;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            if (i < P.nlens) continue;
            if (i >= P.ncomp) break;
            const DComp& cd = P.comp[i];
            const int kind = slot_kind<KS>(i, cd.kind);
            const float* cst = s_c + i * kConstStride;
            const bool src = cd.role == GL_ROLE_SOURCE;
            const float ux = src ? bx : X, uy = src ? by : Y;
            float* acc = st + kStatStride * i;
            if (kind == glm::K_SERSIC || kind == glm::K_SERSIC_ELLIPSE) {
                const glm::SerC C = *reinterpret_cast<const glm::SerC*>(cst);
                glm::SerState ss_;
                const float v = glm::sersic_forward<true>(C, ux, uy, ss_);
                const int c = (M.lstsq || M.famp) ? ch++ : 0;
                if (isnan(v)) continue;  // least squares: this channel was zeroed
                float gdx, gdy;
                glm::sersic_adjoint(C, ss_, g * cp[c], acc, gdx, gdy);
                if (src) { gbx += gdx; gby += gdy; }
            } else if (kind == glm::K_CORE_SERSIC) {
                const glm::CoreSerC C = *reinterpret_cast<const glm::CoreSerC*>(cst);
                glm::CoreSerState cs;
                const float v = glm::core_sersic_forward(C, ux, uy, cs);
                const int c = (M.lstsq || M.famp) ? ch++ : 0;
                if (isnan(v)) continue;
                float gdx, gdy;
                glm::core_sersic_adjoint(C, cs, g * cp[c], acc, gdx, gdy);
                if (src) { gbx += gdx; gby += gdy; }
            } else if (SHP && kind == glm::K_SHAPELETS) {
                const float beta = cst[0];
                const float ib = 1.0f / beta;
                const float u = (ux - cst[1]) * ib, w = (uy - cst[2]) * ib;
                float pu[glm::kMaxShapeletN + 1], pw[glm::kMaxShapeletN + 1];
                float du[glm::kMaxShapeletN + 1], dw[glm::kMaxShapeletN + 1];
                mc_phi(M, u, pu, du);
                mc_phi(M, w, pw, dw);
                float Lu = 0.0f, Lw = 0.0f;
#pragma unroll
                for (int so = 0; so <= glm::kMaxShapeletN; ++so) {
                    if (so > M.shp_nmax) break;
#pragma unroll
                    for (int j = 0; j <= so; ++j) {
                        const float a = cp[ch + so * (so + 1) / 2 + j];
                        Lu = fmaf(a, du[so - j] * pw[j], Lu);
                        Lw = fmaf(a, pu[so - j] * dw[j], Lw);
                    }
                }
                ch += (M.shp_nmax + 1) * (M.shp_nmax + 2) / 2;
                if (isnan(u) || isnan(w)) continue;  // the basis was zeroed
                acc[0] += g * Lu;
                acc[1] += g * Lw;
                acc[2] += g * (u * Lu + w * Lw);
                if (src) { gbx += g * Lu * ib; gby += g * Lw * ib; }
            }
        }
        if (isnan(bx) || isnan(by)) continue;  // no deflection adjoint through a NaN ray
        const float Hx = -gbx, Hy = -gby;
#pragma unroll
        for (int i = 0; i < kMaxLens; ++i) {
            if (i >= P.nlens) break;
            const int kind = slot_kind<KS>(i, P.comp[i].kind);
            const float* cst = s_c + i * kConstStride;
            float* acc = st + kStatStride * i;
            if (lens_in(LM, i, glm::K_EPL) && kind == glm::K_EPL) {
                const glm::EplC C = *reinterpret_cast<const glm::EplC*>(cst);
                glm::epl_adjoint(C, ls.es[i], Hx, Hy, acc);
            } else if (lens_in(LM, i, glm::K_SHEAR) && kind == glm::K_SHEAR) {
                glm::shear_adjoint(X, Y, Hx, Hy, acc);
            } else if (lens_in(LM, i, glm::K_SIS) && kind == glm::K_SIS) {
                const glm::SisC C = *reinterpret_cast<const glm::SisC*>(cst);
                glm::sis_adjoint(C, ls.sis[i], Hx, Hy, acc);
            } else if (lens_in(LM, i, glm::K_SIE) && kind == glm::K_SIE) {
                const glm::SieC C = *reinterpret_cast<const glm::SieC*>(cst);
                glm::sie_adjoint(C, ls.sie[i], Hx, Hy, acc);
            } else if (lens_in(LM, i, glm::K_TNFW) && kind == glm::K_TNFW) {
                const glm::TnfwC C = *reinterpret_cast<const glm::TnfwC*>(cst);
                glm::tnfw_adjoint(C, ls.tn[i], Hx, Hy, acc);
            }
        }
    }
    __shared__ float red[kMcStat][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nst = kStatStride * P.ncomp;
#pragma unroll
    for (int e = 0; e < kStatStride * NC; ++e) {
        if (e < nst && (e % kStatStride) < 11) {
            float v = st[e];
            for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
            if (lane == 0) red[e][wv] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x < nst && (threadIdx.x % kStatStride) < 11) {
        const float v = red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
        atomicAdd(M.stats + (size_t)p * kMcStat + threadIdx.x, v);
    }
}

template <bool MULTI, bool SHP>
__global__ __launch_bounds__(256) void mc_grad(const McTable T) {
    mc_grad_body<MULTI, SHP, kLensAll, kMaxComp>(T);
}

// The narrow instance (EPL + shear lenses, at most 4 components) fits two
// workgroups per CU: 8 waves, up to 256 registers per lane.
template <bool MULTI, bool SHP>
__global__ __launch_bounds__(256, 2) void mc_grad_lean(const McTable T) {
    mc_grad_body<MULTI, SHP, kLensEplShear, 4>(T);
}

// The shapelet configuration's components in their slots: EPL, shear, a Sersic-ellipse
// light and a shapelet light (21 statistics in registers instead of 44).
constexpr int kSlotsEplShearSerShp = slot_kinds(glm::K_EPL, glm::K_SHEAR, glm::K_SERSIC_ELLIPSE, glm::K_SHAPELETS);
template <bool MULTI>
__global__ __launch_bounds__(256, 2) void mc_grad_shp4(const McTable T) {
    mc_grad_body<MULTI, true, kLensEplShear, 4, kSlotsEplShearSerShp>(T);
}

// ---------------------------------------------------------------------------
// Model-specialised basis and gradient for fixed-amplitude plans whose profile
// kinds are those of MODEL (every lens of the launch): the constants come from
// the LDS copy into scalar registers and only MODEL's states exist, through the
// per-sub-pixel code of the fused kernel (eval_forward / eval_adjoint).  The
// statistics land in the generic layout (8 per component) that mc_finish reads.
// ---------------------------------------------------------------------------
template <class M>
__device__ __forceinline__ int mc_stat_slot(int e) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int k = M::lens[i];
        const int n = k == glm::K_EPL ? LensAcc<glm::K_EPL>::n
                    : k == glm::K_SHEAR ? LensAcc<glm::K_SHEAR>::n
                    : k == glm::K_SIS ? LensAcc<glm::K_SIS>::n
                    : k == glm::K_SIE ? LensAcc<glm::K_SIE>::n
                    : k == glm::K_TNFW ? LensAcc<glm::K_TNFW>::n : 0;
        if (e >= M::lens_off(i) && e < M::lens_off(i) + n) return kStatStride * i + (e - M::lens_off(i));
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (M::light[j] == 0) continue;
        const int off = M::light_off(light_acc_slot<M>(j));
        if (e >= off && e < off + 8) return kStatStride * light_comp_index<M>(j) + (e - off);
    }
    return 0;
}

template <bool MULTI, class MODEL>
__global__ __launch_bounds__(256) void mc_basis_m(const McTable T) {
    int local;
    const McLens& E = mc_locate<MULTI>(T, T.pre_basis, blockIdx.x, local);
    const DPlan& P = E.P;
    const McPlan& M = E.M;
    const int bpp = mc_bpp(P);
    const int p = local / bpp;
    const int s = (local - p * bpp) * 256 + threadIdx.x;
    const int nsub = P.nsub;
    extern __shared__ float4 mc_dyn4[];
    const float* pc = mc_stage_consts(M, p, reinterpret_cast<float*>(mc_dyn4));
    Consts<MODEL> C;
    load_consts<MODEL>(pc + ((6 * P.d + 3) & ~3), pc + M.coef_base, P, C);
    if (s >= nsub * nsub) return;
    const int y = s / nsub, x = s - y * nsub;
    const float X = fmaf(P.M00, (float)x, fmaf(P.M01, (float)y, P.ox));
    const float Y = fmaf(P.M10, (float)x, fmaf(P.M11, (float)y, P.oy));
    const float v = eval_forward<MODEL>(C, X, Y);
    float* row = M.B + (size_t)p * nsub * M.BW + (size_t)y * M.BW;
    row[P.pb + x] = isnan(v) ? 0.0f : v;  // tf/simulator.py:81
    mc_zero_margins(M, P, row, x, 1);
}

template <bool MULTI, class MODEL>
__global__ __launch_bounds__(256) void mc_grad_m(const McTable T) {
    int local;
    const McLens& E = mc_locate<MULTI>(T, T.pre_grad, blockIdx.x, local);
    const DPlan& P = E.P;
    const McPlan& M = E.M;
    const int gpp = mc_gpp(P);
    const int p = local / gpp;
    const int blk = local - p * gpp;
    const int nsub = P.nsub;
    extern __shared__ float4 mc_dyn4[];
    const float* pc = mc_stage_consts(M, p, reinterpret_cast<float*>(mc_dyn4));
    Consts<MODEL> C;
    load_consts<MODEL>(pc + ((6 * P.d + 3) & ~3), pc + M.coef_base, P, C);
    const float cf = M.coef[(size_t)p * M.D];  // fixed amplitudes: image = cf * conv
    float acc[MODEL::NACC];
#pragma unroll
    for (int e = 0; e < MODEL::NACC; ++e) acc[e] = 0.0f;
    for (int it = 0; it < kMcGradK; ++it) {
        const int s = (blk * kMcGradK + it) * 256 + threadIdx.x;
        if (s >= nsub * nsub) break;
        const int y = s / nsub, x = s - y * nsub;
        const float g = M.B[(size_t)p * nsub * M.BW + (size_t)y * M.BW + P.pb + x];
        const float X = fmaf(P.M00, (float)x, fmaf(P.M01, (float)y, P.ox));
        const float Y = fmaf(P.M10, (float)x, fmaf(P.M11, (float)y, P.oy));
        eval_adjoint<MODEL>(C, X, Y, g * cf, acc);
    }
    __shared__ float red[MODEL::NACC][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int e = 0; e < MODEL::NACC; ++e) {
        const float v = wave_sum(acc[e]);
        if (lane == 0) red[e][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < MODEL::NACC) {
        const int e = threadIdx.x;
        const float v = red[e][0] + red[e][1] + red[e][2] + red[e][3];
        atomicAdd(M.stats + (size_t)p * kMcStat + mc_stat_slot<MODEL>(e), v);
    }
}

// ---------------------------------------------------------------------------
// mc_ampgrad (fixed per-channel amplitudes, single plans): the image is
// sum_c cf a_c X_c, so dL/da_c = cf sum_pix dL/dimage X_c.  grid (particles x D),
// block 256: one channel of one particle per block, a fixed-order reduction.
// ---------------------------------------------------------------------------
template <bool MULTI>
__global__ __launch_bounds__(256) void mc_ampgrad(const McTable T) {
    static_assert(!MULTI, "single plans only");
    const DPlan& P = T.one.P;
    const McPlan& M = T.one.M;
    const int D = M.D;
    const int p = blockIdx.x / D, c = blockIdx.x - (blockIdx.x / D) * D;
    const int N = P.N, NP = N * N, GW = N + 2 * M.gp;
    const float* Xc = M.X + ((size_t)p * D + c) * NP;
    const float* G = M.gimg + (size_t)p * GW * GW;
    float acc = 0.0f;
    for (int pix = threadIdx.x; pix < NP; pix += blockDim.x) {
        const int I = pix / N, J = pix - (pix / N) * N;
        acc = fmaf(G[(I + M.gp) * GW + (J + M.gp)], Xc[pix], acc);
    }
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    __shared__ float red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) M.gamp[(size_t)p * D + c] = P.cf * (red[0] + red[1] + red[2] + red[3]);
}

// ---------------------------------------------------------------------------
// mc_finish: grid bs, block 64.  Statistics -> dlp/dz (or dL/dx).
// ---------------------------------------------------------------------------
template <bool MULTI>
__global__ void mc_finish(const McTable T, const McArgs A) {
    int p;
    const McLens& E = mc_locate<MULTI>(T, T.pre_part, blockIdx.x, p);
    const int gp = blockIdx.x + E.goff;  // particle of the caller's arrays
    const DPlan& P = E.P;
    const McPlan& M = E.M;
    float* px = pconst(M, p);
    const int d = P.d;
    float* gx = px + 4 * d;
    const float* s_c = px + ((6 * d + 3) & ~3);
    const float* st = M.stats + (size_t)p * kMcStat;
    const float* cp = M.coef + (size_t)p * M.D;
    if (threadIdx.x == 0) {
        for (int j = 0; j < d; ++j) gx[j] = 0.0f;
        int ch = 0;
        for (int i = 0; i < P.ncomp; ++i) {
            const DComp& cd = P.comp[i];
            const float* cst = s_c + i * kConstStride;
            const float* acc = st + kStatStride * i;
            float g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            int np_ = 0;
            bool circ = false;  // circular Sersic: canonical (Rs, n, cx, cy) -> finish order (Rs, n, e1, e2, cx, cy)
            if (cd.kind == glm::K_EPL) {
                glm::epl_finish(*reinterpret_cast<const glm::EplC*>(cst), acc, g);
                np_ = 6;
            } else if (cd.kind == glm::K_SHEAR) {
                g[0] = acc[0]; g[1] = acc[1];
                np_ = 2;
            } else if (cd.kind == glm::K_SIS) {
                glm::sis_finish(acc, g);
                np_ = 3;
            } else if (cd.kind == glm::K_SIE) {
                glm::sie_finish(*reinterpret_cast<const glm::SieC*>(cst), acc, g);
                np_ = 5;
            } else if (cd.kind == glm::K_TNFW) {
                glm::tnfw_finish(*reinterpret_cast<const glm::TnfwC*>(cst), acc, g);
                np_ = 5;
            } else if (cd.kind == glm::K_SERSIC_ELLIPSE || cd.kind == glm::K_SERSIC) {
                glm::sersic_finish(*reinterpret_cast<const glm::SerC*>(cst), acc, g);
                np_ = cd.kind == glm::K_SERSIC_ELLIPSE ? 6 : 4;  // no Ie in least-squares mode
                if (!cd.lstsq && !M.famp) np_ += 1;             // fixed amplitudes: Ie too
                if (M.famp) {                                   // per-channel amplitudes: Ie from mc_ampgrad
                    const int ie = mc_ie_index(cd.kind);
                    if (cd.col[ie] >= 0) gx[cd.col[ie]] += M.gamp[(size_t)p * M.D + ch];
                }
                circ = cd.kind == glm::K_SERSIC;
                ++ch;
            } else if (cd.kind == glm::K_CORE_SERSIC) {
                float gc[GL_MAX_COMPONENT_PARAMS];
                glm::core_sersic_finish(*reinterpret_cast<const glm::CoreSerC*>(cst), acc, gc);
                const int npc = (!cd.lstsq && !M.famp) ? 10 : 9;  // Ie too with fixed amplitudes
                for (int q = 0; q < npc; ++q)
                    if (cd.col[q] >= 0) gx[cd.col[q]] += gc[q];
                if (M.famp && cd.col[9] >= 0) gx[cd.col[9]] += M.gamp[(size_t)p * M.D + ch];
                ++ch;
            } else if (cd.kind == glm::K_SHAPELETS) {
                const float beta = cst[0];
                g[0] = -acc[2] / beta;  // d/d beta
                g[1] = -acc[0] / beta;  // d/d center_x
                g[2] = -acc[1] / beta;  // d/d center_y
                np_ = 3;
                const int nl = (M.shp_nmax + 1) * (M.shp_nmax + 2) / 2;
                if (M.famp)  // sampled amplitudes amp00.. (columns amp_col0..)
                    for (int k = 0; k < nl; ++k) gx[cd.amp_col0 + k] += M.gamp[(size_t)p * M.D + ch + k];
                ch += nl;
            }
            for (int q = 0; q < np_; ++q)
                if (cd.col[q] >= 0) gx[cd.col[q]] += circ ? g[q < 2 ? q : q + 2] : g[q];
        }
    }
    __syncthreads();
    if (threadIdx.x < d && (A.grad || A.map)) {
        const int j = threadIdx.x;
        const float glp = mc_glp(A, E, gp);
        float gz;
        if (!A.use_z) gz = gx[j];
        else gz = (gx[j] + glp * px[2 * d + j]) * px[d + j] + glp * px[3 * d + j];
        if (A.grad) A.grad[(size_t)gp * d + j] = gz;
        if (A.map) {
            // Keras Adam (epsilon outside the sqrt; lr_t carries the bias corrections)
            const size_t i = (size_t)gp * d + j;
            const float m = A.beta1 * A.m[i] + (1.0f - A.beta1) * gz;
            const float v = A.beta2 * A.v[i] + (1.0f - A.beta2) * gz * gz;
            A.m[i] = m;
            A.v[i] = v;
            A.z_io[i] -= A.lr_t * m / (sqrtf(v) + A.eps);
        }
    }
}

}  // namespace gl
