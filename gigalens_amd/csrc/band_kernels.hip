// Banded persistent particle kernel for gfx950 (MI355X): every particle of every
// lens system of a batch -- lenses of different image sizes, supersampling,
// PSFs and particle counts -- in one launch.
//
// The grid is as many workgroups as the CUs hold at once.  Each workgroup takes
// (lens, particle) items from a device work queue (one atomic per item; lenses
// in descending order of cost, so the longest items start first and the launch
// ends on short ones) until the queue is drained, and runs the whole
// log-probability (and gradient, and MAP step) of the item:
//
//   A  prologue   z -> x, log prior, component constants, EPL series tables (LDS)
//   B  ray-trace  per band of BY sub-rows: lens light(theta) + source(beta), NaN -> 0,
//                 into a zero-haloed LDS band
//   C  conv+pool  horizontal pass of each band into T (the workgroup's HBM slot),
//                 then the vertical pass, the Gaussian log-likelihood and dL/dimage
//                 (HBM slot) -- band_passes.h
//   D  adjoint    per band: transposed vertical pass of dL/dimage into LDS, then the
//                 transposed horizontal pass to every sub-pixel of the band (LDS)
//   E  recompute  the ray-trace again per sub-pixel, accumulating the reverse-mode
//                 statistics
//   F  reduce     as the resident-image kernel: chain rule, lp / chi2 / gradient or
//                 the Keras-Adam update
//
// Only one band of the supersampled image is ever in LDS, so an image of any size
// runs at the LDS footprint of the band; the separable (rank <= 8) folded PSF keeps
// the HBM intermediates at R x nsub x N floats per particle, written and read by
// the same workgroup while they are in L2.  The last workgroup to leave resets
// the queue for the next launch (no host work between launches: HIP graphs can
// record it).
//
// Reference semantics as the resident-image kernel (lens_kernels.hip):
// src/gigalens/tf/simulator.py:57-97, src/gigalens/tf/model.py:53-86,
// src/gigalens/tf/inference.py:28-35.
#pragma once
#include <hip/hip_runtime.h>
#include "band_passes.h"
#include "lens_kernels.hip"

namespace gl {

constexpr int kBandBlock = 512;

struct BLens {
    DPlan P;           // geometry, components, prior and observed image of the lens; coef_off and
                       // max_n describe the band kernel's LDS layout; batch_n is null
    band::Geom g;      // band geometry
    int first;         // first particle of the lens in the caller's (lens-major) arrays
    int count;         // its particles
    float glp_map;     // MAP: gradient weight of lp, -1 / (count N^2) (the lens's loss is its mean)
    const float* fac;  // band::fac_floats: vertical factors a_r, horizontal b_r (KFP each, zero padded),
                       // then a_r again with ZP zeros in front (KFV each)
                       // (read-only for the launch: scalar loads through the constant space)
};

struct BArgs {
    KArgs K;           // mode (M_LOGPROB or M_MAP) and the caller's arrays (z, lp, chi2, grad, Adam state)
    const BLens* lens;
    const int* order;  // lens indices in processing order (descending cost)
    const int* pre;    // n + 1 prefix sums of the counts in processing order
    int n, total;      // lenses, items (particles)
    float* scratch;    // slot floats per workgroup
    long slot;
    int* queue;        // [0] next item, [1] workgroups done; both 0 between launches
    int off_x, off_c, off_coef, off_red, off_band;  // LDS layout (floats)
};

// Per-lens values the per-sub-pixel and per-pixel loops use, held in scalar registers.
struct BHot {
    band::Geom g;
    float M00, M01, M10, M11, ox, oy, cf, bkg2, inv_texp;
    int lik;
    float glp_map;
    const float* obs;
    const float* inv_var_obs;
    const float* log_sig_obs;
    const float* fac;
};

// Copy of a lens-uniform struct of 32-bit words into scalar registers (one
// readfirstlane per word: the loads themselves may be vector loads).
template <class T>
__device__ __forceinline__ T uniform_copy(const T& src) {
    static_assert(sizeof(T) % 4 == 0, "32-bit words");
    T out;
    int* o = reinterpret_cast<int*>(&out);
    const int* w = reinterpret_cast<const int*>(&src);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) o[i] = __builtin_amdgcn_readfirstlane(w[i]);
    return out;
}

__device__ __forceinline__ BHot band_hot(const BLens* L) {
    BHot h;
    h.g = L->g;
    h.M00 = L->P.M00; h.M01 = L->P.M01; h.M10 = L->P.M10; h.M11 = L->P.M11; h.ox = L->P.ox; h.oy = L->P.oy;
    h.cf = L->P.cf; h.bkg2 = L->P.bkg2; h.inv_texp = L->P.inv_texp; h.lik = L->P.lik;
    h.glp_map = L->glp_map;
    h.obs = L->P.obs; h.inv_var_obs = L->P.inv_var_obs; h.log_sig_obs = L->P.log_sig_obs;
    h.fac = L->fac;
    return uniform_copy(h);
}

template <int BLOCK, class M, bool GRAD>
__device__ __forceinline__ void band_particle(const BArgs& A, const BLens* Lp, int pid, float* s_x, float* s_c,
                                              float* s_coef, float* s_red, float* s_band, float* slot, int tid) {
    const DPlan& P = Lp->P;  // prologue and epilogue only
    const BHot L = band_hot(Lp);
    const band::Geom& g = L.g;
    const KArgs& K = A.K;
    const int d = P.d;
    const float* fa = L.fac;                  // vertical factors a_r (read by scalar loads)
    const float* fb = L.fac + g.R * g.KFP;    // horizontal factors b_r
    const float* fv = L.fac + 2 * g.R * g.KFP;  // a_r zero padded for the vertical pass
    float* T = slot;
    float* G = slot + band::t_floats(g);

    // ---- A: prologue -------------------------------------------------------
    float zin = 0.0f, m_in = 0.0f, v_in = 0.0f;
    if (tid < d) {
        const size_t o = (size_t)pid * d + tid;
        zin = K.in[o];
        if (K.mode == M_MAP) { m_in = K.m[o]; v_in = K.v[o]; }
    }
    const int nband = band::band_floats(g);
    for (int i = tid; i < nband; i += BLOCK) s_band[i] = 0.0f;  // halo columns of the forward bands
    band::zero_pads(g, tid, BLOCK, T, G);
    prologue_prior(P, true, tid, zin, s_x);
    __syncthreads();
    prologue_components(P, tid, s_x, s_c);
    __syncthreads();
    prologue_tables<M>(P, tid, s_c, s_coef);
    __syncthreads();
    Consts<M> C;
    load_consts<M>(s_c, s_coef, P, C);

    // ---- B + C1: per band, ray-trace then the horizontal pass ------------------
    for (int y0 = 0; y0 < g.nsub; y0 += g.BY) {
        const int rows = g.BY < g.nsub - y0 ? g.BY : g.nsub - y0;
        const int cnt = rows * g.nsub;
        int yy = tid / g.nsub, x = tid - yy * g.nsub;
        for (int s = tid; s < cnt; s += BLOCK) {
            const int y = y0 + yy;
            const float X = fmaf(L.M00, (float)x, fmaf(L.M01, (float)y, L.ox));
            const float Y = fmaf(L.M10, (float)x, fmaf(L.M11, (float)y, L.oy));
            float val = eval_forward<M>(C, X, Y);
            if (isnan(val)) val = 0.0f;
            s_band[yy * g.BW + g.pb + x] = val;
            x += BLOCK;
            while (x >= g.nsub) { x -= g.nsub; ++yy; }
        }
        __syncthreads();
        band::hpass_any(g, tid, BLOCK, y0, rows, s_band, fb, T);
        __syncthreads();
    }

    // ---- C2: vertical pass, likelihood, dL/dimage ---------------------------------
    float ll_acc = 0.0f, chi_acc = 0.0f;
    const float cf = L.cf;
    const float inv_np = 1.0f / (float)(g.N * g.N);
    const float glp = K.mode == M_MAP ? L.glp_map : (K.g_lp ? K.g_lp[pid] : 1.0f);
    const float gch = (K.mode == M_LOGPROB && K.g_chi2) ? K.g_chi2[pid] : 0.0f;
    float* Gi = G + g.GT * g.N;  // row 0 of the image in the padded G
    band::vpass_any(g, tid, BLOCK, T, fv, [&](int I, int J, float conv) {
        const float im = cf * conv;
        const int pix = I * g.N + J;
        float ll, r2, dll, dr2;
        if (L.lik == GL_LIK_FORWARD)
            glm::lik_forward(im, L.obs[pix], L.bkg2, L.inv_texp, ll, r2, dll, dr2);
        else
            glm::lik_backward(im, L.obs[pix], L.inv_var_obs[pix], L.log_sig_obs[pix], ll, r2, dll, dr2);
        ll_acc += ll;
        chi_acc += r2;
        if (GRAD) Gi[pix] = cf * (glp * dll + gch * inv_np * dr2);
    });

    float acc[M::NACC > 0 ? M::NACC : 1];
#pragma unroll
    for (int e = 0; e < M::NACC; ++e) acc[e] = 0.0f;

    if (GRAD) {
        // ---- D + E: per band, the transposed passes into LDS (V, then the sub-pixel
        // adjoint values gb), then the recompute over the band's sub-pixels -------------
        float* gb = s_band + band::g_offset(g);
        __syncthreads();
        for (int i = tid; i < nband; i += BLOCK) s_band[i] = 0.0f;  // V margins
        __syncthreads();
        for (int y0 = 0; y0 < g.nsub; y0 += g.BY) {
            const int rows = g.BY < g.nsub - y0 ? g.BY : g.nsub - y0;
            band::vband_any(g, tid, BLOCK, y0, rows, G, fa, s_band);
            __syncthreads();
            band::gpass_any(g, tid, BLOCK, rows, s_band, fb, gb);
            __syncthreads();
            const int cnt = rows * g.nsub;
            int yy = tid / g.nsub, x = tid - yy * g.nsub;
            for (int s = tid; s < cnt; s += BLOCK) {
                const int y = y0 + yy;
                const float X = fmaf(L.M00, (float)x, fmaf(L.M01, (float)y, L.ox));
                const float Y = fmaf(L.M10, (float)x, fmaf(L.M11, (float)y, L.oy));
                eval_adjoint<M>(C, X, Y, gb[s], acc);
                x += BLOCK;
                while (x >= g.nsub) { x -= g.nsub; ++yy; }
            }
            // the next band's transposed vertical pass writes V only; its barrier orders
            // this band's reads of gb before the next gpass
        }
    }

    // ---- F ------------------------------------------------------------------
    __syncthreads();
    finish_particle<BLOCK, M, GRAD>(P, C, K, tid, pid, ll_acc, chi_acc, acc, glp, false, zin, m_in, v_in, s_x,
                                    s_red);
}

template <int BLOCK, class M, bool GRAD>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void band_kernel(const BArgs A) {
    extern __shared__ float4 smem4[];
    float* smem = reinterpret_cast<float*>(smem4);
    __shared__ int s_item;
    const int tid = threadIdx.x;
    float* slot = A.scratch + (long)blockIdx.x * A.slot;
    for (;;) {
        if (tid == 0) s_item = atomicAdd(A.queue, 1);
        __syncthreads();
        // every thread reads s_item before thread 0 writes it again: band_particle has barriers
        const int item = __builtin_amdgcn_readfirstlane(s_item);
        if (item >= A.total) break;  // the queue is drained: every wave of the workgroup leaves
        // the tables are read-only for the launch: addressed in the constant space, their
        // fields come in through scalar loads
        typedef const int __attribute__((address_space(4))) CInt;
        typedef const BLens __attribute__((address_space(4))) CLens;
        const CInt* pre = (const CInt*)A.pre;
        int lo = 0, hi = A.n;  // pre[lo] <= item < pre[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (pre[mid] <= item) lo = mid;
            else hi = mid;
        }
        const BLens* L = A.lens + ((const CInt*)A.order)[lo];
        const int pid = __builtin_amdgcn_readfirstlane(((const CLens*)A.lens)[((const CInt*)A.order)[lo]].first) +
                        item - pre[lo];
        band_particle<BLOCK, M, GRAD>(A, L, pid, smem + A.off_x, smem + A.off_c, smem + A.off_coef,
                                      smem + A.off_red, smem + A.off_band, slot, tid);
    }
    // every workgroup has taken its last item once the done count reaches the grid size:
    // the last one clears the queue for the next launch on the stream
    if (tid == 0 && atomicAdd(A.queue + 1, 1) == (int)gridDim.x - 1) {
        atomicExch(A.queue, 0);
        atomicExch(A.queue + 1, 0);
    }
}

}  // namespace gl
