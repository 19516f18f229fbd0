// Banded persistent particle kernel for gfx950 (MI355X): every particle of every
// lens system of a batch -- lenses of different image sizes, supersampling,
// PSFs and particle counts -- in one launch.
//
// The grid is as many workgroups as the CUs hold at once.  The (lens, particle)
// items, lenses in descending order of cost, are dealt round-robin over the
// workgroups (the longest items start first and every workgroup ends on short
// ones; hundreds of items per workgroup even out the rest), and each workgroup
// runs the whole log-probability (and gradient, and MAP step) of its items:
//
//   A  prologue   z -> x, log prior, component constants, EPL series tables (LDS)
//   B  ray-trace  per band of BY sub-rows: lens light(theta) + source(beta), NaN -> 0,
//                 into a zero-haloed LDS band
//   C  conv+pool  horizontal pass of each band into T (the workgroup's HBM slot),
//                 then the vertical pass, the Gaussian log-likelihood and dL/dimage
//                 (HBM slot) -- band_passes.h
//   D  adjoint    per band: the rows of dL/dimage it reaches staged in LDS, the transposed
//                 vertical pass into LDS, then the transposed horizontal pass to every
//                 sub-pixel of the band (LDS)
//   E  recompute  the ray-trace again per sub-pixel, accumulating the reverse-mode
//                 statistics
//   F  reduce     as the resident-image kernel: chain rule, lp / chi2 / gradient or
//                 the Keras-Adam update
//
// Only one band of the supersampled image is ever in LDS, so an image of any size
// runs at the LDS footprint of the band; the separable (rank <= 8) folded PSF keeps
// the HBM intermediates at R x nsub x N floats per particle, written and read by
// the same workgroup while they are in L2.  A launch keeps no state between
// calls: HIP graphs can record it.
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
    int off_x, off_c, off_coef, off_red, off_lens, off_band;  // LDS layout (floats)
};

// Copy of a lens-uniform struct of 32-bit words into scalar registers (one
// readfirstlane per word).
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
template <class T>
__device__ __forceinline__ T uniform_get(const T& src) {
    return uniform_copy(src);
}

// Sub-pixel coordinates: theta = transform (x, y) + origin (tf/simulator.py:19-29).
struct BXform {
    float M00, M01, M10, M11, ox, oy;
};
typedef const BLens __attribute__((address_space(4))) CLensT;
__device__ __forceinline__ BXform band_xform(const CLensT& L) {
    BXform t;
    t.M00 = L.P.M00; t.M01 = L.P.M01; t.M10 = L.P.M10; t.M11 = L.P.M11; t.ox = L.P.ox; t.oy = L.P.oy;
    return t;
}

// One particle on the workgroup.  The lens's table entry is copied into LDS (s_lens) and
// every phase reads the few lens values it uses from there into scalar registers, so
// they are not held across the phases: the per-sub-pixel loops keep the scalar
// registers for the profile constants.
template <int BLOCK, class M, bool GRAD>
__device__ __forceinline__ void band_particle(const BArgs& A, const BLens* Lp, int pid, float* s_x, float* s_c,
                                              float* s_coef, float* s_red, float* s_band, float* slot, int tid,
                                              BLens* s_lens) {
    // its previous reader, the last item's epilogue, ended at the barrier after it
    static_assert(sizeof(BLens) % 4 == 0, "lens entry copied as 32-bit words");
    for (int i = tid; i < (int)(sizeof(BLens) / 4); i += BLOCK)
        reinterpret_cast<int*>(s_lens)[i] = reinterpret_cast<const int*>(Lp)[i];
    __syncthreads();
    const DPlan& P = s_lens->P;  // prologue and epilogue
    const KArgs& K = A.K;
    const CLensT& Lc = *(const CLensT*)Lp;  // the same entry, read-only: scalar loads the compiler may repeat

    // ---- A: prologue -------------------------------------------------------
    const int d = uniform_get(P.d);
    float zin = 0.0f, m_in = 0.0f, v_in = 0.0f;
    if (tid < d) {
        const size_t o = (size_t)pid * d + tid;
        zin = K.in[o];
        if (K.mode == M_MAP) { m_in = K.m[o]; v_in = K.v[o]; }
    }
    float4* b4 = reinterpret_cast<float4*>(s_band);
    {
        const band::Geom g = uniform_get(s_lens->g);
        for (int i = tid; i < (band::band_floats(g) + 3) / 4; i += BLOCK)
            b4[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // halo columns of the forward bands
        band::zero_pads(g, tid, BLOCK, slot, slot + band::t_floats(g));
    }
    prologue_prior(P, true, tid, zin, s_x);
    __syncthreads();
    prologue_components(P, tid, s_x, s_c);
    __syncthreads();
    prologue_tables<M>(P, tid, s_c, s_coef);
    __syncthreads();
    Consts<M> C;
    load_consts<M>(s_c, s_coef, P, C);

    // ---- B + C1: per band, ray-trace then the horizontal pass ------------------
    for (int y0 = 0;;) {
        const int nsub = Lc.g.nsub, BY = Lc.g.BY;
        if (y0 >= nsub) break;
        const int rows = BY < nsub - y0 ? BY : nsub - y0;
        {
            const BXform t = band_xform(Lc);
            const int BW = Lc.g.BW, pb = Lc.g.pb;
            const int cnt = rows * nsub;
            int yy = tid / nsub, x = tid - yy * nsub;
            for (int s = tid; s < cnt; s += BLOCK) {
                const int y = y0 + yy;
                const float X = fmaf(t.M00, (float)x, fmaf(t.M01, (float)y, t.ox));
                const float Y = fmaf(t.M10, (float)x, fmaf(t.M11, (float)y, t.oy));
                float val = eval_forward<M>(C, X, Y);
                if (isnan(val)) val = 0.0f;
                s_band[yy * BW + pb + x] = val;
                x += BLOCK;
                while (x >= nsub) { x -= nsub; ++yy; }
            }
        }
        __syncthreads();
        {
            const band::Geom g = uniform_get(s_lens->g);
            const float* fb = uniform_get(s_lens->fac) + g.R * g.KFP;  // horizontal factors b_r
            band::hpass_any(g, tid, BLOCK, y0, rows, s_band, fb, slot);
        }
        __syncthreads();
        y0 += rows;
    }

    // ---- C2: vertical pass, likelihood, dL/dimage ---------------------------------
    float ll_acc = 0.0f, chi_acc = 0.0f;
    const float glp = K.mode == M_MAP ? uniform_get(s_lens->glp_map) : (K.g_lp ? K.g_lp[pid] : 1.0f);
    {
        const band::Geom g = uniform_get(s_lens->g);
        const float cf = uniform_get(P.cf), bkg2 = uniform_get(P.bkg2), inv_texp = uniform_get(P.inv_texp);
        const int lik = uniform_get(P.lik);
        const float* obs = uniform_get(P.obs);
        const float* ivar = uniform_get(P.inv_var_obs);
        const float* lsig = uniform_get(P.log_sig_obs);
        const float* fv = uniform_get(s_lens->fac) + 2 * g.R * g.KFP;  // a_r zero padded
        const float inv_np = 1.0f / (float)(g.N * g.N);
        const float gch = (K.mode == M_LOGPROB && K.g_chi2) ? K.g_chi2[pid] : 0.0f;
        float* Gi = slot + band::t_floats(g) + g.GT * g.N;  // row 0 of the image in the padded G
        band::vpass_any(g, tid, BLOCK, slot, fv, [&](int I, int J, float conv) {
            const float im = cf * conv;
            const int pix = I * g.N + J;
            float ll, r2, dll, dr2;
            if (lik == GL_LIK_FORWARD)
                glm::lik_forward(im, obs[pix], bkg2, inv_texp, ll, r2, dll, dr2);
            else
                glm::lik_backward(im, obs[pix], ivar[pix], lsig[pix], ll, r2, dll, dr2);
            ll_acc += ll;
            chi_acc += r2;
            if (GRAD) Gi[pix] = cf * (glp * dll + gch * inv_np * dr2);
        });
    }

    float acc[M::NACC > 0 ? M::NACC : 1];
#pragma unroll
    for (int e = 0; e < M::NACC; ++e) acc[e] = 0.0f;

    if (GRAD) {
        // ---- D + E: per band, the transposed passes into LDS (the rows of dL/dimage it
        // reaches, V, then the sub-pixel adjoint values gb), then the recompute -------------
        __syncthreads();
        {
            const band::Geom g = uniform_get(s_lens->g);
            for (int i = tid; i < (band::g_offset(g) + 3) / 4; i += BLOCK)
                b4[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // V margins
        }
        for (int y0 = 0;;) {
            const int nsub = Lc.g.nsub, BY = Lc.g.BY;
            if (y0 >= nsub) break;
            const int rows = BY < nsub - y0 ? BY : nsub - y0;
            float* gb;
            {
                const band::Geom g = uniform_get(s_lens->g);
                const float* fac = uniform_get(s_lens->fac);
                float* Gw = s_band + band::gw_offset(g);
                gb = s_band + band::g_offset(g);
                band::stage_g(g, tid, BLOCK, y0, slot + band::t_floats(g), Gw);
                __syncthreads();
                band::vband_any(g, tid, BLOCK, y0, rows, Gw, fac, s_band);
                __syncthreads();
                band::gpass_any(g, tid, BLOCK, rows, s_band, fac + g.R * g.KFP, gb);
                __syncthreads();
            }
            const BXform t = band_xform(Lc);
            const int cnt = rows * nsub;
            int yy = tid / nsub, x = tid - yy * nsub;
            for (int s = tid; s < cnt; s += BLOCK) {
                const int y = y0 + yy;
                const float X = fmaf(t.M00, (float)x, fmaf(t.M01, (float)y, t.ox));
                const float Y = fmaf(t.M10, (float)x, fmaf(t.M11, (float)y, t.oy));
                eval_adjoint<M>(C, X, Y, gb[s], acc);
                x += BLOCK;
                while (x >= nsub) { x -= nsub; ++yy; }
            }
            // the next band's staging writes Gw only and its transposed vertical pass V only;
            // the barrier between them orders this band's reads of gb before the next gpass
            y0 += rows;
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
    const int tid = threadIdx.x;
    float* slot = A.scratch + (long)blockIdx.x * A.slot;
    // items in processing order (lenses longest first), dealt round-robin in two halves
    // that run in opposite directions: even workgroups take the even items from the
    // front, odd ones the odd items from the back, so workgroups that share a CU are
    // seldom in the same phase of equally long items (their latency-bound phases
    // overlap the other's arithmetic) and each half still holds half of every lens.
    // The loop bound is the same for every thread of the workgroup, so every barrier of
    // an item is reached by all of its waves.
    typedef const int __attribute__((address_space(4))) CInt;
    const CInt* pre = (const CInt*)A.pre;
    const CInt* order = (const CInt*)A.order;
    const int split = gridDim.x > 1 ? 1 : 0;     // one workgroup takes every item in order
    const int half = split & blockIdx.x;
    const int nwg = split ? (gridDim.x + 1 - half) >> 1 : 1;  // workgroups of this half
    const int nitems = split ? (A.total + 1 - half) >> 1 : A.total;  // items of this half
    for (int k = blockIdx.x >> split; k < nitems; k += nwg) {
        const int item = !split ? k : half ? 2 * (nitems - 1 - k) + 1 : 2 * k;
        // the tables are read-only for the launch: addressed in the constant space, their
        // fields come in through scalar loads
        int lo = 0, hi = A.n;  // pre[lo] <= item < pre[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (pre[mid] <= item) lo = mid;
            else hi = mid;
        }
        const BLens* L = A.lens + order[lo];
        const int pid = ((const CLensT*)A.lens)[order[lo]].first + item - pre[lo];
        band_particle<BLOCK, M, GRAD>(A, L, pid, smem + A.off_x, smem + A.off_c, smem + A.off_coef,
                                      smem + A.off_red, smem + A.off_band, slot, tid,
                                      reinterpret_cast<BLens*>(smem + A.off_lens));
        __syncthreads();  // the next item's copy of its lens entry overwrites this one's
    }
}

}  // namespace gl
