// HBM-staged pipeline launches: the kernels of lstsq_kernels.hip instantiated per
// (supersampling, PSF tap width) and the stage sequence of one pass over a lens
// table (one plan, or a batch of lenses).
#include "capi_internal.h"

namespace glc {

template <int SS, int KF4, bool MULTI>
static hipError_t launch_mc_conv(int ib, int grid, hipStream_t st, const gl::McTable& T) {
    if (MULTI || ib == 4)
        hipLaunchKernelGGL((gl::mc_conv<SS, KF4, MULTI, 4>), dim3(grid), dim3(256), 0, st, T);
    else if (ib == 2)
        hipLaunchKernelGGL((gl::mc_conv<SS, KF4, MULTI, 2>), dim3(grid), dim3(256), 0, st, T);
    else if (ib == 8)
        hipLaunchKernelGGL((gl::mc_conv<SS, KF4, MULTI, 8>), dim3(grid), dim3(256), 0, st, T);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template <bool MULTI>
static hipError_t mc_conv_dispatch(int ss, int kf4, int ib, int grid, hipStream_t st, const gl::McTable& T) {
#define GL_MCC(S, K) if (ss == S && kf4 == K) return launch_mc_conv<S, K, MULTI>(ib, grid, st, T)
    GL_MCC(0, 8); GL_MCC(0, 16); GL_MCC(0, 22); GL_MCC(0, 32);
    GL_MCC(1, 8); GL_MCC(1, 16); GL_MCC(1, 22); GL_MCC(1, 32);
    GL_MCC(2, 8); GL_MCC(2, 16); GL_MCC(2, 22); GL_MCC(2, 32);
    GL_MCC(3, 8); GL_MCC(3, 16); GL_MCC(3, 22); GL_MCC(3, 32);
    GL_MCC(4, 8); GL_MCC(4, 16); GL_MCC(4, 22); GL_MCC(4, 32);
#undef GL_MCC
    return hipErrorInvalidValue;
}

template <bool MULTI>
static hipError_t mc_adjh_dispatch(int ss, int kf4, int grid, hipStream_t st, const gl::McTable& T) {
#define GL_MCH(S, K)                                                                              \
    if (ss == S && kf4 == K) {                                                                    \
        hipLaunchKernelGGL((gl::mc_adj_h<MULTI, S, K>), dim3(grid), dim3(256), 0, st, T);        \
        return hipGetLastError();                                                                 \
    }
    GL_MCH(1, 8); GL_MCH(1, 16); GL_MCH(1, 22); GL_MCH(1, 32);
    GL_MCH(2, 8); GL_MCH(2, 16); GL_MCH(2, 22); GL_MCH(2, 32);
    GL_MCH(3, 8); GL_MCH(3, 16); GL_MCH(3, 22); GL_MCH(3, 32);
    GL_MCH(4, 8); GL_MCH(4, 16); GL_MCH(4, 22); GL_MCH(4, 32);
#undef GL_MCH
    return hipErrorInvalidValue;
}

template <bool MULTI>
static hipError_t mc_adjv_dispatch(int ss, int kf4, int grid, hipStream_t st, const gl::McTable& T) {
#define GL_MCA(S, K)                                                                              \
    if (ss == S && kf4 == K) {                                                                    \
        hipLaunchKernelGGL((gl::mc_adj_v<MULTI, S, K>), dim3(grid), dim3(256), 0, st, T);        \
        return hipGetLastError();                                                                 \
    }
    GL_MCA(1, 8); GL_MCA(1, 16); GL_MCA(1, 22); GL_MCA(1, 32);
    GL_MCA(2, 8); GL_MCA(2, 16); GL_MCA(2, 22); GL_MCA(2, 32);
    GL_MCA(3, 8); GL_MCA(3, 16); GL_MCA(3, 22); GL_MCA(3, 32);
    GL_MCA(4, 8); GL_MCA(4, 16); GL_MCA(4, 22); GL_MCA(4, 32);
#undef GL_MCA
    return hipErrorInvalidValue;
}

// separable channel conv: T.tot_seph / T.tot_sepv waves (one lens, or one launch group of a batch)
template <bool MULTI, int SS, int KF4>
static hipError_t launch_mc_sep(hipStream_t st, const gl::McTable& T) {
    hipLaunchKernelGGL((gl::mc_sep_h<MULTI, SS, KF4>), dim3((T.tot_seph + 3) / 4), dim3(256), 0, st, T);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // a multiple of 8 workgroups (mc_sep_v's XCD-local block order), one per unit
    hipLaunchKernelGGL((gl::mc_sep_v<MULTI, SS, KF4>), dim3((T.tot_sepv + 7) & ~7), dim3(256), 0, st, T);
    return hipGetLastError();
}

template <bool MULTI>
static hipError_t mc_sep_dispatch(int ss, int kf4, hipStream_t st, const gl::McTable& T) {
#define GL_MCS(S, K) if (ss == S && kf4 == K) return launch_mc_sep<MULTI, S, K>(st, T)
    GL_MCS(1, 8); GL_MCS(1, 16); GL_MCS(1, 22); GL_MCS(1, 32);
    GL_MCS(2, 8); GL_MCS(2, 16); GL_MCS(2, 22); GL_MCS(2, 32);
    GL_MCS(3, 8); GL_MCS(3, 16); GL_MCS(3, 22); GL_MCS(3, 32);
    GL_MCS(4, 8); GL_MCS(4, 16); GL_MCS(4, 22); GL_MCS(4, 32);
#undef GL_MCS
    return hipErrorInvalidValue;
}

static hipError_t mc_gram_dispatch(int nbt, int nb, int ks, hipStream_t st, const gl::McTable& T) {
    switch (nbt) {
        case 1: hipLaunchKernelGGL((gl::mc_gram<1, false>), dim3(nb, ks), dim3(256), 0, st, T, ks); break;
        case 2: hipLaunchKernelGGL((gl::mc_gram<2, false>), dim3(nb, ks), dim3(256), 0, st, T, ks); break;
        case 3: hipLaunchKernelGGL((gl::mc_gram<3, false>), dim3(nb, ks), dim3(256), 0, st, T, ks); break;
        case 4: hipLaunchKernelGGL((gl::mc_gram<4, false>), dim3(nb, ks), dim3(256), 0, st, T, ks); break;
        case 5: hipLaunchKernelGGL((gl::mc_gram<5, false>), dim3(nb, ks), dim3(256), 0, st, T, ks); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// model 1: every lens is a fixed-amplitude plan of the reference default model
// (EPL + shear | Sersic ellipse | Sersic ellipse): model-specialised basis and gradient.
using DefaultModel = gl::ModelT<EPL, SHR, 0, SE, 0, SE, 0>;

template <bool MULTI>
static int mc_pipeline(const gl::McTable& T, const gl::McArgs& A, int ss_launch, int kf4, int lstsq, int gram_nb,
                       int ksplit_cap, size_t solve_lds, bool grad, hipStream_t st, bool shp = false, int model = 0,
                       const std::vector<ConvGroup>* groups = nullptr) {
    const int np = T.tot_part;
    hipLaunchKernelGGL((gl::mc_prologue<MULTI>), dim3(np), dim3(128), 0, st, T, A);
    GL_HIP(hipGetLastError());
    if (model == 1)
        hipLaunchKernelGGL((gl::mc_basis_m<MULTI, DefaultModel>), dim3(T.tot_basis), dim3(256), T.lds_consts, st, T);
    else
        hipLaunchKernelGGL((gl::mc_basis<MULTI>), dim3(T.tot_basis), dim3(256), T.lds_consts, st, T);
    GL_HIP(hipGetLastError());
    if (!MULTI && T.one.M.sep_rank > 0) {
        GL_HIP(mc_sep_dispatch<MULTI>(ss_launch, kf4, st, T));
    } else if (MULTI && groups && !groups->empty()) {
        for (const ConvGroup& g : *groups) {
            gl::McTable Tg = T;
            if (g.tot > 0) {
                Tg.pre_conv = g.pre;
                Tg.tot_conv = g.tot;
                GL_HIP(mc_conv_dispatch<MULTI>(g.ss, g.kf4, gl::kMcIB, (g.tot + 3) / 4, st, Tg));
            }
            if (g.seph_tot > 0) {
                Tg.pre_seph = g.seph_pre;
                Tg.tot_seph = g.seph_tot;
                Tg.pre_sepv = g.sepv_pre;
                Tg.tot_sepv = g.sepv_tot;
                GL_HIP(mc_sep_dispatch<MULTI>(g.ss, g.kf4, st, Tg));
            }
        }
    } else {
        GL_HIP(mc_conv_dispatch<MULTI>(ss_launch, kf4, MULTI ? 4 : T.one.M.conv_ib, (T.tot_conv + 3) / 4, st, T));
    }
    if (lstsq) {
        if (MULTI) return fail(GL_ERR_UNSUPPORTED, "least-squares lenses are not batched");
        const int ks = std::max(1, std::min(ksplit_cap, (1024 + np - 1) / np));
        GL_HIP(mc_gram_dispatch(gram_nb, np, ks, st, T));
        hipLaunchKernelGGL((gl::mc_solve<MULTI>), dim3(np), dim3(1024), solve_lds, st, T, T.one.M.rcond, 4 * ks);
        GL_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL((gl::mc_image<MULTI>), dim3(np), dim3(256), 0, st, T, A, grad ? 1 : 0);
    GL_HIP(hipGetLastError());
    if (grad && !MULTI && T.one.M.famp) {  // amplitude gradients from dL/dimage (single plans)
        hipLaunchKernelGGL(gl::mc_ampgrad<false>, dim3(np * T.one.M.D), dim3(256), 0, st, T);
        GL_HIP(hipGetLastError());
    }
    if (grad) {
        if (T.tot_adjh > 0) {
            // both passes: compile-time (ss, kf4), one launch each per group of lenses
            if (MULTI) {
                if (!groups) return fail(GL_ERR_ARG, "lens batch without launch groups");
                for (const ConvGroup& g : *groups) {
                    if (g.adjv_tot <= 0) continue;
                    gl::McTable Tg = T;
                    Tg.pre_adjh = g.adjh_pre;
                    Tg.tot_adjh = g.adjh_tot;
                    Tg.pre_adjv = g.adjv_pre;
                    Tg.tot_adjv = g.adjv_tot;
                    GL_HIP(mc_adjh_dispatch<MULTI>(g.ss, g.kf4, g.adjh_tot / 4, st, Tg));
                    GL_HIP(mc_adjv_dispatch<MULTI>(g.ss, g.kf4, g.adjv_tot / 4, st, Tg));
                }
            } else {
                GL_HIP(mc_adjh_dispatch<MULTI>(ss_launch, kf4, T.tot_adjh / 4, st, T));
                GL_HIP(mc_adjv_dispatch<MULTI>(ss_launch, kf4, T.tot_adjv / 4, st, T));
            }
        }
        if (T.tot_adj > 0)
            hipLaunchKernelGGL((gl::mc_adjoint<MULTI>), dim3(T.tot_adj / 4), dim3(256), T.lds_w, st, T);
        GL_HIP(hipGetLastError());
        if (model == 1)
            hipLaunchKernelGGL((gl::mc_grad_m<MULTI, DefaultModel>), dim3(T.tot_grad), dim3(256), T.lds_consts, st, T);
        else if (shp) {
            // EPL in slot 0, shear in slot 1, at most 4 components (the shapelet
            // configuration): the instance without the other lens kinds' states
            const gl::DPlan& P1 = T.one.P;
            const bool lean = !MULTI && P1.ncomp <= 4 && P1.nlens >= 1 && P1.nlens <= 2 &&
                              P1.comp[0].kind == glm::K_EPL && (P1.nlens < 2 || P1.comp[1].kind == glm::K_SHEAR);
            const bool shp4 = lean && P1.ncomp == 4 && P1.nlens == 2 && P1.comp[2].kind == glm::K_SERSIC_ELLIPSE &&
                              P1.comp[3].kind == glm::K_SHAPELETS;
            if (shp4)
                hipLaunchKernelGGL((gl::mc_grad_shp4<MULTI>), dim3(T.tot_grad), dim3(256), T.lds_consts, st, T);
            else if (lean)
                hipLaunchKernelGGL((gl::mc_grad_lean<MULTI, true>), dim3(T.tot_grad), dim3(256), T.lds_consts, st, T);
            else
                hipLaunchKernelGGL((gl::mc_grad<MULTI, true>), dim3(T.tot_grad), dim3(256), T.lds_consts, st, T);
        }
        else
            hipLaunchKernelGGL((gl::mc_grad<MULTI, false>), dim3(T.tot_grad), dim3(256), T.lds_consts, st, T);
        GL_HIP(hipGetLastError());
        hipLaunchKernelGGL((gl::mc_finish<MULTI>), dim3(np), dim3(128), 0, st, T, A);
        GL_HIP(hipGetLastError());
    }
    return GL_OK;
}

int mc_pipeline_run(bool multi, const gl::McTable& T, const gl::McArgs& A, int ss_launch, int kf4, int lstsq,
                    int gram_nb, int ksplit_cap, size_t solve_lds, bool grad, hipStream_t st, bool shp, int model,
                    const std::vector<ConvGroup>* groups) {
    if (multi)
        return mc_pipeline<true>(T, A, ss_launch, kf4, lstsq, gram_nb, ksplit_cap, solve_lds, grad, st, shp, model,
                                 groups);
    return mc_pipeline<false>(T, A, ss_launch, kf4, lstsq, gram_nb, ksplit_cap, solve_lds, grad, st, shp, model,
                              groups);
}

hipError_t mc_solve_attr(size_t lds) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&gl::mc_solve<false>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

}  // namespace glc
