// Device-side plan shared by the kernels and the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include "gigalens_amd.h"

namespace gl {

constexpr int kMaxComp = GL_MAX_COMPONENTS;
constexpr int kConstStride = 32;  // floats of LDS per component constant block
constexpr int kMaxLens = 3;
constexpr int kMaxLight = 2;

struct DComp {
    int kind, role, niter, n_max, interpolate, lstsq, amp_col0, pad;
    int col[GL_MAX_COMPONENT_PARAMS];
    float cst[GL_MAX_COMPONENT_PARAMS];
};

// Passed by value as a kernel argument.
struct DPlan {
    int N, ss, nsub, k, kf, pb;
    float M00, M01, M10, M11, ox, oy, cf;
    int d, ncomp, nlens, nll, nsrc;
    int lik, fldj_at_x, batch_series;
    float bkg2, inv_texp;
    // LDS layout of the resident-image kernel (floats)
    int PW, sub_base, sub_size;  // padded sub-image: idx(y,x) = sub_base + (y+pb)*PW + x
    int GW, gp0, g_size;         // padded adjoint image: idx(I,J) = (I+gp0)*GW + (J+gp0)
    int adj_aligned;             // adjoint strip rows start 16-byte aligned (vector loads)
    int off_w, off_x, off_c, off_coef, off_red, lds_floats;
    int coef_off[kMaxLens];      // per lens slot: offset of its coefficient tables in the coef area
    int max_n;                   // max EPL series length supported by the coef tables
    const float* w;              // kf*kf folded PSF+pool weights (correlation form, incl. 1/ss^2)
    const float* obs;            // N*N observed image
    int n_obs;                   // observed images stacked in obs / inv_var_obs / log_sig_obs
    const float* inv_var_obs;    // N*N (backward likelihood)
    const float* log_sig_obs;    // N*N
    const int* prior_kind;       // d
    const float* prior_args;     // d*8
    const int* batch_n;          // device int per EPL slot when batch_series (else null)
    int sep_rank;                // > 0: separable PSF path with this many rank-1 terms
    int urows;                   // rows of the transposed-horizontal buffer U (separable adjoint)
    int ucol;                    // column stride of U (column-major, odd >= urows)
    int u_off;                   // LDS offset of U (after the adjoint image G, no overlap)
    int sep_aligned;             // the transposed-horizontal segments of G start 16-byte aligned
    const float* sa;             // sep_rank x kf vertical factors  (W ~= sum_r sa_r sb_r^T)
    const float* sb;             // sep_rank x kf horizontal factors
    DComp comp[kMaxComp];
};

}  // namespace gl
