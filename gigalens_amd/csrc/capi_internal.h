// Host-side declarations shared by the translation units of the C ABI
// (capi.hip: plans, batches and entry points; fused_*.hip: the resident-image
// particle kernels by profile combination; staged.hip: the HBM-staged pipeline;
// cnn.hip: the learned initialiser; band.hip: the banded persistent kernel).  Kernels are launched only from the unit
// that instantiates them.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>

#include "abi_common.h"
#include "band_kernels.hip"
#include "lens_kernels.hip"
#include "lstsq_kernels.hip"

namespace glc {

constexpr int kBlock = 512;

using LaunchFn = hipError_t (*)(int grid, size_t lds, hipStream_t, const gl::DPlan&, const gl::KArgs&, bool grad);
using AttrFn = hipError_t (*)(size_t lds);

struct ModelEntry {
    int sig[7];  // lens0..2, lens light 0..1, source 0..1
    LaunchFn launch;
    AttrFn attr;
};

constexpr int EPL = glm::K_EPL, SHR = glm::K_SHEAR, SIS = glm::K_SIS, SIE = glm::K_SIE, TNFW = glm::K_TNFW,
              SE = glm::K_SERSIC_ELLIPSE, SC = glm::K_SERSIC, CSE = glm::K_CORE_SERSIC;

// Fused-kernel model tables, one per translation unit (fused_0/1/2.hip): the
// table and its length.
const ModelEntry* models_0(int* n);
const ModelEntry* models_1(int* n);
const ModelEntry* models_2(int* n);

// Banded persistent kernel (band.hip): launch, and LDS attribute + workgroups per CU.
using BandLaunchFn = hipError_t (*)(int grid, size_t lds, hipStream_t, const gl::BArgs&, bool grad);
using BandAttrFn = hipError_t (*)(size_t lds, int* wg_per_cu);
struct BandEntry {
    int sig[7];
    BandLaunchFn launch;
    BandAttrFn attr;
};
const BandEntry* band_models(int* n);

// Auxiliary kernels of the fused path (fused_0.hip).
hipError_t launch_epl_batch_series(const gl::DPlan& P, const float* in, int bs, int use_z, int* out, hipStream_t st);
hipError_t launch_order(const gl::DPlan& P, const float* in, int bs, int use_z, int* order, hipStream_t st);

// Conv launch group of a batch: the lenses of one (ss, kf4); the prefix table
// counts only their conv waves, so the group's launch runs compile-time ss with
// its own tap width.
struct ConvGroup {
    int ss, kf4, tot;
    const int* pre;
    int adjv_tot;         // separable-adjoint vertical-pass units of the group's lenses
    const int* adjv_pre;
    int adjh_tot;         // separable-adjoint horizontal-pass units of the group's lenses
    const int* adjh_pre;
    int seph_tot, sepv_tot;  // separable channel conv units of the group's separable lenses
    const int* seph_pre;
    const int* sepv_pre;
};

// One pass of the staged pipeline over the lenses of table T (staged.hip);
// multi: T is a lens batch addressed through device prefix tables.
int mc_pipeline_run(bool multi, const gl::McTable& T, const gl::McArgs& A, int ss_launch, int kf4, int lstsq,
                    int gram_nb, int ksplit_cap, size_t solve_lds, bool grad, hipStream_t st, bool shp, int model,
                    const std::vector<ConvGroup>* groups);
hipError_t mc_solve_attr(size_t lds);

}  // namespace glc
