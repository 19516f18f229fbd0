// Fused particle-kernel instantiations, part 0 of the profile-combination table
// (split over translation units so they compile in parallel).
#define GL_FUSED_AUX_KERNELS 1
#include "fused_entry.h"

namespace glc {

const ModelEntry* models_0(int* n) {
    static const ModelEntry table[] = {
        GL_ENTRY(EPL, SHR, 0, SE, 0, SE, 0),   // reference default model (tests/conftest.py:76-80)
        GL_ENTRY(EPL, 0, 0, SE, 0, SE, 0),
        GL_ENTRY(EPL, SHR, 0, 0, 0, SE, 0),
        GL_ENTRY(EPL, 0, 0, 0, 0, SE, 0),
    };
    *n = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}

hipError_t launch_epl_batch_series(const gl::DPlan& P, const float* in, int bs, int use_z, int* out, hipStream_t st) {
    hipLaunchKernelGGL(gl::epl_batch_series_kernel, dim3(1), dim3(256), 0, st, P, in, bs, use_z, out);
    return hipGetLastError();
}

hipError_t launch_order(const gl::DPlan& P, const float* in, int bs, int use_z, int* order, hipStream_t st) {
    hipLaunchKernelGGL(gl::order_kernel, dim3(1), dim3(1024), 0, st, P, in, bs, use_z, order);
    return hipGetLastError();
}

}  // namespace glc
