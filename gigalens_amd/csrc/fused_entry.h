// Launch and attribute functions of one fused-kernel instantiation; included by
// the fused_*.hip translation units, each of which instantiates a few profile
// combinations of gl::particle_kernel.
#pragma once
#include "capi_internal.h"

namespace glc {

template <class M>
hipError_t launch_model(int grid, size_t lds, hipStream_t s, const gl::DPlan& P, const gl::KArgs& A, bool grad) {
    if (grad)
        hipLaunchKernelGGL((gl::particle_kernel<kBlock, M, true>), dim3(grid), dim3(kBlock), lds, s, P, A);
    else
        hipLaunchKernelGGL((gl::particle_kernel<kBlock, M, false>), dim3(grid), dim3(kBlock), lds, s, P, A);
    return hipGetLastError();
}

template <class M>
hipError_t attr_model(size_t lds) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&gl::particle_kernel<kBlock, M, true>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&gl::particle_kernel<kBlock, M, false>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

#define GL_ENTRY(L0, L1, L2, A0, A1, S0, S1)                                      \
    {{L0, L1, L2, A0, A1, S0, S1},                                                \
     &launch_model<gl::ModelT<L0, L1, L2, A0, A1, S0, S1>>,                        \
     &attr_model<gl::ModelT<L0, L1, L2, A0, A1, S0, S1>>}

}  // namespace glc
