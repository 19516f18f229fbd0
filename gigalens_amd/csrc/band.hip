// Banded persistent-kernel instantiations (band_kernels.hip) by profile
// combination, with their launch, LDS-attribute and occupancy functions.
#include "capi_internal.h"

namespace glc {

template <class M>
hipError_t band_launch(int grid, size_t lds, hipStream_t s, const gl::BArgs& A, bool grad) {
    if (grad)
        hipLaunchKernelGGL((gl::band_kernel<gl::kBandBlock, M, true>), dim3(grid), dim3(gl::kBandBlock), lds, s, A);
    else
        hipLaunchKernelGGL((gl::band_kernel<gl::kBandBlock, M, false>), dim3(grid), dim3(gl::kBandBlock), lds, s, A);
    return hipGetLastError();
}

// Allows `lds` bytes of dynamic LDS and returns the workgroups per CU of the gradient
// instance at that size (the grid of a persistent launch is CUs x that).
template <class M>
hipError_t band_attr(size_t lds, int* wg_per_cu) {
    const void* kg = reinterpret_cast<const void*>(&gl::band_kernel<gl::kBandBlock, M, true>);
    const void* kn = reinterpret_cast<const void*>(&gl::band_kernel<gl::kBandBlock, M, false>);
    hipError_t e = hipFuncSetAttribute(kg, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) e = hipFuncSetAttribute(kn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, kg, gl::kBandBlock, lds);
    return e;
}

#define GL_BAND_ENTRY(L0, L1, L2, A0, A1, S0, S1)                                 \
    {{L0, L1, L2, A0, A1, S0, S1},                                                \
     &band_launch<gl::ModelT<L0, L1, L2, A0, A1, S0, S1>>,                         \
     &band_attr<gl::ModelT<L0, L1, L2, A0, A1, S0, S1>>}

const BandEntry* band_models(int* n) {
    static const BandEntry table[] = {
        GL_BAND_ENTRY(EPL, SHR, 0, SE, 0, SE, 0),  // reference default model (tests/conftest.py:76-80)
        GL_BAND_ENTRY(EPL, SHR, 0, 0, 0, SE, 0),
    };
    *n = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}

}  // namespace glc
