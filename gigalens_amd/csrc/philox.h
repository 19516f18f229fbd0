// Counter-based random numbers for the device samplers: Philox4x32-10 (Salmon,
// Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC 2011),
// uniforms in (0, 1] from the top 24 bits, and the Box-Muller transform.  Every draw
// is a pure function of (key, counter), so a kernel's output does not depend on its
// launch shape or on scheduling.
#pragma once
#include <hip/hip_runtime.h>

namespace gl {
namespace rng {

struct U4 {
    unsigned x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(U4 c, unsigned long long key) {
    const unsigned M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    unsigned k0 = (unsigned)key, k1 = (unsigned)(key >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const unsigned long long p0 = (unsigned long long)M0 * c.x, p1 = (unsigned long long)M1 * c.z;
        c = U4{(unsigned)(p1 >> 32) ^ c.y ^ k0, (unsigned)p1, (unsigned)(p0 >> 32) ^ c.w ^ k1, (unsigned)p0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// (0, 1] from 32 random bits
__device__ __forceinline__ float unit(unsigned u) { return ((float)(u >> 8) + 1.0f) * (1.0f / 16777216.0f); }

// four standard normals from one Philox block
__device__ __forceinline__ void normals4(U4 r, float out[4]) {
    const float r0 = sqrtf(-2.0f * logf(unit(r.x))), r1 = sqrtf(-2.0f * logf(unit(r.z)));
    const float t0 = 6.283185307179586f * unit(r.y), t1 = 6.283185307179586f * unit(r.w);
    out[0] = r0 * cosf(t0);
    out[1] = r0 * sinf(t0);
    out[2] = r1 * cosf(t1);
    out[3] = r1 * sinf(t1);
}

}  // namespace rng
}  // namespace gl
