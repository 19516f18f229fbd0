// Observation noise on the device: out = Poisson(max(img, 0) t) / t + sigma N(0, 1),
// the shot noise at exposure time t plus Gaussian background of rms sigma (the noise
// model whose variance the reference's likelihood uses, src/gigalens/tf/model.py:73-76).
// It makes the learned initialiser's training images without a host round trip.
//
// Draws per pixel i come from Philox4x32-10 (philox.h) with key = seed and counter
// (i low, i high, j, 0x5A17) for j = 0, 1, ...: block 0 gives the Gaussian (Box-Muller
// on its first two words) and the first two Poisson uniforms, later blocks four each.
// Poisson(lam):
//   lam < 10   inversion: sequential search of the CDF from P(0) = exp(-lam), FP64;
//   lam >= 10  transformed rejection with squeeze, PTRS (W. Hoermann, "The transformed
//              rejection method for generating Poisson random variables", Insurance:
//              Mathematics and Economics 12 (1993) 39-45), FP64.
#include <hip/hip_runtime.h>
#include <math.h>

#include <string>

#include "abi_common.h"
#include "philox.h"

namespace gl {
namespace noise {

struct Uniforms {
    unsigned long long key;
    unsigned i0, i1, j;
    rng::U4 buf;
    int k;
    __device__ double next() {
        if (k == 4) {
            buf = rng::philox4x32_10(rng::U4{i0, i1, j++, 0x5A17u}, key);
            k = 0;
        }
        const unsigned u = k == 0 ? buf.x : (k == 1 ? buf.y : (k == 2 ? buf.z : buf.w));
        ++k;
        return (double)rng::unit(u);
    }
};

__device__ double poisson_small(double lam, Uniforms& U) {
    const double u = U.next();
    double p = exp(-lam), F = p;
    int k = 0;
    while (u > F && k < 200) {
        ++k;
        p *= lam / k;
        F += p;
    }
    return (double)k;
}

__device__ double poisson_ptrs(double lam, Uniforms& U) {
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (int it = 0; it < 1000; ++it) {
        const double u = U.next() - 0.5, v = U.next();
        const double us = 0.5 - fabs(u);
        const double k = floor((2.0 * a / us + b) * u + lam + 0.43);
        if (us >= 0.07 && v <= vr) return k;
        if (k < 0.0 || (us < 0.013 && v > us)) continue;
        if (log(v) + log(inv_alpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0)) return k;
    }
    return floor(lam);  // not reached in practice (acceptance > 0.9 per trial)
}

__global__ void poisson_gauss(const float* img, long long n, float exp_time, float sigma, unsigned long long seed,
                              float* out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Uniforms U{seed, (unsigned)i, (unsigned)(i >> 32), 1u, rng::U4{0, 0, 0, 0}, 2};
    U.buf = rng::philox4x32_10(rng::U4{U.i0, U.i1, 0u, 0x5A17u}, seed);
    const float g = sqrtf(-2.0f * logf(rng::unit(U.buf.x))) * cosf(6.283185307179586f * rng::unit(U.buf.y));
    const float v = img[i];
    const double lam = v > 0.0f ? (double)v * exp_time : 0.0;
    double k = 0.0;
    if (lam > 0.0) k = lam < 10.0 ? poisson_small(lam, U) : poisson_ptrs(lam, U);
    out[i] = (float)(k / exp_time) + sigma * g;
}

}  // namespace noise
}  // namespace gl

using glc::fail;

extern "C" int gl_noise_poisson_gauss(const float* img, long long n, float exp_time, float background_rms,
                                      unsigned long long seed, float* out, hipStream_t stream) {
    if (n < 0 || (n > 0 && (!img || !out))) return fail(GL_ERR_ARG, "gl_noise_poisson_gauss: bad arguments");
    if (!(exp_time > 0.0f) || !(background_rms >= 0.0f))
        return fail(GL_ERR_ARG, "gl_noise_poisson_gauss: exp_time must be > 0 and background_rms >= 0");
    if (n == 0) return GL_OK;
    hipLaunchKernelGGL(gl::noise::poisson_gauss, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, img, n,
                       exp_time, background_rms, seed, out);
    GL_HIP(hipGetLastError());
    return GL_OK;
}
