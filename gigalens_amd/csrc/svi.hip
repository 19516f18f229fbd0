// Diagonal-Gaussian SVI of many lens systems on the device: reparameterised samples
// and the per-lens ELBO gradient + Keras-Adam update, so that a whole SVI step
// (sample -> batched log_prob and gradient -> update) is stream-ordered launches
// without host work and can be recorded into a HIP graph.
//
// The surrogate of lens L is N(loc_L, diag(sigma_L^2)) with sigma = exp(raw) + 1e-6
// (the Exp bijector and diag_shift of the reference's FillScaleTriL,
// src/gigalens/tf/inference.py:56-63, restricted to a diagonal).  The loss is the
// reference's reverse-KL estimate (tfp.vi.fit_surrogate_posterior,
// tf/inference.py:65-71): mean over the lens's samples of log q(z_s) - log p(z_s),
// with z_s = loc + sigma * eps_s.
//
// Normal draws: Philox4x32-10 (philox.h) keyed by the seed, counter (particle,
// dimension group, step, 0), then the Box-Muller transform.
#include <hip/hip_runtime.h>
#include <math.h>

#include <string>

#include "abi_common.h"
#include "philox.h"

namespace gl {
namespace svi {

// Four standard normals for (particle, dimension group g, step).
__device__ __forceinline__ void normals4(unsigned long long seed, unsigned particle, unsigned g, unsigned step,
                                         float out[4]) {
    rng::normals4(rng::philox4x32_10(rng::U4{particle, g, step, 0u}, seed), out);
}

constexpr float kShift = 1e-6f;  // diag_shift of the surrogate scale

// z[L n_vi + s][j] = loc[L][j] + sigma[L][j] eps, eps ~ N(0, 1); one thread per
// (particle, group of 4 dimensions).  step: per-lens step counters (lens L's samples
// use step[L]).
__global__ void sample(const float* loc, const float* raw, int n_lenses, int n_vi, int d, unsigned long long seed,
                       const int* step, float* z, float* eps) {
    const int groups = (d + 3) / 4;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n_lenses * n_vi * groups) return;
    const int part = (int)(t / groups), g = (int)(t - (long)part * groups);
    const int L = part / n_vi;
    float e[4];
    normals4(seed, (unsigned)part, (unsigned)g, (unsigned)step[L], e);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int j = 4 * g + k;
        if (j >= d) break;
        const float sig = expf(raw[(size_t)L * d + j]) + kShift;
        eps[(size_t)part * d + j] = e[k];
        z[(size_t)part * d + j] = loc[(size_t)L * d + j] + sig * e[k];
    }
}

// One workgroup per lens: the ELBO gradient of loc and raw from the lens's samples'
// log posterior and gradient, then Keras Adam (tf/inference.py:34; epsilon outside
// the square root) on both, step[L] advanced.  loss_hist (hist_len x n_lenses, may be
// null) receives the loss of step[L] when step[L] < hist_len.
__global__ __launch_bounds__(256) void update(const float* lp, const float* grad, const float* eps, int n_lenses,
                                              int n_vi, int d, float* loc, float* raw, float* m, float* v, int* step,
                                              float lr, float b1, float b2, float adam_eps, float* loss_hist,
                                              int hist_len) {
    const int L = blockIdx.x;
    const int t = step[L] + 1;
    const float lr_t = lr * sqrtf(1.0f - powf(b2, (float)t)) / (1.0f - powf(b1, (float)t));
    const float inv_n = 1.0f / (float)n_vi;
    const size_t P = (size_t)n_lenses * d;  // moments: [loc | raw]
    if (loss_hist) {
        // loss = mean_s [log q(z_s) - lp_s], log q = sum_j (-eps^2/2 - log sigma_j) - d/2 log 2 pi
        __shared__ double red[256];
        double acc = 0.0;
        for (int i = threadIdx.x; i < n_vi * d; i += blockDim.x) {
            const int s = i / d, j = i - (i / d) * d;
            const float e = eps[((size_t)L * n_vi + s) * d + j];
            acc += -0.5 * (double)e * e;
        }
        for (int j = threadIdx.x; j < d; j += blockDim.x)
            acc -= (double)n_vi * log((double)expf(raw[(size_t)L * d + j]) + kShift);
        for (int s = threadIdx.x; s < n_vi; s += blockDim.x) acc -= (double)lp[(size_t)L * n_vi + s];
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int w = blockDim.x / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0 && step[L] < hist_len)
            loss_hist[(size_t)step[L] * n_lenses + L] = (float)(red[0] * inv_n - 0.5 * d * 1.8378770664093453);
    }
    __syncthreads();  // the loss reads raw before the update
    for (int j = threadIdx.x; j < d; j += blockDim.x) {
        double sg = 0.0, sge = 0.0;
        for (int s = 0; s < n_vi; ++s) {
            const size_t i = ((size_t)L * n_vi + s) * d + j;
            sg += grad[i];
            sge += (double)grad[i] * eps[i];
        }
        const size_t k = (size_t)L * d + j;
        const float sig = expf(raw[k]) + kShift, ds = sig - kShift;  // d sigma / d raw
        const float g_loc = -(float)(sg * inv_n);
        const float g_raw = -((float)(sge * inv_n) * ds + ds / sig);
        const float gs[2] = {g_loc, g_raw};
        float* par[2] = {loc + k, raw + k};
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const size_t mi = q * P + k;
            const float mm = b1 * m[mi] + (1.0f - b1) * gs[q];
            const float vv = b2 * v[mi] + (1.0f - b2) * gs[q] * gs[q];
            m[mi] = mm;
            v[mi] = vv;
            *par[q] -= lr_t * mm / (sqrtf(vv) + adam_eps);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) step[L] = t;
}

}  // namespace svi
}  // namespace gl

using glc::fail;

extern "C" int gl_svi_diag_sample(const float* loc, const float* raw, int n_lenses, int n_vi, int d,
                                  unsigned long long seed, const int* step, float* z, float* eps,
                                  hipStream_t stream) {
    if (!loc || !raw || !step || !z || !eps) return fail(GL_ERR_ARG, "gl_svi_diag_sample: null argument");
    if (n_lenses <= 0 || n_vi <= 0 || d <= 0 || d > GL_MAX_PARAMS)
        return fail(GL_ERR_ARG, "gl_svi_diag_sample: bad sizes");
    const long threads = (long)n_lenses * n_vi * ((d + 3) / 4);
    hipLaunchKernelGGL(gl::svi::sample, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, stream, loc, raw,
                       n_lenses, n_vi, d, seed, step, z, eps);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

extern "C" int gl_svi_diag_update(const float* lp, const float* grad_z, const float* eps, int n_lenses, int n_vi,
                                  int d, float* loc, float* raw, float* m, float* v, int* step, float lr,
                                  float beta1, float beta2, float epsilon, float* loss_hist, int hist_len,
                                  hipStream_t stream) {
    if (!lp || !grad_z || !eps || !loc || !raw || !m || !v || !step)
        return fail(GL_ERR_ARG, "gl_svi_diag_update: null argument");
    if (n_lenses <= 0 || n_vi <= 0 || d <= 0 || d > GL_MAX_PARAMS)
        return fail(GL_ERR_ARG, "gl_svi_diag_update: bad sizes");
    hipLaunchKernelGGL(gl::svi::update, dim3(n_lenses), dim3(256), 0, stream, lp, grad_z, eps, n_lenses, n_vi, d, loc,
                       raw, m, v, step, lr, beta1, beta2, epsilon, loss_hist, hist_len);
    GL_HIP(hipGetLastError());
    return GL_OK;
}
