// Preconditioned HMC with step-size and trajectory-length adaptation on the device
// (the kernel stack of src/gigalens/tf/inference.py:98-131: PreconditionedHamiltonian-
// MonteCarlo inside GradientBasedTrajectoryLengthAdaptation inside
// DualAveragingStepSizeAdaptation).  One iteration is a fixed sequence of launches --
// begin, max_leapfrog x (position, target log-density + gradient, momentum), end,
// adapt -- whose data-dependent trip count L lives in device memory: position and
// momentum updates (and, through gl_log_prob_masked, the target) do nothing for leapfrog
// steps k >= L.  So a whole iteration, adaptation included, can be recorded into a HIP
// graph once and replayed, with no host work or synchronisation.
//
// Momentum p ~ N(0, Sigma^-1) for the mass matrix Sigma (the surrogate covariance,
// :98-105): p = L^-T xi with Sigma = L L^T; kinetic energy p^T Sigma p / 2, so the
// position moves by eps Sigma p.  Normals and uniforms: Philox4x32-10 (Salmon et al.,
// SC 2011) keyed by the seed, Box-Muller for the normals.
//
// Adaptation during the first n_adapt iterations (80% of the burn-in, :112):
//   step size -- dual averaging (Hoffman & Gelman 2014, algorithm 5) on the mean
//     acceptance probability over chains, target 0.75, gamma 0.05, t0 10, kappa 0.75,
//     mu = log(10 eps0); after adaptation the averaged step size;
//   trajectory length T -- ChEES (Hoffman, Radul & Sountsov, AISTATS 2021): the
//     acceptance-weighted gradient with respect to T of E[(|z' - <z'>|^2 - |z - <z>|^2)^2] / 4
//     (chains with acceptance <= 1e-4 and non-finite terms left out), the proposal
//     moved along its velocity for the jittered time u T; an RMS-normalised step of
//     0.025 on log T (running squared gradient with weight 0.5, bias-corrected); T
//     at most max_leapfrog eps; log T iterate-averaged with weight m^-1/2 and
//     the average used after adaptation.  Each trajectory runs L = ceil(u T / eps)
//     steps, u the base-2 Halton sequence.
#include <hip/hip_runtime.h>
#include <math.h>

#include <string>

#include "abi_common.h"
#include "philox.h"

namespace gl {
namespace hmc {

// integer state: iteration, leapfrog steps of the current trajectory
enum { SI_ITER = 0, SI_L = 1, SI_N = 4 };
// float state
enum { SF_EPS = 0, SF_JITTER, SF_LOGT, SF_HBAR, SF_LOGEPSBAR, SF_MU, SF_SQGRAD, SF_LOGTBAR, SF_N = 8 };

using rng::U4;
using rng::unit;

__device__ __forceinline__ U4 philox(U4 c, unsigned long long seed) { return rng::philox4x32_10(c, seed); }

// base-2 radical inverse of i >= 1, in (0, 1)
__device__ __forceinline__ float halton2(unsigned i) {
    float f = 1.0f, r = 0.0f;
    while (i > 0) {
        f *= 0.5f;
        r += f * (float)(i & 1u);
        i >>= 1;
    }
    return r;
}

// One thread per chain: momentum draw, keep the start state, first half step.
__global__ void begin(int n, int d, const float* chol_inv, unsigned long long seed, const int* si, const float* sf,
                      const float* z, const float* lp, const float* g, float* p, float* z0, float* lp0, float* g0,
                      float* p0) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int it = si[SI_ITER];
    const float eps = sf[SF_EPS];
    float xi[4];
    for (int i = 0; i < d; ++i) p0[(size_t)c * d + i] = 0.0f;
    for (int k = 0; k < d; ++k) {
        if ((k & 3) == 0) {
            rng::normals4(philox(U4{(unsigned)c, (unsigned)(k >> 2), (unsigned)it, 0u}, seed), xi);
        }
        const float x = xi[k & 3];
        for (int i = 0; i < d; ++i) p0[(size_t)c * d + i] = fmaf(x, chol_inv[(size_t)k * d + i], p0[(size_t)c * d + i]);
    }
    for (int i = 0; i < d; ++i) {
        const size_t o = (size_t)c * d + i;
        z0[o] = z[o];
        g0[o] = g[o];
        p[o] = fmaf(0.5f * eps, g[o], p0[o]);
    }
    lp0[c] = lp[c];
}

// z += eps Sigma p for leapfrog step k < L; one thread per (chain, dimension).
__global__ void position(int n, int d, const float* cov, const int* si, const float* sf, int k, float* z,
                         const float* p) {
    if (k >= si[SI_L]) return;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n * d) return;
    const int c = (int)(t / d), i = (int)(t - (long)c * d);
    float v = 0.0f;
    for (int j = 0; j < d; ++j) v = fmaf(p[(size_t)c * d + j], cov[(size_t)j * d + i], v);
    z[t] = fmaf(sf[SF_EPS], v, z[t]);
}

// p += eps g (the last step of the trajectory: eps / 2) for leapfrog step k < L.
__global__ void momentum(int n, int d, const int* si, const float* sf, int k, float* p, const float* g) {
    const int L = si[SI_L];
    if (k >= L) return;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n * d) return;
    const float h = (k == L - 1 ? 0.5f : 1.0f) * sf[SF_EPS];
    p[t] = fmaf(h, g[t], p[t]);
}

// One thread per chain: Metropolis correction, keep the proposal and its velocity
// Sigma p for the trajectory-length gradient, store the sample after the burn-in.
__global__ void end(int n, int d, const float* cov, unsigned long long seed, const int* si, float* z, float* lp,
                    float* g, const float* p, const float* z0, const float* lp0, const float* g0, const float* p0,
                    float* zp, float* vel, float* acc, float* samples, int burnin, int n_results) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int it = si[SI_ITER];
    double k0 = 0.0, k1 = 0.0;
    for (int i = 0; i < d; ++i) {
        float v0 = 0.0f, v1 = 0.0f;
        for (int j = 0; j < d; ++j) {
            v0 = fmaf(p0[(size_t)c * d + j], cov[(size_t)j * d + i], v0);
            v1 = fmaf(p[(size_t)c * d + j], cov[(size_t)j * d + i], v1);
        }
        k0 += 0.5 * (double)v0 * p0[(size_t)c * d + i];
        k1 += 0.5 * (double)v1 * p[(size_t)c * d + i];
        vel[(size_t)c * d + i] = v1;
        zp[(size_t)c * d + i] = z[(size_t)c * d + i];
    }
    double la = ((double)lp[c] - k1) - ((double)lp0[c] - k0);
    if (!(la == la)) la = -INFINITY;  // NaN energy: reject
    if (la > 0.0) la = 0.0;
    const U4 r = philox(U4{(unsigned)c, 0x7FFFFFFFu, (unsigned)it, 1u}, seed);
    const bool accept = logf(unit(r.x)) < (float)la;
    acc[c] = (float)exp(la);
    if (!accept) {
        for (int i = 0; i < d; ++i) {
            const size_t o = (size_t)c * d + i;
            z[o] = z0[o];
            g[o] = g0[o];
        }
        lp[c] = lp0[c];
    }
    if (it >= burnin && it < burnin + n_results)
        for (int i = 0; i < d; ++i)
            samples[((size_t)(it - burnin) * n + c) * d + i] = z[(size_t)c * d + i];
}

__device__ double block_sum(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// One block of 256 threads: adaptation of the step size and trajectory length,
// histories (step size, mean acceptance, L and T used by this iteration), and the
// next trajectory's L.  adapt = 0: fixed eps and L.
__global__ __launch_bounds__(256) void adapt(int n, int d, int* si, float* sf, const float* z0, const float* zp,
                                             const float* vel, const float* acc, int adapt_on, int n_adapt,
                                             float target, int max_l, float* eps_hist, float* acc_hist,
                                             int* l_hist, float* t_hist, int hist_len) {
    __shared__ double red[256];
    const int it = si[SI_ITER];
    double sa = 0.0;
    for (int c = threadIdx.x; c < n; c += blockDim.x) sa += acc[c];
    const double acc_sum = block_sum(sa, red);
    const float mean_acc = (float)(acc_sum / n);
    const float eps_used = sf[SF_EPS], logt_used = sf[SF_LOGT];
    double grad = 0.0;
    if (adapt_on && it < n_adapt) {
        // ChEES gradient with respect to T: per chain (s1 - s0) (z' - <z'>) . v u, with
        // s0 = |z - <z>|^2, s1 = |z' - <z'>|^2, <z'> acceptance-weighted
        const double wsum = acc_sum + 1e-20;
        __shared__ float m0[128], m1[128];  // d <= 128
        for (int i = threadIdx.x; i < d; i += blockDim.x) {
            double a = 0.0, b = 0.0;
            for (int c = 0; c < n; ++c) {
                a += z0[(size_t)c * d + i];
                const float x = zp[(size_t)c * d + i];
                if (isfinite(x)) b += (double)acc[c] * x;
            }
            m0[i] = (float)(a / n);
            m1[i] = (float)(b / wsum);
        }
        __syncthreads();
        double gpart = 0.0;
        for (int c = threadIdx.x; c < n; c += blockDim.x) {
            if (!(acc[c] > 1e-4f)) continue;
            double s0 = 0.0, s1 = 0.0, dot = 0.0;
            for (int i = 0; i < d; ++i) {
                const double a = z0[(size_t)c * d + i] - m0[i], b = zp[(size_t)c * d + i] - m1[i];
                s0 += a * a;
                s1 += b * b;
                dot += b * vel[(size_t)c * d + i];
            }
            const double term = (double)acc[c] * (s1 - s0) * dot;
            if (isfinite(term)) gpart += term;
        }
        grad = block_sum(gpart, red) / wsum * sf[SF_JITTER];
    }
    if (threadIdx.x != 0) return;
    float eps = eps_used;
    if (adapt_on && it < n_adapt) {
        const float m = (float)(it + 1);
        // dual averaging (Hoffman & Gelman 2014, algorithm 5)
        const float t0 = 10.0f, gamma = 0.05f, kappa = 0.75f;
        const float hb = (1.0f - 1.0f / (m + t0)) * sf[SF_HBAR] + (target - mean_acc) / (m + t0);
        const float log_eps = sf[SF_MU] - sqrtf(m) / gamma * hb;
        const float w = powf(m, -kappa);
        sf[SF_HBAR] = hb;
        sf[SF_LOGEPSBAR] = w * log_eps + (1.0f - w) * sf[SF_LOGEPSBAR];
        eps = expf(log_eps);
        // trajectory length: RMS-normalised step on log T along the gradient in T
        float logt = logt_used;
        if (isfinite(grad)) {
            const double sq = 0.5 * (double)sf[SF_SQGRAD] + 0.5 * grad * grad;
            sf[SF_SQGRAD] = (float)sq;
            const double sq_hat = sq / (1.0 - pow(0.5, (double)m));
            logt = (float)(logt + 0.025 * grad / sqrt(sq_hat + 1e-20));
        }
        logt = fminf(logt, logf((float)max_l * eps));  // no lower bound: L >= 1 anyway
        const float wt = rsqrtf(m);
        sf[SF_LOGT] = logt;
        sf[SF_LOGTBAR] = wt * logt + (1.0f - wt) * sf[SF_LOGTBAR];
    } else if (adapt_on && it == n_adapt) {
        eps = expf(sf[SF_LOGEPSBAR]);
        sf[SF_LOGT] = sf[SF_LOGTBAR];
    }
    if (it < hist_len) {
        eps_hist[it] = eps_used;
        acc_hist[it] = mean_acc;
        l_hist[it] = si[SI_L];
        t_hist[it] = expf(logt_used);
    }
    sf[SF_EPS] = eps;
    si[SI_ITER] = it + 1;
    if (adapt_on) {
        const float u = halton2((unsigned)(it + 2));
        sf[SF_JITTER] = u;
        const int L = (int)ceilf(u * expf(sf[SF_LOGT]) / eps);
        si[SI_L] = L < 1 ? 1 : (L > max_l ? max_l : L);
    }
}

// p += h g, element-wise over n d
__global__ void kick(long nd, float h, const float* g, float* p) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nd) p[t] = fmaf(h, g[t], p[t]);
}

// z += eps Sigma p (Sigma = identity when cov is null); one thread per (chain, dimension)
__global__ void drift(int n, int d, const float* cov, float eps, const float* p, float* z) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n * d) return;
    const int c = (int)(t / d), i = (int)(t - (long)c * d);
    float v = 0.0f;
    if (cov) {
        for (int j = 0; j < d; ++j) v = fmaf(p[(size_t)c * d + j], cov[(size_t)j * d + i], v);
    } else {
        v = p[t];
    }
    z[t] = fmaf(eps, v, z[t]);
}

}  // namespace hmc
}  // namespace gl

using glc::fail;

namespace {
constexpr int kT = 256;
unsigned blocks(long n) { return (unsigned)((n + kT - 1) / kT); }
}  // namespace

extern "C" int gl_hmc_begin(int n, int d, const float* chol_inv, unsigned long long seed, const int* si,
                            const float* sf, const float* z, const float* lp, const float* g, float* p, float* z0,
                            float* lp0, float* g0, float* p0, hipStream_t stream) {
    if (n <= 0 || d <= 0 || d > 128) return fail(GL_ERR_ARG, "gl_hmc_begin: bad sizes (d <= 128)");
    hipLaunchKernelGGL(gl::hmc::begin, dim3(blocks(n)), dim3(kT), 0, stream, n, d, chol_inv, seed, si, sf, z, lp, g, p,
                       z0, lp0, g0, p0);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

extern "C" int gl_hmc_position(int n, int d, const float* cov, const int* si, const float* sf, int k, float* z,
                               const float* p, hipStream_t stream) {
    hipLaunchKernelGGL(gl::hmc::position, dim3(blocks((long)n * d)), dim3(kT), 0, stream, n, d, cov, si, sf, k, z, p);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

extern "C" int gl_hmc_momentum(int n, int d, const int* si, const float* sf, int k, float* p, const float* g,
                               hipStream_t stream) {
    hipLaunchKernelGGL(gl::hmc::momentum, dim3(blocks((long)n * d)), dim3(kT), 0, stream, n, d, si, sf, k, p, g);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

extern "C" int gl_hmc_end(int n, int d, const float* cov, unsigned long long seed, const int* si, float* z,
                          float* lp, float* g, const float* p, const float* z0, const float* lp0, const float* g0,
                          const float* p0, float* zp, float* vel, float* acc, float* samples, int burnin,
                          int n_results, hipStream_t stream) {
    hipLaunchKernelGGL(gl::hmc::end, dim3(blocks(n)), dim3(kT), 0, stream, n, d, cov, seed, si, z, lp, g, p, z0, lp0,
                       g0, p0, zp, vel, acc, samples, burnin, n_results);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

extern "C" int gl_hmc_adapt(int n, int d, int* si, float* sf, const float* z0, const float* zp, const float* vel,
                            const float* acc, int adapt, int n_adapt, float target, int max_leapfrog,
                            float* eps_hist, float* acc_hist, int* l_hist, float* t_hist, int hist_len,
                            hipStream_t stream) {
    if (d > 128) return fail(GL_ERR_ARG, "gl_hmc_adapt: d <= 128");
    hipLaunchKernelGGL(gl::hmc::adapt, dim3(1), dim3(256), 0, stream, n, d, si, sf, z0, zp, vel, acc, adapt, n_adapt,
                       target, max_leapfrog, eps_hist, acc_hist, l_hist, t_hist, hist_len);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

extern "C" int gl_leapfrog_step(const gl_plan* plan, int n, const float* cov, float eps, float* z, float* p,
                                float* lp, float* chi2, float* grad_z, hipStream_t stream) {
    if (!plan || !z || !p || !lp || !chi2 || !grad_z) return fail(GL_ERR_ARG, "gl_leapfrog_step: null argument");
    gl_plan_info info;
    int rc = gl_plan_query(plan, &info);
    if (rc) return rc;
    const int d = info.n_params;
    if (n <= 0 || d <= 0) return fail(GL_ERR_ARG, "gl_leapfrog_step: bad sizes");
    const long nd = (long)n * d;
    hipLaunchKernelGGL(gl::hmc::kick, dim3(blocks(nd)), dim3(kT), 0, stream, nd, 0.5f * eps, grad_z, p);
    GL_HIP(hipGetLastError());
    hipLaunchKernelGGL(gl::hmc::drift, dim3(blocks(nd)), dim3(kT), 0, stream, n, d, cov, eps, p, z);
    GL_HIP(hipGetLastError());
    if ((rc = gl_log_prob(plan, z, n, lp, chi2, nullptr, nullptr, grad_z, stream))) return rc;
    hipLaunchKernelGGL(gl::hmc::kick, dim3(blocks(nd)), dim3(kT), 0, stream, nd, 0.5f * eps, grad_z, p);
    GL_HIP(hipGetLastError());
    return GL_OK;
}
