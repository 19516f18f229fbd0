// C ABI of the lens-model hot path: plan creation and launches.
// Declarations and reference citations: include/gigalens_amd.h.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "capi_internal.h"

using namespace glc;

namespace {
thread_local std::string g_err;
}  // namespace

int glc::fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

static const ModelEntry* find_model(const int sig[7]) {
    int ns[3];
    const ModelEntry* tabs[3] = {models_0(&ns[0]), models_1(&ns[1]), models_2(&ns[2])};
    for (int t = 0; t < 3; ++t)
        for (int i = 0; i < ns[t]; ++i) {
            bool ok = true;
            for (int j = 0; j < 7; ++j) ok = ok && tabs[t][i].sig[j] == sig[j];
            if (ok) return &tabs[t][i];
        }
    return nullptr;
}

struct gl_plan {
    gl::DPlan P;
    const ModelEntry* model;
    int mc_model;            // staged fixed-amplitude plan of the reference default model: 1 (specialised kernels)
    size_t lds_bytes;
    int epl_batch;
    int* order_buf;          // fused plans with an EPL lens: launch order of up to gl::kOrderCap particles
    int mc;                  // 1: multi-channel least-squares plan
    gl::McPlan M;
    int max_batch;
    int kf4;                 // conv template parameter of the multi-channel path
    size_t solve_lds;
    int ksplit_cap;          // Gram pixel-range splits allowed by the image size
    std::vector<float> W_host;  // folded PSF (kf x kf), for batches
    int sig[7];              // profile kinds: lens 0..2, lens light 0..1, source 0..1
    std::vector<void*> dev;  // owned device buffers
    mutable StreamGuard guard;  // orders calls across streams (the scratch above is shared)
};

extern "C" const char* gl_last_error(void) { return g_err.c_str(); }

static int upload(gl_plan* p, const void* host, size_t bytes, void** out) {
    void* d = nullptr;
    GL_HIP(hipMalloc(&d, bytes ? bytes : 4));
    p->dev.push_back(d);
    if (bytes) GL_HIP(hipMemcpy(d, host, bytes, hipMemcpyHostToDevice));
    *out = d;
    return GL_OK;
}

static int n_params_of(int kind) {
    switch (kind) {
    case glm::K_EPL: return 6;
    case glm::K_SHEAR: return 2;
    case glm::K_SIS: return 3;
    case glm::K_SIE: return 5;
    case glm::K_TNFW: return 5;
    case glm::K_SERSIC: return 5;
    case glm::K_SERSIC_ELLIPSE: return 7;
    case glm::K_CORE_SERSIC: return 10;
    case glm::K_SHAPELETS: return 3;
    default: return -1;
    }
}

// One-sided Jacobi SVD of a small n x n matrix (row-major, double):
// A = U diag(s) V^T.  Returns s (descending), U columns, V columns.
static void jacobi_svd(int n, const std::vector<double>& A, std::vector<double>& s, std::vector<double>& U,
                       std::vector<double>& V) {
    std::vector<double> W(A);  // columns of W converge to U * s
    V.assign(n * n, 0.0);
    for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double a = 0, b = 0, c = 0;
                for (int i = 0; i < n; ++i) {
                    a += W[i * n + p] * W[i * n + p];
                    b += W[i * n + q] * W[i * n + q];
                    c += W[i * n + p] * W[i * n + q];
                }
                if (fabs(c) <= 1e-300 || fabs(c) <= 1e-17 * sqrt(a * b)) continue;
                off = fmax(off, fabs(c) / sqrt(a * b));
                double zeta = (b - a) / (2.0 * c);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
                for (int i = 0; i < n; ++i) {
                    double wp = W[i * n + p], wq = W[i * n + q];
                    W[i * n + p] = cs * wp - sn * wq;
                    W[i * n + q] = sn * wp + cs * wq;
                    double vp = V[i * n + p], vq = V[i * n + q];
                    V[i * n + p] = cs * vp - sn * vq;
                    V[i * n + q] = sn * vp + cs * vq;
                }
            }
        if (off < 1e-15) break;
    }
    std::vector<int> order(n);
    std::vector<double> norms(n);
    for (int j = 0; j < n; ++j) {
        double acc = 0;
        for (int i = 0; i < n; ++i) acc += W[i * n + j] * W[i * n + j];
        norms[j] = sqrt(acc);
        order[j] = j;
    }
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (norms[order[j]] > norms[order[i]]) std::swap(order[i], order[j]);
    s.assign(n, 0.0);
    U.assign(n * n, 0.0);
    std::vector<double> V2(n * n, 0.0);
    for (int k = 0; k < n; ++k) {
        int j = order[k];
        s[k] = norms[j];
        for (int i = 0; i < n; ++i) {
            U[i * n + k] = norms[j] > 0 ? W[i * n + j] / norms[j] : 0.0;
            V2[i * n + k] = V[i * n + j];
        }
    }
    V.swap(V2);
}

extern "C" int gl_plan_destroy(gl_plan* plan) {
    if (!plan) return GL_OK;
    for (void* d : plan->dev) (void)hipFree(d);
    plan->guard.release();
    delete plan;
    return GL_OK;
}

// ---------------------------------------------------------------------------
// HBM-staged plans: least squares (linear amplitudes) and fixed amplitudes
// for images / models the LDS-resident kernel does not take
// ---------------------------------------------------------------------------
static int upload_common(const gl_plan_desc* desc, gl_plan* p);

static int pick_kf4(int kf) {
    const int need = (kf + 3) / 4;
    for (int c : {8, 16, 22, 32})
        if (need <= c) return c;
    return -1;
}

// Orthonormal Hermite functions on linspace(-5, 5, 6000) (lenstronomy Shapelets.phi_n,
// shapelets.py:39-40) by the stable three-term recurrence in FP64.
static void hermite_table(int n_max, std::vector<float>& tab) {
    const int T = glm::kShapeletTable;
    tab.assign((size_t)(n_max + 1) * T, 0.0f);
    for (int i = 0; i < T; ++i) {
        const double x = -5.0 + 10.0 * i / (T - 1);
        double pm2 = 0.0, pm1 = pow(M_PI, -0.25) * exp(-0.5 * x * x);
        tab[i] = (float)pm1;
        for (int n = 1; n <= n_max; ++n) {
            const double pn = sqrt(2.0 / n) * x * pm1 - sqrt((n - 1.0) / n) * pm2;
            tab[(size_t)n * T + i] = (float)pn;
            pm2 = pm1;
            pm1 = pn;
        }
    }
}

static int dev_alloc(std::vector<void*>& owned, size_t bytes, void** out, bool zero) {
    void* d = nullptr;
    GL_HIP(hipMalloc(&d, bytes ? bytes : 4));
    owned.push_back(d);
    if (zero) GL_HIP(hipMemset(d, 0, bytes ? bytes : 4));
    *out = d;
    return GL_OK;
}

// Where the scratch buffers of a staged workspace go: device allocations owned by a
// plan or batch (zeroed when asked), or consecutive 256-byte aligned slices of a
// shared arena.  Arena memory is not cleared: the kernels write every element they
// later read (the conv's column margins and the padded adjoint image's margins
// included), so lenses of different shapes may use it one after the other.
struct ScratchAlloc {
    std::vector<void*>* owned = nullptr;
    char* arena = nullptr;
    size_t off = 0, cap = 0;
    bool measure = false;  // count the bytes only
};
static int scratch_alloc(ScratchAlloc& a, size_t bytes, void** out, bool zero) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (a.measure) {
        a.off += bytes;
        *out = nullptr;
        return GL_OK;
    }
    if (a.arena) {
        if (a.off + bytes > a.cap) return fail(GL_ERR_ARG, "staged workspace exceeds its arena");
        *out = a.arena + a.off;
        a.off += bytes;
        return GL_OK;
    }
    return dev_alloc(*a.owned, bytes, out, zero);
}

// Constants of a staged lens whose conv is launched at template width kf4 (>= the
// lens's own): the staged row width BW and the zero-padded weight rows follow it.
// W: the lens's folded weights (kf x kf, host).
static int mc_consts_upload(const gl::DPlan& P, gl::McPlan& M, int kf4, const std::vector<float>& W,
                            std::vector<void*>& owned) {
    const int N = P.N, ss = P.ss, nsub = P.nsub, kf = P.kf, pb = P.pb;
    int rc;
    void* dptr = nullptr;
    M.BW = ((ss * (N - 1) + 4 * kf4 > pb + nsub ? ss * (N - 1) + 4 * kf4 : pb + nsub) + 3) & ~3;
    M.wpad_w = 4 * kf4;
    std::vector<float> wp((size_t)kf * M.wpad_w, 0.0f);
    for (int a = 0; a < kf; ++a)
        for (int b = 0; b < kf; ++b) wp[(size_t)a * M.wpad_w + b] = W[a * kf + b];
    if ((rc = dev_alloc(owned, wp.size() * sizeof(float), &dptr, false))) return rc;
    GL_HIP(hipMemcpy(dptr, wp.data(), wp.size() * sizeof(float), hipMemcpyHostToDevice));
    M.wpad = (const float*)dptr;
    return GL_OK;
}

// Scratch of a staged lens for `mb` particles (after mc_consts_upload).
static int mc_scratch(const gl::DPlan& P, gl::McPlan& M, int mb, int ksplit_cap, ScratchAlloc& sa) {
    const int N = P.N, ss = P.ss, nsub = P.nsub, D = M.D;
    int rc;
    void* dptr = nullptr;
    const int GW = N + 2 * M.gp;
    const int gram_nt = M.gram_nb * (M.gram_nb + 1) / 2;
    if ((rc = scratch_alloc(sa, (size_t)mb * M.pstride * 4, &dptr, true))) return rc;
    M.consts = (float*)dptr;
    if ((rc = scratch_alloc(sa, (size_t)mb * D * nsub * M.BW * 4, &dptr, true))) return rc;
    M.B = (float*)dptr;
    if ((rc = scratch_alloc(sa, (size_t)mb * D * N * N * 4, &dptr, false))) return rc;
    M.X = (float*)dptr;
    M.gpart = nullptr;
    if (M.lstsq) {
        if ((rc = scratch_alloc(sa, (size_t)std::min(mb * ksplit_cap, 1024 + mb) * 4 * gram_nt * 256 * 8, &dptr,
                                false)))
            return rc;
        M.gpart = (double*)dptr;
    }
    if ((rc = scratch_alloc(sa, (size_t)mb * D * 4, &dptr, false))) return rc;
    M.coef = (float*)dptr;
    if ((rc = scratch_alloc(sa, (size_t)mb * GW * GW * 4, &dptr, true))) return rc;
    M.gimg = (float*)dptr;
    if ((rc = scratch_alloc(sa, (size_t)mb * gl::kMcStat * 4, &dptr, true))) return rc;
    M.stats = (float*)dptr;
    M.gamp = nullptr;
    if (M.famp) {
        if ((rc = scratch_alloc(sa, (size_t)mb * D * 4, &dptr, true))) return rc;
        M.gamp = (float*)dptr;
    }
    M.U = nullptr;
    if (M.adj_rank > 0) {
        if ((rc = scratch_alloc(sa, (size_t)mb * M.adj_rank * N * ss * N * 4, &dptr, false))) return rc;
        M.U = (float*)dptr;
    }
    M.H = nullptr;
    if (M.sep_rank > 0) {  // separable channel conv: horizontal-pass intermediate
        if ((rc = scratch_alloc(sa, (size_t)mb * D * M.sep_rank * nsub * N * 4, &dptr, false))) return rc;
        M.H = (float*)dptr;
    }
    return GL_OK;
}

static int mc_create(const gl_plan_desc* desc, gl_plan* p, const std::vector<float>& W, int max_niter, int lstsq,
                     int famp) {
    gl::DPlan& P = p->P;
    gl::McPlan& M = p->M;
    memset(&M, 0, sizeof(M));
    M.lstsq = lstsq;
    M.famp = famp;
    M.conv_ib = desc->conv_rows > 0 ? desc->conv_rows : gl::kMcIB;
    if (M.conv_ib != 2 && M.conv_ib != 4 && M.conv_ib != 8)
        return fail(GL_ERR_ARG, "conv_rows must be 0, 2, 4 or 8");
    const int N = P.N, ss = P.ss, nsub = P.nsub, kf = P.kf, pb = P.pb;
    int rc;
    void* dptr = nullptr;
    if (ss > 4) return fail(GL_ERR_UNSUPPORTED, "the staged path supports supersample <= 4");
    p->kf4 = pick_kf4(kf);
    if (p->kf4 < 0) return fail(GL_ERR_UNSUPPORTED, "PSF too large for the staged path (kf > 128)");
    if (kf * kf * 4 > 64 * 1024) return fail(GL_ERR_UNSUPPORTED, "folded PSF too large for the staged adjoint (kf > 128)");
    if (P.ncomp > gl::kMaxComp) return fail(GL_ERR_UNSUPPORTED, "too many components");
    // channels
    int D = 0, nshp = 0;
    M.shp_comp = -1;
    for (int i = P.nlens; i < P.ncomp; ++i) {
        const gl::DComp& c = P.comp[i];
        const bool sersic = c.kind == glm::K_SERSIC || c.kind == glm::K_SERSIC_ELLIPSE || c.kind == glm::K_CORE_SERSIC;
        if (!lstsq && !famp) {
            if (!sersic) return fail(GL_ERR_UNSUPPORTED, "fixed-amplitude staged plans take Sersic-family lights");
            continue;
        }
        if (famp && !sersic && c.kind != glm::K_SHAPELETS)
            return fail(GL_ERR_UNSUPPORTED, "plans with a fixed-amplitude shapelet light take Sersic and shapelet lights");
        if (c.kind == glm::K_SHAPELETS) {
            if (++nshp > 1) return fail(GL_ERR_UNSUPPORTED, "at most one shapelet component");
            if (c.n_max < 0 || c.n_max > glm::kMaxShapeletN)
                return fail(GL_ERR_UNSUPPORTED, "shapelet n_max must be in [0, " + std::to_string(glm::kMaxShapeletN) + "]");
            M.shp_comp = i;
            M.shp_nmax = c.n_max;
            M.shp_interp = c.interpolate;
            M.shp_chan0 = D;
            if (famp && (c.amp_col0 < 0 || c.amp_col0 + (c.n_max + 1) * (c.n_max + 2) / 2 > P.d))
                return fail(GL_ERR_ARG, "fixed-amplitude shapelets need amp_col0 with every amplitude column in range");
            D += (c.n_max + 1) * (c.n_max + 2) / 2;
        } else {
            D += 1;
        }
    }
    if (!lstsq && !famp) D = 1;
    if (D < 1 || D > gl::kMcMaxChan) return fail(GL_ERR_UNSUPPORTED, "number of linear channels out of range");
    M.D = D;
    // per-particle constants: prior arrays, component blocks, EPL coefficient tables
    P.max_n = max_niter;
    const int coef_block = glm::epl_tab_floats(max_niter);
    const int prior_f = (6 * P.d + 3) & ~3;
    M.coef_base = prior_f + gl::kMaxComp * gl::kConstStride;
    int coff = 0;
    for (int i = 0; i < 3; ++i) {
        P.coef_off[i] = coff;
        if (i < P.nlens && P.comp[i].kind == glm::K_EPL) coff += (coef_block + 3) & ~3;
    }
    M.pstride = (M.coef_base + coff + 63) & ~63;
    // padded dL/dimage
    int gp = 0;
    for (int a = 0; a < ss; ++a) {
        const int ra = (a + pb) % ss, da = (a + pb) / ss, Ma = (kf - ra + ss - 1) / ss;
        if (Ma - 1 - da > gp) gp = Ma - 1 - da;
        if (da > gp) gp = da;
    }
    // mc_adj_h loads a window of kfm + 1 columns per lane: margin >= kfm + 1
    const int kfm_pad = (4 * p->kf4 + ss - 1) / ss;
    if (gp < kfm_pad + 1) gp = kfm_pad + 1;
    M.gp = gp;
    const int GW = N + 2 * gp;
    // Gram tiles and pixel-range splits
    M.gram_nb = lstsq ? (D + 1 + 15) / 16 : 0;
    const int gram_nt = M.gram_nb * (M.gram_nb + 1) / 2;
    const int nchunk = (N * N + 63) / 64;
    p->ksplit_cap = std::max(1, std::min(gl::kMcGramMaxSplit, nchunk / 4));
    // separable channel convolution (mc_sep_h + mc_sep_v, single plans and lens
    // batches): when the folded PSF has rank <= 8 at the tolerance (0 selects 1e-6;
    // psf_rank_tol < 0 keeps the direct conv) and the flop count is lower.
    // Measured: 100x100 ss 4 with the 83x83 Moffat (rank 7), 52 ms against 104 ms of
    // direct conv per 512 particles.
    M.sep_rank = 0;
    std::vector<float> sa, sbp;
    std::vector<double> Wd(kf * kf), sv, Uv, Vv;
    for (int i = 0; i < kf * kf; ++i) Wd[i] = W[i];
    jacobi_svd(kf, Wd, sv, Uv, Vv);
    // separable adjoint conv (default; psf_rank_tol < 0 keeps the direct form): rank at
    // 1e-6 (or the given tolerance), at most 8, when its two passes cost under 0.8 of the
    // direct taps (measured: rank 1 at 100x100 ss 4 cuts the adjoint 5x, rank 7 of the
    // 83x83 Moffat 6x; rank 5 at ss 3 with a 35x35 folded PSF, flop ratio 0.56, 2x)
    M.adj_rank = 0;
    std::vector<float> asa, asb;
    if (desc->psf_rank_tol >= 0.0f && sv.size() && sv[0] > 0.0) {
        const double tol = desc->psf_rank_tol > 0.0f ? desc->psf_rank_tol : 1e-6;
        int R = 0;
        while (R < kf && sv[R] >= tol * sv[0]) ++R;
        const double mt = (double)(kf + ss - 1) / ss;  // taps per phase, rounded up
        const double direct = (double)nsub * nsub * mt * mt;
        const double sep = (double)R * ((double)N * nsub * mt + (double)nsub * nsub * mt);
        if (R >= 1 && R <= gl::kMcMaxRank && sep < 0.8 * direct) {
            M.adj_rank = R;
            asa.assign((size_t)R * kf, 0.0f);
            asb.assign((size_t)R * kf, 0.0f);
            for (int r = 0; r < R; ++r)
                for (int i = 0; i < kf; ++i) {
                    asa[(size_t)r * kf + i] = (float)(sv[r] * Uv[i * kf + r]);
                    asb[(size_t)r * kf + i] = (float)Vv[i * kf + r];
                }
        }
    }
    if (desc->psf_rank_tol >= 0.0f) {
        const double tol = desc->psf_rank_tol > 0.0f ? desc->psf_rank_tol : 1e-6;
        int R = 0;
        while (R < kf && sv[R] >= tol * sv[0]) ++R;
        const double ratio = (double)R * (nsub + N) / ((double)N * kf);
        if (R >= 1 && R <= gl::kMcMaxRank && ratio < 0.8) {
            M.sep_rank = R;
            const int wpw = 4 * p->kf4;
            sa.assign((size_t)R * kf, 0.0f);
            sbp.assign((size_t)R * wpw, 0.0f);
            for (int r = 0; r < R; ++r)
                for (int i = 0; i < kf; ++i) {
                    sa[(size_t)r * kf + i] = (float)(sv[r] * Uv[i * kf + r]);
                    sbp[(size_t)r * wpw + i] = (float)Vv[i * kf + r];
                }
        }
    }
    const int BW = ((ss * (N - 1) + 4 * p->kf4 > pb + nsub ? ss * (N - 1) + 4 * p->kf4 : pb + nsub) + 3) & ~3;
    const double per_particle = 4.0 * ((double)M.pstride + (double)D * nsub * BW + (double)D * N * N + D +
                                       (double)GW * GW + gl::kMcStat +
                                       (double)D * M.sep_rank * nsub * N + (double)M.adj_rank * N * nsub) +
                                8.0 * 4 * p->ksplit_cap * gram_nt * 256;
    // particles per chunk: the caller's max_batch, at most what a 64 GiB workspace holds;
    // by default as many as fit 32 GiB, at most 1024
    int mb = desc->max_batch;
    if (mb <= 0) mb = (int)std::max(1.0, std::min(1024.0, 32.0 * 1073741824.0 / per_particle));
    else mb = (int)std::max(1.0, std::min((double)mb, 64.0 * 1073741824.0 / per_particle));
    p->max_batch = mb;
    M.rcond = 1e-6f;
    if ((rc = upload(p, W.data(), W.size() * sizeof(float), &dptr))) return rc;
    P.w = (const float*)dptr;
    p->W_host = W;
    if ((rc = upload_common(desc, p))) return rc;
    if (lstsq) {
        // W = 1 / err and obs / err (tf/simulator.py:157-158), err from the noise map or tf/model.py:113-115
        std::vector<float> sw(N * N, 0.0f), swo(N * N, 0.0f);
        for (int i = 0; i < N * N; ++i) {
            const double o = desc->observed ? desc->observed[i] : 0.0;
            const double var = desc->noise_map ? (double)desc->noise_map[i] * desc->noise_map[i]
                                               : (double)P.bkg2 + std::max(o, 0.0) / desc->exp_time;
            const float w = (float)(1.0 / sqrt(var));
            sw[i] = w;
            swo[i] = (float)o * w;
        }
        if ((rc = upload(p, sw.data(), sw.size() * sizeof(float), &dptr))) return rc;
        M.sw = (const float*)dptr;
        if ((rc = upload(p, swo.data(), swo.size() * sizeof(float), &dptr))) return rc;
        M.swobs = (const float*)dptr;
    }
    // shapelet tables
    if (M.shp_comp >= 0) {
        std::vector<float> tab;
        hermite_table(M.shp_nmax, tab);
        if ((rc = upload(p, tab.data(), tab.size() * sizeof(float), &dptr))) return rc;
        M.shp_table = (const float*)dptr;
        std::vector<float> pref(M.shp_nmax + 2);
        for (int n = 0; n <= M.shp_nmax + 1; ++n)
            pref[n] = (float)(1.0 / sqrt(pow(2.0, n) * sqrt(M_PI) * tgamma(n + 1.0)));
        if ((rc = upload(p, pref.data(), pref.size() * sizeof(float), &dptr))) return rc;
        M.shp_pref = (const float*)dptr;
    }
    // workspace for max_batch particles (allocated here once; calls never allocate)
    if ((rc = mc_consts_upload(P, M, p->kf4, W, p->dev))) return rc;
    {
        ScratchAlloc sa;
        sa.owned = &p->dev;
        if ((rc = mc_scratch(P, M, mb, p->ksplit_cap, sa))) return rc;
    }
    if (M.adj_rank > 0) {
        if ((rc = upload(p, asa.data(), asa.size() * sizeof(float), &dptr))) return rc;
        M.asa = (const float*)dptr;
        // polyphase vertical factors for mc_adj_v: asap[(r ss + a) kfm + m] = asa_r[ss m + ra], m < Ma
        const int kfm = (4 * p->kf4 + ss - 1) / ss;
        std::vector<float> asap((size_t)M.adj_rank * ss * kfm, 0.0f);
        for (int r = 0; r < M.adj_rank; ++r)
            for (int a = 0; a < ss; ++a) {
                const int ra = (a + pb) % ss, Ma = (kf - ra + ss - 1) / ss;
                for (int m = 0; m < Ma && m < kfm; ++m)
                    asap[((size_t)r * ss + a) * kfm + m] = asa[(size_t)r * kf + ss * m + ra];
            }
        if ((rc = upload(p, asap.data(), asap.size() * sizeof(float), &dptr))) return rc;
        M.asap = (const float*)dptr;
        // horizontal factors at mc_adj_h's window positions: phase b's tap nn sits at
        // j = kfm - 1 + (db(b) - db(0)) - nn of the lane's kfm + 1 loaded columns
        std::vector<float> asbq((size_t)M.adj_rank * ss * (kfm + 1), 0.0f);
        const int db0 = pb / ss;
        for (int r = 0; r < M.adj_rank; ++r)
            for (int b2 = 0; b2 < ss; ++b2) {
                const int rb = (b2 + pb) % ss, db = (b2 + pb) / ss, Mb = (kf - rb + ss - 1) / ss;
                for (int nn = 0; nn < Mb; ++nn) {
                    const int j = kfm - 1 + (db - db0) - nn;
                    if (j >= 0 && j <= kfm) asbq[((size_t)r * ss + b2) * (kfm + 1) + j] = asb[(size_t)r * kf + ss * nn + rb];
                }
            }
        if ((rc = upload(p, asbq.data(), asbq.size() * sizeof(float), &dptr))) return rc;
        M.asbq = (const float*)dptr;
        if ((rc = upload(p, asb.data(), asb.size() * sizeof(float), &dptr))) return rc;
        M.asb = (const float*)dptr;
    }
    if (M.sep_rank > 0) {
        if ((rc = upload(p, sa.data(), sa.size() * sizeof(float), &dptr))) return rc;
        M.sa = (const float*)dptr;
        // polyphase vertical factors for mc_sep_v: sap[(r ss + a) kfm + m] = sa_r[ss m + a]
        const int kfm = (4 * p->kf4 + ss - 1) / ss;
        std::vector<float> sap((size_t)M.sep_rank * ss * kfm, 0.0f);
        for (int r = 0; r < M.sep_rank; ++r)
            for (int a = 0; a < ss; ++a)
                for (int m = 0; m < kfm; ++m)
                    if (ss * m + a < kf) sap[((size_t)r * ss + a) * kfm + m] = sa[(size_t)r * kf + ss * m + a];
        if ((rc = upload(p, sap.data(), sap.size() * sizeof(float), &dptr))) return rc;
        M.sap = (const float*)dptr;
        if ((rc = upload(p, sbp.data(), sbp.size() * sizeof(float), &dptr))) return rc;
        M.sbp = (const float*)dptr;
    }
    const int n2 = D + (D & 1);
    // A, V (n2 x n2), rotations (n2), pairs (n2 ints), 2 reduction slots, rhs (n2)
    p->solve_lds = (size_t)(2 * n2 * (n2 + 1) + n2) * 8 + (size_t)n2 * 4 + 16 + (size_t)n2 * 8 + 64;  // A, V: n2 x (n2 + 1)
    if (lstsq)
        GL_HIP(mc_solve_attr(p->solve_lds));
    p->epl_batch = 0;
    p->lds_bytes = 0;
    P.sep_rank = 0;
    GL_HIP(hipDeviceSynchronize());
    return GL_OK;
}

// Work-unit totals of a lens with `count` particles (the prefix increments of a table).
struct McUnits {
    long part, basis, conv, adj, grad, adjh, adjv, seph, sepv;
};
static McUnits mc_units(const gl::DPlan& P, const gl::McPlan& M, int count) {
    McUnits u;
    const long njb = (P.N + 63) / 64;
    u.part = count;
    u.basis = (long)count * ((P.nsub * P.nsub + 255) / 256);
    u.grad = (long)count * (((long)P.nsub * P.nsub + 256 * gl::kMcGradK - 1) / (256 * gl::kMcGradK));
    // channel conv: direct (mc_conv) or separable (mc_sep_h + mc_sep_v)
    u.conv = M.sep_rank > 0 ? 0 : (long)count * M.D * ((P.N + M.conv_ib - 1) / M.conv_ib) * njb;
    u.seph = M.sep_rank > 0 ? (long)count * M.D * (((long)P.nsub * P.N + 255) / 256) : 0;
    u.sepv = M.sep_rank > 0 ? (long)count * M.D * ((P.N + gl::kSepRows - 1) / gl::kSepRows) * njb : 0;
    // adjoint conv: direct (mc_adjoint) or separable (mc_adj_h + mc_adj_v)
    const long nyb = (P.N + gl::kAdjKY - 1) / gl::kAdjKY;
    u.adj = M.adj_rank > 0 ? 0 : (long)count * (((long)P.nsub * P.ss * njb + 3) & ~3L);
    u.adjh = M.adj_rank > 0 ? (long)count * (((long)P.N * njb + 3) & ~3L) : 0;
    u.adjv = M.adj_rank > 0 ? (long)count * (((long)P.ss * nyb * P.ss * njb + 3) & ~3L) : 0;
    return u;
}

// Run a staged plan over bs particles in chunks of max_batch.  A carries the
// full-batch pointers; map: MAP step with loss weight 1/bs per particle.
static int run_mc(const gl_plan* plan, const gl::McArgs& A0, bool grad, hipStream_t st) {
    const gl::DPlan& P = plan->P;
    const gl::McPlan& M = plan->M;
    const int N = P.N, D = M.D, d = P.d, bs = A0.bs;
    if (bs <= 0 || !A0.in) return fail(GL_ERR_ARG, "bad batch");
    for (int off = 0; off < bs; off += plan->max_batch) {
        const int nb = bs - off < plan->max_batch ? bs - off : plan->max_batch;
        gl::McArgs A = A0;
        A.in = A0.in + (size_t)off * d;
        A.bs = nb;
        A.lp = A0.lp ? A0.lp + off : nullptr;
        A.chi2 = A0.chi2 ? A0.chi2 + off : nullptr;
        A.img = A0.img ? A0.img + (size_t)off * N * N : nullptr;
        A.coef_out = A0.coef_out ? A0.coef_out + (size_t)off * D : nullptr;
        A.stacked = A0.stacked ? A0.stacked + (size_t)off * N * N * D : nullptr;
        A.g_lp = A0.g_lp ? A0.g_lp + off : nullptr;
        A.g_chi2 = A0.g_chi2 ? A0.g_chi2 + off : nullptr;
        A.g_img = A0.g_img ? A0.g_img + (size_t)off * N * N : nullptr;
        A.grad = A0.grad ? A0.grad + (size_t)off * d : nullptr;
        A.z_io = A0.z_io ? A0.z_io + (size_t)off * d : nullptr;
        A.m = A0.m ? A0.m + (size_t)off * d : nullptr;
        A.v = A0.v ? A0.v + (size_t)off * d : nullptr;
        gl::McTable T;
        memset(&T, 0, sizeof(T));
        T.one.P = P;
        T.one.M = M;
        T.one.first = 0;
        T.one.count = nb;
        T.one.weight = 1.0f / (float)bs;
        T.n = 1;
        const McUnits u = mc_units(P, M, nb);
        if (u.basis > INT32_MAX || u.conv > INT32_MAX || u.adj > INT32_MAX || u.adjh > INT32_MAX ||
            u.adjv > INT32_MAX || u.seph > INT32_MAX || u.sepv > INT32_MAX)
            return fail(GL_ERR_UNSUPPORTED, "staged launch too large; lower max_batch");
        T.tot_part = nb; T.tot_basis = (int)u.basis; T.tot_conv = (int)u.conv; T.tot_adj = (int)u.adj;
        T.tot_grad = (int)u.grad;
        T.tot_adjh = (int)u.adjh;
        T.tot_adjv = (int)u.adjv;
        T.tot_seph = (int)u.seph;
        T.tot_sepv = (int)u.sepv;
        T.lds_consts = M.pstride * 4;
        T.lds_w = P.kf * P.kf * 4;
        int rc = mc_pipeline_run(false, T, A, P.ss, plan->kf4, M.lstsq, M.gram_nb, plan->ksplit_cap, plan->solve_lds,
                                 grad, st, M.shp_comp >= 0, plan->mc_model, nullptr);
        if (rc) return rc;
    }
    return GL_OK;
}

// Observed image, its noise map (tf/model.py:113-115) and the prior table.
static int upload_common(const gl_plan_desc* desc, gl_plan* p) {
    gl::DPlan& P = p->P;
    const int N = P.N;
    void* dptr = nullptr;
    int rc;
    const int nobs = desc->n_observed > 0 ? desc->n_observed : 1;
    P.n_obs = nobs;
    std::vector<float> obs((size_t)nobs * N * N, 0.0f), ivar((size_t)nobs * N * N, 0.0f),
        lsig((size_t)nobs * N * N, 0.0f);
    if (desc->observed) {
        for (size_t i = 0; i < (size_t)nobs * N * N; ++i) {
            obs[i] = desc->observed[i];
            // sqrt(bkg^2 + clip(obs, 0, inf) / t_exp), tf/model.py:113-115, or the given map
            float var = desc->noise_map ? desc->noise_map[i] * desc->noise_map[i]
                                        : P.bkg2 + fmaxf(obs[i], 0.0f) / desc->exp_time;
            ivar[i] = 1.0f / var;
            lsig[i] = logf(sqrtf(var));
        }
    }
    if ((rc = upload(p, obs.data(), obs.size() * sizeof(float), &dptr))) return rc;
    P.obs = (const float*)dptr;
    if ((rc = upload(p, ivar.data(), ivar.size() * sizeof(float), &dptr))) return rc;
    P.inv_var_obs = (const float*)dptr;
    if ((rc = upload(p, lsig.data(), lsig.size() * sizeof(float), &dptr))) return rc;
    P.log_sig_obs = (const float*)dptr;
    std::vector<int> pk(P.d, 0);
    std::vector<float> pa(8 * P.d, 0.0f);
    if (desc->prior_kind) {
        for (int j = 0; j < P.d; ++j) {
            pk[j] = desc->prior_kind[j];
            for (int t = 0; t < 4; ++t) pa[8 * j + t] = desc->prior_args[4 * j + t];
            if (pk[j] == GL_PRIOR_TRUNCNORMAL) {
                double loc = pa[8 * j], sc = pa[8 * j + 1], lo = pa[8 * j + 2], hi = pa[8 * j + 3];
                double Z = 0.5 * (erf((hi - loc) / sc / sqrt(2.0)) - erf((lo - loc) / sc / sqrt(2.0)));
                pa[8 * j + 4] = (float)log(Z);
            }
        }
    }
    if ((rc = upload(p, pk.data(), pk.size() * sizeof(int), &dptr))) return rc;
    P.prior_kind = (const int*)dptr;
    if ((rc = upload(p, pa.data(), pa.size() * sizeof(float), &dptr))) return rc;
    P.prior_args = (const float*)dptr;
    return GL_OK;
}

extern "C" int gl_plan_create(const gl_plan_desc* desc, gl_plan** out) {
    if (out) *out = nullptr;
    if (!desc || !out) return fail(GL_ERR_ARG, "gl_plan_create: null argument");
    const int N = desc->num_pix, ss = desc->supersample;
    if (N <= 0 || ss <= 0) return fail(GL_ERR_ARG, "num_pix and supersample must be positive");
    if (desc->n_params <= 0 || desc->n_params > GL_MAX_PARAMS)
        return fail(GL_ERR_ARG, "n_params must be in [1, " + std::to_string(GL_MAX_PARAMS) + "]");
    if (desc->n_components <= 0 || desc->n_components > GL_MAX_COMPONENTS)
        return fail(GL_ERR_ARG, "n_components must be in [1, 8]");

    gl_plan* p = new gl_plan();
    gl::DPlan& P = p->P;
    memset(&P, 0, sizeof(P));
    P.N = N; P.ss = ss; P.nsub = N * ss;
    P.M00 = desc->transform_sub[0]; P.M01 = desc->transform_sub[1];
    P.M10 = desc->transform_sub[2]; P.M11 = desc->transform_sub[3];
    P.ox = desc->origin[0]; P.oy = desc->origin[1];
    P.cf = desc->conversion_factor;
    P.d = desc->n_params;
    P.lik = desc->likelihood;
    P.fldj_at_x = desc->fldj_at_x;
    P.bkg2 = desc->background_rms * desc->background_rms;
    P.inv_texp = 1.0f / desc->exp_time;

    // ---- components and model signature -----------------------------------
    int sig[7] = {0, 0, 0, 0, 0, 0, 0};
    int nl = 0, nll = 0, ns = 0, last_role = 0, max_niter = 0;
    P.ncomp = desc->n_components;
    // least squares: some light is linear (use_lstsq), and then every light must be;
    // fixed per-channel amplitudes: a shapelet light with sampled amplitudes
    int any_lstsq = 0, any_shp = 0;
    for (int i = 0; i < desc->n_components; ++i) {
        any_lstsq |= desc->components[i].use_lstsq ? 1 : 0;
        any_shp |= desc->components[i].kind == glm::K_SHAPELETS ? 1 : 0;
    }
    const int famp = any_shp && !any_lstsq;
    p->mc = (any_lstsq || any_shp) ? 1 : 0;
    for (int i = 0; i < desc->n_components; ++i) {
        const gl_component_desc& c = desc->components[i];
        if (c.role < last_role) {
            gl_plan_destroy(p);
            return fail(GL_ERR_ARG, "components must be ordered lenses, lens light, source");
        }
        last_role = c.role;
        if (n_params_of(c.kind) < 0) {
            gl_plan_destroy(p);
            return fail(GL_ERR_UNSUPPORTED, "profile kind " + std::to_string(c.kind) +
                                                " is not supported by the resident-image kernel");
        }
        for (int j = 0; j < n_params_of(c.kind); ++j)
            if (c.param_col[j] >= desc->n_params) {
                gl_plan_destroy(p);
                return fail(GL_ERR_ARG, "param_col out of range");
            }
        if (c.role == GL_ROLE_LENS) {
            if (c.kind != glm::K_EPL && c.kind != glm::K_SHEAR && c.kind != glm::K_SIS && c.kind != glm::K_SIE &&
                c.kind != glm::K_TNFW) {
                gl_plan_destroy(p);
                return fail(GL_ERR_ARG, "light profile given as a lens");
            }
            if (nl >= 3) { gl_plan_destroy(p); return fail(GL_ERR_UNSUPPORTED, "at most 3 lens components"); }
            sig[nl++] = c.kind;
        } else {
            if (c.kind != glm::K_SERSIC && c.kind != glm::K_SERSIC_ELLIPSE && c.kind != glm::K_SHAPELETS &&
                c.kind != glm::K_CORE_SERSIC) {
                gl_plan_destroy(p);
                return fail(GL_ERR_ARG, "mass profile given as a light component");
            }
            if (any_lstsq && !c.use_lstsq) {
                gl_plan_destroy(p);
                return fail(GL_ERR_UNSUPPORTED, "least-squares plans need every light component linear (use_lstsq)");
            }
            if (c.role == GL_ROLE_LENS_LIGHT) {
                if (nll >= 2) { gl_plan_destroy(p); return fail(GL_ERR_UNSUPPORTED, "at most 2 lens-light components"); }
                sig[3 + nll++] = c.kind;
            } else {
                if (ns >= 2) { gl_plan_destroy(p); return fail(GL_ERR_UNSUPPORTED, "at most 2 source components"); }
                sig[5 + ns++] = c.kind;
            }
        }
        gl::DComp& dc = P.comp[i];
        dc.kind = c.kind; dc.role = c.role; dc.niter = c.niter; dc.n_max = c.n_max;
        dc.interpolate = c.interpolate; dc.lstsq = c.use_lstsq; dc.amp_col0 = c.amp_col0;
        for (int j = 0; j < GL_MAX_COMPONENT_PARAMS; ++j) { dc.col[j] = c.param_col[j]; dc.cst[j] = c.param_const[j]; }
        if (c.kind == glm::K_EPL) {
            if (c.niter < 0 || c.niter > GL_MAX_EPL_ITER) {
                gl_plan_destroy(p);
                return fail(GL_ERR_ARG, "EPL niter must be in [0, " + std::to_string(GL_MAX_EPL_ITER) + "]");
            }
            if (c.niter > max_niter) max_niter = c.niter;
        }
    }
    P.nlens = nl; P.nll = nll; P.nsrc = ns;
    for (int j = 0; j < 7; ++j) p->sig[j] = sig[j];
    p->model = nullptr;
    const int lstsq = any_lstsq;
    bool staged = p->mc || desc->staged;
    if (!staged) {
        p->model = find_model(sig);
        staged = !p->model;  // no fused kernel for this combination: the staged pipeline
    }

    // ---- PSF folded with the pooling box (correlation weights) -----------------
    const int k = desc->kernel_size > 0 ? desc->kernel_size : 1;
    std::vector<double> K(k * k, 0.0);
    if (desc->kernel_size > 0) {
        if (!desc->kernel) { gl_plan_destroy(p); return fail(GL_ERR_ARG, "kernel_size > 0 but kernel is null"); }
        for (int i = 0; i < k * k; ++i) K[i] = desc->kernel[i];
    } else {
        K[0] = 1.0;
    }
    const int kf = k + ss - 1, pb = (k - 1) / 2;
    P.k = k; P.kf = kf; P.pb = pb;
    std::vector<float> W(kf * kf, 0.0f);
    for (int pp = 0; pp < kf; ++pp)
        for (int qq = 0; qq < kf; ++qq) {
            double acc = 0.0;
            for (int a = 0; a < ss; ++a)
                for (int b = 0; b < ss; ++b) {
                    int u = pp - a, v = qq - b;
                    if (u >= 0 && u < k && v >= 0 && v < k) acc += K[(k - 1 - u) * k + (k - 1 - v)];  // flipped
                }
            W[pp * kf + qq] = (float)(acc / (ss * ss));
        }

    // HBM-staged pipeline: least squares, forced, no fused kernel, or (below) an
    // image the LDS-resident kernel cannot hold
    auto create_staged = [&]() -> int {
        if (desc->n_observed > 1) {
            gl_plan_destroy(p);
            return fail(GL_ERR_UNSUPPORTED, "staged plans take one observed image (use gl_batch_* for many lenses)");
        }
        p->mc = 1;
        p->model = nullptr;
        p->mc_model = 0;
        if (!lstsq && !famp) {
            const int dflt[7] = {EPL, SHR, 0, SE, 0, SE, 0};
            bool ok = true;
            for (int j = 0; j < 7; ++j) ok = ok && sig[j] == dflt[j];
            p->mc_model = ok ? 1 : 0;
        }
        int rc = mc_create(desc, p, W, max_niter, lstsq, famp);
        if (rc) { gl_plan_destroy(p); return rc; }
        if (p->guard.init() != hipSuccess) { gl_plan_destroy(p); return fail(GL_ERR_HIP, "gl_plan_create: event"); }
        *out = p;
        return GL_OK;
    };
    if (staged) return create_staged();

    // ---- LDS layout --------------------------------------------------------
    const int nsub = P.nsub;
    const int nstr = (N + gl::kRB - 1) / gl::kRB;
    const int overshoot = ss * (nstr * gl::kRB - N);
    int H = pb > (k - 1 - pb + overshoot) ? pb : (k - 1 - pb + overshoot);
    int PW = (nsub + H + 3) & ~3;
    P.PW = PW;
    P.sub_base = pb;
    const int sub_img = pb + (nsub + k - 1) * PW;
    // adjoint image padding
    int gp0 = 0, dmax = 0;
    for (int a = 0; a < ss; ++a) {
        int ra = (a + pb) % ss, da = (a + pb) / ss, Ma = (kf - ra + ss - 1) / ss;
        if (Ma - 1 - da > gp0) gp0 = Ma - 1 - da;
        if (da > dmax) dmax = da;
    }
    const int xstr = (N + gl::kRB2 - 1) / gl::kRB2;
    // width covers the widest vector read of an adjoint strip (rounded to 4 floats)
    const int mmax = (kf + ss - 1) / ss;
    const int sega4 = ((gl::kRB2 + mmax - 1 + 3) / 4) * 4;
    const int GW = (gp0 + xstr * gl::kRB2 + dmax + sega4 + 3) & ~3;
    const int grows = gp0 + N + dmax;
    P.GW = GW; P.gp0 = gp0; P.g_size = GW * grows;
    {
        bool al = true;
        for (int b = 0; b < ss; ++b) {
            int rb = (b + pb) % ss, db = (b + pb) / ss, Mb = (kf - rb + ss - 1) / ss;
            al = al && ((db - (Mb - 1) + gp0) % 4 + 4) % 4 == 0;
        }
        P.adj_aligned = al ? 1 : 0;
    }
    P.sub_size = sub_img > P.g_size ? sub_img : P.g_size;
    int off = (P.sub_size + 3) & ~3;
    P.off_w = off;
    P.off_x = off; off += (6 * P.d + 3) & ~3;
    P.off_c = off; off += gl::kMaxComp * gl::kConstStride;
    P.off_coef = off;
    P.max_n = max_niter;
    const int coef_block = glm::epl_tab_floats(max_niter);
    for (int i = 0; i < 3; ++i) {
        P.coef_off[i] = (sig[i] == glm::K_EPL) ? (off - P.off_coef) : 0;
        if (sig[i] == glm::K_EPL) off += (coef_block + 3) & ~3;
    }
    P.off_red = off;
    const int nred = 2 + 8 * 3 + 11 * 4;
    off += (kBlock / 64 + 1) * nred;

    // ---- bounds proof of every LDS index the kernel forms --------------------
    {
        // forward conv: base = sub_base - pb; reads base + (ss I + p) PW + ss J0 + c, c < ss(RB-1)+kf
        const long j0max = (long)(nstr - 1) * gl::kRB;
        const long seg = (((long)ss * (gl::kRB - 1) + kf + 3) / 4) * 4;  // vector reads round up
        const long rmax = (long)P.sub_base - pb + (long)(ss * (N - 1) + kf - 1) * PW + ss * j0max + seg - 1;
        const long wmax = (long)P.sub_base + (long)(nsub - 1 + pb) * PW + nsub - 1;
        bool ok = rmax < P.sub_size && wmax < P.sub_size && P.sub_base - pb >= 0;
        // image column reach must stay inside the zero halo of the same row
        ok = ok && (ss * j0max + seg - 1 - pb) < PW && pb <= PW - nsub;
        // adjoint: reads (Y + da - m + gp0) GW + X0 + db - (Mb-1) + gp0 + j, j < RB2 + Mb - 1
        const long x0max = (long)(xstr - 1) * gl::kRB2;
        for (int a = 0; a < ss && ok; ++a) {
            int ra = (a + pb) % ss, da = (a + pb) / ss, Ma = (kf - ra + ss - 1) / ss;
            long rlo = (long)(0 + da - (Ma - 1) + gp0), rhi = (long)(N - 1 + da + gp0);
            long clo = (long)(0 + da - (Ma - 1) + gp0), chi = x0max + da - (Ma - 1) + gp0 + sega4 - 1;
            if (chi < x0max + da + gp0 + gl::kRB2 - 1) chi = x0max + da + gp0 + gl::kRB2 - 1;
            ok = ok && rlo >= 0 && clo >= 0 && rhi < grows && chi < GW;
        }
        ok = ok && (long)(N - 1 + gp0) * GW + (N - 1 + gp0) < P.g_size && P.g_size <= P.off_w;
        if (!ok) return create_staged();
    }
    P.lds_floats = off;
    p->lds_bytes = (size_t)off * sizeof(float);
    if (p->lds_bytes > 160 * 1024 || N * nstr > gl::kMaxT * kBlock) return create_staged();

    // ---- separable (low-rank) factorisation of the folded PSF ----------------
    std::vector<float> sa, sb;
    P.sep_rank = 0;
    {
        const float tol_in = desc->psf_rank_tol;
        const bool have_sep = (ss == 2 && kf == 22) || (ss == 1 && kf == 21) || (ss == 3 && kf == 23);
        if (tol_in >= 0.0f && have_sep) {
            const double tol = tol_in > 0.0f ? tol_in : 1e-6;
            std::vector<double> Wd(kf * kf), sv, Uv, Vv;
            for (int i = 0; i < kf * kf; ++i) Wd[i] = W[i];
            jacobi_svd(kf, Wd, sv, Uv, Vv);
            int R = 0;
            while (R < kf && sv[R] >= tol * sv[0]) ++R;
            const int TR = nsub + k - 1;
            const int xs = (N + gl::kTB - 1) / gl::kTB;
            const int urows = gp0 + N + dmax;
            const int ucol = urows | 1;  // U column-major with an odd column stride
            // transposed horizontal pass (gl::SepGeom): phase b reads G columns from
            // X0 + off_b + gp0, off_b = db - (Mb - 1); one segment covers every phase
            int omin = 1 << 30, omax = -(1 << 30);
            for (int b = 0; b < ss; ++b) {
                const int rb = (b + pb) % ss, db = (b + pb) / ss, Mb = (kf - rb + ss - 1) / ss;
                const int o = db - (Mb - 1);
                omin = o < omin ? o : omin;
                omax = o > omax ? o : omax;
            }
            const int sega4 = ((gl::kTB + mmax - 1 + omax - omin + 3) / 4) * 4;
            const int u_off = (P.g_size + 3) & ~3;
            const bool fits = R >= 1 && R <= gl::kMaxRank &&
                              // forward: R planes of T plus the rows read for discarded outputs
                              (long)(R * TR + ss * (gl::kRV - 1)) * N <= P.sub_size &&
                              (long)TR * nstr <= (long)gl::kMaxTT * kBlock && (long)N * nstr <= (long)gl::kMaxT * kBlock &&
                              // adjoint: G segments stay inside their row, U after G
                              omin + gp0 >= 0 && (long)(xs - 1) * gl::kTB + omin + gp0 + sega4 - 1 < GW &&
                              (long)u_off + (long)R * nsub * ucol <= P.sub_size;
            if (fits) {
                P.sep_rank = R;
                P.urows = urows;
                P.ucol = ucol;
                P.u_off = u_off;
                P.sep_aligned = ((omin + gp0) % 4 == 0) ? 1 : 0;
                sa.resize(R * kf);
                sb.resize(R * kf);
                for (int r = 0; r < R; ++r)
                    for (int i = 0; i < kf; ++i) {
                        sa[r * kf + i] = (float)(sv[r] * Uv[i * kf + r]);
                        sb[r * kf + i] = (float)Vv[i * kf + r];
                    }
            }
        }
    }

    // ---- device constants ---------------------------------------------------
    void* dptr = nullptr;
    int rc;
    if ((rc = upload(p, W.data(), W.size() * sizeof(float), &dptr))) { gl_plan_destroy(p); return rc; }
    P.w = (const float*)dptr;
    p->W_host = W;
    if (P.sep_rank > 0) {
        if ((rc = upload(p, sa.data(), sa.size() * sizeof(float), &dptr))) { gl_plan_destroy(p); return rc; }
        P.sa = (const float*)dptr;
        if ((rc = upload(p, sb.data(), sb.size() * sizeof(float), &dptr))) { gl_plan_destroy(p); return rc; }
        P.sb = (const float*)dptr;
    }
    if ((rc = upload_common(desc, p))) { gl_plan_destroy(p); return rc; }
    p->epl_batch = desc->epl_batch_series;
    P.batch_n = nullptr;
    if (p->epl_batch) {
        std::vector<int> z8(8, -1);
        if ((rc = upload(p, z8.data(), z8.size() * sizeof(int), &dptr))) { gl_plan_destroy(p); return rc; }
        P.batch_n = (const int*)dptr;
    }
    p->order_buf = nullptr;
    bool has_epl = false;
    for (int i = 0; i < P.nlens; ++i) has_epl = has_epl || P.comp[i].kind == glm::K_EPL;
    if (has_epl && !p->epl_batch) {
        if ((rc = dev_alloc(p->dev, (size_t)gl::kOrderCap * sizeof(int), &dptr, false))) { gl_plan_destroy(p); return rc; }
        p->order_buf = (int*)dptr;
    }
    hipError_t e = p->model->attr(p->lds_bytes);
    if (e != hipSuccess) {
        gl_plan_destroy(p);
        return fail(GL_ERR_HIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
    }
    if (p->guard.init() != hipSuccess) { gl_plan_destroy(p); return fail(GL_ERR_HIP, "gl_plan_create: event"); }
    *out = p;
    return GL_OK;
}

extern "C" int gl_plan_query(const gl_plan* plan, gl_plan_info* info) {
    if (!plan || !info) return fail(GL_ERR_ARG, "gl_plan_query: null argument");
    const gl::DPlan& P = plan->P;
    info->num_pix = P.N;
    info->supersample = P.ss;
    info->n_params = P.d;
    info->kernel_size = P.k;
    info->folded_size = P.kf;
    info->sep_rank = plan->mc ? plan->M.sep_rank : P.sep_rank;
    info->lds_bytes = (int)plan->lds_bytes;
    info->block = kBlock;
    info->adj_rank = plan->mc ? plan->M.adj_rank : 0;
    return GL_OK;
}

// Scratch of the fused path.  Plans own it (the batch-coupled EPL series counts and the
// launch-order buffer); gl_map_adam_step_noalloc takes it from `scratch` instead: the
// first kNoAllocSeries ints hold the series counts, the next gl::kOrderCap the launch
// order.  Every element a launch reads is written earlier in the same launch sequence
// (the series kernel writes every count, the order kernel a full permutation of the
// particles), so nothing is cleared between calls.
constexpr int kNoAllocSeries = 8;
static __device__ __align__(256) int scratch[kNoAllocSeries + gl::kOrderCap];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on the current device, looked up once per device (each
// GPU holds its own copy of the array at its own address).
static int scratch_ptr(int** out) {
    int dev = -1;
    GL_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(GL_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        GL_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = (int*)ptr;
    }
    *out = g_scratch_ptr[dev];
    return GL_OK;
}

// One pass of a plan over A.bs particles.  scratch (fused plans only): device memory of
// kNoAllocSeries + gl::kOrderCap ints used in place of the plan's own buffers.
static int run(const gl_plan* plan, gl::KArgs& A, bool grad, hipStream_t stream, int* scratch_dev = nullptr) {
    if (!plan) return fail(GL_ERR_ARG, "null plan");
    if (plan->mc && scratch_dev) return fail(GL_ERR_UNSUPPORTED, "staged plans run on their own workspace");
    GL_HIP(plan->guard.enter(stream));
    if (plan->mc) {
        if (A.lens) return fail(GL_ERR_UNSUPPORTED, "staged plans take one observed image (use gl_batch_*)");
        gl::McArgs B;
        memset(&B, 0, sizeof(B));
        B.in = A.in;
        B.bs = A.bs;
        B.use_z = (A.mode == gl::M_LOGPROB || A.mode == gl::M_MAP) ? 1 : 0;
        if (A.mode == gl::M_SIM) {
            if (plan->M.lstsq) return fail(GL_ERR_UNSUPPORTED, "least-squares plans: use gl_lstsq_simulate");
            B.img = A.img;
        } else if (A.mode == gl::M_VJP) {
            if (plan->M.lstsq) return fail(GL_ERR_UNSUPPORTED, "least-squares plans have no simulate VJP");
            B.g_img = A.g_img;
            B.grad = A.grad;
        } else {
            B.lp = A.lp;
            B.chi2 = A.chi2;
            B.g_lp = A.g_lp;
            B.g_chi2 = A.g_chi2;
            B.grad = A.grad;
            if (A.mode == gl::M_MAP) {
                if (A.weight) return fail(GL_ERR_UNSUPPORTED, "staged plans: per-particle MAP weights need gl_batch_*");
                B.map = 1;
                B.z_io = A.z_io;
                B.m = A.m;
                B.v = A.v;
                B.lr_t = A.lr_t;
                B.beta1 = A.beta1;
                B.beta2 = A.beta2;
                B.eps = A.eps;
            }
        }
        return run_mc(plan, B, grad, stream);
    }
    if (A.bs <= 0) return fail(GL_ERR_ARG, "bs must be positive");
    if (!A.in) return fail(GL_ERR_ARG, "null parameter pointer");
    gl::DPlan P = plan->P;
    const int use_z = (A.mode == gl::M_LOGPROB || A.mode == gl::M_MAP) ? 1 : 0;
    if (plan->epl_batch) {
        int* counts = scratch_dev ? scratch_dev : const_cast<int*>(plan->P.batch_n);
        GL_HIP(launch_epl_batch_series(P, A.in, A.bs, use_z, counts, stream));
        P.batch_n = counts;
    }
    // launch order by cost when the launch has several rounds of workgroups (gradient launches)
    A.order = nullptr;
    int* order = plan->order_buf ? (scratch_dev ? scratch_dev + kNoAllocSeries : plan->order_buf) : nullptr;
    if (order && grad && !A.limit && A.bs > 1024 && A.bs <= gl::kOrderCap) {
        GL_HIP(launch_order(P, A.in, A.bs, use_z, order, stream));
        A.order = order;
    }
    GL_HIP(plan->model->launch(A.bs, plan->lds_bytes, stream, P, A, grad));
    return GL_OK;
}

static gl::KArgs blank(int mode, int bs) {
    gl::KArgs A;
    memset(&A, 0, sizeof(A));
    A.mode = mode;
    A.bs = bs;
    return A;
}

extern "C" int gl_simulate(const gl_plan* plan, const float* x, int bs, float* img, hipStream_t stream) {
    if (!img) return fail(GL_ERR_ARG, "null image pointer");
    gl::KArgs A = blank(gl::M_SIM, bs);
    A.in = x;
    A.img = img;
    return run(plan, A, false, stream);
}

extern "C" int gl_simulate_vjp(const gl_plan* plan, const float* x, int bs, const float* g_img, float* grad_x,
                               hipStream_t stream) {
    if (!g_img || !grad_x) return fail(GL_ERR_ARG, "null gradient pointer");
    gl::KArgs A = blank(gl::M_VJP, bs);
    A.in = x;
    A.g_img = g_img;
    A.grad = grad_x;
    return run(plan, A, true, stream);
}

extern "C" int gl_log_prob(const gl_plan* plan, const float* z, int bs, float* lp, float* chi2, const float* g_lp,
                           const float* g_chi2, float* grad_z, hipStream_t stream) {
    gl::KArgs A = blank(gl::M_LOGPROB, bs);
    A.in = z;
    A.lp = lp;
    A.chi2 = chi2;
    A.g_lp = g_lp;
    A.g_chi2 = g_chi2;
    A.grad = grad_z;
    return run(plan, A, grad_z != nullptr, stream);
}

extern "C" int gl_log_prob_masked(const gl_plan* plan, const float* z, int bs, float* lp, float* chi2,
                                  const float* g_lp, const float* g_chi2, float* grad_z, const int* limit, int k,
                                  hipStream_t stream) {
    if (!plan) return fail(GL_ERR_ARG, "null plan");
    if (plan->mc) return fail(GL_ERR_UNSUPPORTED, "gl_log_prob_masked takes resident-image plans (lds_bytes > 0)");
    if (!limit) return fail(GL_ERR_ARG, "gl_log_prob_masked: null limit");
    gl::KArgs A = blank(gl::M_LOGPROB, bs);
    A.in = z;
    A.lp = lp;
    A.chi2 = chi2;
    A.g_lp = g_lp;
    A.g_chi2 = g_chi2;
    A.grad = grad_z;
    A.limit = limit;
    A.limit_k = k;
    return run(plan, A, grad_z != nullptr, stream);
}

extern "C" int gl_log_prob_lenses(const gl_plan* plan, const float* z, const int* lens, int bs, float* lp,
                                  float* chi2, const float* g_lp, const float* g_chi2, float* grad_z,
                                  hipStream_t stream) {
    if (plan && plan->mc && lens) return fail(GL_ERR_UNSUPPORTED, "least-squares plans take one observed image");
    gl::KArgs A = blank(gl::M_LOGPROB, bs);
    A.in = z;
    A.lp = lp;
    A.chi2 = chi2;
    A.g_lp = g_lp;
    A.g_chi2 = g_chi2;
    A.grad = grad_z;
    A.lens = lens;
    return run(plan, A, grad_z != nullptr, stream);
}

extern "C" int gl_lstsq_simulate(const gl_plan* plan, const float* x, int bs, float* img, float* coeffs,
                                 float* stacked, hipStream_t stream) {
    if (!plan) return fail(GL_ERR_ARG, "null plan");
    if (!plan->mc) return fail(GL_ERR_ARG, "gl_lstsq_simulate needs a plan with linear (use_lstsq) light components");
    if (!plan->M.lstsq) return fail(GL_ERR_ARG, "gl_lstsq_simulate needs a plan with linear (use_lstsq) light components");
    gl::McArgs B;
    memset(&B, 0, sizeof(B));
    B.in = x;
    B.bs = bs;
    B.img = img;
    B.coef_out = coeffs;
    B.stacked = stacked;
    GL_HIP(plan->guard.enter(stream));
    return run_mc(plan, B, false, stream);
}

static int map_step(const gl_plan* plan, float* z, float* m, float* v, const int* lens, const float* weight, int bs,
                    int step, float lr, float beta1, float beta2, float epsilon, float* lp, float* chi2,
                    hipStream_t stream, int* scratch_dev) {
    if (!z || !m || !v) return fail(GL_ERR_ARG, "null optimizer state");
    if (step < 1) return fail(GL_ERR_ARG, "step is 1-based");
    gl::KArgs A = blank(gl::M_MAP, bs);
    A.in = z;
    A.z_io = z;
    A.m = m;
    A.v = v;
    A.lp = lp;
    A.chi2 = chi2;
    double b1t = pow((double)beta1, step), b2t = pow((double)beta2, step);
    A.lr_t = (float)((double)lr * sqrt(1.0 - b2t) / (1.0 - b1t));
    A.beta1 = beta1;
    A.beta2 = beta2;
    A.eps = epsilon;
    A.gscale = weight ? -1.0f / (float)(plan ? plan->P.N * plan->P.N : 1)
                      : -1.0f / ((float)bs * (float)(plan ? plan->P.N * plan->P.N : 1));
    A.lens = lens;
    A.weight = weight;
    return run(plan, A, true, stream, scratch_dev);
}

extern "C" int gl_map_adam_step_lenses(const gl_plan* plan, float* z, float* m, float* v, const int* lens,
                                       const float* weight, int bs, int step, float lr, float beta1, float beta2,
                                       float epsilon, float* lp, float* chi2, hipStream_t stream) {
    return map_step(plan, z, m, v, lens, weight, bs, step, lr, beta1, beta2, epsilon, lp, chi2, stream, nullptr);
}

extern "C" int gl_map_adam_step(const gl_plan* plan, float* z, float* m, float* v, int bs, int step, float lr,
                                float beta1, float beta2, float epsilon, float* lp, float* chi2, hipStream_t stream) {
    return gl_map_adam_step_lenses(plan, z, m, v, nullptr, nullptr, bs, step, lr, beta1, beta2, epsilon, lp, chi2,
                                   stream);
}

extern "C" int gl_map_adam_step_noalloc(const gl_plan* plan, float* z, float* m, float* v, int bs, int step,
                                        float lr, float beta1, float beta2, float epsilon, float* lp, float* chi2,
                                        hipStream_t stream) {
    if (!plan) return fail(GL_ERR_ARG, "null plan");
    if (plan->mc)
        return fail(GL_ERR_UNSUPPORTED, "gl_map_adam_step_noalloc takes resident-image plans (gl_plan_info.lds_bytes > 0); "
                                        "staged plans need their own workspace");
    if (bs > gl::kOrderCap)
        return fail(GL_ERR_ARG, "gl_map_adam_step_noalloc: bs " + std::to_string(bs) + " exceeds the static scratch (" +
                                    std::to_string(gl::kOrderCap) + " particles)");
    int* sp = nullptr;
    int rc = scratch_ptr(&sp);
    if (rc) return rc;
    return map_step(plan, z, m, v, nullptr, nullptr, bs, step, lr, beta1, beta2, epsilon, lp, chi2, stream, sp);
}

// ---------------------------------------------------------------------------
// Batches of lens systems with their own image sizes, PSFs and particle counts
// (no reference counterpart).  Staged lenses are grouped into chunks whose
// workspaces share one arena of bounded size; each chunk runs every stage of the
// staged pipeline as one launch over its lenses, addressed through a device lens
// table.  Lenses whose plan runs on the resident-image (fused) kernel take one
// fused launch each.  All launches of a call are ordered on its stream.
// ---------------------------------------------------------------------------
struct BatchChunk {
    gl::McTable T;
    std::vector<ConvGroup> groups;  // conv launches by (ss, kf4)
    int ss_launch, kf4;
    int mc_model;  // 1: every lens is of the reference default model (specialised kernels)
};
struct FusedLens {
    const gl_plan* plan;
    int first, count;
};
// Lenses on the banded persistent kernel: one launch over all of them.
struct BandSet {
    const BandEntry* model = nullptr;
    gl::BArgs A;            // device tables and LDS layout; A.K is set per call
    int n = 0;              // lenses
    int grid = 0;           // workgroups (CUs x workgroups per CU at the LDS size)
    size_t lds = 0;         // dynamic LDS bytes per workgroup
    int rows_min = 0, rows_max = 0;
    size_t scratch_bytes = 0;
};
struct gl_batch {
    int n, d, total;
    std::vector<BatchChunk> chunks;
    std::vector<FusedLens> fused;
    BandSet band;
    int n_staged = 0;
    std::vector<int> counts;
    size_t arena_bytes;
    std::vector<void*> dev;
    mutable StreamGuard guard;  // orders calls across streams (the workspace is shared)
};

static const BandEntry* find_band_model(const int sig[7]) {
    int nm = 0;
    const BandEntry* tab = band_models(&nm);
    for (int i = 0; i < nm; ++i) {
        bool ok = true;
        for (int j = 0; j < 7; ++j) ok = ok && tab[i].sig[j] == sig[j];
        if (ok) return &tab[i];
    }
    return nullptr;
}

// Rank of a folded PSF at relative singular value 1e-6 and its factors for the banded
// kernel (gl::band::fac_floats layout: a_r = sv_r u_r and b_r = v_r, KFP floats each,
// zero padded, then a_r with ZP zeros in front, KFV floats each).  0 when the rank
// exceeds gl::band::kMaxRank.
static int band_factors(const std::vector<float>& W, int k, int ss, std::vector<float>& fac) {
    const int kf = k + ss - 1;
    std::vector<double> Wd(kf * kf), sv, Uv, Vv;
    for (int i = 0; i < kf * kf; ++i) Wd[i] = W[i];
    jacobi_svd(kf, Wd, sv, Uv, Vv);
    if (sv.empty() || !(sv[0] > 0.0)) return 0;
    int R = 0;
    while (R < kf && sv[R] >= 1e-6 * sv[0]) ++R;
    if (R > gl::band::kMaxRank) return 0;
    const gl::band::Geom g = gl::band::make_geom(1, ss, k, R, 1);
    fac.assign((size_t)gl::band::fac_floats(g), 0.0f);
    for (int r = 0; r < R; ++r)
        for (int i = 0; i < kf; ++i) {
            const float a = (float)(sv[r] * Uv[i * kf + r]);
            fac[(size_t)r * g.KFP + i] = a;
            fac[(size_t)(R + r) * g.KFP + i] = (float)Vv[i * kf + r];
            fac[(size_t)2 * R * g.KFP + (size_t)r * g.KFV + g.ZP + i] = a;
        }
    return R;
}

// Puts the lenses marked in `want` that the banded kernel can take on it; clears
// want[i] for the others.  gfirst: lens-major particle offsets.
static int band_setup(gl_batch* b, const gl_plan* const* plans, const int* counts, const std::vector<int>& gfirst,
                      const gl_batch_opts& o, std::vector<char>& want) {
    const int n = b->n;
    BandSet& B = b->band;
    std::vector<int> R(n, 0);
    std::vector<std::vector<float>> fac(n);
    for (int i = 0; i < n; ++i) {
        if (!want[i]) continue;
        const gl_plan* q = plans[i];
        const BandEntry* e = find_band_model(q->sig);
        want[i] = 0;
        if (!e || q->epl_batch || q->P.n_obs != 1 || (B.model && e != B.model)) continue;
        R[i] = band_factors(q->W_host, q->P.k, q->P.ss, fac[i]);
        if (R[i] < 1) continue;
        B.model = e;
        want[i] = 1;
    }
    // LDS layout shared by every lens (floats)
    int max_n = 0;
    for (int i = 0; i < n; ++i)
        if (want[i]) max_n = std::max(max_n, plans[i]->P.max_n);
    if (!B.model) return GL_OK;
    const int* sig = B.model->sig;
    gl::BArgs& A = B.A;
    memset(&A, 0, sizeof(A));
    int off = 0;
    A.off_x = off; off += (6 * b->d + 3) & ~3;
    A.off_c = off; off += gl::kMaxComp * gl::kConstStride;
    A.off_coef = off;
    int coef_off[3] = {0, 0, 0};
    const int coef_block = (glm::epl_tab_floats(max_n) + 3) & ~3;
    for (int i = 0; i < 3; ++i)
        if (sig[i] == glm::K_EPL) { coef_off[i] = off - A.off_coef; off += coef_block; }
    A.off_red = off; off += (gl::kBandBlock / 64 + 1) * (2 + 8 * 3 + 11 * 4);
    off = (off + 3) & ~3;
    A.off_lens = off; off += (int)((sizeof(gl::BLens) / 4 + 3) & ~(size_t)3);
    A.off_band = off;
    const int lds_cap = o.band_lds_bytes > 0 ? o.band_lds_bytes : 80 * 1024 - 256;
    const int region = lds_cap / 4 - off - 3;  // room to round the band up to whole float4s
    if (region <= 0) return fail(GL_ERR_ARG, "gl_batch_create: band_lds_bytes too small for the banded kernel");
    // per lens: band rows that fit the region; lenses with none stay where they were
    std::vector<gl::BLens> L;
    std::vector<int> idx;
    long slot = 0;
    int band_max = 0;
    B.rows_min = 1 << 30;
    B.rows_max = 0;
    for (int i = 0; i < n; ++i) {
        if (!want[i]) continue;
        const gl_plan* q = plans[i];
        // the most sub-rows per band (at most band_rows, 64 by default) whose LDS fits the region
        int rows = std::min(o.band_rows > 0 ? o.band_rows : 64, q->P.N * q->P.ss);
        gl::band::Geom g = gl::band::make_geom(q->P.N, q->P.ss, q->P.k, R[i], rows);
        while (rows > 1 && gl::band::band_floats(g) > region)
            g = gl::band::make_geom(q->P.N, q->P.ss, q->P.k, R[i], --rows);
        if (gl::band::band_floats(g) > region) { want[i] = 0; continue; }
        gl::BLens E;
        memset(&E, 0, sizeof(E));
        E.P = q->P;
        E.P.batch_n = nullptr;
        E.P.max_n = max_n;
        for (int j = 0; j < 3; ++j) E.P.coef_off[j] = coef_off[j];
        E.g = g;
        E.first = gfirst[i];
        E.count = counts[i];
        E.glp_map = -1.0f / ((float)counts[i] * (float)q->P.N * (float)q->P.N);
        void* dptr = nullptr;
        int rc = dev_alloc(b->dev, fac[i].size() * sizeof(float), &dptr, false);
        if (rc) return rc;
        if (hipMemcpy(dptr, fac[i].data(), fac[i].size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            return fail(GL_ERR_HIP, "gl_batch_create: upload");
        E.fac = (const float*)dptr;
        slot = std::max(slot, gl::band::slot_floats(g));
        band_max = std::max(band_max, gl::band::band_floats(g));
        B.rows_min = std::min(B.rows_min, rows);
        B.rows_max = std::max(B.rows_max, rows);
        L.push_back(E);
        idx.push_back(i);
    }
    B.n = (int)L.size();
    if (!B.n) { B.model = nullptr; return GL_OK; }
    // the kernel clears the band with 16-byte stores: whole float4s
    B.lds = (size_t)(A.off_band + ((band_max + 3) & ~3)) * sizeof(float);
    int wg_per_cu = 0, dev = 0, cus = 0;
    GL_HIP(B.model->attr(B.lds, &wg_per_cu));
    GL_HIP(hipGetDevice(&dev));
    GL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (wg_per_cu < 1) return fail(GL_ERR_UNSUPPORTED, "gl_batch_create: the banded kernel does not fit a CU");
    // lenses by descending cost (sub-pixels per particle): the longest items start first
    std::vector<int> order(B.n);
    for (int j = 0; j < B.n; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return L[a].g.nsub > L[c].g.nsub; });
    std::vector<int> pre(B.n + 1, 0);
    for (int j = 0; j < B.n; ++j) pre[j + 1] = pre[j] + L[order[j]].count;
    A.n = B.n;
    A.total = pre[B.n];
    B.grid = std::max(1, std::min(cus * wg_per_cu, A.total));
    slot = (slot + 63) & ~63L;
    A.slot = slot;
    B.scratch_bytes = (size_t)B.grid * (size_t)slot * sizeof(float);
    void* dptr = nullptr;
    int rc;
    if ((rc = dev_alloc(b->dev, B.scratch_bytes, &dptr, false))) return rc;
    A.scratch = (float*)dptr;
    auto put = [&](const void* host, size_t bytes, void** out) -> int {
        int r = dev_alloc(b->dev, bytes, out, false);
        if (r) return r;
        if (hipMemcpy(*out, host, bytes, hipMemcpyHostToDevice) != hipSuccess)
            return fail(GL_ERR_HIP, "gl_batch_create: upload");
        return GL_OK;
    };
    if ((rc = put(L.data(), L.size() * sizeof(gl::BLens), &dptr))) return rc;
    A.lens = (const gl::BLens*)dptr;
    if ((rc = put(order.data(), order.size() * sizeof(int), &dptr))) return rc;
    A.order = (const int*)dptr;
    if ((rc = put(pre.data(), pre.size() * sizeof(int), &dptr))) return rc;
    A.pre = (const int*)dptr;
    (void)idx;
    return GL_OK;
}

// One launch of the banded kernel over every lens on it.
static int band_run(const gl_batch* b, const gl::KArgs& K, bool grad, hipStream_t stream) {
    gl::BArgs A = b->band.A;
    A.K = K;
    GL_HIP(b->band.model->launch(b->band.grid, b->band.lds, stream, A, grad));
    return GL_OK;
}

extern "C" int gl_batch_destroy(gl_batch* b) {
    if (!b) return GL_OK;
    for (void* p : b->dev) (void)hipFree(p);
    b->guard.release();
    delete b;
    return GL_OK;
}

// Device tables of one chunk: its lenses (global particle offsets in goff), the
// prefix sums of every stage's work units, and the conv launch groups.
static int batch_chunk(gl_batch* b, const std::vector<gl::McLens>& lens, const std::vector<int>& gfirst,
                       const std::vector<int>& kf4s, BatchChunk& C) {
    const int n = (int)lens.size();
    std::vector<gl::McLens> L(lens);
    std::vector<int> pre_part(n + 1, 0), pre_basis(n + 1, 0), pre_conv(n + 1, 0), pre_adj(n + 1, 0),
        pre_grad(n + 1, 0), pre_adjh(n + 1, 0), pre_adjv(n + 1, 0);
    long acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<McUnits> units(n);
    C.kf4 = 0;
    C.ss_launch = L[0].P.ss;
    for (int i = 0; i < n; ++i) {
        gl::McLens& E = L[i];
        E.first = (int)acc[0];
        E.goff = gfirst[i] - (int)acc[0];
        const McUnits u = mc_units(E.P, E.M, E.count);
        units[i] = u;
        acc[0] += u.part; acc[1] += u.basis; acc[2] += u.conv; acc[3] += u.adj; acc[4] += u.grad;
        acc[5] += u.adjh; acc[6] += u.adjv; acc[7] += u.seph; acc[8] += u.sepv;
        for (int k = 0; k < 9; ++k)
            if (acc[k] > INT32_MAX) return fail(GL_ERR_UNSUPPORTED, "gl_batch_create: chunk too large");
        pre_part[i + 1] = (int)acc[0];
        pre_basis[i + 1] = (int)acc[1];
        pre_conv[i + 1] = (int)acc[2];
        pre_adj[i + 1] = (int)acc[3];
        pre_grad[i + 1] = (int)acc[4];
        pre_adjh[i + 1] = (int)acc[5];
        pre_adjv[i + 1] = (int)acc[6];
        C.kf4 = std::max(C.kf4, kf4s[i]);
        if (E.P.ss != C.ss_launch) C.ss_launch = 0;  // mixed supersampling: runtime ss
    }
    int rc = GL_OK;
    void* dptr = nullptr;
    auto put = [&](const void* host, size_t bytes) -> void* {
        if ((rc = dev_alloc(b->dev, bytes, &dptr, false))) return nullptr;
        if (hipMemcpy(dptr, host, bytes, hipMemcpyHostToDevice) != hipSuccess) {
            rc = fail(GL_ERR_HIP, "gl_batch_create: upload");
            return nullptr;
        }
        return dptr;
    };
    memset(&C.T, 0, sizeof(C.T));
    C.T.lens = (const gl::McLens*)put(L.data(), L.size() * sizeof(gl::McLens));
    if (!rc) C.T.pre_part = (const int*)put(pre_part.data(), pre_part.size() * sizeof(int));
    if (!rc) C.T.pre_basis = (const int*)put(pre_basis.data(), pre_basis.size() * sizeof(int));
    if (!rc) C.T.pre_conv = (const int*)put(pre_conv.data(), pre_conv.size() * sizeof(int));
    if (!rc) C.T.pre_adj = (const int*)put(pre_adj.data(), pre_adj.size() * sizeof(int));
    if (!rc) C.T.pre_grad = (const int*)put(pre_grad.data(), pre_grad.size() * sizeof(int));
    if (!rc) C.T.pre_adjh = (const int*)put(pre_adjh.data(), pre_adjh.size() * sizeof(int));
    if (!rc) C.T.pre_adjv = (const int*)put(pre_adjv.data(), pre_adjv.size() * sizeof(int));
    // conv launch groups by (ss, kf4): each group's prefix tables count only its lenses' units
    for (int i = 0; i < n && !rc; ++i) {
        bool seen = false;
        for (const ConvGroup& g : C.groups) seen = seen || (g.ss == L[i].P.ss && g.kf4 == kf4s[i]);
        if (seen) continue;
        ConvGroup g;
        g.ss = L[i].P.ss;
        g.kf4 = kf4s[i];
        std::vector<int> pc(n + 1, 0), pv(n + 1, 0), pah(n + 1, 0), psh(n + 1, 0), psv(n + 1, 0);
        long ac = 0, av = 0, aah = 0, ash = 0, asv = 0;
        for (int j = 0; j < n; ++j) {
            if (L[j].P.ss == g.ss && kf4s[j] == g.kf4) {
                ac += units[j].conv;
                av += units[j].adjv;
                aah += units[j].adjh;
                ash += units[j].seph;
                asv += units[j].sepv;
            }
            pc[j + 1] = (int)ac;
            pv[j + 1] = (int)av;
            pah[j + 1] = (int)aah;
            psh[j + 1] = (int)ash;
            psv[j + 1] = (int)asv;
        }
        g.tot = (int)ac;
        g.adjv_tot = (int)av;
        g.adjh_tot = (int)aah;
        g.seph_tot = (int)ash;
        g.sepv_tot = (int)asv;
        g.pre = (const int*)put(pc.data(), pc.size() * sizeof(int));
        if (!rc) g.adjv_pre = (const int*)put(pv.data(), pv.size() * sizeof(int));
        if (!rc) g.adjh_pre = (const int*)put(pah.data(), pah.size() * sizeof(int));
        if (!rc) g.seph_pre = (const int*)put(psh.data(), psh.size() * sizeof(int));
        if (!rc) g.sepv_pre = (const int*)put(psv.data(), psv.size() * sizeof(int));
        C.groups.push_back(g);
    }
    if (rc) return rc;
    C.T.n = n;
    C.T.tot_part = (int)acc[0];
    C.T.tot_basis = (int)acc[1];
    C.T.tot_conv = (int)acc[2];
    C.T.tot_adj = (int)acc[3];
    C.T.tot_grad = (int)acc[4];
    C.T.tot_adjh = (int)acc[5];
    C.T.tot_adjv = (int)acc[6];
    C.T.tot_seph = (int)acc[7];
    C.T.tot_sepv = (int)acc[8];
    C.T.lds_consts = 0;
    C.T.lds_w = 0;
    for (int i = 0; i < n; ++i) {
        C.T.lds_consts = std::max(C.T.lds_consts, L[i].M.pstride * 4);
        C.T.lds_w = std::max(C.T.lds_w, L[i].P.kf * L[i].P.kf * 4);
    }
    if (C.T.lds_consts > 64 * 1024) return fail(GL_ERR_UNSUPPORTED, "gl_batch_create: constant block too large");
    return GL_OK;
}

extern "C" int gl_batch_create_ex(const gl_plan* const* plans, int n, const int* counts, const gl_batch_opts* opts,
                                  gl_batch** out) {
    if (out) *out = nullptr;
    if (!plans || !counts || !out || n <= 0) return fail(GL_ERR_ARG, "gl_batch_create: bad arguments");
    gl_batch_opts o;
    memset(&o, 0, sizeof(o));
    if (opts) o = *opts;
    if (o.band < 0 || o.band > 2 || o.band_rows < 0 || o.band_lds_bytes < 0 || o.band_lds_bytes > 160 * 1024)
        return fail(GL_ERR_ARG, "gl_batch_create: band must be 0, 1 or 2, band_rows >= 0, band_lds_bytes in [0, 160 KiB]");
    for (int i = 0; i < n; ++i) {
        const gl_plan* q = plans[i];
        if (!q) return fail(GL_ERR_ARG, "gl_batch_create: null plan");
        if (q->mc && (q->M.lstsq || q->M.famp))
            return fail(GL_ERR_UNSUPPORTED, "gl_batch_create: least-squares and shapelet lights are not batched");
        if (counts[i] <= 0) return fail(GL_ERR_ARG, "gl_batch_create: particle counts must be positive");
        if (q->P.d != plans[0]->P.d || q->P.lik != plans[0]->P.lik)
            return fail(GL_ERR_ARG, "gl_batch_create: plans must share the parameter count and likelihood");
    }
    const size_t budget = o.workspace_bytes ? o.workspace_bytes : (size_t)32 << 30;
    gl_batch* b = new gl_batch();
    b->n = n;
    b->d = plans[0]->P.d;
    b->counts.assign(counts, counts + n);
    b->arena_bytes = 0;
    int rc = GL_OK;
    // lens-major particle offsets; fused lenses run on their own plans
    std::vector<int> gfirst(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        if ((long)gfirst[i] + counts[i] > INT32_MAX) { gl_batch_destroy(b); return fail(GL_ERR_ARG, "gl_batch_create: too many particles"); }
        gfirst[i + 1] = gfirst[i] + counts[i];
    }
    b->total = gfirst[n];
    // lenses on the banded kernel (band 1: any lens it takes; 2: staged lenses only)
    std::vector<char> on_band(n, 0);
    if (o.band)
        for (int i = 0; i < n; ++i) on_band[i] = o.band == 1 || plans[i]->mc;
    if ((rc = band_setup(b, plans, counts, gfirst, o, on_band))) { gl_batch_destroy(b); return rc; }
    // staged lenses: constants (per lens, kept) and scratch sizes; chunks in lens order
    std::vector<gl::McLens> lens;
    std::vector<int> firsts, kf4s, mcm;
    std::vector<size_t> bytes;
    for (int i = 0; i < n && !rc; ++i) {
        if (on_band[i]) continue;
        if (!plans[i]->mc) {
            if (counts[i] > gl::kOrderCap) {
                gl_batch_destroy(b);
                return fail(GL_ERR_ARG, "gl_batch_create: a resident-image lens takes at most 16384 particles");
            }
            b->fused.push_back(FusedLens{plans[i], gfirst[i], counts[i]});
            continue;
        }
        ++b->n_staged;
        gl::McLens E;
        memset(&E, 0, sizeof(E));
        E.P = plans[i]->P;
        E.M = plans[i]->M;
        E.M.conv_ib = gl::kMcIB;  // batches launch the direct conv at the default rows per lane
        E.count = counts[i];
        E.weight = 1.0f / (float)counts[i];
        if ((rc = mc_consts_upload(E.P, E.M, plans[i]->kf4, plans[i]->W_host, b->dev))) break;
        ScratchAlloc m;
        m.measure = true;
        if ((rc = mc_scratch(E.P, E.M, counts[i], 1, m))) break;
        lens.push_back(E);
        firsts.push_back(gfirst[i]);
        kf4s.push_back(plans[i]->kf4);
        mcm.push_back(plans[i]->mc_model);
        bytes.push_back(m.off);
    }
    if (rc) { gl_batch_destroy(b); return rc; }
    // a chunk stays within the byte budget and keeps every stage's launch under 2^31
    // work-items (a dispatch holds at most 2^32)
    auto stage_threads = [](const McUnits& u) {
        return std::max({u.basis * 256, u.grad * 256, u.conv * 64, u.adj * 64, u.adjh * 64, u.adjv * 64,
                         u.seph * 64, u.sepv * 256, u.part * 256});
    };
    const long kMaxThreads = 1L << 31;
    std::vector<std::vector<int>> groups;
    size_t cur = 0;
    McUnits acc_u{};
    for (size_t i = 0; i < lens.size(); ++i) {
        const McUnits u = mc_units(lens[i].P, lens[i].M, lens[i].count);
        McUnits sum = acc_u;
        sum.part += u.part; sum.basis += u.basis; sum.conv += u.conv; sum.adj += u.adj; sum.grad += u.grad;
        sum.adjh += u.adjh; sum.adjv += u.adjv; sum.seph += u.seph; sum.sepv += u.sepv;
        if (groups.empty() || cur + bytes[i] > budget || stage_threads(sum) > kMaxThreads) {
            groups.emplace_back();
            cur = 0;
            sum = u;
        }
        if (stage_threads(u) > kMaxThreads) { gl_batch_destroy(b); return fail(GL_ERR_UNSUPPORTED, "gl_batch_create: a lens's launch exceeds 2^31 work-items; lower its particle count"); }
        groups.back().push_back((int)i);
        cur += bytes[i];
        acc_u = sum;
        b->arena_bytes = std::max(b->arena_bytes, cur);
    }
    char* arena = nullptr;
    if (b->arena_bytes) {
        void* dptr = nullptr;
        if ((rc = dev_alloc(b->dev, b->arena_bytes, &dptr, false))) { gl_batch_destroy(b); return rc; }
        arena = (char*)dptr;
    }
    for (const std::vector<int>& g : groups) {
        ScratchAlloc sa;
        sa.arena = arena;
        sa.cap = b->arena_bytes;
        std::vector<gl::McLens> cl;
        std::vector<int> cf, ck;
        for (int i : g) {
            gl::McLens E = lens[i];
            if ((rc = mc_scratch(E.P, E.M, E.count, 1, sa))) break;
            cl.push_back(E);
            cf.push_back(firsts[i]);
            ck.push_back(kf4s[i]);
        }
        if (rc) break;
        BatchChunk C;
        if ((rc = batch_chunk(b, cl, cf, ck, C))) break;
        C.mc_model = 1;
        for (int i : g) C.mc_model = C.mc_model && mcm[i] == 1;
        b->chunks.push_back(C);
    }
    if (rc) { gl_batch_destroy(b); return rc; }
    if (hipDeviceSynchronize() != hipSuccess) { gl_batch_destroy(b); return fail(GL_ERR_HIP, "gl_batch_create: sync"); }
    if (b->guard.init() != hipSuccess) { gl_batch_destroy(b); return fail(GL_ERR_HIP, "gl_batch_create: event"); }
    *out = b;
    return GL_OK;
}

extern "C" int gl_batch_create_ws(const gl_plan* const* plans, int n, const int* counts, size_t workspace_bytes,
                                  gl_batch** out) {
    gl_batch_opts o;
    memset(&o, 0, sizeof(o));
    o.workspace_bytes = workspace_bytes;
    return gl_batch_create_ex(plans, n, counts, &o, out);
}

extern "C" int gl_batch_create(const gl_plan* const* plans, int n, const int* counts, gl_batch** out) {
    return gl_batch_create_ws(plans, n, counts, 0, out);
}

extern "C" int gl_batch_query(const gl_batch* b, gl_batch_info* info) {
    if (!b || !info) return fail(GL_ERR_ARG, "gl_batch_query: null argument");
    memset(info, 0, sizeof(*info));
    info->total = b->total;
    info->lenses_fused = (int)b->fused.size();
    info->lenses_staged = b->n_staged;
    info->lenses_band = b->band.n;
    info->chunks = (int)b->chunks.size();
    info->band_grid = b->band.grid;
    info->band_lds_bytes = (int)b->band.lds;
    info->band_rows_min = b->band.n ? b->band.rows_min : 0;
    info->band_rows_max = b->band.rows_max;
    info->band_scratch_bytes = b->band.scratch_bytes;
    info->workspace_bytes = b->arena_bytes;
    return GL_OK;
}

extern "C" int gl_batch_size(const gl_batch* b) { return b ? b->total : -1; }

extern "C" size_t gl_batch_workspace_bytes(const gl_batch* b) { return b ? b->arena_bytes : 0; }

extern "C" int gl_batch_num_chunks(const gl_batch* b) { return b ? (int)b->chunks.size() : -1; }

extern "C" int gl_batch_log_prob(const gl_batch* b, const float* z, float* lp, float* chi2, float* grad_z,
                                 hipStream_t stream) {
    if (!b || !z) return fail(GL_ERR_ARG, "gl_batch_log_prob: null argument");
    GL_HIP(b->guard.enter(stream));
    gl::McArgs A;
    memset(&A, 0, sizeof(A));
    A.in = z;
    A.use_z = 1;
    A.bs = b->total;
    A.lp = lp;
    A.chi2 = chi2;
    A.grad = grad_z;
    if (b->band.n) {
        gl::KArgs K = blank(gl::M_LOGPROB, b->total);
        K.in = z;
        K.lp = lp;
        K.chi2 = chi2;
        K.grad = grad_z;
        int rc = band_run(b, K, grad_z != nullptr, stream);
        if (rc) return rc;
    }
    for (const BatchChunk& C : b->chunks) {
        int rc = mc_pipeline_run(true, C.T, A, C.ss_launch, C.kf4, 0, 0, 1, 0, grad_z != nullptr, stream, false,
                                 C.mc_model, &C.groups);
        if (rc) return rc;
    }
    const int d = b->d;
    for (const FusedLens& F : b->fused) {
        gl::KArgs K = blank(gl::M_LOGPROB, F.count);
        K.in = z + (size_t)F.first * d;
        K.lp = lp ? lp + F.first : nullptr;
        K.chi2 = chi2 ? chi2 + F.first : nullptr;
        K.grad = grad_z ? grad_z + (size_t)F.first * d : nullptr;
        int rc = run(F.plan, K, grad_z != nullptr, stream);
        if (rc) return rc;
    }
    return GL_OK;
}

extern "C" int gl_batch_map_adam_step(const gl_batch* b, float* z, float* m, float* v, int step, float lr,
                                      float beta1, float beta2, float epsilon, float* lp, float* chi2,
                                      hipStream_t stream) {
    if (!b || !z || !m || !v) return fail(GL_ERR_ARG, "gl_batch_map_adam_step: null argument");
    if (step < 1) return fail(GL_ERR_ARG, "step is 1-based");
    GL_HIP(b->guard.enter(stream));
    gl::McArgs A;
    memset(&A, 0, sizeof(A));
    A.in = z;
    A.use_z = 1;
    A.bs = b->total;
    A.lp = lp;
    A.chi2 = chi2;
    A.map = 1;
    A.z_io = z;
    A.m = m;
    A.v = v;
    const double b1t = pow((double)beta1, step), b2t = pow((double)beta2, step);
    A.lr_t = (float)((double)lr * sqrt(1.0 - b2t) / (1.0 - b1t));
    A.beta1 = beta1;
    A.beta2 = beta2;
    A.eps = epsilon;
    if (b->band.n) {
        gl::KArgs K = blank(gl::M_MAP, b->total);
        K.in = z;
        K.z_io = z;
        K.m = m;
        K.v = v;
        K.lp = lp;
        K.chi2 = chi2;
        K.lr_t = A.lr_t;
        K.beta1 = beta1;
        K.beta2 = beta2;
        K.eps = epsilon;
        int rc = band_run(b, K, true, stream);
        if (rc) return rc;
    }
    for (const BatchChunk& C : b->chunks) {
        int rc = mc_pipeline_run(true, C.T, A, C.ss_launch, C.kf4, 0, 0, 1, 0, true, stream, false, C.mc_model,
                                 &C.groups);
        if (rc) return rc;
    }
    const int d = b->d;
    for (const FusedLens& F : b->fused) {
        // a lens's loss is the mean over its particles of -lp / N^2 (tf/inference.py:28-35)
        int rc = map_step(F.plan, z + (size_t)F.first * d, m + (size_t)F.first * d, v + (size_t)F.first * d, nullptr,
                          nullptr, F.count, step, lr, beta1, beta2, epsilon, lp ? lp + F.first : nullptr,
                          chi2 ? chi2 + F.first : nullptr, stream, nullptr);
        if (rc) return rc;
    }
    return GL_OK;
}
