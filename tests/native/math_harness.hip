// Host build of the per-sub-pixel device math (gigalens_amd/csrc/lens_math.h)
// for CPU unit tests.  Test infrastructure only.
#include <hip/hip_runtime.h>
#include <vector>
#include "../../gigalens_amd/csrc/lens_math.h"

using namespace glm;

static void coef_tables(const EplC& C, std::vector<float>& a, std::vector<float4>& b) {
    a.assign(epl_pad(C.N) + 1, 0.0f);  // zero padding above N, as on the device
    b.assign(epl_pad(C.N) + 1, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    for (int n = 0; n <= C.N; ++n) {
        float c, dt, df;
        epl_coef(n, (double)C.f, (double)C.t, c, dt, df);
        a[n] = c;
        b[n] = make_float4(c, dt, df, 0.0f);
    }
}

extern "C" int h_series_len(double f, int niter) { return epl_series_len(f, niter); }

extern "C" void h_epl_alpha(const float* p, int niter, int n, const float* x, const float* y, float* ax, float* ay,
                            int* N_out) {
    EplC C;
    epl_prepare(p[0], p[1], p[2], p[3], p[4], p[5], niter, -1, C);
    std::vector<float> a;
    std::vector<float4> b;
    coef_tables(C, a, b);
    EplState S;
    for (int i = 0; i < n; ++i) epl_forward<false>(C, a.data(), nullptr, x[i], y[i], ax[i], ay[i], S);
    *N_out = C.N;
}

// sum_i H_i . d alpha(x_i)/d params -> g[6]; also returns alpha of the adjoint path
extern "C" void h_epl_grad(const float* p, int niter, int n, const float* x, const float* y, const float* Hx,
                           const float* Hy, float* g, float* ax, float* ay) {
    EplC C;
    epl_prepare(p[0], p[1], p[2], p[3], p[4], p[5], niter, -1, C);
    std::vector<float> a;
    std::vector<float4> b;
    coef_tables(C, a, b);
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
        EplState S;
        epl_forward<true>(C, a.data(), b.data(), x[i], y[i], ax[i], ay[i], S);
        epl_adjoint(C, S, Hx[i], Hy[i], acc);
    }
    epl_finish(C, acc, g);
}

extern "C" void h_sersic(const float* p, int elliptical, int n, const float* x, const float* y, const float* gup,
                         float* I, float* gpar, float* gdx, float* gdy) {
    SerC C;
    if (elliptical) sersic_prepare(p[0], p[1], p[2], p[3], p[4], p[5], p[6], 1, C);
    else sersic_prepare(p[0], p[1], 0.f, 0.f, p[2], p[3], p[4], 0, C);
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
        SerState S;
        I[i] = sersic_forward<true>(C, x[i], y[i], S);
        sersic_adjoint(C, S, gup[i], acc, gdx[i], gdy[i]);
    }
    sersic_finish(C, acc, gpar);
}

extern "C" void h_sis(const float* p, int n, const float* x, const float* y, const float* Hx, const float* Hy,
                      float* ax, float* ay, float* g) {
    SisC C{p[0], p[1], p[2]};
    float acc[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) {
        float st[3];
        sis_forward(C, x[i], y[i], ax[i], ay[i], st);
        sis_adjoint(C, st, Hx[i], Hy[i], acc);
    }
    sis_finish(acc, g);
}

extern "C" void h_sie(const float* p, int n, const float* x, const float* y, const float* Hx, const float* Hy,
                      float* ax, float* ay, float* g) {
    SieC C;
    sie_prepare(p[0], p[1], p[2], p[3], p[4], C);
    float acc[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
        SieState S;
        sie_forward<true>(C, x[i], y[i], ax[i], ay[i], S);
        sie_adjoint(C, S, Hx[i], Hy[i], acc);
    }
    sie_finish(C, acc, g);
}

extern "C" void h_tnfw(const float* p, int n, const float* x, const float* y, const float* Hx, const float* Hy,
                       float* ax, float* ay, float* g) {
    TnfwC C;
    tnfw_prepare(p[0], p[1], p[2], p[3], p[4], C);
    float acc[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
        TnfwState S;
        tnfw_forward<true>(C, x[i], y[i], ax[i], ay[i], S);
        tnfw_adjoint(C, S, Hx[i], Hy[i], acc);
    }
    tnfw_finish(C, acc, g);
}

extern "C" void h_core_sersic(const float* p, int n, const float* x, const float* y, const float* gup, float* I,
                              float* gpar, float* gdx, float* gdy) {
    CoreSerC C;
    core_sersic_prepare(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], C);
    float acc[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
        CoreSerState S;
        I[i] = core_sersic_forward(C, x[i], y[i], S);
        core_sersic_adjoint(C, S, gup[i], acc, gdx[i], gdy[i]);
    }
    core_sersic_finish(C, acc, gpar);
}

extern "C" void h_prior(int kind, const float* args, float z, int fldj_at_x, float* out) {
    PriorOut o;
    prior_eval(kind, args, z, fldj_at_x, o);
    out[0] = o.x; out[1] = o.dxdz; out[2] = o.lp; out[3] = o.dlpdx; out[4] = o.fldj; out[5] = o.dfldjdz;
}

extern "C" void h_lik(int mode, float im, float obs, float a, float b, float* out) {
    float ll, r2, dll, dr2;
    if (mode == 0) lik_forward(im, obs, a, b, ll, r2, dll, dr2);
    else lik_backward(im, obs, a, b, ll, r2, dll, dr2);
    out[0] = ll; out[1] = r2; out[2] = dll; out[3] = dr2;
}
