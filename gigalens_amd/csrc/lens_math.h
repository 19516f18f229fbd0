// Per-particle and per-sub-pixel arithmetic of the lens model, forward and
// adjoint.  All functions are __host__ __device__ so the same code is unit
// tested on the CPU and used by the gfx950 kernels.
//
// Reference semantics (file:line in the reference repository):
//   EPL            src/gigalens/tf/profiles/mass/epl.py:23-68
//   Shear          src/gigalens/tf/profiles/mass/shear.py:12-14
//   SIS            src/gigalens/tf/profiles/mass/sis.py:14-18
//   Sersic(Ellip.) src/gigalens/tf/profiles/light/sersic.py:21-72
//   likelihood     src/gigalens/tf/model.py:77-86, 113-118, 159-163
//   prior/bijector src/gigalens/tf/model.py:40-51,75,81-83 (TFP default event-space bijectors)
//
// Formulation (own design; the reference evaluates these as unfused tensor ops):
//  * EPL: the series sum_n Omega_n = e^{i phi} sum_n c_n w^n with w = e^{2 i phi}
//    and per-particle real coefficients c_n is evaluated by a Clenshaw
//    recurrence (2 flops / term); cos/sin phi come from q x / R and y / R
//    instead of atan2/cos/sin; (b/R)^(t-1) is one exp2/log2 pair.
//  * Adjoint: reverse mode with per-particle "sufficient statistics": every
//    sub-pixel adds a handful of scalars (e.g. h . alpha, h . alpha log R) to
//    per-thread accumulators, and the per-particle Jacobians (d/de1, d/dq, ...)
//    are applied once in the epilogue.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define GLM __host__ __device__ __forceinline__

namespace glm {

constexpr float kLog2e = 1.4426950408889634f;
constexpr float kLn2 = 0.6931471805599453f;
constexpr float kHalfLn2Pi = 0.91893853320467274f;

// ---- transcendental wrappers: hardware instructions on the device --------
GLM float ex2(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_exp2f(x);
#else
    return exp2f(x);
#endif
}
GLM float lg2(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_logf(x);
#else
    return log2f(x);
#endif
}
GLM float rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
GLM float rsq(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}

// ---------------------------------------------------------------------------
// Component constants (prepared once per particle)
// ---------------------------------------------------------------------------
enum Kind { K_EPL = 1, K_SHEAR = 2, K_SIS = 3, K_SIE = 4, K_TNFW = 5, K_SERSIC = 10, K_SERSIC_ELLIPSE = 11,
            K_SHAPELETS = 12, K_CORE_SERSIC = 13 };

// float slots of a component's constant block
struct EplC {
    float cx, cy, cphi, sphi, q, f, t, tm1, K0;
    float thetaE, e1, e2, c, lnb, craw;
    int N;
};
struct SerC {
    float cx, cy, Axx, Axy, Ayy, k1, k0, m1, m0, Ie, k1m1;
    float bn, n, Rs, e1, e2, q, cphi, sphi, c;
    int elliptical;
};
struct ShearC { float g1, g2; };
struct SisC { float thetaE, cx, cy; };
struct SieC {
    float cx, cy, b, q, k2, cphi, sphi, thetaE, e1, e2, c, craw, sq;
    int clipped;
};
struct TnfwC {
    float cx, cy, Rs, alpha_Rs, rt, tau, T, P, dP, K0, K, xmin;
};
struct CoreSerC {
    float cx, cy, Axx, Axy, Ayy, q, cphi, sphi, c, e1, e2;
    float Rs, n, Rb, al, gam, Ie, lnRs, lnRb, Rba, iRsa, kexp;
};

// Series length rule of epl.py:41,51-58 for a given f: the loop runs for
// n = 1, 2, ... while n < log(1e-12)/log(f) + 2, at most niter times.
GLM int epl_series_len(double f, int niter) {
    if (!(f > 0.0)) {
        // log(0) = -inf -> limit 2 -> one iteration; NaN -> zero iterations
        if (f == 0.0) return niter < 1 ? niter : 1;
        return 0;
    }
    double lim = log(1e-12) / log(f) + 2.0;  // f<1: positive; f>=1: -inf / negative
    if (!(lim > 1.0)) return 0;
    double n = ceil(lim) - 1.0;
    if (n > niter) n = niter;
    return (int)n;
}

// EPL constants, epl.py:25-30,37,59.  b = theta_E sqrt(q) is the algebraic
// simplification of epl.py:28-29.  The prefactor 2b/(1+q) (b/R)^(t-1) is evaluated
// as exp2(K0 + (1-t) log2 R); at b = 0 (|e| >= 1 clips q to 0) the reference's product
// 0 * 0^(t-1) is NaN for t < 1 (gamma < 2) and 0 otherwise, so K0 is NaN there too.
GLM void epl_prepare(float thetaE, float gamma, float e1, float e2, float cx, float cy, int niter,
                     int n_override, EplC& C) {
    float phi = atan2f(e2, e1) * 0.5f;
    float craw = sqrtf(e1 * e1 + e2 * e2);
    float c = fminf(fmaxf(craw, 0.0f), 1.0f);
    float q = (1.0f - c) / (1.0f + c);
    float b = thetaE * sqrtf(q);
    float t = gamma - 1.0f;
    C.cx = cx; C.cy = cy;
    C.cphi = cosf(phi); C.sphi = sinf(phi);
    C.q = q;
    C.f = (1.0f - q) / (1.0f + q);
    C.t = t; C.tm1 = t - 1.0f;
    C.lnb = logf(b);
    C.K0 = log2f(2.0f / (1.0f + q)) + t * log2f(b);
    if (b == 0.0f && C.tm1 < 0.0f) C.K0 = __builtin_nanf("");
    C.thetaE = thetaE; C.e1 = e1; C.e2 = e2; C.c = c; C.craw = craw;
    C.N = n_override >= 0 ? n_override : epl_series_len((double)C.f, niter);
}

// Coefficient n of the three series (value, d/dt, d/df):
//   c_n   = prod_{k<=n} (-f)(2k-2+t)/(2k+2-t),  c_0 = 1
//   dt_n  = d c_n / dt,  df_n = d c_n / df  (recurrences avoid 1/f, 1/(2k-2+t)).
GLM void epl_coef(int n, double f, double t, float& c, float& dt, float& df) {
    double a = 1.0, da = 0.0, pw = 1.0;  // a_n (without f), d a_n/dt, (-f)^(n-1)
    for (int k = 1; k <= n; ++k) {
        double den = 2.0 * k + 2.0 - t;
        double r = (2.0 * k - 2.0 + t) / den;
        double dr = 4.0 * k / (den * den);
        da = da * r + a * dr;
        a = a * r;
        if (k > 1) pw *= -f;
    }
    if (n == 0) { c = 1.0f; dt = 0.0f; df = 0.0f; return; }
    c = (float)(-f * pw * a);
    dt = (float)(-f * pw * da);
    df = (float)(-(double)n * pw * a);
}

// Sersic constants, sersic.py:33-59 (no clip on |e|), sersic.py:66-72.
GLM void sersic_prepare(float Rs, float n, float e1, float e2, float cx, float cy, float Ie, int elliptical,
                        SerC& C) {
    C.cx = cx; C.cy = cy; C.Rs = Rs; C.n = n; C.Ie = Ie; C.e1 = e1; C.e2 = e2;
    C.elliptical = elliptical;
    if (elliptical) {
        float phi = atan2f(e2, e1) * 0.5f;
        float c = sqrtf(e1 * e1 + e2 * e2);
        float q = (1.0f - c) / (1.0f + c);
        float cp = cosf(phi), sp = sinf(phi);
        C.Axx = q * cp * cp + sp * sp / q;
        C.Ayy = q * sp * sp + cp * cp / q;
        C.Axy = cp * sp * (q - 1.0f / q);
        C.q = q; C.cphi = cp; C.sphi = sp; C.c = c;
    } else {
        C.Axx = 1.0f; C.Ayy = 1.0f; C.Axy = 0.0f;
        C.q = 1.0f; C.cphi = 1.0f; C.sphi = 0.0f; C.c = 0.0f;
    }
    C.bn = 1.9992f * n - 0.3271f;
    C.k1 = 0.5f / n;
    C.k1m1 = C.k1 - 1.0f;
    C.k0 = -log2f(Rs) / n;
    C.m1 = -C.bn * kLog2e;
    C.m0 = C.bn * kLog2e;
}

// ---------------------------------------------------------------------------
// Per sub-pixel forward
// ---------------------------------------------------------------------------
struct EplState {
    float u, v, qu, ir, r2c, L, A, cp, sp, inrange;
    float Fr, Fi, aLu, aLv;
    float Etr, Eti, Efr, Efi;
};

// Coefficient tables of one EPL lens: c_n (floats), then (c_n, dc_n/dt, dc_n/df, 0)
// (float4s), each holding indices 0..epl_pad(N) with zeros above N.  The Clenshaw
// loops below step four terms at a time from epl_pad(N) down; a zero coefficient
// above the last nonzero one keeps the recurrence state exactly 0, so the result
// equals the unpadded N-term recurrence bit for bit, without a remainder loop.
GLM int epl_pad(int n) { return (n + 3) & ~3; }
// floats of the c_n table for series lengths up to max_n (a multiple of 4, so the
// float4 table after it stays 16-byte aligned)
GLM int epl_tab_a_floats(int max_n) { return epl_pad(max_n) + 4; }
// floats of both tables of one lens
GLM int epl_tab_floats(int max_n) { return epl_tab_a_floats(max_n) + 4 * (epl_pad(max_n) + 1); }

// Clenshaw evaluation of sum_{n=0..N} c_n w^n, w = cos th + i sin th.
// coef: c_n (float), zero-padded to epl_pad(N).  Returns real/imag parts.
GLM void clenshaw1(const float* __restrict__ coef, int N, float c2, float s2, float& re, float& im) {
    float x2 = c2 + c2;
    float b1 = 0.0f, b2 = 0.0f;
#pragma unroll 1
    for (int k = epl_pad(N); k >= 4; k -= 4) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float b0 = fmaf(x2, b1, coef[k - j]) - b2;
            b2 = b1;
            b1 = b0;
        }
    }
    re = fmaf(c2, b1, coef[0]) - b2;
    im = s2 * b1;
}

// Three series at once; coef4[n] = (c_n, dc_n/dt, dc_n/df, 0), zero-padded to epl_pad(N).
GLM void clenshaw3(const float4* __restrict__ coef4, int N, float c2, float s2, float& re, float& im,
                   float& tre, float& tim, float& fre, float& fim) {
    float x2 = c2 + c2;
    float a1 = 0.f, a2 = 0.f, d1 = 0.f, d2 = 0.f, g1 = 0.f, g2 = 0.f;
#pragma unroll 1
    for (int k = epl_pad(N); k >= 4; k -= 4) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float4 cc = coef4[k - j];
            float a0 = fmaf(x2, a1, cc.x) - a2;
            float d0 = fmaf(x2, d1, cc.y) - d2;
            float g0 = fmaf(x2, g1, cc.z) - g2;
            a2 = a1; a1 = a0; d2 = d1; d1 = d0; g2 = g1; g1 = g0;
        }
    }
    re = fmaf(c2, a1, coef4[0].x) - a2;
    im = s2 * a1;
    tre = c2 * d1 - d2;  // dc_0/dt = 0
    tim = s2 * d1;
    fre = c2 * g1 - g2;
    fim = s2 * g1;
}

// EPL deflection (image frame).  Fills the state needed by the adjoint when
// coef4 != nullptr (then the two derivative series are evaluated too).
template <bool ADJ>
GLM void epl_forward(const EplC& C, const float* __restrict__ coef, const float4* __restrict__ coef4, float X,
                     float Y, float& ax, float& ay, EplState& S) {
    float dx = X - C.cx, dy = Y - C.cy;
    float u = fmaf(dx, C.cphi, dy * C.sphi);
    float v = fmaf(-dx, C.sphi, dy * C.cphi);
    float qu = C.q * u;
    float r2 = fmaf(qu, qu, v * v);
    float cp, sp, ir = 0.0f;
    if (r2 > 0.0f) {
        ir = rsq(r2);
        cp = qu * ir;
        sp = v * ir;
    } else {  // atan2(0, +-0)
        cp = signbit(qu) ? -1.0f : 1.0f;
        sp = 0.0f;
    }
    float r2c = fminf(fmaxf(r2, 1e-20f), 1e20f);
    float L = lg2(r2c);
    float A = ex2(fmaf(-C.tm1 * 0.5f, L, C.K0));
    float c2 = fmaf(cp, cp, -sp * sp);
    float s2 = 2.0f * cp * sp;
    float gr, gi;
    if (ADJ) {
        float tr, ti, fr, fi;
        clenshaw3(coef4, C.N, c2, s2, gr, gi, tr, ti, fr, fi);
        S.Etr = cp * tr - sp * ti; S.Eti = sp * tr + cp * ti;
        S.Efr = cp * fr - sp * fi; S.Efi = sp * fr + cp * fi;
    } else {
        clenshaw1(coef, C.N, c2, s2, gr, gi);
    }
    float Fr = cp * gr - sp * gi;
    float Fi = sp * gr + cp * gi;
    float aLu = A * Fr, aLv = A * Fi;
    ax = aLu * C.cphi - aLv * C.sphi;
    ay = aLu * C.sphi + aLv * C.cphi;
    if (ADJ) {
        S.u = u; S.v = v; S.qu = qu; S.ir = ir; S.r2c = r2c; S.L = L; S.A = A; S.cp = cp; S.sp = sp;
        S.inrange = (r2 >= 1e-20f && r2 <= 1e20f) ? 1.0f : 0.0f;
        S.Fr = Fr; S.Fi = Fi; S.aLu = aLu; S.aLv = aLv;
    }
}

// EPL adjoint accumulation; H = dL/d alpha_image at this sub-pixel.
// acc[0..7] = (sum h.a, sum h.a log2 R^2, S_t, S_f, S_q, S_u, S_v, S_phi).
GLM void epl_adjoint(const EplC& C, const EplState& S, float Hx, float Hy, float* acc) {
    float hu = fmaf(Hx, C.cphi, Hy * C.sphi);
    float hv = fmaf(-Hx, C.sphi, Hy * C.cphi);
    float sa = fmaf(hu, S.aLu, hv * S.aLv);
    acc[0] += sa;
    acc[1] = fmaf(sa, S.L, acc[1]);
    acc[2] = fmaf(S.A, fmaf(hu, S.Etr, hv * S.Eti), acc[2]);
    float sf = fmaf(hu, S.Efr, hv * S.Efi);
    acc[3] = fmaf(S.A, sf, acc[3]);
    float f2 = C.f + C.f;
    float Yr = fmaf(f2, S.Efr, S.Fr);
    float Yi = fmaf(f2, S.Efi, S.Fi);
    float sphi = S.A * fmaf(-hu, Yi, hv * Yr);
    float ir2 = S.ir * S.ir;  // 1/r^2, 0 at r = 0
    float sR = (1.0f - C.t) * sa * S.inrange;
    float dLdu = (sR * C.q * S.qu - sphi * C.q * S.v) * ir2;
    float dLdv = (sR * S.v + sphi * S.qu) * ir2;
    acc[4] = fmaf((sR * S.qu - sphi * S.v) * S.u, ir2, acc[4]);
    acc[5] += dLdu;
    acc[6] += dLdv;
    acc[7] += dLdu * S.v - dLdv * S.u + hv * S.aLu - hu * S.aLv;
}

// Epilogue: statistics -> d/d(theta_E, gamma, e1, e2, center_x, center_y).
GLM void epl_finish(const EplC& C, const float* acc, float* g) {
    float t = C.t, q = C.q;
    float dK = acc[0];
    g[0] = dK * t / C.thetaE;
    g[1] = dK * C.lnb - 0.5f * kLn2 * acc[1] + acc[2];
    float dq = dK * (t / (2.0f * q) - 1.0f / (1.0f + q)) + acc[4];
    float dc = (C.craw <= 1.0f) ? (dq * (-2.0f / ((1.0f + C.c) * (1.0f + C.c))) + acc[3]) : 0.0f;
    float dphi = acc[7];
    float cc = C.e1 * C.e1 + C.e2 * C.e2;
    float inv2c2 = 0.5f / cc;
    float ic = 1.0f / C.craw;
    g[2] = dphi * (-C.e2 * inv2c2) + dc * C.e1 * ic;
    g[3] = dphi * (C.e1 * inv2c2) + dc * C.e2 * ic;
    g[4] = -(acc[5] * C.cphi - acc[6] * C.sphi);
    g[5] = -(acc[5] * C.sphi + acc[6] * C.cphi);
}

GLM void shear_forward(const ShearC& C, float X, float Y, float& ax, float& ay) {
    ax = fmaf(C.g1, X, C.g2 * Y);
    ay = fmaf(C.g2, X, -C.g1 * Y);
}
GLM void shear_adjoint(float X, float Y, float Hx, float Hy, float* acc) {
    acc[0] += Hx * X - Hy * Y;
    acc[1] += Hx * Y + Hy * X;
}

GLM void sis_forward(const SisC& C, float X, float Y, float& ax, float& ay, float* st) {
    float dx = X - C.cx, dy = Y - C.cy;
    float r2 = dx * dx + dy * dy;
    float ir = r2 > 0.0f ? rsq(r2) : 0.0f;
    float a = C.thetaE * ir;
    ax = a * dx; ay = a * dy;
    st[0] = dx * ir; st[1] = dy * ir; st[2] = ir;
}
// acc: (sum H.n, S_dx, S_dy) with dL/d(dx,dy) = thetaE/R (H - n (n.H))
GLM void sis_adjoint(const SisC& C, const float* st, float Hx, float Hy, float* acc) {
    float nh = st[0] * Hx + st[1] * Hy;
    acc[0] += nh;
    float s = C.thetaE * st[2];
    acc[1] += s * (Hx - st[0] * nh);
    acc[2] += s * (Hy - st[1] * nh);
}
GLM void sis_finish(const float* acc, float* g) {
    g[0] = acc[0];
    g[1] = -acc[1];
    g[2] = -acc[2];
}

struct SerState { float dx, dy, gx, gy, Pq, LQ, P, I; };

// ADJ: also P/Q for sersic_adjoint, as one exp2 of (k1 - 1) log2 Q + k0 with P = Q (P/Q)
// (no reciprocal of Q).
template <bool ADJ = false>
GLM float sersic_forward(const SerC& C, float X, float Y, SerState& S) {
    float dx = X - C.cx, dy = Y - C.cy;
    float gx = fmaf(C.Axx, dx, C.Axy * dy);
    float gy = fmaf(C.Axy, dx, C.Ayy * dy);
    float Q = fmaf(dx, gx, dy * gy);
    float LQ = lg2(Q);
    float P, Pq = 0.0f;
    if (ADJ) {
        // Q > 0: P = Q Pq.  Q < 0 (|e| > 1, q < 0): NaN like the forward pass, so the
        // sub-pixel is skipped as zeroed.  Q = 0 (a sub-pixel on the centre): P = 0 and the
        // radial terms vanish -- R is even in (dx, dy), so the symmetric derivative there
        // is 0 (its one-sided derivatives are infinite for n > 1/2).
        Pq = Q > 0.0f ? ex2(fmaf(C.k1m1, LQ, C.k0)) : 0.0f;
        P = Q > 0.0f ? Q * Pq : (Q == 0.0f ? 0.0f : __builtin_nanf(""));
        LQ = Q > 0.0f ? LQ : 0.0f;
    } else {
        P = ex2(fmaf(C.k1, LQ, C.k0));
    }
    float I = C.Ie * ex2(fmaf(C.m1, P, C.m0));
    S.dx = dx; S.dy = dy; S.gx = gx; S.gy = gy; S.Pq = Pq; S.LQ = LQ; S.P = P; S.I = I;
    return I;
}

// acc[0..7] = (S_I, S_P, S_PL, S_xx, S_xy, S_yy, S_dx, S_dy); returns dL/d(dx,dy)
// of this sub-pixel in (gdx, gdy) (needed to chain a source into the lenses).
// S comes from sersic_forward<true>.
GLM void sersic_adjoint(const SerC& C, const SerState& S, float g, float* acc, float& gdx, float& gdy) {
    float gI = g * S.I;
    float gP = gI * S.P;
    acc[0] += gI;
    acc[1] += gP;
    acc[2] = fmaf(gP, S.LQ, acc[2]);
    float w = gI * S.Pq;  // g I P / Q
    acc[3] = fmaf(w * S.dx, S.dx, acc[3]);
    acc[4] = fmaf(w * S.dx, S.dy, acc[4]);
    acc[5] = fmaf(w * S.dy, S.dy, acc[5]);
    acc[6] = fmaf(w, S.gx, acc[6]);
    acc[7] = fmaf(w, S.gy, acc[7]);
    float Ks2 = -2.0f * C.bn * C.k1;
    gdx = Ks2 * w * S.gx;
    gdy = Ks2 * w * S.gy;
}

// d/d(e1, e2) from d/d(Axx, Axy, Ayy) of the quadratic form Q = A_xx dx^2 +
// 2 A_xy dx dy + A_yy dy^2 of the elliptical distance (sersic.py:33-59), through
// q = (1 - c)/(1 + c), c = |e| and phi = atan2(e2, e1)/2.
GLM void ellipse_finish(float q, float cp, float sp, float c, float e1, float e2, float dAxx, float dAxy,
                        float dAyy, float& ge1, float& ge2) {
    float s2p = 2.0f * sp * cp, c2p = cp * cp - sp * sp;
    float iq = 1.0f / q, iq2 = iq * iq;
    float dq = dAxx * (cp * cp - sp * sp * iq2) + dAyy * (sp * sp - cp * cp * iq2) + dAxy * 0.5f * s2p * (1.0f + iq2);
    float dphi = dAxx * s2p * (iq - q) + dAyy * s2p * (q - iq) + dAxy * c2p * (q - iq);
    float dc = dq * (-2.0f / ((1.0f + c) * (1.0f + c)));
    float cc = e1 * e1 + e2 * e2;
    ge1 = dphi * (-e2 * 0.5f / cc) + dc * e1 / c;
    ge2 = dphi * (e1 * 0.5f / cc) + dc * e2 / c;
}

// Epilogue: g receives d/d(R_sersic, n_sersic, e1, e2, center_x, center_y, Ie)
// (e1,e2 only for the elliptical profile).
GLM void sersic_finish(const SerC& C, const float* acc, float* g) {
    float bn = C.bn, n = C.n;
    float dbn = -(acc[1] - acc[0]);
    float dk0 = -bn * kLn2 * acc[1];
    float dk1 = -bn * kLn2 * acc[2];
    float Ks = -bn * C.k1;
    g[0] = dk0 * (-1.0f / (n * C.Rs * kLn2));
    g[1] = 1.9992f * dbn + dk1 * (-0.5f / (n * n)) + dk0 * (log2f(C.Rs) / (n * n));
    if (C.elliptical) {
        ellipse_finish(C.q, C.cphi, C.sphi, C.c, C.e1, C.e2, Ks * acc[3], 2.0f * Ks * acc[4], Ks * acc[5], g[2],
                       g[3]);
    } else {
        g[2] = 0.0f; g[3] = 0.0f;
    }
    g[4] = -2.0f * Ks * acc[6];
    g[5] = -2.0f * Ks * acc[7];
    g[6] = acc[0] / C.Ie;
}

// ---------------------------------------------------------------------------
// SIE, src/gigalens/tf/profiles/mass/sie.py:13-43 (s_scale = 0).
//   c = min(|e|, 0.9999), q = (1-c)/(1+c), b = theta_E sqrt(q) (the algebraic form of
//   sie.py:19-21), k^2 = 1 - q^2 = 4c/(1+c)^2, (u, v) = R(phi)(x - c),
//   psi = sqrt(q^2 u^2 + v^2), f_x = b/k atan(k u/psi), f_y = b/k atanh(k v/psi).
// With s = u/psi (resp. v/psi) the two are written b s (1 + k^2 P_k) with
//   P_k = (atan(k s)/(k s) - 1)/k^2,  Q_k = (atanh(k s)/(k s) - 1)/k^2,
// which have the finite limits -s^2/3 and s^2/3 at k -> 0 (q = 1, where the
// reference's expression is 0/0); power series below |k s| = 1/4.
// ---------------------------------------------------------------------------
GLM void sie_prepare(float thetaE, float e1, float e2, float cx, float cy, SieC& C) {
    float phi = atan2f(e2, e1) * 0.5f;
    float craw = sqrtf(e1 * e1 + e2 * e2);
    float c = fminf(craw, 0.9999f);
    float q = (1.0f - c) / (1.0f + c);
    C.cx = cx; C.cy = cy;
    C.q = q;
    C.sq = sqrtf(q);
    C.b = thetaE * C.sq;
    C.k2 = 4.0f * c / ((1.0f + c) * (1.0f + c));
    C.cphi = cosf(phi); C.sphi = sinf(phi);
    C.thetaE = thetaE; C.e1 = e1; C.e2 = e2; C.c = c; C.craw = craw;
    C.clipped = craw > 0.9999f ? 1 : 0;
}

// r = atan(A)/A (sign = -1) or atanh(A)/A (sign = +1) at A^2 = k2 s2, and
// pk = (r - 1)/k2.
GLM void sie_ratio(float k2, float s2, float sign, float& r, float& pk) {
    const float A2 = k2 * s2;
    if (A2 < 0.0625f) {
        // (r - 1)/A^2 = sum_{n>=1} sign^n A^(2n-2)/(2n+1), through n = 7
        const float a = sign * A2;
        float t = fmaf(a, 1.0f / 15.0f, 1.0f / 13.0f);
        t = fmaf(a, t, 1.0f / 11.0f);
        t = fmaf(a, t, 1.0f / 9.0f);
        t = fmaf(a, t, 1.0f / 7.0f);
        t = fmaf(a, t, 1.0f / 5.0f);
        t = fmaf(a, t, 1.0f / 3.0f);
        pk = sign * s2 * t;
        r = fmaf(k2, pk, 1.0f);
    } else {
        const float A = sqrtf(A2);
        r = (sign < 0.0f ? atanf(A) : atanhf(A)) / A;
        pk = (r - 1.0f) / k2;
    }
}

struct SieState { float u, v, psi, fx, fy, dqx, dqy; };

template <bool ADJ>
GLM void sie_forward(const SieC& C, float X, float Y, float& ax, float& ay, SieState& S) {
    float dx = X - C.cx, dy = Y - C.cy;
    float u = fmaf(dx, C.cphi, dy * C.sphi);
    float v = fmaf(-dx, C.sphi, dy * C.cphi);
    float psi = sqrtf(fmaf(C.q * u, C.q * u, v * v));
    float ip = 1.0f / psi;
    float su = u * ip, sv = v * ip;
    float ru, pu, rv, pv;
    sie_ratio(C.k2, su * su, -1.0f, ru, pu);
    sie_ratio(C.k2, sv * sv, 1.0f, rv, pv);
    float fx = C.b * su * ru, fy = C.b * sv * rv;
    ax = fx * C.cphi - fy * C.sphi;
    ay = fx * C.sphi + fy * C.cphi;
    if (ADJ) {
        S.u = u; S.v = v; S.psi = psi; S.fx = fx; S.fy = fy;
        // d f / d q at fixed b
        S.dqx = C.q * C.b * su * pu;
        S.dqy = C.q * C.b * sv * pv - C.b * sv / C.q;
    }
}

// acc[0..4] = (sum h.f, S_u, S_v, S_phi, S_q), h = R(phi) H:
//   J_f = b/(psi r^2) [[v^2, -uv], [-uv, u^2]],  (S_u, S_v) += J_f^T h,
//   S_phi += h_v f_x - h_u f_y + b (v h_u - u h_v)/psi,  S_q += h . df/dq.
GLM void sie_adjoint(const SieC& C, const SieState& S, float Hx, float Hy, float* acc) {
    float hu = fmaf(Hx, C.cphi, Hy * C.sphi);
    float hv = fmaf(-Hx, C.sphi, Hy * C.cphi);
    acc[0] += fmaf(hu, S.fx, hv * S.fy);
    float cr = fmaf(S.v, hu, -S.u * hv);
    float bp = C.b / S.psi;
    float w = bp * cr / fmaf(S.u, S.u, S.v * S.v);
    acc[1] = fmaf(w, S.v, acc[1]);
    acc[2] = fmaf(-w, S.u, acc[2]);
    acc[3] += fmaf(hv, S.fx, fmaf(-hu, S.fy, bp * cr));
    acc[4] += fmaf(hu, S.dqx, hv * S.dqy);
}

// g = d/d(theta_E, e1, e2, center_x, center_y)
GLM void sie_finish(const SieC& C, const float* acc, float* g) {
    float q = C.q;
    float dLdb = acc[0] / C.b;
    float dq = acc[4] + dLdb * C.b / (2.0f * q);
    g[0] = dLdb * C.sq;
    float dc = C.clipped ? 0.0f : dq * (-2.0f / ((1.0f + C.c) * (1.0f + C.c)));
    float dphi = acc[3];
    float cc = C.e1 * C.e1 + C.e2 * C.e2;
    g[1] = dphi * (-C.e2 * 0.5f / cc) + dc * C.e1 / C.craw;
    g[2] = dphi * (C.e1 * 0.5f / cc) + dc * C.e2 / C.craw;
    g[3] = -(acc[1] * C.cphi - acc[2] * C.sphi);
    g[4] = -(acc[1] * C.sphi + acc[2] * C.cphi);
}

// ---------------------------------------------------------------------------
// Truncated NFW, src/gigalens/tf/profiles/mass/tnfw.py:14-58:
//   alpha = K phi(x, tau) (dx, dy),  phi = g/x^2,  x = max(R, 0.001 Rs)/Rs,
//   tau = r_trunc/Rs, K = alpha_Rs/(Rs (1 + ln 1/2)),  g = P(tau) Gb,
//   P = tau^2/(tau^2+1)^2,
//   Gb = (tau^2+1+2(x^2-1)) F + tau pi + (tau^2-1) ln tau + Sq (-pi + L (tau^2-1)/tau),
//   Sq = sqrt(tau^2+x^2), L = ln(x/(tau+Sq)).
// The terms of Gb are O(tau^2 ln x) and cancel to O(x^2 ln x) at small x, so
// this evaluates the same function in the regrouped form
//   Gb = (tau^2-1) B1 + x^2 C1,
//   B1 = [F + ln(x/2)] + ln(2 tau/(tau+Sq)) + D L,   D = Sq/tau - 1 = x^2/(tau (tau+Sq)),
//   C1 = 2F - pi/(tau+Sq),
// with F + ln(x/2) = [ln((1+s)/2) + x^2 ln(2/x)/(1+s)]/s for x < 1 (s^2 = 1-x^2),
// and the derivatives
//   dB1/dx = x [(F-1)/s^2 + L/(tau Sq)],  dB1/dtau = -x^2 L/(tau^2 Sq),
//   dC1/dx = 2F' + pi x/(Sq (tau+Sq)^2),  dC1/dtau = pi/(Sq (tau+Sq)).
// F(x) = sum_n s2^n/(2n+1) (s2 = 1 - x^2) near x = 1, else atanh/atan forms.
// ---------------------------------------------------------------------------
GLM void tnfw_prepare(float Rs, float alpha_Rs, float rt, float cx, float cy, TnfwC& C) {
    C.cx = cx; C.cy = cy; C.Rs = Rs; C.alpha_Rs = alpha_Rs; C.rt = rt;
    float tau = rt / Rs;
    float T = tau * tau;
    C.tau = tau; C.T = T;
    C.P = T / ((T + 1.0f) * (T + 1.0f));
    C.dP = 2.0f * tau * (1.0f - T) / ((T + 1.0f) * (T + 1.0f) * (T + 1.0f));
    C.K0 = 1.0f / (Rs * (1.0f - kLn2));
    C.K = alpha_Rs * C.K0;
    C.xmin = 0.001f;
}

// F(x), (F - 1)/s2 and F'(x) with s2 = 1 - x^2.
GLM void tnfw_F(float x, float& F, float& Fm1, float& dF, float& FL2) {
    const float s2 = (1.0f - x) * (1.0f + x);
    if (fabsf(s2) < 0.1f) {
        // F = sum_{n>=0} s2^n/(2n+1); (F-1)/s2 and dF/ds2 through n = 9
        float t = fmaf(s2, 1.0f / 19.0f, 1.0f / 17.0f);
        t = fmaf(s2, t, 1.0f / 15.0f);
        t = fmaf(s2, t, 1.0f / 13.0f);
        t = fmaf(s2, t, 1.0f / 11.0f);
        t = fmaf(s2, t, 1.0f / 9.0f);
        t = fmaf(s2, t, 1.0f / 7.0f);
        t = fmaf(s2, t, 1.0f / 5.0f);
        t = fmaf(s2, t, 1.0f / 3.0f);
        Fm1 = t;
        F = fmaf(s2, t, 1.0f);
        float d = fmaf(s2, 9.0f / 19.0f, 8.0f / 17.0f);
        d = fmaf(s2, d, 7.0f / 15.0f);
        d = fmaf(s2, d, 6.0f / 13.0f);
        d = fmaf(s2, d, 5.0f / 11.0f);
        d = fmaf(s2, d, 4.0f / 9.0f);
        d = fmaf(s2, d, 3.0f / 7.0f);
        d = fmaf(s2, d, 2.0f / 5.0f);
        d = fmaf(s2, d, 1.0f / 3.0f);
        dF = -2.0f * x * d;
        FL2 = F + logf(0.5f * x);
    } else if (x < 1.0f) {
        const float s = sqrtf(s2);
        const float l2x = logf(2.0f / x);
        F = (logf(1.0f + s) - logf(x)) / s;
        Fm1 = (F - 1.0f) / s2;
        dF = -(1.0f - x * x * F) / (x * s2);
        FL2 = (log1pf(-x * x / (2.0f * (1.0f + s))) + x * x * l2x / (1.0f + s)) / s;
    } else {
        const float s = sqrtf(-s2);
        F = atanf(s) / s;
        Fm1 = (F - 1.0f) / s2;
        dF = -(1.0f - x * x * F) / (x * s2);
        FL2 = F + logf(0.5f * x);
    }
}

struct TnfwState { float dx, dy, R, x, phi, phix, phit; int clamped; };

template <bool ADJ>
GLM void tnfw_forward(const TnfwC& C, float X, float Y, float& ax, float& ay, TnfwState& S) {
    const float kPi = 3.14159265358979f;
    float dx = X - C.cx, dy = Y - C.cy;
    float R = sqrtf(fmaf(dx, dx, dy * dy));
    float x = R / C.Rs;
    int clamped = 0;
    if (!(x >= C.xmin)) { x = C.xmin; clamped = 1; }
    float F, Fm1, dF, FL2;
    tnfw_F(x, F, Fm1, dF, FL2);
    const float tau = C.tau, x2 = x * x;
    const float Sq = sqrtf(fmaf(x, x, C.T));
    const float ts = tau + Sq;
    const float L = logf(x) - logf(ts);
    const float D = x2 / (tau * ts);
    const float B1 = FL2 - log1pf(x2 / (2.0f * tau * ts)) + D * L;
    const float C1 = 2.0f * F - kPi / ts;
    const float Gb = (C.T - 1.0f) * B1 + x2 * C1;
    const float ix2 = 1.0f / x2;
    const float phi = C.P * Gb * ix2;
    const float a = C.K * phi;
    ax = a * dx;
    ay = a * dy;
    if (ADJ) {
        const float dB1x = x * (Fm1 + L / (tau * Sq));
        const float dC1x = 2.0f * dF + kPi * x / (Sq * ts * ts);
        const float Gbx = (C.T - 1.0f) * dB1x + 2.0f * x * C1 + x2 * dC1x;
        const float dB1t = -x2 * L / (C.T * Sq);
        const float Gbt = 2.0f * tau * B1 + (C.T - 1.0f) * dB1t + x2 * kPi / (Sq * ts);
        S.dx = dx; S.dy = dy; S.R = R; S.x = x; S.phi = phi;
        S.phix = clamped ? 0.0f : C.P * ix2 * (Gbx - 2.0f * Gb / x);
        S.phit = (C.dP * Gb + C.P * Gbt) * ix2;
        S.clamped = clamped;
    }
}

// acc[0..4] = (sum phi H.d, sum x phi_x H.d, sum phi_tau H.d, S_cx, S_cy), all
// without the factor K; (S_cx, S_cy) = d(H.alpha)/d(center) / K.
GLM void tnfw_adjoint(const TnfwC& C, const TnfwState& S, float Hx, float Hy, float* acc) {
    float Hd = fmaf(Hx, S.dx, Hy * S.dy);
    acc[0] = fmaf(S.phi, Hd, acc[0]);
    acc[1] = fmaf(S.x * S.phix, Hd, acc[1]);
    acc[2] = fmaf(S.phit, Hd, acc[2]);
    float w = S.clamped ? 0.0f : S.phix * Hd / (S.R * C.Rs);
    acc[3] -= fmaf(S.phi, Hx, w * S.dx);
    acc[4] -= fmaf(S.phi, Hy, w * S.dy);
}

// g = d/d(Rs, alpha_Rs, r_trunc, center_x, center_y)
GLM void tnfw_finish(const TnfwC& C, const float* acc, float* g) {
    g[0] = -C.K * (acc[0] + acc[1] + C.tau * acc[2]) / C.Rs;
    g[1] = C.K0 * acc[0];
    g[2] = C.K * acc[2] / C.Rs;
    g[3] = C.K * acc[3];
    g[4] = C.K * acc[4];
}

// ---------------------------------------------------------------------------
// Core-Sersic light, src/gigalens/tf/profiles/light/sersic.py:75-121, as the
// reference writes it (R_sersic ** alpha ** 1.0 == R_sersic^alpha):
//   I = Ie (1 + (Rb/R)^a)^(g/a) exp(-bn (R^a + Rb^a) / (Rs^a a n) - 1),
//   bn = 1.9992 n - 0.3271, R the elliptical distance of sersic.py:33-59.
// ln I = ln Ie + (g/a) ln(1 + A) - kexp (R^a + Rb^a) Rs^-a - 1 with A = (Rb/R)^a,
// kexp = bn/(a n).
// ---------------------------------------------------------------------------
GLM void core_sersic_prepare(float Rs, float n, float Rb, float al, float gam, float e1, float e2, float cx,
                             float cy, float Ie, CoreSerC& C) {
    C.cx = cx; C.cy = cy; C.e1 = e1; C.e2 = e2;
    float phi = atan2f(e2, e1) * 0.5f;
    float c = sqrtf(e1 * e1 + e2 * e2);
    float q = (1.0f - c) / (1.0f + c);
    float cp = cosf(phi), sp = sinf(phi);
    C.Axx = q * cp * cp + sp * sp / q;
    C.Ayy = q * sp * sp + cp * cp / q;
    C.Axy = cp * sp * (q - 1.0f / q);
    C.q = q; C.cphi = cp; C.sphi = sp; C.c = c;
    C.Rs = Rs; C.n = n; C.Rb = Rb; C.al = al; C.gam = gam; C.Ie = Ie;
    C.lnRs = logf(Rs);
    C.lnRb = logf(Rb);
    C.Rba = expf(al * C.lnRb);
    C.iRsa = expf(-al * C.lnRs);
    C.kexp = (1.9992f * n - 0.3271f) / (al * n);
}

struct CoreSerState { float dx, dy, gx, gy, Q, lnR, Ra, A, L1, I; };

GLM float core_sersic_forward(const CoreSerC& C, float X, float Y, CoreSerState& S) {
    float dx = X - C.cx, dy = Y - C.cy;
    float gx = fmaf(C.Axx, dx, C.Axy * dy);
    float gy = fmaf(C.Axy, dx, C.Ayy * dy);
    float Q = fmaf(dx, gx, dy * gy);
    float lnR = 0.5f * logf(Q);
    float Ra = expf(C.al * lnR);
    float A = C.Rba / Ra;
    float L1 = log1pf(A);
    float lnI = fmaf(C.gam / C.al, L1, fmaf(-C.kexp * C.iRsa, Ra + C.Rba, -1.0f));
    float I = C.Ie * expf(lnI);
    S.dx = dx; S.dy = dy; S.gx = gx; S.gy = gy; S.Q = Q; S.lnR = lnR; S.Ra = Ra; S.A = A; S.L1 = L1; S.I = I;
    return I;
}

// acc[0..10] = (S_I, S_I L1, S_I A/(1+A), S_I R^a, S_I A lnR/(1+A), S_I R^a lnR,
//               S_xx, S_xy, S_yy, S_dx, S_dy) with S_I = sum g I; the geometric sums
// carry w = g I d lnI/dQ.  Returns dL/d(dx, dy) in (gdx, gdy).
GLM void core_sersic_adjoint(const CoreSerC& C, const CoreSerState& S, float g, float* acc, float& gdx,
                             float& gdy) {
    float gI = g * S.I;
    float t = S.A / (1.0f + S.A);
    acc[0] += gI;
    acc[1] = fmaf(gI, S.L1, acc[1]);
    acc[2] = fmaf(gI, t, acc[2]);
    acc[3] = fmaf(gI, S.Ra, acc[3]);
    acc[4] = fmaf(gI * t, S.lnR, acc[4]);
    acc[5] = fmaf(gI * S.Ra, S.lnR, acc[5]);
    // d lnI / d lnR = -g t - a kexp R^a Rs^-a;  dQ: d lnR = dQ / (2 Q)
    float dlnR = -C.gam * t - C.al * C.kexp * S.Ra * C.iRsa;
    float w = gI * dlnR * 0.5f / S.Q;
    acc[6] = fmaf(w * S.dx, S.dx, acc[6]);
    acc[7] = fmaf(w * S.dx, S.dy, acc[7]);
    acc[8] = fmaf(w * S.dy, S.dy, acc[8]);
    acc[9] = fmaf(w, S.gx, acc[9]);
    acc[10] = fmaf(w, S.gy, acc[10]);
    gdx = 2.0f * w * S.gx;
    gdy = 2.0f * w * S.gy;
}

// g = d/d(R_sersic, n_sersic, Rb, alpha, gamma, e1, e2, center_x, center_y, Ie)
GLM void core_sersic_finish(const CoreSerC& C, const float* acc, float* g) {
    const float al = C.al, ia = 1.0f / al;
    const float base = C.iRsa * (acc[3] + C.Rba * acc[0]);  // sum g I (R^a + Rb^a) Rs^-a
    g[0] = C.kexp * al * base / C.Rs;
    g[1] = -0.3271f * ia / (C.n * C.n) * base;
    g[2] = (C.gam * acc[2] - al * C.kexp * C.Rba * C.iRsa * acc[0]) / C.Rb;
    g[3] = -C.gam * ia * ia * acc[1] + C.gam * ia * (C.lnRb * acc[2] - acc[4]) +
           C.kexp * (base * (ia + C.lnRs) - C.iRsa * (acc[5] + C.Rba * C.lnRb * acc[0]));
    g[4] = acc[1] * ia;
    ellipse_finish(C.q, C.cphi, C.sphi, C.c, C.e1, C.e2, acc[6], 2.0f * acc[7], acc[8], g[5], g[6]);
    g[7] = -2.0f * acc[9];
    g[8] = -2.0f * acc[10];
    g[9] = acc[0] / C.Ie;
}

// ---------------------------------------------------------------------------
// Prior + bijector (per parameter), TFP semantics
// ---------------------------------------------------------------------------
// args: a0..a3, a4 = precomputed log normaliser (TruncatedNormal: log Z).
struct PriorOut { float x, dxdz, lp, dlpdx, fldj, dfldjdz; };

GLM float softplusf(float v) { return v > 20.0f ? v : log1pf(expf(v)); }
GLM float sigmoidf_(float v) { return 1.0f / (1.0f + expf(-v)); }

GLM void prior_eval(int kind, const float* a, float z, int fldj_at_x, PriorOut& o) {
    switch (kind) {
    case 1: {  // Normal, Identity
        o.x = z; o.dxdz = 1.0f;
        float r = (z - a[0]) / a[1];
        o.lp = -0.5f * r * r - logf(a[1]) - kHalfLn2Pi;
        o.dlpdx = -r / a[1];
        o.fldj = 0.0f; o.dfldjdz = 0.0f;
        break;
    }
    case 2: {  // LogNormal, Exp
        float x = expf(z);
        o.x = x; o.dxdz = x;
        float lx = logf(x);
        float r = (lx - a[0]) / a[1];
        o.lp = -0.5f * r * r - logf(a[1]) - kHalfLn2Pi - lx;
        o.dlpdx = (-r / a[1] - 1.0f) / x;
        if (fldj_at_x) { o.fldj = x; o.dfldjdz = x; }
        else { o.fldj = z; o.dfldjdz = 1.0f; }
        break;
    }
    case 3:
    case 4: {  // Uniform / TruncatedNormal, Sigmoid(low, high)
        float lo = kind == 3 ? a[0] : a[2];
        float hi = kind == 3 ? a[1] : a[3];
        float s = sigmoidf_(z), sm = sigmoidf_(-z);
        float x = hi * s + lo * sm;
        o.x = x;
        o.dxdz = (hi - lo) * s * sm;
        if (kind == 3) {
            bool in = (x >= lo) && (x < hi);
            o.lp = in ? -logf(hi - lo) : -INFINITY;
            o.dlpdx = 0.0f;
        } else {
            bool in = (x >= lo) && (x <= hi);
            float r = (x - a[0]) / a[1];
            o.lp = in ? (-0.5f * r * r - logf(a[1]) - kHalfLn2Pi - a[4]) : -INFINITY;
            o.dlpdx = -r / a[1];
        }
        float w = fldj_at_x ? x : z;
        o.fldj = logf(hi - lo) - softplusf(-w) - softplusf(w);
        float dw = 1.0f - 2.0f * sigmoidf_(w);
        o.dfldjdz = fldj_at_x ? dw * o.dxdz : dw;
        break;
    }
    default:
        o.x = z; o.dxdz = 1.0f; o.lp = 0.0f; o.dlpdx = 0.0f; o.fldj = 0.0f; o.dfldjdz = 0.0f;
    }
}

// ---------------------------------------------------------------------------
// Likelihood per pixel
// ---------------------------------------------------------------------------
// Forward model (sigma from the model image, tf/model.py:77-80):
// returns log-pdf without the -1/2 ln 2 pi constant; r2 = r^2; dll / dchi
// are d(log-pdf)/d im and d(r^2)/d im.
GLM void lik_forward(float im, float obs, float bkg2, float inv_texp, float& ll, float& r2, float& dll,
                     float& dr2) {
    float d = obs - im;
    float pos = im >= 0.0f ? 1.0f : 0.0f;
    float var = fmaf(fmaxf(im, 0.0f), inv_texp, bkg2);
    float iv = 1.0f / var;
    float d2iv = d * d * iv;
    r2 = d2iv;
    ll = -0.5f * d2iv - 0.5f * logf(var);
    float dv = pos * inv_texp;
    dll = d * iv + 0.5f * (d2iv * iv - iv) * dv;
    dr2 = -2.0f * d * iv - d2iv * iv * dv;
}
// Backward model (sigma from the observed image, tf/model.py:113-118):
GLM void lik_backward(float im, float obs, float iv, float log_sig, float& ll, float& r2, float& dll, float& dr2) {
    float d = im - obs;
    r2 = d * d * iv;
    ll = -0.5f * r2 - log_sig;
    dll = -d * iv;
    dr2 = 2.0f * d * iv;
}


// ---------------------------------------------------------------------------
// Shapelets (src/gigalens/tf/profiles/light/shapelets.py:53-85)
// ---------------------------------------------------------------------------
constexpr int kShapeletTable = 6000;  // points of linspace(-5, 5, 6000) (shapelets.py:39-40)
constexpr int kMaxShapeletN = 12;     // n_max supported

// 1-D Hermite functions phi_n(u), n = 0..n_max, and their derivatives.
// Table mode: linear interpolation on the regular [-5, 5] grid with zero
// fill outside (tfp.math.interp_regular_1d_grid, shapelets.py:55-60); the
// derivative is the slope of the segment (zero outside).
GLM void shapelet_1d_table(const float* __restrict__ table, int n_max, float u, float* phi, float* dphi) {
    const float scale = (kShapeletTable - 1) / 10.0f;
    const float xi = (u + 5.0f) * scale;
    const bool out = !(xi >= 0.0f && xi <= (float)(kShapeletTable - 1));
    float xc = fminf(fmaxf(xi, 0.0f), (float)(kShapeletTable - 1));
    int lo = (int)floorf(xc);
    int hi = lo + 1 < kShapeletTable ? lo + 1 : kShapeletTable - 1;
    lo = hi - 1 > 0 ? hi - 1 : 0;
    const float a = xc - (float)lo;
#pragma unroll
    for (int n = 0; n <= kMaxShapeletN; ++n) {
        if (n <= n_max) {
            const float y0 = table[n * kShapeletTable + lo];
            const float y1 = table[n * kShapeletTable + hi];
            phi[n] = out ? 0.0f : fmaf(a, y1 - y0, y0);
            if (dphi) dphi[n] = out ? 0.0f : (y1 - y0) * scale;
        }
    }
}

// Direct mode: phi_n(u) = H_n(u) exp(-u^2/2) / sqrt(2^n sqrt(pi) n!) by the
// recurrence H_n = 2 (u H_{n-1} - (n-1) H_{n-2}) (shapelets.py:66-85);
// phi_n' = sqrt(n/2) phi_{n-1} - sqrt((n+1)/2) phi_{n+1}.
GLM void shapelet_1d_direct(const float* __restrict__ pref, int n_max, float u, float* phi, float* dphi) {
    const float g = expf(-0.5f * u * u);
    float h[kMaxShapeletN + 2];
    h[0] = 1.0f;
    h[1] = 2.0f * u;
#pragma unroll
    for (int n = 2; n <= kMaxShapeletN + 1; ++n) h[n] = 2.0f * (u * h[n - 1] - (float)(n - 1) * h[n - 2]);
#pragma unroll
    for (int n = 0; n <= kMaxShapeletN; ++n)
        if (n <= n_max) phi[n] = pref[n] * h[n] * g;
    if (dphi) {
#pragma unroll
        for (int n = 0; n <= kMaxShapeletN; ++n) {
            if (n <= n_max) {
                const float up = pref[n + 1] * h[n + 1] * g;  // phi_{n+1}
                const float dn = n > 0 ? phi[n - 1] : 0.0f;
                dphi[n] = sqrtf(0.5f * n) * dn - sqrtf(0.5f * (n + 1)) * up;
            }
        }
    }
}

// Basis order (n1, n2): (0,0),(1,0),(0,1),(2,0),(1,1),(0,2),... (shapelets.py:27-46)
GLM void shapelet_order(int k, int& n1, int& n2) {
    int s = 0;
    while ((s + 1) * (s + 2) / 2 <= k) ++s;
    const int j = k - s * (s + 1) / 2;  // position within the shell of total order s
    n1 = s - j;
    n2 = j;
}

}  // namespace glm
