// Host build of the banded kernel's separable-PSF passes (gigalens_amd/csrc/band_passes.h)
// for CPU unit tests: every pass runs task by task over a 512-thread "workgroup" on
// buffers of exactly the sizes the kernel uses, surrounded by guard zones of NaN, and the
// results are compared with a direct correlation (forward) and its transpose (adjoint).
// A read outside a buffer returns NaN into the result; a write outside one changes a
// guard word.  Test infrastructure only.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>
#include "../../gigalens_amd/csrc/band_passes.h"

using gl::band::Geom;

namespace {
constexpr int kGuard = 4096;  // floats on each side of every buffer
constexpr int kThreads = 512;

struct Guarded {
    std::vector<float> mem;
    size_t n;
    explicit Guarded(size_t count) : mem(count + 2 * kGuard, std::nanf("")), n(count) {}
    float* data() { return mem.data() + kGuard; }
    bool guards_intact() const {
        for (size_t i = 0; i < kGuard; ++i)
            if (!std::isnan(mem[i]) || !std::isnan(mem[kGuard + n + i])) return false;
        return true;
    }
};
}  // namespace

// Forward and adjoint of one geometry.  Returns 0 on success, and the largest errors
// relative to the largest reference value; 1: a guard word changed; 2: a bad geometry.
// mutate plants a geometry mistake (for the test that the checks respond): 1 V rows
// without their zero margins, 2 band rows without their halo columns, 3 one tap per
// phase too few, 4 the PSF centre one sub-pixel off; 5 runs the generic passes on a
// geometry that has compile-time ones (no mistake).
extern "C" int h_band_check_m(int N, int ss, int k, int R, int BY, unsigned seed, int mutate, double* err_fwd,
                              double* err_adj) {
    if (N < 1 || ss < 1 || k < 1 || R < 1 || R > gl::band::kMaxRank || BY < 1) return 2;
    Geom g = gl::band::make_geom(N, ss, k, R, BY);
    if (g.BY > g.nsub) g.BY = g.nsub;
    const Geom g0 = g;  // the reference sums use the geometry as built
    if (mutate == 1) { g.vm = 0; g.VW = g.N; }
    if (mutate == 2) g.BW = g.nsub & ~3;
    if (mutate == 3) { g.Mt -= 1; g.fixed = 0; }  // the compile-time passes take their tap count from (SS, KF)
    if (mutate == 4) g.pb += 1;
    if (mutate == 5) g.fixed = 0;  // not a mistake: the generic passes on a compile-time geometry
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> U(-1.0f, 1.0f);
    // factors, zero above kf; W[p][q] = sum_r a_r[p] b_r[q]
    std::vector<float> fac(gl::band::fac_floats(g0), 0.0f);
    for (int r = 0; r < R; ++r)
        for (int i = 0; i < g.kf; ++i) {
            fac[r * g.KFP + i] = U(rng);
            fac[(R + r) * g.KFP + i] = U(rng);
            fac[2 * R * g.KFP + r * g.KFV + g.ZP + i] = fac[r * g.KFP + i];
        }
    Guarded facg(fac.size());
    memcpy(facg.data(), fac.data(), fac.size() * sizeof(float));
    const float* fa = facg.data();
    const float* fb = facg.data() + R * g.KFP;
    const float* fv = facg.data() + 2 * R * g.KFP;
    std::vector<double> W(g.kf * g.kf, 0.0);
    for (int r = 0; r < R; ++r)
        for (int p = 0; p < g.kf; ++p)
            for (int q = 0; q < g.kf; ++q) W[p * g.kf + q] += (double)fac[r * g.KFP + p] * fac[(R + r) * g.KFP + q];
    const int ns = g.nsub;
    std::vector<float> S((size_t)ns * ns);
    for (float& v : S) v = U(rng);
    // ---- forward ---------------------------------------------------------------
    std::vector<double> ref((size_t)N * N, 0.0);
    double ref_max = 0.0;
    for (int I = 0; I < N; ++I)
        for (int J = 0; J < N; ++J) {
            double acc = 0.0;
            for (int p = 0; p < g.kf; ++p)
                for (int q = 0; q < g.kf; ++q) {
                    const int y = ss * I + p - g0.pb, x = ss * J + q - g0.pb;
                    if (y >= 0 && y < ns && x >= 0 && x < ns) acc += W[p * g.kf + q] * S[(size_t)y * ns + x];
                }
            ref[(size_t)I * N + J] = acc;
            ref_max = std::fmax(ref_max, std::fabs(acc));
        }
    Guarded band(gl::band::band_floats(g0));
    // the HBM slot, stale values from a previous particle everywhere (the pads must be cleared)
    Guarded slot((size_t)gl::band::slot_floats(g0));
    for (size_t i = 0; i < slot.n; ++i) slot.data()[i] = 1e30f;
    float* T = slot.data();
    float* G = slot.data() + gl::band::t_floats(g0);
    for (int t = 0; t < kThreads; ++t) gl::band::zero_pads(g, t, kThreads, T, G);
    for (size_t i = 0; i < band.n; ++i) band.data()[i] = 0.0f;  // as the kernel: halo columns zero
    for (int y0 = 0; y0 < ns; y0 += g.BY) {
        const int rows = g.BY < ns - y0 ? g.BY : ns - y0;
        for (int yy = 0; yy < rows; ++yy)
            for (int x = 0; x < ns; ++x) band.data()[yy * g.BW + g.pb + x] = S[(size_t)(y0 + yy) * ns + x];
        for (int t = 0; t < kThreads; ++t) gl::band::hpass_any(g, t, kThreads, y0, rows, band.data(), fb, T);
    }
    std::vector<float> out((size_t)N * N, std::nanf(""));
    std::vector<int> hits((size_t)N * N, 0);
    for (int t = 0; t < kThreads; ++t)
        gl::band::vpass_any(g, t, kThreads, T, fv, [&](int I, int J, float v) {
            out[(size_t)I * N + J] = v;
            hits[(size_t)I * N + J] += 1;
        });
    double ef = 0.0;
    for (size_t i = 0; i < out.size(); ++i) {
        if (hits[i] != 1) return 3;  // every output exactly once
        const double e = std::fabs((double)out[i] - ref[i]);
        ef = std::isnan(out[i]) ? INFINITY : std::fmax(ef, e);
    }
    *err_fwd = ef / (ref_max > 0 ? ref_max : 1.0);
    // ---- adjoint -----------------------------------------------------------------
    std::vector<float> Gh((size_t)N * N);
    for (float& v : Gh) v = U(rng);
    for (int I = 0; I < N; ++I)
        for (int J = 0; J < N; ++J) G[(size_t)(I + g.GT) * N + J] = Gh[(size_t)I * N + J];
    std::vector<double> aref((size_t)ns * ns, 0.0);
    double aref_max = 0.0;
    for (int I = 0; I < N; ++I)
        for (int J = 0; J < N; ++J)
            for (int p = 0; p < g.kf; ++p)
                for (int q = 0; q < g.kf; ++q) {
                    const int y = ss * I + p - g0.pb, x = ss * J + q - g0.pb;
                    if (y >= 0 && y < ns && x >= 0 && x < ns)
                        aref[(size_t)y * ns + x] += W[p * g.kf + q] * Gh[(size_t)I * N + J];
                }
    for (double v : aref) aref_max = std::fmax(aref_max, std::fabs(v));
    for (size_t i = 0; i < band.n; ++i) band.data()[i] = 0.0f;  // V margins
    float* gbuf = band.data() + gl::band::g_offset(g0);
    // the staged window in a buffer of its own (exact size, guarded); the kernel keeps it in the band
    Guarded gw((size_t)g0.GWR * N);
    double ea = 0.0;
    for (int y0 = 0; y0 < ns; y0 += g.BY) {
        const int rows = g.BY < ns - y0 ? g.BY : ns - y0;
        for (int t = 0; t < kThreads; ++t) gl::band::stage_g(g, t, kThreads, y0, G, gw.data());
        for (int t = 0; t < kThreads; ++t) gl::band::vband_any(g, t, kThreads, y0, rows, gw.data(), fa, band.data());
        for (int t = 0; t < kThreads; ++t) gl::band::gpass_any(g, t, kThreads, rows, band.data(), fb, gbuf);
        for (int yy = 0; yy < rows; ++yy)
            for (int x = 0; x < ns; ++x) {
                const float v = gbuf[yy * ns + x];
                const double e = std::fabs((double)v - aref[(size_t)(y0 + yy) * ns + x]);
                ea = std::isnan(v) ? INFINITY : std::fmax(ea, e);
            }
    }
    *err_adj = ea / (aref_max > 0 ? aref_max : 1.0);
    if (!facg.guards_intact() || !band.guards_intact() || !slot.guards_intact() || !gw.guards_intact()) return 1;
    return 0;
}

extern "C" int h_band_check(int N, int ss, int k, int R, int BY, unsigned seed, double* err_fwd, double* err_adj) {
    return h_band_check_m(N, ss, k, R, BY, seed, 0, err_fwd, err_adj);
}

// Geometry of the kernel, for the host-side checks of the batch's LDS and slot sizes.
extern "C" void h_band_geom(int N, int ss, int k, int R, int BY, int* out) {
    const Geom g = gl::band::make_geom(N, ss, k, R, BY);
    const int v[] = {g.N, g.ss, g.nsub, g.kf, g.pb, g.R, g.BY, g.BW, g.KFP, g.Mt, g.vm, g.VW, g.TR, g.GT, g.GR,
                     g.fixed, gl::band::band_floats(g), (int)gl::band::slot_floats(g)};
    for (int i = 0; i < (int)(sizeof(v) / sizeof(v[0])); ++i) out[i] = v[i];
}
