// Separable-PSF passes of the banded particle kernel (band_kernels.hip).
//
// The supersampled image of one particle is produced and consumed in bands of
// BY sub-rows, so only one band sits in LDS at a time and images of any size
// run with a fixed LDS footprint.  With the folded PSF+pool weights
// W[p][q] ~= sum_r a_r[p] b_r[q] (kf x kf, correlation form, see gl_plan_create):
//
//   forward   T_r[y][J] = sum_q b_r[q] S[y][ss J + q - pb]              (per band, -> HBM)
//             out[I][J] = sum_r sum_p a_r[p] T_r[ss I + p - pb][J]       (rows outside the image: 0)
//   adjoint   V_r[y][J] = sum_I a_r[y + pb - ss I] G[I][J]              (per band, -> LDS)
//             g[y][x]   = sum_r sum_J b_r[x + pb - ss J] V_r[y][J]       (per band, -> LDS)
//
// out = conv2d(S, PSF, SAME) followed by avg_pool(ss) (tf/simulator.py:83-96);
// g is its adjoint applied to G = dL/dout.  Every intermediate carries zero
// padding where a tap reaches outside the image -- halo columns of a band, pad
// rows of T above and below the image, pad rows of G, margin columns of V -- so
// no pass tests a bound.  Factor rows are zero padded to KFP = kf + ss floats, so
// every tap loop has the fixed trip count Mt = ceil(kf / ss) per phase.
//
// Each pass comes twice: for compile-time (SS, KF, R) -- every tap index a
// constant, the factors read through the constant address space (scalar loads),
// one lane per 4 outputs or per output column with all SS phases -- and generic
// (runtime loops) for other PSFs.  Every pass is a loop over tasks `for (task =
// tid; task < count; task += nthr)` written for the host and the device alike:
// the kernel runs it with its workgroup; the CPU harness
// (tests/native/band_harness.hip) runs it task by task against a direct
// correlation, with guard values around every buffer.
#pragma once
#include <hip/hip_runtime.h>

namespace gl {
namespace band {

constexpr int kMaxRank = 8;  // rank-1 terms of the folded PSF
constexpr int kRows = 4;     // output rows per task of the vertical pass
constexpr int kCols = 4;     // T columns per task of the fixed horizontal pass

struct Geom {
    int N, ss, nsub, kf, pb, R;
    int BY;   // sub-rows per band
    int BW;   // band row width (floats, multiple of 4); column c holds sub-column c - pb
    int KFP;  // factor row length: kf + ss, zero above kf
    int Mt;   // taps per phase: ceil(kf / ss)
    int TR;   // rows of T per rank: sub-row y at row y + pb, zero rows above and below
    int GT;   // zero rows of G above the image (G row I at I + GT)
    int GR;   // rows of G
    int vm;   // zero columns on each side of a V row
    int VW;   // V row width: N + 2 vm
    int GWR;  // rows of G one band's transposed vertical pass reads (staged in LDS)
    int ZP;   // zeros before the vertical pass's copy of a_r: ss (kRows - 1)
    int KFV;  // its row length: kf + 2 ZP (a_r[p] at ZP + p, zero for p outside [0, kf))
    int fixed;  // compile-time pass set (fixed_code), 0: generic
};

// Compile-time pass sets: (SS, KF, R) of the folded 11 x 11 PSFs of the uneven-batch
// workload and the 60 x 60 workload (k = 21, 33, 43 sub-grid kernels).
__host__ __device__ inline int fixed_code(int ss, int kf, int R) {
    if (ss == 2 && kf == 22 && R == 1) return 1;
    if (ss == 3 && kf == 35 && R == 5) return 2;
    if (ss == 4 && kf == 46 && R == 1) return 3;
    return 0;
}

// Geometry of an N x N image at supersampling ss with a folded PSF of width kf
// (kf = k + ss - 1 for a k x k PSF, pb = (k - 1) / 2) and rank R, bands of BY rows.
__host__ __device__ inline Geom make_geom(int N, int ss, int k, int R, int BY) {
    Geom g;
    g.N = N;
    g.ss = ss;
    g.nsub = N * ss;
    g.kf = k + ss - 1;
    g.pb = (k - 1) / 2;
    g.R = R;
    g.BY = BY;
    const int NB = (N + kCols - 1) / kCols * kCols;  // columns of the last horizontal task group
    const int need = ss * (NB - 1) + g.kf;            // reach of a band row
    const int w = need > g.pb + g.nsub ? need : g.pb + g.nsub;
    g.BW = (w + 3) & ~3;
    g.KFP = g.kf + ss;
    g.Mt = (g.kf + ss - 1) / ss;
    const int NR = (N + kRows - 1) / kRows * kRows;  // rows of the last vertical task
    g.TR = ss * (NR - 1) + g.kf;
    const int db_min = g.pb / ss, db_max = (ss - 1 + g.pb) / ss;
    g.GT = g.Mt - 1 - db_min > 0 ? g.Mt - 1 - db_min : 0;
    const int gb = (g.nsub - 1 + g.pb) / ss - (N - 1);
    g.GR = g.GT + N + (gb > 0 ? gb : 0);
    g.vm = g.Mt - 1 - db_min > db_max ? g.Mt - 1 - db_min : db_max;
    if (g.vm < 1) g.vm = 1;
    g.VW = g.N + 2 * g.vm;
    g.GWR = (BY + ss - 2) / ss + g.Mt + 1;
    g.ZP = ss * (kRows - 1);
    g.KFV = g.kf + 2 * g.ZP;
    g.fixed = fixed_code(ss, g.kf, R);
    return g;
}

// Floats of a lens's factor table: a_r and b_r (R rows of KFP each), then the vertical
// pass's zero-padded copy of a_r (R rows of KFV).
__host__ __device__ inline int fac_floats(const Geom& g) { return 2 * g.R * g.KFP + g.R * g.KFV; }

// Floats of LDS per band: the forward band (BY rows of BW), or the adjoint's R planes of V
// (BY rows of VW), the band's adjoint sub-pixels g (BY rows of nsub) and the window of G
// rows its transposed vertical pass reads (GWR rows of N).
__host__ __device__ inline int band_floats(const Geom& g) {
    const int f = g.BY * g.BW, a = g.R * g.BY * g.VW + g.BY * g.nsub + g.GWR * g.N;
    return f > a ? f : a;
}
__host__ __device__ inline int g_offset(const Geom& g) { return g.R * g.BY * g.VW; }
__host__ __device__ inline int gw_offset(const Geom& g) { return g.R * g.BY * g.VW + g.BY * g.nsub; }

// First padded G row (index I + GT) the band from sub-row y0 reads.
__host__ __device__ inline int gw_first(const Geom& g, int y0) { return (y0 + g.pb) / g.ss - g.Mt + 1 + g.GT; }

// Copy the band's window of padded G rows from HBM into LDS (rows past GR are not read).
__host__ __device__ inline void stage_g(const Geom& g, int tid, int nthr, int y0, const float* __restrict__ G,
                                        float* __restrict__ Gw) {
    const int w0 = gw_first(g, y0);
    const int nr = g.GR - w0 < g.GWR ? g.GR - w0 : g.GWR;
    const float* src = G + (long)w0 * g.N;
    const int total = nr * g.N;
    constexpr int kIn = 8;  // loads in flight per thread
    for (int base = tid; base < total; base += kIn * nthr) {
        float v[kIn];
#pragma unroll
        for (int k = 0; k < kIn; ++k) v[k] = base + k * nthr < total ? src[base + k * nthr] : 0.0f;
#pragma unroll
        for (int k = 0; k < kIn; ++k)
            if (base + k * nthr < total) Gw[base + k * nthr] = v[k];
    }
}

// HBM floats of one particle: T (R x TR x N) then G (GR x N).
__host__ __device__ inline long t_floats(const Geom& g) { return (long)g.R * g.TR * g.N; }
__host__ __device__ inline long slot_floats(const Geom& g) { return t_floats(g) + (long)g.GR * g.N; }

// Factor reads: uniform addresses into read-only memory, through the constant address
// space on the device (scalar loads).
__host__ __device__ inline float ldc(const float* p, int i) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const float __attribute__((address_space(4)))*)p)[i];
#else
    return p[i];
#endif
}

// A value the caller knows to be equal across the wave, in a scalar register on the device.
__host__ __device__ inline int uniform_int(int v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_readfirstlane(v);
#else
    return v;
#endif
}

// Zero the pad rows of T and G (every particle: the slot is shared by every lens).
__host__ __device__ inline void zero_pads(const Geom& g, int tid, int nthr, float* __restrict__ T,
                                          float* __restrict__ G) {
    const int top = g.pb, bot = g.TR - g.pb - g.nsub;
    const int per_rank = (top + bot) * g.N;
    for (int s = tid; s < g.R * per_rank; s += nthr) {
        const int r = s / per_rank, o = s - r * per_rank;
        const int row = o / g.N, J = o - row * g.N;
        const int y = row < top ? row : row - top + g.pb + g.nsub;
        T[((long)r * g.TR + y) * g.N + J] = 0.0f;
    }
    const int gtop = g.GT * g.N, gbot = (g.GR - g.GT - g.N) * g.N;
    for (int s = tid; s < gtop + gbot; s += nthr) G[s < gtop ? s : s - gtop + (g.GT + g.N) * g.N] = 0.0f;
}

// ---- generic passes (runtime ss, kf, R) ---------------------------------------------

// Horizontal pass over `rows` band rows from sub-row y0: T[r][y0 + yy + pb][J].
__host__ __device__ inline void hpass(const Geom& g, int tid, int nthr, int y0, int rows,
                                      const float* __restrict__ band, const float* __restrict__ fb,
                                      float* __restrict__ T) {
    const int cnt = rows * g.N;
    for (int s = tid; s < cnt; s += nthr) {
        const int yy = s / g.N, J = s - yy * g.N;
        const float* src = band + yy * g.BW + g.ss * J;
        float t[kMaxRank];
#pragma unroll
        for (int r = 0; r < kMaxRank; ++r) t[r] = 0.0f;
        for (int q = 0; q < g.kf; ++q) {
            const float v = src[q];
#pragma unroll
            for (int r = 0; r < kMaxRank; ++r)
                if (r < g.R) t[r] = fmaf(ldc(fb, r * g.KFP + q), v, t[r]);
        }
#pragma unroll
        for (int r = 0; r < kMaxRank; ++r)
            if (r < g.R) T[((long)r * g.TR + y0 + yy + g.pb) * g.N + J] = t[r];
    }
}

// Vertical pass: every output pixel (I, J) once, kRows rows per task; calls
// out(I, J, value) for I < N.  fv: the zero-padded copy of a_r (KFV per rank).
template <class Out>
__host__ __device__ inline void vpass(const Geom& g, int tid, int nthr, const float* __restrict__ T,
                                      const float* __restrict__ fv, Out&& out) {
    const int NI = (g.N + kRows - 1) / kRows;
    const int cnt = NI * g.N;
    const int span = g.ss * (kRows - 1) + g.kf;
    for (int s = tid; s < cnt; s += nthr) {
        const int ib = s / g.N, J = s - ib * g.N;
        const int I0 = ib * kRows;
        float o[kRows];
#pragma unroll
        for (int i = 0; i < kRows; ++i) o[i] = 0.0f;
        for (int r = 0; r < g.R; ++r) {
            const float* col = T + ((long)r * g.TR + g.ss * I0) * g.N + J;  // padded row ss I0 + c
            const float* a = fv + r * g.KFV + g.ZP;                           // a[p] = a_r[p], 0 outside
            for (int c = 0; c < span; ++c) {
                const float v = col[(long)c * g.N];
#pragma unroll
                for (int i = 0; i < kRows; ++i) o[i] = fmaf(ldc(a, c - g.ss * i), v, o[i]);
            }
        }
#pragma unroll
        for (int i = 0; i < kRows; ++i)
            if (I0 + i < g.N) out(I0 + i, J, o[i]);
    }
}

// Transposed vertical pass for the `rows` band rows from sub-row y0:
// V[r][yy][vm + J] (the vm margins stay as they are: zero).  Gw: the band's window of G
// rows (stage_g).
__host__ __device__ inline void vband(const Geom& g, int tid, int nthr, int y0, int rows,
                                      const float* __restrict__ Gw, const float* __restrict__ fa,
                                      float* __restrict__ V) {
    const int w0 = gw_first(g, y0);
    const int cnt = rows * g.N;
    for (int s = tid; s < cnt; s += nthr) {
        const int yy = s / g.N, J = s - yy * g.N;
        const int y = y0 + yy;
        const int Ihi = (y + g.pb) / g.ss;  // the I of the smallest tap, y + pb - ss Ihi in [0, ss)
        const int q0 = y + g.pb - g.ss * Ihi;
        const float* gcol = Gw + (Ihi + g.GT - w0) * g.N + J;
        float v[kMaxRank];
#pragma unroll
        for (int r = 0; r < kMaxRank; ++r) v[r] = 0.0f;
        for (int n = 0; n < g.Mt; ++n) {
            const float gv = gcol[-n * g.N];
            const int q = q0 + g.ss * n;  // < ss Mt <= KFP - 1; zero above kf
#pragma unroll
            for (int r = 0; r < kMaxRank; ++r)
                if (r < g.R) v[r] = fmaf(ldc(fa, r * g.KFP + q), gv, v[r]);
        }
#pragma unroll
        for (int r = 0; r < kMaxRank; ++r)
            if (r < g.R) V[(r * g.BY + yy) * g.VW + g.vm + J] = v[r];
    }
}

// Transposed horizontal pass: the adjoint value of every band sub-pixel into
// gb[yy * nsub + x].
__host__ __device__ inline void gpass(const Geom& g, int tid, int nthr, int rows, const float* __restrict__ V,
                                      const float* __restrict__ fb, float* __restrict__ gb) {
    const int cnt = rows * g.nsub;
    for (int s = tid; s < cnt; s += nthr) {
        const int yy = s / g.nsub, x = s - yy * g.nsub;
        const int Jhi = (x + g.pb) / g.ss;  // J of the smallest tap; Jhi - n >= -vm, Jhi <= N - 1 + vm
        const int q0 = x + g.pb - g.ss * Jhi;
        float acc = 0.0f;
        for (int r = 0; r < g.R; ++r) {
            const float* vrow = V + (r * g.BY + yy) * g.VW + g.vm + Jhi;
            for (int n = 0; n < g.Mt; ++n) acc = fmaf(fb[r * g.KFP + q0 + g.ss * n], vrow[-n], acc);
        }
        gb[s] = acc;
    }
}

// ---- passes for compile-time (SS, KF, R) ---------------------------------------------

template <int SS, int KF>
struct Fix {
    static constexpr int PB = (KF - SS) / 2;  // (k - 1) / 2 with k = KF - SS + 1
    static constexpr int MT = (KF + SS - 1) / SS;
    static constexpr int KFP = KF + SS;
    static constexpr int SEG = SS * (kCols - 1) + KF;  // band floats one horizontal task reads
    static constexpr int SEG4 = (SEG + 3) / 4;
    static constexpr int SPAN = SS * (kRows - 1) + KF;  // T rows one vertical task reads
    static constexpr int DB0 = PB / SS;                  // J offset of phase 0's smallest tap
    static constexpr int db(int b) { return (b + PB) / SS; }
    static constexpr int qb(int b) { return (b + PB) % SS; }
    static constexpr int WIN = MT + db(SS - 1) - DB0;  // V values one column task reads
};

// Horizontal pass: a task is kCols consecutive T columns of one band row, from one
// 16-byte-aligned segment of the row held in registers.
template <int SS, int KF, int R>
__host__ __device__ inline void hpass_fixed(const Geom& g, int tid, int nthr, int y0, int rows,
                                            const float* __restrict__ band, const float* __restrict__ fb,
                                            float* __restrict__ T) {
    using F = Fix<SS, KF>;
    const int NG = (g.N + kCols - 1) / kCols;
    const int cnt = rows * NG;
    for (int s = tid; s < cnt; s += nthr) {
        const int yy = s / NG, J0 = (s - yy * NG) * kCols;
        const float4* row4 = reinterpret_cast<const float4*>(band + yy * g.BW + SS * J0);
        float seg[F::SEG4 * 4];
#pragma unroll
        for (int c = 0; c < F::SEG4; ++c) {
            const float4 v = row4[c];
            seg[4 * c] = v.x; seg[4 * c + 1] = v.y; seg[4 * c + 2] = v.z; seg[4 * c + 3] = v.w;
        }
        float* dst = T + ((long)y0 + yy + F::PB) * g.N + J0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float t[kCols];
#pragma unroll
            for (int j = 0; j < kCols; ++j) t[j] = 0.0f;
#pragma unroll
            for (int q = 0; q < KF; ++q) {
                const float w = ldc(fb, r * F::KFP + q);
#pragma unroll
                for (int j = 0; j < kCols; ++j) t[j] = fmaf(w, seg[SS * j + q], t[j]);
            }
#pragma unroll
            for (int j = 0; j < kCols; ++j)
                if (J0 + j < g.N) dst[(long)r * g.TR * g.N + j] = t[j];
        }
    }
}

// Vertical pass: a task is kRows output rows of one column; the T rows it reads are
// constant offsets from one address (padded rows: no bounds tests), the factors of a
// row come from the zero-padded copy (no tap tests), one rank at a time.
template <int SS, int KF, int R, class Out>
__host__ __device__ inline void vpass_fixed(const Geom& g, int tid, int nthr, const float* __restrict__ T,
                                            const float* __restrict__ fv, Out&& out) {
    using F = Fix<SS, KF>;
    constexpr int ZP = SS * (kRows - 1), KFV = KF + 2 * ZP;
    const int NI = (g.N + kRows - 1) / kRows;
    const int cnt = NI * g.N;
    for (int s = tid; s < cnt; s += nthr) {
        const int ib = s / g.N, J = s - ib * g.N;
        const int I0 = ib * kRows;
        float o[kRows];
#pragma unroll
        for (int i = 0; i < kRows; ++i) o[i] = 0.0f;
#pragma unroll 1
        for (int r = 0; r < R; ++r) {
            // the column window first (every load in flight at once), then the taps
            const float* col = T + ((long)r * g.TR + SS * I0) * g.N + J;
            const float* a = fv + r * KFV + ZP;
            float win[F::SPAN];
#pragma unroll
            for (int c = 0; c < F::SPAN; ++c) win[c] = col[(long)c * g.N];
#pragma unroll
            for (int c = 0; c < F::SPAN; ++c)
#pragma unroll
                for (int i = 0; i < kRows; ++i)
                    if (c - SS * i >= 0 && c - SS * i < KF) o[i] = fmaf(ldc(a, c - SS * i), win[c], o[i]);
        }
#pragma unroll
        for (int i = 0; i < kRows; ++i)
            if (I0 + i < g.N) out(I0 + i, J, o[i]);
    }
}

// Transposed vertical pass: a wave task is one band row and 64 columns, so the tap
// phase of the row and its factors are wave-uniform.
template <int SS, int KF, int R>
__host__ __device__ inline void vband_fixed(const Geom& g, int tid, int nthr, int y0, int rows,
                                            const float* __restrict__ Gw, const float* __restrict__ fa,
                                            float* __restrict__ V) {
    using F = Fix<SS, KF>;
    const int w0 = gw_first(g, y0);
    const int NC = (g.N + 63) / 64;
    const int lane = tid & 63;
    for (int w = uniform_int(tid >> 6); w < rows * NC; w += nthr >> 6) {
        const int yy = w / NC, J = (w - yy * NC) * 64 + lane;
        const int y = y0 + yy;
        const int Ihi = (y + F::PB) / SS;
        const int q0 = y + F::PB - SS * Ihi;
        if (J < g.N) {
            const float* gcol = Gw + (Ihi + g.GT - w0) * g.N + J;
            float v[R];
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = 0.0f;
#pragma unroll
            for (int n = 0; n < F::MT; ++n) {
                const float gv = gcol[-n * g.N];
#pragma unroll
                for (int r = 0; r < R; ++r) v[r] = fmaf(ldc(fa, r * F::KFP + q0 + SS * n), gv, v[r]);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) V[(r * g.BY + yy) * g.VW + g.vm + J] = v[r];
        }
    }
}

// Transposed horizontal pass: a task is the SS sub-pixels of one output column X of
// one band row; the V values they share are read once.
template <int SS, int KF, int R>
__host__ __device__ inline void gpass_fixed(const Geom& g, int tid, int nthr, int rows, const float* __restrict__ V,
                                            const float* __restrict__ fb, float* __restrict__ gb) {
    using F = Fix<SS, KF>;
    const int cnt = rows * g.N;
    for (int s = tid; s < cnt; s += nthr) {
        const int yy = s / g.N, X = s - yy * g.N;
        float acc[SS];
#pragma unroll
        for (int b = 0; b < SS; ++b) acc[b] = 0.0f;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            // window J = X + DB0 - MT + 1 + j, j < WIN
            const float* vrow = V + (r * g.BY + yy) * g.VW + g.vm + X + F::DB0 - F::MT + 1;
            float win[F::WIN];
#pragma unroll
            for (int j = 0; j < F::WIN; ++j) win[j] = vrow[j];
#pragma unroll
            for (int b = 0; b < SS; ++b)
#pragma unroll
                for (int n = 0; n < F::MT; ++n)
                    acc[b] = fmaf(ldc(fb, r * F::KFP + F::qb(b) + SS * n),
                                  win[F::db(b) - F::DB0 + F::MT - 1 - n], acc[b]);
        }
        float* dst = gb + yy * g.nsub + SS * X;
#pragma unroll
        for (int b = 0; b < SS; ++b) dst[b] = acc[b];
    }
}

// ---- dispatch ---------------------------------------------------------------------------
__host__ __device__ inline void hpass_any(const Geom& g, int tid, int nthr, int y0, int rows, const float* band,
                                          const float* fb, float* T) {
    switch (g.fixed) {
    case 1: hpass_fixed<2, 22, 1>(g, tid, nthr, y0, rows, band, fb, T); break;
    case 2: hpass_fixed<3, 35, 5>(g, tid, nthr, y0, rows, band, fb, T); break;
    case 3: hpass_fixed<4, 46, 1>(g, tid, nthr, y0, rows, band, fb, T); break;
    default: hpass(g, tid, nthr, y0, rows, band, fb, T);
    }
}
template <class Out>
__host__ __device__ inline void vpass_any(const Geom& g, int tid, int nthr, const float* T, const float* fv,
                                          Out&& out) {
    switch (g.fixed) {
    case 1: vpass_fixed<2, 22, 1>(g, tid, nthr, T, fv, out); break;
    case 2: vpass_fixed<3, 35, 5>(g, tid, nthr, T, fv, out); break;
    case 3: vpass_fixed<4, 46, 1>(g, tid, nthr, T, fv, out); break;
    default: vpass(g, tid, nthr, T, fv, out);
    }
}
__host__ __device__ inline void vband_any(const Geom& g, int tid, int nthr, int y0, int rows, const float* Gw,
                                          const float* fa, float* V) {
    switch (g.fixed) {
    case 1: vband_fixed<2, 22, 1>(g, tid, nthr, y0, rows, Gw, fa, V); break;
    case 2: vband_fixed<3, 35, 5>(g, tid, nthr, y0, rows, Gw, fa, V); break;
    case 3: vband_fixed<4, 46, 1>(g, tid, nthr, y0, rows, Gw, fa, V); break;
    default: vband(g, tid, nthr, y0, rows, Gw, fa, V);
    }
}
__host__ __device__ inline void gpass_any(const Geom& g, int tid, int nthr, int rows, const float* V,
                                          const float* fb, float* gb) {
    switch (g.fixed) {
    case 1: gpass_fixed<2, 22, 1>(g, tid, nthr, rows, V, fb, gb); break;
    case 2: gpass_fixed<3, 35, 5>(g, tid, nthr, rows, V, fb, gb); break;
    case 3: gpass_fixed<4, 46, 1>(g, tid, nthr, rows, V, fb, gb); break;
    default: gpass(g, tid, nthr, rows, V, fb, gb);
    }
}

}  // namespace band
}  // namespace gl
