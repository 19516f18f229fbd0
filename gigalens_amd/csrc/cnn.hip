// C ABI of the learned initialiser's network (gl_cnn_*): workspace, parameter
// packing and the launch sequences of the BF16 MFMA kernels (cnn_conv.hip: wave-streaming
// and LDS-transposed convolutions; cnn_kernels.hip: input transform, head, reductions, AdamW).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "abi_common.h"
#include "cnn_conv.hip"

using glc::fail;
using glc::StreamGuard;

// ---------------------------------------------------------------------------
// Learned initialiser (no reference counterpart)
// ---------------------------------------------------------------------------
// One wave-streaming convolution pass (gl::cnn::conv_wave): its kernel instance, LDS
// bytes (the weight tiles) and the most workgroups that are resident at once.
struct WavePass {
    const void* fn;
    int kind, nmt, nk, lds, grid;
};

struct gl_cnn {
    struct Layer {
        int cin, cout, s, H, W, Ho, Wo, kp;
        int rw;                    // band rows of the weight-gradient pass
        WavePass fwd, dgrad;       // wave-streaming forward and data-gradient passes
        int wblk, wncg, wnp, wchunks;  // weight-gradient grid: blocks over images, tile chunks
        size_t woff, boff;         // offsets in the parameter vector
        gl::cnn::u16* wt;          // [cout][kp]
        gl::cnn::u16* w2t;         // [cin][9 cout] (layers > 0)
        gl::cnn::u16* Y;           // [B][Ho][Wo][cout]
        gl::cnn::u16* Z;
        gl::cnn::u16* dZ;
    };
    int H, W, d, hidden, maxb, P, cus;
    float in_scale;
    Layer L[4];
    size_t w1, b1, w2, b2;
    float *params, *m, *v, *grad;
    gl::cnn::u16* x0;
    float *h0, *h1, *da1, *out, *dout, *lossb, *part;
    size_t part_elems;
    std::vector<void*> dev;
    mutable StreamGuard guard;  // orders calls across streams (the activations are shared)
};

extern "C" int gl_cnn_destroy(gl_cnn* net) {
    if (!net) return GL_OK;
    for (void* p : net->dev) (void)hipFree(p);
    net->guard.release();
    delete net;
    return GL_OK;
}

static int cnn_alloc(gl_cnn* net, size_t bytes, void** out) {
    void* d = nullptr;
    GL_HIP(hipMalloc(&d, bytes ? bytes : 4));
    net->dev.push_back(d);
    GL_HIP(hipMemset(d, 0, bytes ? bytes : 4));
    *out = d;
    return GL_OK;
}

// LDS budget of one band workgroup: two fit in a CU's 160 KB
constexpr size_t kCnnLds = 80 * 1024;
// zeros after each convolution operand tensor: at least the most channels of a layer
constexpr size_t kZeroTail = gl::cnn::kHeadMaxC;

// first layer: rows of a forward band (about one pixel per thread)
static int conv1_rows(const gl_cnn::Layer& L) { return std::max(1, std::min(L.H, (256 + L.W - 1) / L.W)); }
static size_t conv1_fwd_lds(const gl_cnn::Layer& L) {
    return ((size_t)L.cout * 10 + (size_t)(conv1_rows(L) + 2) * (L.W + 2)) * 4;
}
static size_t conv1_wgrad_lds(const gl_cnn::Layer& L) { return (size_t)(L.H + 2) * (L.W + 2) * 4; }
// weight gradient: the transposed dZ band and the three column planes of the input band
static int wgrad_wp(const gl_cnn::Layer& L) { return (L.Wo + 7) / 8 * 8; }
static size_t wgrad_lds(const gl_cnn::Layer& L, int R) {
    return (size_t)gl::cnn::wgrad_lds_elems(L.s, R, wgrad_wp(L), L.cin, L.cout) * 2;
}
// largest band (rows <= n) meeting the budget, then evened out over the bands
static int pick_rows(int n, bool (*ok)(const gl_cnn::Layer&, int), const gl_cnn::Layer& L) {
    int R = 0;
    for (int r = 1; r <= n; ++r)
        if (ok(L, r)) R = r;
    if (R == 0) return 0;
    const int bands = (n + R - 1) / R;
    return (n + bands - 1) / bands;
}
// Wave-streaming pass of K operand channels into Nout output channels: 64-channel
// output tiles when the weights of all K channels fit 72 KB (two workgroups per CU),
// 32-channel tiles otherwise (at most 144 KB for K = 256).  K steps of 64 channels
// only for the stride-2 data gradient (at most 4 taps per group); the 9-tap passes
// take 32-channel steps, which keeps them at 128 VGPRs (two workgroups per CU).
typedef void (*WaveKernelFn)(gl::cnn::WaveArgs);
static WaveKernelFn wave_fn(int kind, int s, int nmt, int nk) {
    using gl::cnn::conv_wave;
    if (kind == 1 && s == 2) {
        static const WaveKernelFn tab[2][2] = {{conv_wave<1, 2, 2, 1>, conv_wave<1, 2, 2, 2>},
                                               {conv_wave<1, 2, 4, 1>, conv_wave<1, 2, 4, 2>}};
        return tab[nmt == 4][nk == 2];
    }
    static const WaveKernelFn tab[3][2] = {{conv_wave<0, 1, 2, 1>, conv_wave<0, 1, 4, 1>},
                                           {conv_wave<0, 2, 2, 1>, conv_wave<0, 2, 4, 1>},
                                           {conv_wave<1, 1, 2, 1>, conv_wave<1, 1, 4, 1>}};
    return tab[kind == 1 ? 2 : s - 1][nmt == 4];
}
static hipError_t wave_setup(WavePass& P, int kind, int s, int K, int Nout, int cus) {
    P.kind = kind;
    P.nmt = (Nout % 64 == 0 && gl::cnn::wave_lds(K, 4) <= 72 * 1024) ? 4 : 2;
    P.nk = (kind == 1 && s == 2 && K % 64 == 0) ? 2 : 1;
    P.lds = gl::cnn::wave_lds(K, P.nmt);
    P.fn = (const void*)wave_fn(kind, s, P.nmt, P.nk);
    hipError_t e = hipFuncSetAttribute(P.fn, hipFuncAttributeMaxDynamicSharedMemorySize, P.lds);
    if (e != hipSuccess) return e;
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, P.fn, 512, P.lds);
    if (e != hipSuccess) return e;
    P.grid = std::max(1, per_cu) * cus;
    return hipSuccess;
}
// Items of R output rows of one image: bands short enough that the batch makes about
// four items per resident workgroup.
static int wave_launch(const WavePass& P, gl::cnn::WaveArgs A, hipStream_t st) {
    const long target = 4L * P.grid;
    const int nb = (int)std::min<long>(A.Ho, std::max<long>(1, (target + A.B - 1) / A.B));
    A.R = (A.Ho + nb - 1) / nb;
    const long items = (long)A.B * ((A.Ho + A.R - 1) / A.R);
    const dim3 grid((unsigned)std::min<long>(items, P.grid));
    void* args[] = {&A};
    GL_HIP(hipLaunchKernel(P.fn, grid, dim3(512), args, (size_t)P.lds, st));
    return GL_OK;
}
static bool wgrad_ok(const gl_cnn::Layer& L, int R) { return wgrad_lds(L, R) <= kCnnLds; }

static hipError_t cnn_lds_attrs() {
    const int lds = (int)kCnnLds;
    hipError_t e;
    if ((e = hipFuncSetAttribute((const void*)&gl::cnn::conv_wgrad<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 lds)))
        return e;
    if ((e = hipFuncSetAttribute((const void*)&gl::cnn::conv1_fwd, hipFuncAttributeMaxDynamicSharedMemorySize, lds)))
        return e;
    if ((e = hipFuncSetAttribute((const void*)&gl::cnn::conv1_wgrad, hipFuncAttributeMaxDynamicSharedMemorySize, lds)))
        return e;
    return hipFuncSetAttribute((const void*)&gl::cnn::conv_wgrad<2>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
}

extern "C" int gl_cnn_create(const gl_cnn_desc* desc, gl_cnn** out) {
    if (out) *out = nullptr;
    if (!desc || !out) return fail(GL_ERR_ARG, "gl_cnn_create: null argument");
    if (desc->height < 4 || desc->width < 4 || desc->max_batch <= 0)
        return fail(GL_ERR_ARG, "gl_cnn_create: bad image size or max_batch");
    for (int i = 0; i < 4; ++i)
        if (desc->channels[i] <= 0 || desc->channels[i] % 32 || desc->channels[i] > gl::cnn::kHeadMaxC)
            return fail(GL_ERR_ARG, "gl_cnn_create: channels must be positive multiples of 32, at most 256");
    if (desc->hidden <= 0 || desc->hidden > gl::cnn::kHeadMaxHd) return fail(GL_ERR_ARG, "gl_cnn_create: hidden must be in [1, 512]");
    if (desc->n_out <= 0 || 2 * desc->n_out > gl::cnn::kHeadMaxOut) return fail(GL_ERR_ARG, "gl_cnn_create: n_out must be in [1, 128]");
    if (!(desc->in_scale > 0.0f)) return fail(GL_ERR_ARG, "gl_cnn_create: in_scale must be positive");
    gl_cnn* net = new gl_cnn();
    if (net->guard.init() != hipSuccess) {
        delete net;
        return fail(GL_ERR_HIP, "gl_cnn_create: event creation failed");
    }
    net->H = desc->height; net->W = desc->width; net->d = desc->n_out; net->hidden = desc->hidden;
    net->maxb = desc->max_batch; net->in_scale = desc->in_scale;
    {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
            gl_cnn_destroy(net);
            return fail(GL_ERR_HIP, "gl_cnn_create: device properties");
        }
        net->cus = std::max(1, prop.multiProcessorCount);
    }
    const int strides[4] = {1, 2, 1, 2};
    size_t off = 0;
    int H = net->H, W = net->W, cin = 1;
    for (int l = 0; l < 4; ++l) {
        gl_cnn::Layer& L = net->L[l];
        L.cin = cin; L.cout = desc->channels[l]; L.s = strides[l];
        L.H = H; L.W = W;
        L.Ho = (H - 1) / L.s + 1; L.Wo = (W - 1) / L.s + 1;
        L.kp = (9 * cin + 31) / 32 * 32;
        L.rw = 1;
        L.fwd = L.dgrad = WavePass{nullptr, 0, 0, 0, 0, 0};
        L.wblk = L.wncg = L.wnp = L.wchunks = 0;
        if (l == 0 && (conv1_fwd_lds(L) > kCnnLds || conv1_wgrad_lds(L) > kCnnLds)) {
            gl_cnn_destroy(net);
            return fail(GL_ERR_UNSUPPORTED, "gl_cnn_create: image too large for the first layer's LDS tiles");
        }
        if (l > 0) {
            L.rw = pick_rows(L.Ho, wgrad_ok, L);
            if (!L.rw) {
                gl_cnn_destroy(net);
                return fail(GL_ERR_UNSUPPORTED, "gl_cnn_create: image rows too wide for the LDS band of layer " +
                                                    std::to_string(l + 1));
            }
            if (wave_setup(L.fwd, 0, L.s, cin, L.cout, net->cus) != hipSuccess ||
                wave_setup(L.dgrad, 1, L.s, L.cout, cin, net->cus) != hipSuccess) {
                gl_cnn_destroy(net);
                return fail(GL_ERR_HIP, "gl_cnn_create: convolution kernel attributes");
            }
            L.wblk = std::min(desc->max_batch, 512);
            L.wncg = std::min(L.cout / 16, 4);
            L.wnp = gl::cnn::kWaves * gl::cnn::kMaxPairs;
            const int pairs = 9 * (cin / 16) + 1;
            L.wchunks = ((L.cout / 16 + L.wncg - 1) / L.wncg) * ((pairs + L.wnp - 1) / L.wnp);
        }
        L.woff = off; off += (size_t)L.cout * 9 * cin;
        L.boff = off; off += L.cout;
        H = L.Ho; W = L.Wo; cin = L.cout;
    }
    // the convolution passes index activations with 32-bit element offsets
    for (int l = 0; l < 4; ++l)
        if ((size_t)net->maxb * net->L[l].H * net->L[l].W * std::max(net->L[l].cin, net->L[l].cout) >= (1u << 31)) {
            gl_cnn_destroy(net);
            return fail(GL_ERR_UNSUPPORTED, "gl_cnn_create: max_batch x height x width x channels must stay below 2^31");
        }
    const int c4 = net->L[3].cout, Hd = net->hidden, d2 = 2 * net->d;
    net->w1 = off; off += (size_t)Hd * c4;
    net->b1 = off; off += Hd;
    net->w2 = off; off += (size_t)d2 * Hd;
    net->b2 = off; off += d2;
    net->P = (int)off;
    const size_t B = net->maxb;
    void* p = nullptr;
    int rc;
#define CNN_ALLOC(field, type, count)                                                  \
    if ((rc = cnn_alloc(net, (size_t)(count) * sizeof(type), &p))) { gl_cnn_destroy(net); return rc; } \
    field = (type*)p;
    CNN_ALLOC(net->params, float, net->P);
    CNN_ALLOC(net->m, float, net->P);
    CNN_ALLOC(net->v, float, net->P);
    CNN_ALLOC(net->grad, float, net->P);
    CNN_ALLOC(net->x0, gl::cnn::u16, B * net->H * net->W);
    size_t part = 0;
    for (int l = 0; l < 4; ++l) {
        gl_cnn::Layer& L = net->L[l];
        const size_t act = B * L.Ho * L.Wo * L.cout;
        CNN_ALLOC(L.wt, gl::cnn::u16, (size_t)L.cout * L.kp);
        L.w2t = nullptr;
        if (l > 0) { CNN_ALLOC(L.w2t, gl::cnn::u16, (size_t)L.cin * 9 * L.cout); }
        // Y and dZ, the operands of the convolution passes, end in kZeroTail zeros that
        // out-of-image taps read (never written: every pass writes the first bs images)
        CNN_ALLOC(L.Y, gl::cnn::u16, act + kZeroTail);
        CNN_ALLOC(L.Z, gl::cnn::u16, act);
        CNN_ALLOC(L.dZ, gl::cnn::u16, act + kZeroTail);
        const size_t sp = l > 0 ? (size_t)L.wblk : B;
        part = std::max(part, sp * L.cout * (9 * L.cin + 1));
    }
    // the head's weight-gradient chunks share the partial buffer
    part = std::max(part, (size_t)gl::cnn::kHeadChunks * ((size_t)Hd * c4 + Hd + (size_t)d2 * Hd + d2));
    net->part_elems = part;
    CNN_ALLOC(net->part, float, part);
    CNN_ALLOC(net->h0, float, B * c4);
    CNN_ALLOC(net->h1, float, B * Hd);
    CNN_ALLOC(net->da1, float, B * Hd);
    CNN_ALLOC(net->out, float, B * d2);
    CNN_ALLOC(net->dout, float, B * d2);
    CNN_ALLOC(net->lossb, float, B);
#undef CNN_ALLOC
    if (cnn_lds_attrs() != hipSuccess) { gl_cnn_destroy(net); return fail(GL_ERR_HIP, "gl_cnn_create: LDS attribute"); }
    if (hipDeviceSynchronize() != hipSuccess) { gl_cnn_destroy(net); return fail(GL_ERR_HIP, "gl_cnn_create: sync"); }
    *out = net;
    return GL_OK;
}

extern "C" int gl_cnn_num_params(const gl_cnn* net) { return net ? net->P : -1; }

static int cnn_pack(gl_cnn* net, hipStream_t st) {
    for (int l = 0; l < 4; ++l) {
        const gl_cnn::Layer& L = net->L[l];
        const int n = L.cout * L.kp;
        hipLaunchKernelGGL(gl::cnn::pack_weights, dim3((n + 255) / 256), dim3(256), 0, st, net->params + L.woff, L.cin,
                           L.cout, L.kp, L.wt, L.w2t);
        GL_HIP(hipGetLastError());
    }
    return GL_OK;
}

extern "C" int gl_cnn_set_params(gl_cnn* net, const float* params, hipStream_t stream) {
    if (!net || !params) return fail(GL_ERR_ARG, "gl_cnn_set_params: null argument");
    GL_HIP(net->guard.enter(stream));
    GL_HIP(hipMemcpyAsync(net->params, params, (size_t)net->P * 4, hipMemcpyDeviceToDevice, stream));
    GL_HIP(hipMemsetAsync(net->m, 0, (size_t)net->P * 4, stream));
    GL_HIP(hipMemsetAsync(net->v, 0, (size_t)net->P * 4, stream));
    return cnn_pack(net, stream);
}

extern "C" int gl_cnn_get_params(const gl_cnn* net, float* params, hipStream_t stream) {
    if (!net || !params) return fail(GL_ERR_ARG, "gl_cnn_get_params: null argument");
    GL_HIP(net->guard.enter(stream));
    GL_HIP(hipMemcpyAsync(params, net->params, (size_t)net->P * 4, hipMemcpyDeviceToDevice, stream));
    return GL_OK;
}

static gl::cnn::HeadArgs cnn_head_args(gl_cnn* net, const float* target, int bs) {
    gl::cnn::HeadArgs A;
    const gl_cnn::Layer& L = net->L[3];
    A.y4 = L.Y; A.z4 = L.Z; A.P = L.Ho * L.Wo; A.C = L.cout; A.Hd = net->hidden; A.d = net->d;
    A.w1 = net->params + net->w1; A.b1 = net->params + net->b1;
    A.w2 = net->params + net->w2; A.b2 = net->params + net->b2;
    A.h0 = net->h0; A.h1 = net->h1; A.da1 = net->da1; A.out = net->out; A.dout = net->dout;
    A.lossb = net->lossb; A.dz4 = L.dZ; A.target = target; A.inv_bs = 1.0f / (float)bs;
    return A;
}

static int cnn_forward_layers(gl_cnn* net, const float* images, int bs, hipStream_t st) {
    using namespace gl::cnn;
    const long n = (long)bs * net->H * net->W;
    hipLaunchKernelGGL(input_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, images, net->x0, n,
                       1.0f / net->in_scale);
    GL_HIP(hipGetLastError());
    {
        const gl_cnn::Layer& L = net->L[0];
        const int R = conv1_rows(L);
        hipLaunchKernelGGL(conv1_fwd, dim3((L.H + R - 1) / R, bs), dim3(256), conv1_fwd_lds(L), st, net->x0, L.wt,
                           L.kp, net->params + L.boff, L.H, L.W, L.cout, R, L.Z, L.Y);
        GL_HIP(hipGetLastError());
    }
    for (int l = 1; l < 4; ++l) {
        const gl_cnn::Layer& L = net->L[l];
        gl::cnn::WaveArgs A{net->L[l - 1].Y, L.wt, net->params + L.boff, nullptr, L.Z, L.Y,
                            bs, L.Ho, L.Wo, L.H, L.W, L.cin, L.cout, 0,
                            (unsigned)((size_t)net->maxb * L.H * L.W * L.cin)};
        int rc;
        if ((rc = wave_launch(L.fwd, A, st))) return rc;
        GL_HIP(hipGetLastError());
    }
    return GL_OK;
}

extern "C" int gl_cnn_forward(gl_cnn* net, const float* images, int bs, float* out, hipStream_t stream) {
    if (!net || !images || !out) return fail(GL_ERR_ARG, "gl_cnn_forward: null argument");
    if (bs <= 0 || bs > net->maxb) return fail(GL_ERR_ARG, "gl_cnn_forward: bs must be in [1, max_batch]");
    GL_HIP(net->guard.enter(stream));
    int rc;
    if ((rc = cnn_forward_layers(net, images, bs, stream))) return rc;
    gl::cnn::HeadArgs A = cnn_head_args(net, nullptr, bs);
    A.out = out;
    hipLaunchKernelGGL(gl::cnn::head_kernel, dim3(bs), dim3(256), 0, stream, A);
    GL_HIP(hipGetLastError());
    return GL_OK;
}

extern "C" int gl_cnn_loss_grad(gl_cnn* net, const float* images, const float* target, int bs, float* loss,
                                float* grad, hipStream_t stream) {
    using namespace gl::cnn;
    if (!net || !images || !target) return fail(GL_ERR_ARG, "gl_cnn_loss_grad: null argument");
    if (bs <= 0 || bs > net->maxb) return fail(GL_ERR_ARG, "gl_cnn_loss_grad: bs must be in [1, max_batch]");
    GL_HIP(net->guard.enter(stream));
    int rc;
    if ((rc = cnn_forward_layers(net, images, bs, stream))) return rc;
    HeadArgs A = cnn_head_args(net, target, bs);
    hipLaunchKernelGGL(head_kernel, dim3(bs), dim3(256), 0, stream, A);
    GL_HIP(hipGetLastError());
    {
        const int n = net->hidden * net->L[3].cout + net->hidden + 2 * net->d * net->hidden + 2 * net->d;
        // the layers' weight-gradient partials reuse net->part after this (stream order)
        hipLaunchKernelGGL(head_wgrad, dim3((n + 255) / 256, kHeadChunks), dim3(256), 0, stream, A, bs, net->part);
        GL_HIP(hipGetLastError());
        hipLaunchKernelGGL(head_wgrad_reduce, dim3((n + 255) / 256), dim3(256), 0, stream, A, net->part,
                           net->grad + net->w1, net->grad + net->b1, net->grad + net->w2, net->grad + net->b2);
        GL_HIP(hipGetLastError());
    }
    if (loss) {
        hipLaunchKernelGGL(loss_mean, dim3(1), dim3(256), 0, stream, net->lossb, bs, loss);
        GL_HIP(hipGetLastError());
    }
    for (int l = 3; l >= 0; --l) {
        const gl_cnn::Layer& L = net->L[l];
        const int ncol = 9 * L.cin + 1;
        // weight (and bias) gradient: per-block sums over the pixels, then an ordered reduction
        int splits;
        if (l == 0) {
            splits = bs;
            hipLaunchKernelGGL(conv1_wgrad, dim3(bs), dim3(256), conv1_wgrad_lds(L), stream, net->x0, L.dZ, L.H, L.W,
                               L.cout, net->part);
        } else {
            splits = std::min(L.wblk, bs);
            WgradArgs A{L.dZ, net->L[l - 1].Y, net->part, bs, L.H, L.W, L.cin, L.cout, L.Ho, L.Wo, L.rw, L.wncg, L.wnp,
                        wgrad_wp(L)};
            const dim3 grid(splits, L.wchunks);
            const size_t lds = wgrad_lds(L, L.rw);
            if (L.s == 1) hipLaunchKernelGGL(conv_wgrad<1>, grid, dim3(512), lds, stream, A);
            else hipLaunchKernelGGL(conv_wgrad<2>, grid, dim3(512), lds, stream, A);
        }
        GL_HIP(hipGetLastError());
        if ((size_t)splits * L.cout * ncol > net->part_elems) return fail(GL_ERR_ARG, "gl_cnn_loss_grad: partial buffer");
        hipLaunchKernelGGL(wgrad_reduce, dim3((L.cout * ncol + kWrOut - 1) / kWrOut), dim3(256), 0, stream, net->part,
                           splits, L.cout, ncol, net->grad + L.woff, net->grad + L.boff);
        GL_HIP(hipGetLastError());
        if (l == 0) break;
        // data gradient into the previous layer: dZprev = conv^T(dZ, W) * gelu'(Zprev)
        const gl_cnn::Layer& Lp = net->L[l - 1];
        gl::cnn::WaveArgs A{L.dZ, L.w2t, nullptr, Lp.Z, Lp.dZ, nullptr, bs, L.H, L.W, L.Ho, L.Wo, L.cout, L.cin, 0,
                            (unsigned)((size_t)net->maxb * L.Ho * L.Wo * L.cout)};
        if ((rc = wave_launch(L.dgrad, A, stream))) return rc;
    }
    if (grad) GL_HIP(hipMemcpyAsync(grad, net->grad, (size_t)net->P * 4, hipMemcpyDeviceToDevice, stream));
    return GL_OK;
}

extern "C" int gl_cnn_adamw_step(gl_cnn* net, int step, float lr, float beta1, float beta2, float eps,
                                 float weight_decay, hipStream_t stream) {
    if (!net) return fail(GL_ERR_ARG, "gl_cnn_adamw_step: null network");
    if (step < 1) return fail(GL_ERR_ARG, "gl_cnn_adamw_step: step is 1-based");
    GL_HIP(net->guard.enter(stream));
    const float bc1 = (float)(1.0 - pow((double)beta1, step));
    const float bc2s = (float)sqrt(1.0 - pow((double)beta2, step));
    hipLaunchKernelGGL(gl::cnn::adamw, dim3((net->P + 255) / 256), dim3(256), 0, stream, net->params, net->m, net->v,
                       net->grad, net->P, lr, beta1, beta2, eps, weight_decay, bc1, bc2s);
    GL_HIP(hipGetLastError());
    return cnn_pack(net, stream);
}
