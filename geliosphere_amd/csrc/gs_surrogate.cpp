// gs_surrogate.cpp - host side of the MLP-ensemble surrogate (C ABI in
// include/gs_surrogate.h).  All tensors live on the device; the host only
// draws initial weights and minibatch orders and launches kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/gs_surrogate.h"
#include "gs_surrogate_kernels.h"

namespace {

constexpr int kPad = 32;   // padded width of the input and output layers

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct gs_surrogate {
    int device = 0;
    hipStream_t stream = nullptr;
    gs_surrogate_config cfg{};
    int E = 0, L = 0;                  // members, weight layers (hidden layers + 1)
    std::vector<int> dims;             // dims[0] = kPad, hidden..., dims[L] = kPad
    std::vector<size_t> woff, boff;    // per-member offsets (floats)
    size_t Pm = 0;                     // floats per member
    float *P = nullptr, *G = nullptr, *Mo = nullptr, *Ve = nullptr;
    __bf16 *Pb = nullptr;
    uint64_t step = 0;
    bool normalized = false;
    float mu[GS_SG_MAX_INPUTS], sd[GS_SG_MAX_INPUTS], lo[GS_SG_MAX_INPUTS], hi[GS_SG_MAX_INPUTS];
    // work buffers, grown on demand
    std::vector<DevBuf> bufs;
    std::string err;
};

namespace {

gs_status fail(gs_surrogate *s, gs_status st, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (s) s->err = buf;
    return st;
}

#define SG_TRY(s, call)                                                                       \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail((s), e_ == hipErrorOutOfMemory ? GS_ERR_OUT_OF_MEMORY : GS_ERR_HIP,   \
                        "%s failed: %s", #call, hipGetErrorString(e_));                       \
    } while (0)

// Named work buffers (index into s->bufs).
enum Buf {
    B_X = 0,      // data set inputs, bf16 [n][kPad]
    B_XF,         // staging fp32 [n][n_inputs]
    B_Y,          // data set targets fp32 [n][kPad]
    B_W,          // data set weights fp32 [n][kPad]
    B_IDX,        // minibatch rows uint32 [E][B]
    B_BX, B_BY, B_BW,     // minibatch per member
    B_Z,          // pre-activations fp32 [L-1][E][B][H]
    B_ACT,        // activations bf16 [L-1][E][B][H]
    B_OUT,        // outputs fp32 [E][B][kPad]
    B_DYF, B_DYB, // output gradient
    B_DZF0, B_DZB0, B_DZF1, B_DZB1,   // hidden gradients, ping-pong
    B_WSUM, B_LOSS,
    B_MEAN, B_SD,
    B_FLAGS,      // predict: fallback flags uint8 [n]
    B_COUNT_
};

template <class T>
gs_status need(gs_surrogate *s, int which, size_t count, T **out) {
    DevBuf &b = s->bufs[which];
    const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    if (b.bytes < bytes) {
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
        SG_TRY(s, hipMalloc(&b.p, bytes));
        b.bytes = bytes;
    }
    *out = (T *)b.p;
    return GS_OK;
}

#define SG_NEED(s, which, count, ptr)                       \
    do {                                                    \
        gs_status st_ = need((s), (which), (count), &(ptr)); \
        if (st_ != GS_OK) return st_;                       \
    } while (0)

// Forward pass for `rows` rows per member.  X: bf16 [rows][kPad] with
// member stride xstride (0: all members read the same rows).  Zbuf null
// skips the pre-activations (inference).
gs_status forward(gs_surrogate *s, const __bf16 *X, size_t xstride, int rows, float *Zbuf, __bf16 *Act,
                  float *Out) {
    const int H = (int)s->cfg.hidden;
    const size_t layer_stride = (size_t)s->E * rows * H;
    for (int l = 0; l < s->L; ++l) {
        GemmArgs g{};
        const bool last = l == s->L - 1;
        g.A = l == 0 ? X : Act + (size_t)(l - 1) * layer_stride;
        g.strideA = l == 0 ? xstride : (size_t)rows * H;
        g.lda = s->dims[l];
        g.B = s->Pb + s->woff[l];
        g.strideB = s->Pm;
        g.ldb = s->dims[l];
        g.M = rows;
        g.N = s->dims[l + 1];
        g.K = s->dims[l];
        g.ldc = s->dims[l + 1];
        g.strideC = (size_t)rows * s->dims[l + 1];
        g.bias = s->P + s->boff[l];
        g.strideBias = s->Pm;
        if (last) {
            g.Cf = Out;
            SG_TRY(s, sg_launch_gemm(false, false, EPI_OUT, g, s->E, s->stream));
        } else {
            g.Zf = Zbuf ? Zbuf + (size_t)l * layer_stride : nullptr;
            g.Cb = Act + (size_t)l * layer_stride;
            SG_TRY(s, sg_launch_gemm(false, false, EPI_HIDDEN, g, s->E, s->stream));
        }
    }
    return GS_OK;
}

// Inference forward pass into Out [E][rows][kPad]: the fused kernel when the
// network has its shape (hidden 256, at most kMlpMaxLayers weight layers),
// otherwise the layer-by-layer GEMMs through the activation buffer.
bool fused_shape(const gs_surrogate *s) {
    return s->cfg.hidden == 256 && s->L >= 2 && s->L <= kMlpMaxLayers && s->dims[0] == 32 && s->dims[s->L] == 32;
}

gs_status forward_infer(gs_surrogate *s, const __bf16 *X, int rows, float *Out) {
    if (fused_shape(s)) {
        MlpArgs g{};
        g.X = X;
        g.rows = rows;
        g.L = s->L;
        g.Pb = s->Pb;
        g.P = s->P;
        g.Pm = s->Pm;
        for (int l = 0; l < s->L; ++l) {
            g.woff[l] = s->woff[l];
            g.boff[l] = s->boff[l];
        }
        g.Out = Out;
        SG_TRY(s, sg_launch_mlp_forward(g, s->E, s->stream));
        return GS_OK;
    }
    __bf16 *act;
    SG_NEED(s, B_ACT, (size_t)(s->L - 1) * s->E * rows * s->cfg.hidden, act);
    return forward(s, X, 0, rows, nullptr, act, Out);
}

// Backward pass from dY (fp32 and bf16 [E][rows][kPad]) into s->G.
gs_status backward(gs_surrogate *s, const __bf16 *X, size_t xstride, int rows, const float *Zbuf,
                   const __bf16 *Act, const float *dYf, const __bf16 *dYb) {
    const int H = (int)s->cfg.hidden;
    const size_t layer_stride = (size_t)s->E * rows * H;
    float *dzf[2];
    __bf16 *dzb[2];
    SG_NEED(s, B_DZF0, layer_stride, dzf[0]);
    SG_NEED(s, B_DZB0, layer_stride, dzb[0]);
    SG_NEED(s, B_DZF1, layer_stride, dzf[1]);
    SG_NEED(s, B_DZB1, layer_stride, dzb[1]);
    const float *cur_f = dYf;
    const __bf16 *cur_b = dYb;
    for (int l = s->L - 1; l >= 0; --l) {
        const int out_w = s->dims[l + 1], in_w = s->dims[l];
        const __bf16 *in_act = l == 0 ? X : Act + (size_t)(l - 1) * layer_stride;
        const size_t in_stride = l == 0 ? xstride : (size_t)rows * H;
        // dW[o][i] = sum_b dZ[b][o] in[b][i]
        GemmArgs gw{};
        gw.A = cur_b;
        gw.strideA = (size_t)rows * out_w;
        gw.lda = out_w;
        gw.B = in_act;
        gw.strideB = in_stride;
        gw.ldb = in_w;
        gw.M = out_w;
        gw.N = in_w;
        gw.K = rows;
        gw.ldc = in_w;
        gw.strideC = s->Pm;
        gw.Cf = s->G + s->woff[l];
        SG_TRY(s, sg_launch_gemm(true, true, EPI_STORE, gw, s->E, s->stream));
        SG_TRY(s, sg_launch_colsum(cur_f, rows, out_w, (size_t)rows * out_w, s->G + s->boff[l], s->Pm, s->E,
                                   s->stream));
        if (l == 0) break;
        // dZ_prev[b][i] = (sum_o dZ[b][o] W[o][i]) silu'(Z_prev[b][i])
        const int nxt = (l & 1);
        GemmArgs gd{};
        gd.A = cur_b;
        gd.strideA = (size_t)rows * out_w;
        gd.lda = out_w;
        gd.B = s->Pb + s->woff[l];
        gd.strideB = s->Pm;
        gd.ldb = in_w;
        gd.M = rows;
        gd.N = in_w;
        gd.K = out_w;
        gd.ldc = in_w;
        gd.strideC = (size_t)rows * in_w;
        gd.Zf = const_cast<float *>(Zbuf) + (size_t)(l - 1) * layer_stride;
        gd.Cf = dzf[nxt];
        gd.Cb = dzb[nxt];
        SG_TRY(s, sg_launch_gemm(false, true, EPI_DSILU, gd, s->E, s->stream));
        cur_f = dzf[nxt];
        cur_b = dzb[nxt];
    }
    return GS_OK;
}

// fp32 rows [n][n_inputs] (host memory, or device memory if x_on_device) ->
// standardised, padded bf16 rows [n][kPad] on the device.
gs_status upload_inputs(gs_surrogate *s, const float *x, uint32_t n, __bf16 **out, bool x_on_device = false) {
    const uint32_t ni = s->cfg.n_inputs;
    float *xf;
    __bf16 *xb;
    SG_NEED(s, B_XF, (size_t)n * ni, xf);
    SG_NEED(s, B_X, (size_t)n * kPad, xb);
    SG_TRY(s, hipMemcpyAsync(xf, x, (size_t)n * ni * sizeof(float),
                             x_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s->stream));
    SG_TRY(s, sg_launch_standardise(xf, (int)n, (int)ni, kPad, s->mu, s->sd, xb, s->stream));
    SG_TRY(s, hipStreamSynchronize(s->stream));
    *out = xb;
    return GS_OK;
}

std::vector<float> pad_rows(const float *v, uint32_t n, uint32_t width) {
    std::vector<float> h((size_t)n * kPad, 0.0f);
    for (uint32_t r = 0; r < n; ++r)
        for (uint32_t c = 0; c < width; ++c) h[(size_t)r * kPad + c] = v[(size_t)r * width + c];
    return h;
}

bool check_cfg(const gs_surrogate_config *c, std::string &msg) {
    if (c->n_members < 1 || c->n_members > 64) { msg = "n_members must be 1..64"; return false; }
    if (c->n_inputs < 1 || c->n_inputs > GS_SG_MAX_INPUTS) { msg = "n_inputs must be 1..32"; return false; }
    if (c->n_outputs < 1 || c->n_outputs > GS_SG_MAX_OUTPUTS) { msg = "n_outputs must be 1..32"; return false; }
    if (c->hidden < 64 || c->hidden % 64 || c->hidden > 4096) { msg = "hidden must be a multiple of 64 up to 4096"; return false; }
    if (c->n_hidden_layers < 1 || c->n_hidden_layers > 16) { msg = "n_hidden_layers must be 1..16"; return false; }
    if (c->batch < 64 || c->batch % 64 || c->batch > 65536) { msg = "batch must be a multiple of 64 up to 65536"; return false; }
    if (!(c->lr > 0) || !(c->beta1 >= 0 && c->beta1 < 1) || !(c->beta2 >= 0 && c->beta2 < 1) || !(c->eps > 0)) {
        msg = "bad Adam hyper-parameters";
        return false;
    }
    return true;
}

}  // namespace

extern "C" {

gs_status gs_surrogate_default_config(gs_surrogate_config *c) {
    if (!c) return GS_ERR_INVALID_ARG;
    std::memset(c, 0, sizeof *c);
    c->n_members = 4;
    c->n_inputs = 8;
    c->n_outputs = 30;
    c->hidden = 256;
    c->n_hidden_layers = 4;
    c->batch = 1024;
    c->seed = 1;
    c->lr = 1e-3f;
    c->beta1 = 0.9f;
    c->beta2 = 0.999f;
    c->eps = 1e-8f;
    c->weight_decay = 0.0f;
    c->std_limit = 0.01f;
    c->std_outputs = 0xffffffffu;
    return GS_OK;
}

const char *gs_surrogate_last_error(const gs_surrogate *s) { return s ? s->err.c_str() : "no surrogate"; }

gs_status gs_surrogate_create(int device, const gs_surrogate_config *cfg, gs_surrogate **out) {
    if (!out || !cfg) return GS_ERR_INVALID_ARG;
    *out = nullptr;
    std::string msg;
    if (!check_cfg(cfg, msg)) return GS_ERR_INVALID_ARG;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0) return GS_ERR_NO_DEVICE;
    if (device < 0 || device >= nd) return GS_ERR_INVALID_ARG;
    gs_surrogate *s = new gs_surrogate();
    s->device = device;
    s->cfg = *cfg;
    s->E = (int)cfg->n_members;
    s->L = (int)cfg->n_hidden_layers + 1;
    s->bufs.resize(B_COUNT_);
    s->dims.push_back(kPad);
    for (uint32_t i = 0; i < cfg->n_hidden_layers; ++i) s->dims.push_back((int)cfg->hidden);
    s->dims.push_back(kPad);
    size_t off = 0;
    for (int l = 0; l < s->L; ++l) {
        s->woff.push_back(off);
        off += (size_t)s->dims[l + 1] * s->dims[l];
        s->boff.push_back(off);
        off += (size_t)s->dims[l + 1];
    }
    s->Pm = (off + 7) / 8 * 8;   // keeps every member's block 16-byte aligned in bf16
    for (int i = 0; i < GS_SG_MAX_INPUTS; ++i) { s->mu[i] = 0.f; s->sd[i] = 1.f; s->lo[i] = -INFINITY; s->hi[i] = INFINITY; }
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    const size_t np = (size_t)s->E * s->Pm;
    if (e == hipSuccess) e = hipMalloc((void **)&s->P, np * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&s->G, np * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&s->Mo, np * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&s->Ve, np * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&s->Pb, np * sizeof(__bf16));
    if (e != hipSuccess) {
        gs_surrogate_destroy(s);
        return e == hipErrorOutOfMemory ? GS_ERR_OUT_OF_MEMORY : GS_ERR_HIP;
    }
    // He-uniform weights over the real fan-in, zero biases and padding;
    // the output layer starts ten times smaller.
    std::vector<float> h(np, 0.0f);
    for (int m = 0; m < s->E; ++m) {
        std::mt19937_64 rng(cfg->seed * 0x9E3779B97F4A7C15ull + (uint64_t)m * 1000003ull + 17);
        for (int l = 0; l < s->L; ++l) {
            const int out_real = l == s->L - 1 ? (int)cfg->n_outputs : s->dims[l + 1];
            const int in_real = l == 0 ? (int)cfg->n_inputs : s->dims[l];
            const float a = std::sqrt(6.0f / in_real) * (l == s->L - 1 ? 0.1f : 1.0f);
            std::uniform_real_distribution<float> U(-a, a);
            for (int o = 0; o < out_real; ++o)
                for (int i = 0; i < in_real; ++i)
                    h[(size_t)m * s->Pm + s->woff[l] + (size_t)o * s->dims[l] + i] = U(rng);
        }
    }
    gs_status st = gs_surrogate_set_params(s, h.data(), h.size());
    if (st != GS_OK) {
        gs_surrogate_destroy(s);
        return st;
    }
    *out = s;
    return GS_OK;
}

void gs_surrogate_destroy(gs_surrogate *s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (DevBuf &b : s->bufs)
        if (b.p) (void)hipFree(b.p);
    for (void *p : {(void *)s->P, (void *)s->G, (void *)s->Mo, (void *)s->Ve, (void *)s->Pb})
        if (p) (void)hipFree(p);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

gs_status gs_surrogate_layout(const gs_surrogate *s, uint64_t *per_member, uint64_t *layout, uint32_t cap) {
    if (!s) return GS_ERR_INVALID_ARG;
    if (per_member) *per_member = s->Pm;
    if (layout) {
        if (cap < 4u * (uint32_t)s->L) return GS_ERR_INVALID_ARG;
        for (int l = 0; l < s->L; ++l) {
            layout[4 * l + 0] = s->woff[l];
            layout[4 * l + 1] = s->boff[l];
            layout[4 * l + 2] = (uint64_t)s->dims[l + 1];
            layout[4 * l + 3] = (uint64_t)s->dims[l];
        }
    }
    return GS_OK;
}

gs_status gs_surrogate_get_params(gs_surrogate *s, float *p, uint64_t count) {
    if (!s || !p) return GS_ERR_INVALID_ARG;
    if (count != (uint64_t)s->E * s->Pm) return fail(s, GS_ERR_INVALID_ARG, "count must be %zu", (size_t)s->E * s->Pm);
    SG_TRY(s, hipSetDevice(s->device));
    SG_TRY(s, hipMemcpyAsync(p, s->P, count * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    SG_TRY(s, hipStreamSynchronize(s->stream));
    return GS_OK;
}

gs_status gs_surrogate_set_params(gs_surrogate *s, const float *p, uint64_t count) {
    if (!s || !p) return GS_ERR_INVALID_ARG;
    if (count != (uint64_t)s->E * s->Pm) return fail(s, GS_ERR_INVALID_ARG, "count must be %zu", (size_t)s->E * s->Pm);
    SG_TRY(s, hipSetDevice(s->device));
    SG_TRY(s, hipMemcpyAsync(s->P, p, count * sizeof(float), hipMemcpyHostToDevice, s->stream));
    SG_TRY(s, hipMemsetAsync(s->Mo, 0, count * sizeof(float), s->stream));
    SG_TRY(s, hipMemsetAsync(s->Ve, 0, count * sizeof(float), s->stream));
    SG_TRY(s, sg_launch_to_bf16(s->P, s->Pb, count, s->stream));
    SG_TRY(s, hipStreamSynchronize(s->stream));
    s->step = 0;
    return GS_OK;
}

gs_status gs_surrogate_get_normalization(const gs_surrogate *s, float *mean, float *sd, float *lo, float *hi) {
    if (!s) return GS_ERR_INVALID_ARG;
    for (uint32_t i = 0; i < s->cfg.n_inputs; ++i) {
        if (mean) mean[i] = s->mu[i];
        if (sd) sd[i] = s->sd[i];
        if (lo) lo[i] = s->lo[i];
        if (hi) hi[i] = s->hi[i];
    }
    return GS_OK;
}

gs_status gs_surrogate_set_normalization(gs_surrogate *s, const float *mean, const float *sd, const float *lo,
                                         const float *hi) {
    if (!s || !mean || !sd || !lo || !hi) return GS_ERR_INVALID_ARG;
    for (uint32_t i = 0; i < s->cfg.n_inputs; ++i) {
        if (!(sd[i] > 0) || !std::isfinite(mean[i])) return fail(s, GS_ERR_INVALID_ARG, "bad normalization");
        s->mu[i] = mean[i];
        s->sd[i] = sd[i];
        s->lo[i] = lo[i];
        s->hi[i] = hi[i];
    }
    s->normalized = true;
    return GS_OK;
}

gs_status gs_surrogate_train(gs_surrogate *s, const float *x, const float *y, const float *w, uint32_t n,
                             uint32_t epochs, float *last_loss) {
    if (!s) return GS_ERR_INVALID_ARG;
    if (!x || !y || !w) return fail(s, GS_ERR_INVALID_ARG, "NULL data");
    const int B = (int)s->cfg.batch;
    if (n < (uint32_t)B) return fail(s, GS_ERR_INVALID_ARG, "need at least one minibatch (%d rows)", B);
    SG_TRY(s, hipSetDevice(s->device));
    const uint32_t ni = s->cfg.n_inputs, no = s->cfg.n_outputs;
    for (size_t i = 0; i < (size_t)n * no; ++i)
        if (!std::isfinite(y[i]) && w[i] != 0.0f) return fail(s, GS_ERR_INVALID_ARG, "non-finite target with weight");
    if (!s->normalized) {
        for (uint32_t c = 0; c < ni; ++c) {
            double m = 0, m2 = 0;
            float lo = INFINITY, hi = -INFINITY;
            for (uint32_t r = 0; r < n; ++r) {
                const double v = x[(size_t)r * ni + c];
                m += v;
                m2 += v * v;
                lo = std::min(lo, (float)v);
                hi = std::max(hi, (float)v);
            }
            m /= n;
            const double var = std::max(0.0, m2 / n - m * m);
            s->mu[c] = (float)m;
            s->sd[c] = var > 1e-12 ? (float)std::sqrt(var) : 1.0f;
            s->lo[c] = lo;
            s->hi[c] = hi;
        }
        s->normalized = true;
    }
    __bf16 *X;
    gs_status st = upload_inputs(s, x, n, &X);
    if (st != GS_OK) return st;
    std::vector<float> yh = pad_rows(y, n, no), wh = pad_rows(w, n, no);
    for (size_t i = 0; i < yh.size(); ++i)
        if (wh[i] == 0.0f) yh[i] = 0.0f;   // unweighted targets may be NaN (empty bins)
    float *Y, *W;
    SG_NEED(s, B_Y, yh.size(), Y);
    SG_NEED(s, B_W, wh.size(), W);
    SG_TRY(s, hipMemcpyAsync(Y, yh.data(), yh.size() * sizeof(float), hipMemcpyHostToDevice, s->stream));
    SG_TRY(s, hipMemcpyAsync(W, wh.data(), wh.size() * sizeof(float), hipMemcpyHostToDevice, s->stream));
    const int E = s->E, H = (int)s->cfg.hidden;
    uint32_t *idx;
    __bf16 *bx, *act, *dyb;
    float *by, *bw, *Z, *out, *dyf, *wsum, *loss;
    const int steps = (int)(n / (uint32_t)B);
    const int parts = 8;
    SG_NEED(s, B_IDX, (size_t)E * B * steps, idx);
    SG_NEED(s, B_BX, (size_t)E * B * kPad, bx);
    SG_NEED(s, B_BY, (size_t)E * B * kPad, by);
    SG_NEED(s, B_BW, (size_t)E * B * kPad, bw);
    SG_NEED(s, B_Z, (size_t)(s->L - 1) * E * B * H, Z);
    SG_NEED(s, B_ACT, (size_t)(s->L - 1) * E * B * H, act);
    SG_NEED(s, B_OUT, (size_t)E * B * kPad, out);
    SG_NEED(s, B_DYF, (size_t)E * B * kPad, dyf);
    SG_NEED(s, B_DYB, (size_t)E * B * kPad, dyb);
    SG_NEED(s, B_WSUM, (size_t)E, wsum);
    SG_NEED(s, B_LOSS, (size_t)E * parts * steps, loss);
    std::vector<uint32_t> order(n), hidx((size_t)E * B * steps);
    std::vector<float> hl((size_t)E * parts * steps);
    const gs_surrogate_config &c = s->cfg;
    double last = 0.0;
    for (uint32_t ep = 0; ep < epochs; ++ep) {
        for (int m = 0; m < E; ++m) {
            std::mt19937_64 rng(c.seed ^ (0xD1B54A32D192ED03ull * (s->step + 1)) ^ ((uint64_t)m << 48));
            for (uint32_t i = 0; i < n; ++i) order[i] = i;
            std::shuffle(order.begin(), order.end(), rng);
            for (int k = 0; k < steps; ++k)
                std::copy(order.begin() + (size_t)k * B, order.begin() + (size_t)(k + 1) * B,
                          hidx.begin() + ((size_t)k * E + m) * B);
        }
        SG_TRY(s, hipMemcpyAsync(idx, hidx.data(), hidx.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s->stream));
        for (int k = 0; k < steps; ++k) {
            SG_TRY(s, sg_launch_gather(X, Y, W, idx + (size_t)k * E * B, B, kPad, kPad, bx, by, bw, E, s->stream));
            st = forward(s, bx, (size_t)B * kPad, B, Z, act, out);
            if (st != GS_OK) return st;
            SG_TRY(s, sg_launch_loss(out, by, bw, B * kPad, (size_t)B * kPad, (size_t)B * kPad, wsum, dyf, dyb,
                                     loss + (size_t)k * E * parts, parts, E, s->stream));
            st = backward(s, bx, (size_t)B * kPad, B, Z, act, dyf, dyb);
            if (st != GS_OK) return st;
            ++s->step;
            const float c1 = 1.0f - std::pow(c.beta1, (float)s->step), c2 = 1.0f - std::pow(c.beta2, (float)s->step);
            SG_TRY(s, sg_launch_adam(s->P, s->G, s->Mo, s->Ve, s->Pb, (size_t)E * s->Pm, c.lr, c.beta1, c.beta2,
                                     c.eps, c.weight_decay, c1, c2, s->stream));
        }
        if (ep + 1 == epochs) {
            SG_TRY(s, hipMemcpyAsync(hl.data(), loss, hl.size() * sizeof(float), hipMemcpyDeviceToHost, s->stream));
            SG_TRY(s, hipStreamSynchronize(s->stream));
            double tot = 0;
            for (float v : hl) tot += v;
            last = tot / ((double)steps * E);
        }
    }
    SG_TRY(s, hipStreamSynchronize(s->stream));
    if (last_loss) *last_loss = (float)last;
    return GS_OK;
}

gs_status gs_surrogate_gradients(gs_surrogate *s, const float *x, const float *y, const float *w, uint32_t n,
                                 float *grads, float *loss_out) {
    if (!s) return GS_ERR_INVALID_ARG;
    if (!x || !y || !w || !grads || n == 0 || n > 65536) return fail(s, GS_ERR_INVALID_ARG, "bad arguments");
    SG_TRY(s, hipSetDevice(s->device));
    const uint32_t rows = (n + 63) / 64 * 64;   // padded rows carry zero weight
    std::vector<float> xp((size_t)rows * s->cfg.n_inputs, 0.0f), yp((size_t)rows * s->cfg.n_outputs, 0.0f),
        wp((size_t)rows * s->cfg.n_outputs, 0.0f);
    std::copy(x, x + (size_t)n * s->cfg.n_inputs, xp.begin());
    std::copy(y, y + (size_t)n * s->cfg.n_outputs, yp.begin());
    std::copy(w, w + (size_t)n * s->cfg.n_outputs, wp.begin());
    __bf16 *X;
    gs_status st = upload_inputs(s, xp.data(), rows, &X);
    if (st != GS_OK) return st;
    const int E = s->E, H = (int)s->cfg.hidden;
    std::vector<float> yh = pad_rows(yp.data(), rows, s->cfg.n_outputs), wh = pad_rows(wp.data(), rows, s->cfg.n_outputs);
    // the same rows for every member
    std::vector<float> ye, we;
    for (int m = 0; m < E; ++m) {
        ye.insert(ye.end(), yh.begin(), yh.end());
        we.insert(we.end(), wh.begin(), wh.end());
    }
    float *by, *bw, *Z, *out, *dyf, *wsum, *loss;
    __bf16 *act, *dyb;
    const int parts = 8;
    SG_NEED(s, B_BY, ye.size(), by);
    SG_NEED(s, B_BW, we.size(), bw);
    SG_NEED(s, B_Z, (size_t)(s->L - 1) * E * rows * H, Z);
    SG_NEED(s, B_ACT, (size_t)(s->L - 1) * E * rows * H, act);
    SG_NEED(s, B_OUT, (size_t)E * rows * kPad, out);
    SG_NEED(s, B_DYF, (size_t)E * rows * kPad, dyf);
    SG_NEED(s, B_DYB, (size_t)E * rows * kPad, dyb);
    SG_NEED(s, B_WSUM, (size_t)E, wsum);
    SG_NEED(s, B_LOSS, (size_t)E * parts, loss);
    SG_TRY(s, hipMemcpyAsync(by, ye.data(), ye.size() * sizeof(float), hipMemcpyHostToDevice, s->stream));
    SG_TRY(s, hipMemcpyAsync(bw, we.data(), we.size() * sizeof(float), hipMemcpyHostToDevice, s->stream));
    st = forward(s, X, 0, (int)rows, Z, act, out);
    if (st != GS_OK) return st;
    SG_TRY(s, sg_launch_loss(out, by, bw, (int)rows * kPad, (size_t)rows * kPad, (size_t)rows * kPad, wsum, dyf, dyb,
                             loss, parts, E, s->stream));
    // The first layer reads the shared inputs (member stride 0).
    st = backward(s, X, 0, (int)rows, Z, act, dyf, dyb);
    if (st != GS_OK) return st;
    SG_TRY(s, hipMemcpyAsync(grads, s->G, (size_t)E * s->Pm * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    std::vector<float> hl((size_t)E * parts);
    SG_TRY(s, hipMemcpyAsync(hl.data(), loss, hl.size() * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    SG_TRY(s, hipStreamSynchronize(s->stream));
    if (loss_out)
        for (int m = 0; m < E; ++m) {
            double t = 0;
            for (int p = 0; p < parts; ++p) t += hl[(size_t)m * parts + p];
            loss_out[m] = (float)t;
        }
    return GS_OK;
}

// Both predict entry points: host buffers (copied through the surrogate's
// stream) or device buffers (the statistics and flags kernels write into the
// caller's rows directly).
static gs_status predict_rows(gs_surrogate *s, const float *x, uint32_t n, float *mean, float *sd,
                              uint8_t *fallback, bool on_device) {
    if (!s) return GS_ERR_INVALID_ARG;
    if (!x && n) return fail(s, GS_ERR_INVALID_ARG, "NULL inputs");
    if (n == 0) return GS_OK;
    SG_TRY(s, hipSetDevice(s->device));
    const uint32_t no = s->cfg.n_outputs, ni = s->cfg.n_inputs;
    const int E = s->E;
    const uint32_t chunk = 65536;
    for (uint32_t r0 = 0; r0 < n; r0 += chunk) {
        const uint32_t rows = std::min(chunk, n - r0);
        __bf16 *X;
        float *out, *dm, *ds;
        gs_status st = upload_inputs(s, x + (size_t)r0 * ni, rows, &X, on_device);
        if (st != GS_OK) return st;
        SG_NEED(s, B_OUT, (size_t)E * rows * kPad, out);
        if (on_device && mean) dm = mean + (size_t)r0 * no;
        else SG_NEED(s, B_MEAN, (size_t)rows * no, dm);
        if (on_device && sd) ds = sd + (size_t)r0 * no;
        else SG_NEED(s, B_SD, (size_t)rows * no, ds);
        st = forward_infer(s, X, (int)rows, out);
        if (st != GS_OK) return st;
        SG_TRY(s, sg_launch_stats(out, E, (int)rows, kPad, (int)no, (size_t)rows * kPad, dm, ds, s->stream));
        // Flags on the device (from the raw inputs still in B_XF).
        const size_t bytes = (size_t)rows * no * sizeof(float);
        if (fallback) {
            const float *xraw;
            uint8_t *fl;
            SG_NEED(s, B_XF, (size_t)rows * ni, xraw);
            if (on_device) fl = fallback + r0;
            else SG_NEED(s, B_FLAGS, (size_t)rows, fl);
            SG_TRY(s, sg_launch_flags(ds, xraw, (int)rows, (int)no, (int)ni, s->cfg.std_outputs, s->cfg.std_limit,
                                      s->lo, s->hi, fl, s->stream));
            if (!on_device) SG_TRY(s, hipMemcpyAsync(fallback + r0, fl, rows, hipMemcpyDeviceToHost, s->stream));
        }
        if (!on_device) {
            if (mean) SG_TRY(s, hipMemcpyAsync(mean + (size_t)r0 * no, dm, bytes, hipMemcpyDeviceToHost, s->stream));
            if (sd) SG_TRY(s, hipMemcpyAsync(sd + (size_t)r0 * no, ds, bytes, hipMemcpyDeviceToHost, s->stream));
        }
        // B_XF and the scratch outputs are reused by the next chunk
        SG_TRY(s, hipStreamSynchronize(s->stream));
    }
    return GS_OK;
}

gs_status gs_surrogate_predict(gs_surrogate *s, const float *x, uint32_t n, float *mean, float *sd,
                               uint8_t *fallback) {
    return predict_rows(s, x, n, mean, sd, fallback, false);
}

gs_status gs_surrogate_predict_device(gs_surrogate *s, const float *x, uint32_t n, float *mean, float *sd,
                                      uint8_t *fallback) {
    return predict_rows(s, x, n, mean, sd, fallback, true);
}

}  // extern "C"
