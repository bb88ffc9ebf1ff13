// lz_abi.cpp — the C ABI (include/leela_amd.h): context, layer tables,
// weight packing and per-precision forward dispatch.
//
// Reference seam being replaced: OpenCL_Network / OpenCL (OpenCL.h:47-123,
// OpenCL.cpp:474-744). The reference keeps one in-order queue and one set
// of buffers per search thread and evaluates one position per forward; this
// context owns one stream and batch-sized workspaces, and a forward covers
// n positions in one pass over the layer table.
#include "leela_amd.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "lz_internal.h"
#include "lz_ctx.h"
#include "lz_net.h"

namespace {

std::mutex g_init_err_mu;
std::string g_init_err;

void set_init_err(const std::string &s) {
    std::lock_guard<std::mutex> g(g_init_err_mu);
    g_init_err = s;
}

}  // namespace

namespace lz {

int forward_locked(lz_ctx *ctx, int net, int precision, int n,
                   const uint8_t *d_planes, float *d_raw, float *d_out,
                   hipStream_t stream, LaunchTimer *timer) {
    if (!ctx->finalized)
        return ctx->fail(LZ_ERR_STATE, "forward before lz_finalize");
    if (net != LZ_NET_POLICY && net != LZ_NET_VALUE)
        return ctx->fail(LZ_ERR_ARG, "bad net id %d", net);
    if (precision < LZ_PREC_FP32 || precision > LZ_PREC_FP16)
        return ctx->fail(LZ_ERR_ARG, "bad precision %d", precision);
    if (n < 0) return ctx->fail(LZ_ERR_ARG, "negative batch");
    if (n == 0) return LZ_OK;
    if (!d_planes) return ctx->fail(LZ_ERR_ARG, "planes is NULL");
    std::string msg;
    int rc = ensure_workspace(ctx->ws, ctx->net, n, &msg);
    if (rc != LZ_OK) return ctx->fail(rc, "%s", msg.c_str());
    rc = run_net(ctx->net[net], net, precision, n, d_planes, d_raw, d_out,
                 ctx->temperature, ctx->ws, stream, &msg, timer);
    if (rc != LZ_OK) return ctx->fail(rc, "%s", msg.c_str());
    return LZ_OK;
}

}  // namespace lz

using lz::forward_locked;

extern "C" {

int lz_init(lz_ctx **out, int device) {
    if (!out) return LZ_ERR_ARG;
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        set_init_err(std::string("no HIP device: ") + hipGetErrorString(e));
        return LZ_ERR_HIP;
    }
    if (device < 0 || device >= count) {
        set_init_err("device ordinal out of range");
        return LZ_ERR_ARG;
    }
    if ((e = hipSetDevice(device)) != hipSuccess) {
        set_init_err(std::string("hipSetDevice: ") + hipGetErrorString(e));
        return LZ_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) {
        set_init_err(std::string("hipGetDeviceProperties: ") +
                     hipGetErrorString(e));
        return LZ_ERR_HIP;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_init_err(std::string("kernels are built for gfx950, device is ") +
                     prop.gcnArchName);
        return LZ_ERR_UNSUPPORTED;
    }
    // The no-allocation entry point's scratch array: its address on this
    // device, looked up once per device (the status is kept and reported
    // by lz_forward_device_noalloc; other entry points do not need it).
    (void)lz::noalloc_init(device);
    lz_ctx *ctx = new (std::nothrow) lz_ctx();
    if (!ctx) return LZ_ERR_NOMEM;
    ctx->device = device;
    ctx->name = std::string(prop.name) + " (" + prop.gcnArchName + ", " +
                std::to_string(prop.multiProcessorCount) + " CUs)";
    if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) !=
        hipSuccess) {
        set_init_err(std::string("hipStreamCreate: ") + hipGetErrorString(e));
        delete ctx;
        return LZ_ERR_HIP;
    }
    *out = ctx;
    return LZ_OK;
}

int lz_push_convolve(lz_ctx *ctx, int net, int filter, int cin, int cout,
                     const float *w, const float *b) {
    if (!ctx) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (ctx->finalized)
        return ctx->fail(LZ_ERR_STATE, "push after lz_finalize");
    if (net != LZ_NET_POLICY && net != LZ_NET_VALUE)
        return ctx->fail(LZ_ERR_ARG, "net must be LZ_NET_POLICY or LZ_NET_VALUE");
    if (!w || !b || cin <= 0 || cout <= 0 || (filter != 3 && filter != 5))
        return ctx->fail(LZ_ERR_ARG, "bad conv layer (filter %d, %d->%d)",
                         filter, cin, cout);
    lz::Layer L;
    L.kind = lz::Layer::kConv;
    L.filter = filter;
    L.cin = cin;
    L.cout = cout;
    L.w.assign(w, w + (size_t)cout * cin * filter * filter);
    L.b.assign(b, b + cout);
    ctx->net[net].layers.push_back(std::move(L));
    return LZ_OK;
}

static_assert(LZ_ACT_ELU == lz::kActElu && LZ_ACT_RELU == lz::kActRelu &&
                  LZ_ACT_NONE == lz::kActNone,
              "activation ids");

int lz_push_activation(lz_ctx *ctx, int net, int kind) {
    if (!ctx) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (ctx->finalized)
        return ctx->fail(LZ_ERR_STATE, "push after lz_finalize");
    if (net != LZ_NET_POLICY && net != LZ_NET_VALUE)
        return ctx->fail(LZ_ERR_ARG, "net must be LZ_NET_POLICY or LZ_NET_VALUE");
    if (kind != LZ_ACT_ELU && kind != LZ_ACT_RELU && kind != LZ_ACT_NONE)
        return ctx->fail(LZ_ERR_ARG, "bad activation %d", kind);
    auto &layers = ctx->net[net].layers;
    if (layers.empty() || layers.back().kind != lz::Layer::kConv)
        return ctx->fail(LZ_ERR_STATE, "activation must follow a conv");
    if (!layers.back().bn_mean.empty())
        return ctx->fail(LZ_ERR_STATE,
                         "activation after the conv's batch normalisation");
    layers.back().act = kind;
    return LZ_OK;
}

int lz_push_batchnorm(lz_ctx *ctx, int net, int channels, const float *means,
                      const float *variances, float scale) {
    if (!ctx) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (ctx->finalized)
        return ctx->fail(LZ_ERR_STATE, "push after lz_finalize");
    if (net != LZ_NET_POLICY && net != LZ_NET_VALUE)
        return ctx->fail(LZ_ERR_ARG, "net must be LZ_NET_POLICY or LZ_NET_VALUE");
    auto &layers = ctx->net[net].layers;
    if (layers.empty() || layers.back().kind != lz::Layer::kConv)
        return ctx->fail(LZ_ERR_STATE, "batch normalisation must follow a conv");
    lz::Layer &L = layers.back();
    if (!L.bn_mean.empty())
        return ctx->fail(LZ_ERR_STATE, "conv already normalised");
    if (!means || !variances || channels != L.cout || !(scale != 0.0f) ||
        !std::isfinite(scale))
        return ctx->fail(LZ_ERR_ARG, "bad batch normalisation (%d channels "
                         "for a %d-channel conv, scale %g)", channels, L.cout,
                         (double)scale);
    // Network.cpp:436-441 (OpenCL.cpp:397-399): mean and variance over the
    // stored scale factor, epsilon 1e-5, 1 / sqrt in FP32.
    const float epsilon = 1e-5f;
    std::vector<float> m(channels), inv(channels);
    for (int c = 0; c < channels; c++) {
        const float mean = means[c] / scale;
        float variance = variances[c] / scale;
        variance += epsilon;
        if (!(variance > 0.0f) || !std::isfinite(mean))
            return ctx->fail(LZ_ERR_ARG, "channel %d: mean %g, variance %g",
                             c, (double)mean, (double)variance);
        m[c] = mean;
        inv[c] = 1.0f / std::sqrt(variance);
    }
    L.bn_mean = std::move(m);
    L.bn_inv_std = std::move(inv);
    return LZ_OK;
}

int lz_push_innerproduct(lz_ctx *ctx, int net, int inputs, int outputs,
                         const float *w, const float *b) {
    if (!ctx) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (ctx->finalized)
        return ctx->fail(LZ_ERR_STATE, "push after lz_finalize");
    if (net != LZ_NET_POLICY && net != LZ_NET_VALUE)
        return ctx->fail(LZ_ERR_ARG, "net must be LZ_NET_POLICY or LZ_NET_VALUE");
    if (!w || !b || inputs <= 0 || outputs <= 0)
        return ctx->fail(LZ_ERR_ARG, "bad inner product %d->%d", inputs,
                         outputs);
    lz::Layer L;
    L.kind = lz::Layer::kFC;
    L.filter = 0;
    L.cin = inputs;
    L.cout = outputs;
    L.w.assign(w, w + (size_t)outputs * inputs);
    L.b.assign(b, b + outputs);
    ctx->net[net].layers.push_back(std::move(L));
    return LZ_OK;
}

int lz_finalize(lz_ctx *ctx) {
    if (!ctx) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (ctx->finalized) return ctx->fail(LZ_ERR_STATE, "already finalized");
    (void)hipSetDevice(ctx->device);
    for (int k = 0; k < 2; k++) {
        std::string msg;
        int rc = lz::validate_net(ctx->net[k], k, &msg);
        if (rc != LZ_OK) return ctx->fail(rc, "%s net: %s",
                                          k == 0 ? "policy" : "value",
                                          msg.c_str());
    }
    for (int k = 0; k < 2; k++) {
        std::string msg;
        int rc = lz::upload_net(ctx->net[k], k, &msg);
        if (rc != LZ_OK) return ctx->fail(rc, "%s", msg.c_str());
    }
    ctx->finalized = true;
    return LZ_OK;
}

int lz_set_layered_max(lz_ctx *ctx, int max_boards) {
    if (!ctx) return LZ_ERR_ARG;
    if (max_boards < 0 || max_boards > lz::kLayeredMaxBoards)
        return ctx->fail(LZ_ERR_ARG, "layered batch limit %d outside 0..%d",
                         max_boards, lz::kLayeredMaxBoards);
    std::lock_guard<std::mutex> g(ctx->mu);
    for (auto &n : ctx->net) n.layered_max = max_boards;
    return LZ_OK;
}

int lz_set_softmax_temp(lz_ctx *ctx, float temperature) {
    if (!ctx) return LZ_ERR_ARG;
    if (!(temperature > 0.0f) || !std::isfinite(temperature))
        return ctx->fail(LZ_ERR_ARG, "temperature must be positive");
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->temperature = temperature;
    return LZ_OK;
}

int lz_forward_ragged_device(lz_ctx *ctx, int net, int precision, int n,
                             const int32_t *sides, const uint8_t *d_planes,
                             float *d_raw, float *d_out, void *stream) {
    if (!ctx) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (!ctx->finalized)
        return ctx->fail(LZ_ERR_STATE, "forward before lz_finalize");
    if (n < 0 || (n > 0 && (!sides || !d_planes)))
        return ctx->fail(LZ_ERR_ARG, "bad ragged batch");
    if (n == 0) return LZ_OK;
    if (net != LZ_NET_POLICY && net != LZ_NET_VALUE)
        return ctx->fail(LZ_ERR_ARG, "bad net id %d", net);
    (void)hipSetDevice(ctx->device);
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: default stream
    std::string msg;
    int rc = lz::build_ragged_layout(ctx->ragged, n, sides, s, &msg);
    if (rc == LZ_OK) rc = lz::ensure_workspace(ctx->ws, ctx->net, n, &msg);
    if (rc == LZ_OK)
        rc = lz::run_net(ctx->net[net], net, precision, n, d_planes, d_raw,
                         d_out, ctx->temperature, ctx->ws, s, &msg, nullptr,
                         &ctx->ragged);
    if (rc != LZ_OK) return ctx->fail(rc, "%s", msg.c_str());
    return LZ_OK;
}

int lz_forward_ragged(lz_ctx *ctx, int net, int precision, int n,
                      const int32_t *sides, const uint8_t *planes, float *raw,
                      float *out) {
    if (!ctx) return LZ_ERR_ARG;
    if (n < 0 || (n > 0 && (!sides || !planes))) return LZ_ERR_ARG;
    if (n == 0) return LZ_OK;
    size_t pbytes = 0, pfloats = 0;
    for (int i = 0; i < n; i++) {
        if (sides[i] < 5 || sides[i] > 19) return LZ_ERR_ARG;
        pbytes += (size_t)32 * sides[i] * sides[i];
        pfloats += (size_t)sides[i] * sides[i];
    }
    const size_t outs = net == LZ_NET_POLICY ? pfloats : (size_t)n;
    uint8_t *d_planes = nullptr;
    float *d_raw = nullptr, *d_out = nullptr;
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&d_planes), pbytes);
    if (e == hipSuccess)
        e = hipMalloc(reinterpret_cast<void **>(&d_raw), sizeof(float) * outs);
    if (e == hipSuccess)
        e = hipMalloc(reinterpret_cast<void **>(&d_out), sizeof(float) * outs);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_planes, planes, pbytes, hipMemcpyHostToDevice,
                           ctx->stream);
    int rc = e == hipSuccess ? LZ_OK : LZ_ERR_HIP;
    if (rc == LZ_OK)
        rc = lz_forward_ragged_device(ctx, net, precision, n, sides, d_planes,
                                      d_raw, d_out, ctx->stream);
    if (rc == LZ_OK && raw)
        e = hipMemcpyAsync(raw, d_raw, sizeof(float) * outs,
                           hipMemcpyDeviceToHost, ctx->stream);
    if (rc == LZ_OK && e == hipSuccess && out)
        e = hipMemcpyAsync(out, d_out, sizeof(float) * outs,
                           hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (rc == LZ_OK && e != hipSuccess) {
        std::lock_guard<std::mutex> g(ctx->mu);
        rc = ctx->fail(LZ_ERR_HIP, "ragged forward: %s", hipGetErrorString(e));
    }
    (void)hipFree(d_planes);
    (void)hipFree(d_raw);
    (void)hipFree(d_out);
    return rc;
}

int lz_forward_timed(lz_ctx *ctx, int net, int precision, int n,
                     const uint8_t *d_planes, float *d_raw, float *d_out,
                     void *stream, float *launch_ms, int max_launches,
                     int *n_launches) {
    if (!ctx) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: default stream
    lz::LaunchTimer timer;
    int rc = forward_locked(ctx, net, precision, n, d_planes, d_raw, d_out, s,
                            &timer);
    if (rc != LZ_OK) return rc;
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return ctx->fail(LZ_ERR_HIP, "timed forward: %s", hipGetErrorString(e));
    const int count = timer.events.empty() ? 0 : (int)timer.events.size() - 1;
    if (n_launches) *n_launches = count;
    for (int i = 0; i < count && i < max_launches && launch_ms; i++) {
        float ms = 0.0f;
        e = hipEventElapsedTime(&ms, timer.events[i], timer.events[i + 1]);
        if (e != hipSuccess)
            return ctx->fail(LZ_ERR_HIP, "event time: %s", hipGetErrorString(e));
        launch_ms[i] = ms;
    }
    return LZ_OK;
}

int lz_forward_device(lz_ctx *ctx, int net, int precision, int n,
                      const uint8_t *d_planes, float *d_raw, float *d_out,
                      void *stream) {
    if (!ctx) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: default stream
    return forward_locked(ctx, net, precision, n, d_planes, d_raw, d_out, s);
}

}  // extern "C"

// A caller-owned workspace (lz_workspace_create): its own activation
// buffers, value-head planes and layered tiles, so that forwards on
// different streams with different workspaces can overlap.
struct lz_workspace {
    lz_ctx *ctx = nullptr;
    std::mutex mu;  // guards ws (growth) against concurrent calls
    lz::Workspace ws;
};

extern "C" {

int lz_workspace_create(lz_ctx *ctx, int max_boards, lz_workspace **out) {
    if (!ctx || !out || max_boards < 1) return LZ_ERR_ARG;
    *out = nullptr;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (!ctx->finalized)
        return ctx->fail(LZ_ERR_STATE, "workspace before lz_finalize");
    (void)hipSetDevice(ctx->device);
    lz_workspace *w = new (std::nothrow) lz_workspace();
    if (!w) return LZ_ERR_NOMEM;
    w->ctx = ctx;
    std::string msg;
    const int rc = lz::ensure_workspace(w->ws, ctx->net, max_boards, &msg);
    if (rc != LZ_OK) {
        lz::free_workspace(w->ws);
        delete w;
        return ctx->fail(rc, "%s", msg.c_str());
    }
    *out = w;
    return LZ_OK;
}

void lz_workspace_destroy(lz_workspace *w) {
    if (!w) return;
    (void)hipSetDevice(w->ctx->device);
    (void)hipDeviceSynchronize();  // no launch may still use the buffers
    lz::free_workspace(w->ws);
    delete w;
}

int lz_forward_device_ws(lz_ctx *ctx, lz_workspace *w, int net,
                         int precision, int n, const uint8_t *d_planes,
                         float *d_raw, float *d_out, void *stream) {
    if (!ctx || !w || w->ctx != ctx) return LZ_ERR_ARG;
    float temperature;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        if (!ctx->finalized)
            return ctx->fail(LZ_ERR_STATE, "forward before lz_finalize");
        temperature = ctx->temperature;
    }
    if (net != LZ_NET_POLICY && net != LZ_NET_VALUE) return LZ_ERR_ARG;
    if (precision < LZ_PREC_FP32 || precision > LZ_PREC_FP16) return LZ_ERR_ARG;
    if (n < 0 || (n > 0 && !d_planes)) return LZ_ERR_ARG;
    if (n == 0) return LZ_OK;
    (void)hipSetDevice(ctx->device);
    std::lock_guard<std::mutex> g(w->mu);
    std::string msg;
    int rc = lz::ensure_workspace(w->ws, ctx->net, n, &msg);
    if (rc == LZ_OK)
        rc = lz::run_net(ctx->net[net], net, precision, n, d_planes, d_raw,
                         d_out, temperature, w->ws,
                         static_cast<hipStream_t>(stream), &msg);
    return rc;
}

int lz_forward(lz_ctx *ctx, int net, int precision, int n,
               const uint8_t *planes, float *raw, float *out) {
    if (!ctx) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (n < 0) return ctx->fail(LZ_ERR_ARG, "negative batch");
    if (n == 0) return LZ_OK;
    if (!planes) return ctx->fail(LZ_ERR_ARG, "planes is NULL");
    if (!ctx->finalized)
        return ctx->fail(LZ_ERR_STATE, "forward before lz_finalize");
    (void)hipSetDevice(ctx->device);
    std::string msg;
    int rc = lz::ensure_workspace(ctx->ws, ctx->net, n, &msg);
    if (rc != LZ_OK) return ctx->fail(rc, "%s", msg.c_str());
    lz::Workspace &ws = ctx->ws;
    const lz::Net &N = ctx->net[net < 0 || net > 1 ? 0 : net];
    const int cin0 = N.layers.empty() ? 0 : N.layers[0].cin;
    const size_t in_bytes = (size_t)n * cin0 * N.side * N.side;
    hipError_t e = hipMemcpyAsync(ws.planes, planes, in_bytes,
                                  hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess)
        return ctx->fail(LZ_ERR_HIP, "H2D planes: %s", hipGetErrorString(e));
    rc = forward_locked(ctx, net, precision, n, ws.planes, ws.raw, ws.out,
                        ctx->stream);
    if (rc != LZ_OK) return rc;
    const size_t per = net == LZ_NET_POLICY ? (size_t)N.side * N.side : 1;
    if (raw &&
        (e = hipMemcpyAsync(raw, ws.raw, sizeof(float) * per * n,
                            hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess)
        return ctx->fail(LZ_ERR_HIP, "D2H raw: %s", hipGetErrorString(e));
    if (out &&
        (e = hipMemcpyAsync(out, ws.out, sizeof(float) * per * n,
                            hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess)
        return ctx->fail(LZ_ERR_HIP, "D2H out: %s", hipGetErrorString(e));
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
        return ctx->fail(LZ_ERR_HIP, "forward: %s", hipGetErrorString(e));
    return LZ_OK;
}

int lz_normal_fill(uint32_t seed, size_t n, float stddev, float *out) {
    if (!out && n) return LZ_ERR_ARG;
    lz::normal_fill(seed, n, stddev, out);
    return LZ_OK;
}

const char *lz_device_name(lz_ctx *ctx) {
    return ctx ? ctx->name.c_str() : "";
}

const char *lz_last_error(lz_ctx *ctx) {
    if (ctx) return ctx->err.c_str();
    std::lock_guard<std::mutex> g(g_init_err_mu);
    return g_init_err.c_str();
}

void lz_destroy(lz_ctx *ctx) {
    if (!ctx) return;
    lz::destroy_queue(ctx);
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    lz::free_net(ctx->net[0]);
    lz::free_net(ctx->net[1]);
    lz::free_workspace(ctx->ws);
    lz::free_ragged_layout(ctx->ragged);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

}  // extern "C"
