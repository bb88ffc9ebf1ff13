// lz_noalloc.cpp — lz_forward_device_noalloc: the fused 16-bit and FP8
// forwards of up to LZ_NOALLOC_MAX_BOARDS boards with no allocation.
//
// The only device memory these forms use beyond the caller's inputs and
// outputs is the value net's head plane h0 [n][361]: the fused kernel writes
// each board's ELU'd 1-channel conv output there and value_fc_head reads it.
// Here h0 lives in a file-scope device array instead of the context's
// workspace, so the call makes no hipMalloc / hipFree (nor any other
// allocating call) and can be made where allocation is not allowed (under
// stream capture, in serving loops). Every board's row is written before it
// is read, so the array needs no clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "leela_amd.h"
#include "lz_ctx.h"
#include "lz_net.h"

namespace lz {
namespace {

constexpr int kMaxDevices = 64;
constexpr int kPlane = 361;

// Value-head planes of one call: [LZ_NOALLOC_MAX_BOARDS][361] floats,
// 2.96 MB. Each device holds its own copy until the process ends. Host code
// never passes this name to a kernel: the address on a device comes from
// hipGetSymbolAddress (scratch_ptr below).
__device__ __align__(256) float scratch[LZ_NOALLOC_MAX_BOARDS * kPlane];

std::mutex g_mu;
float *g_ptr[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};
bool g_looked_up[kMaxDevices] = {};

}  // namespace

// Looks up the device address of `scratch` on `device`, which must be the
// current device; once per device (lz_init calls it). Returns the lookup's
// status, which later calls repeat.
hipError_t noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_looked_up[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_err[device] = e;
        g_ptr[device] = e == hipSuccess ? static_cast<float *>(p) : nullptr;
        g_looked_up[device] = true;
    }
    return g_err[device];
}

// The device address of `scratch` on `device`, or nullptr (status in *err)
// when the lookup failed or has not been made.
static float *scratch_ptr(int device, hipError_t *err) {
    if (device < 0 || device >= kMaxDevices) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> g(g_mu);
    *err = g_looked_up[device] ? g_err[device] : hipErrorNotInitialized;
    return g_ptr[device];
}

}  // namespace lz

extern "C" int lz_forward_device_noalloc(lz_ctx *ctx, int net, int precision,
                                         int n, const uint8_t *d_planes,
                                         float *d_raw, float *d_out,
                                         void *stream) {
    if (!ctx) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (!ctx->finalized)
        return ctx->fail(LZ_ERR_STATE, "forward before lz_finalize");
    if (net != LZ_NET_POLICY && net != LZ_NET_VALUE)
        return ctx->fail(LZ_ERR_ARG, "bad net id %d", net);
    if (precision == LZ_PREC_FP32)
        return ctx->fail(LZ_ERR_UNSUPPORTED,
                         "the FP32 form keeps 128 x 361 floats of activations "
                         "per board in a workspace; use lz_forward_device");
    if (precision != LZ_PREC_BF16 && precision != LZ_PREC_FP16 &&
        precision != LZ_PREC_FP8)
        return ctx->fail(LZ_ERR_ARG, "bad precision %d", precision);
    if (n < 0 || n > LZ_NOALLOC_MAX_BOARDS)
        return ctx->fail(LZ_ERR_ARG,
                         "batch of %d boards outside 0..%d (the size of the "
                         "static scratch array)",
                         n, LZ_NOALLOC_MAX_BOARDS);
    if (n == 0) return LZ_OK;
    if (!d_planes) return ctx->fail(LZ_ERR_ARG, "planes is NULL");
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipSuccess;
    float *h0 = lz::scratch_ptr(ctx->device, &e);
    if (!h0)
        return ctx->fail(LZ_ERR_HIP, "scratch array address: %s",
                         hipGetErrorString(e));
    // Only h0 is read by the fused forms; nothing here is allocated or
    // freed.
    lz::Workspace ws;
    ws.h0 = h0;
    std::string msg;
    const int rc = lz::run_net(ctx->net[net], net, precision, n, d_planes,
                               d_raw, d_out, ctx->temperature, ws,
                               static_cast<hipStream_t>(stream), &msg);
    if (rc != LZ_OK) return ctx->fail(rc, "%s", msg.c_str());
    return LZ_OK;
}
