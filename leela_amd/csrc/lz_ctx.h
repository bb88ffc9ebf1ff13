// lz_ctx.h — the context behind the C ABI's opaque lz_ctx (private).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <memory>
#include <mutex>
#include <string>

#include "lz_net.h"

namespace lz {
class LeafQueue;
}

struct lz_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string name;
    std::string err;
    std::mutex mu;  // guards the nets, workspace and err
    lz::Net net[2];
    lz::Workspace ws;
    lz::RaggedLayout ragged;  // layout of the last ragged batch
    float temperature = 0.75f;
    bool finalized = false;
    lz::LeafQueue *queue = nullptr;  // owned; see lz::destroy_queue

    int fail(int code, const char *fmt, ...) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
};

namespace lz {
// Forward of `net` on device buffers with ctx->mu held (lz_abi.cpp).
int forward_locked(lz_ctx *ctx, int net, int precision, int n,
                   const uint8_t *d_planes, float *d_raw, float *d_out,
                   hipStream_t stream, LaunchTimer *timer = nullptr);
// Stops and frees the leaf queue of ctx, if any (lz_queue.cpp).
void destroy_queue(lz_ctx *ctx);
// The running leaf queue's precision, -1 when none runs.
int queue_precision(const lz_ctx *ctx);
// Looks up, once per device, the address of the static scratch array of
// lz_forward_device_noalloc on `device` (current); lz_noalloc.cpp.
hipError_t noalloc_init(int device);
}  // namespace lz
