// lz_encoder_abi.cpp — C ABI of the device feature encoder: packs boards
// into records in pinned memory, sends them in one copy and launches the
// encoder kernel (encoder.hip) on the caller's stream.
//
// Two record slots alternate, so packing the next batch overlaps the copy
// and kernel of the previous one; a slot is reused once the event recorded
// after its kernel has completed.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "go_board.h"
#include "input_planes.h"
#include "leela_amd.h"
#include "lz_encoder.h"
#include "lz_position.h"

struct lz_encoder {
    int device = 0;
    int max_batch = 0;
    std::mutex mu;
    std::string err;
    struct Slot {
        lz::EncRecord *host = nullptr;  // pinned
        lz::EncRecord *dev = nullptr;
        hipEvent_t done = nullptr;
        bool used = false;
    } slot[2];
    int next = 0;

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


extern "C" {

int lz_encoder_create(lz_encoder **out, int device, int max_batch) {
    if (!out || max_batch < 1) return LZ_ERR_ARG;
    *out = nullptr;
    lz_encoder *e = new (std::nothrow) lz_encoder();
    if (!e) return LZ_ERR_NOMEM;
    e->device = device;
    e->max_batch = max_batch;
    if (hipSetDevice(device) != hipSuccess) {
        delete e;
        return LZ_ERR_HIP;
    }
    const size_t bytes = sizeof(lz::EncRecord) * (size_t)max_batch;
    for (auto &s : e->slot) {
        if (hipHostMalloc(reinterpret_cast<void **>(&s.host), bytes,
                          hipHostMallocDefault) != hipSuccess ||
            hipMalloc(reinterpret_cast<void **>(&s.dev), bytes) != hipSuccess ||
            hipEventCreateWithFlags(&s.done, hipEventDisableTiming) !=
                hipSuccess) {
            lz_encoder_destroy(e);
            return LZ_ERR_NOMEM;
        }
    }
    *out = e;
    return LZ_OK;
}

void lz_encoder_destroy(lz_encoder *e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    for (auto &s : e->slot) {
        if (s.done) {
            (void)hipEventSynchronize(s.done);
            (void)hipEventDestroy(s.done);
        }
        if (s.dev) (void)hipFree(s.dev);
        if (s.host) (void)hipHostFree(s.host);
    }
    delete e;
}

const char *lz_encoder_last_error(const lz_encoder *e) {
    return e ? e->err.c_str() : "null encoder";
}

}  // extern "C"

namespace {

int encode(lz_encoder *e, const lz_position *const *pos, int n, int symmetry,
           uint8_t *d_policy, uint8_t *d_value, void *stream) {
    if (n < 0 || n > e->max_batch)
        return e->fail(LZ_ERR_ARG, "batch %d outside 0..%d", n, e->max_batch);
    if (symmetry < 0 || symmetry > 7)
        return e->fail(LZ_ERR_ARG, "symmetry %d", symmetry);
    if (n == 0) return LZ_OK;
    if (!pos || (!d_policy && !d_value))
        return e->fail(LZ_ERR_ARG, "null buffer");
    for (int i = 0; i < n; i++)
        if (!pos[i]) return e->fail(LZ_ERR_ARG, "position %d is NULL", i);
    if (hipSetDevice(e->device) != hipSuccess)
        return e->fail(LZ_ERR_HIP, "hipSetDevice failed");
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: default stream
    lz_encoder::Slot &sl = e->slot[e->next];
    e->next ^= 1;
    if (sl.used) {
        const hipError_t err = hipEventSynchronize(sl.done);
        if (err != hipSuccess)
            return e->fail(LZ_ERR_HIP, "encoder slot wait: %s",
                           hipGetErrorString(err));
    }
    int64_t off = 0;
    for (int i = 0; i < n; i++) {
        lz::pack_record(pos[i]->st, off, sl.host + i);
        const int S = pos[i]->st.board.side();
        off += (int64_t)lz::kPlanes * S * S;
    }
    hipError_t err = hipMemcpyAsync(sl.dev, sl.host,
                                    sizeof(lz::EncRecord) * (size_t)n,
                                    hipMemcpyHostToDevice, s);
    if (err != hipSuccess)
        return e->fail(LZ_ERR_HIP, "record copy: %s", hipGetErrorString(err));
    err = lz::launch_encode(sl.dev, n, symmetry, d_policy, d_value, s);
    if (err != hipSuccess)
        return e->fail(LZ_ERR_HIP, "encoder launch: %s", hipGetErrorString(err));
    err = hipEventRecord(sl.done, s);
    if (err != hipSuccess)
        return e->fail(LZ_ERR_HIP, "encoder event: %s", hipGetErrorString(err));
    sl.used = true;
    return LZ_OK;
}

}  // namespace

extern "C" {

int lz_encode_device(lz_encoder *e, const lz_position *const *pos, int n,
                     int net, int symmetry, uint8_t *d_planes, void *stream) {
    if (!e) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(e->mu);
    if (net != LZ_NET_POLICY && net != LZ_NET_VALUE)
        return e->fail(LZ_ERR_ARG, "net %d", net);
    if (n > 0 && !d_planes) return e->fail(LZ_ERR_ARG, "null buffer");
    return encode(e, pos, n, symmetry, net == LZ_NET_POLICY ? d_planes : nullptr,
                  net == LZ_NET_VALUE ? d_planes : nullptr, stream);
}

int lz_encode_device_both(lz_encoder *e, const lz_position *const *pos, int n,
                          int symmetry, uint8_t *d_policy, uint8_t *d_value,
                          void *stream) {
    if (!e) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(e->mu);
    if (n > 0 && (!d_policy || !d_value))
        return e->fail(LZ_ERR_ARG, "null buffer");
    return encode(e, pos, n, symmetry, d_policy, d_value, stream);
}

}  // extern "C"
