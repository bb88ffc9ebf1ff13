// lz_encoder_abi.cpp — C ABI of the device feature encoder: packs boards
// into records in pinned memory, sends them in one copy and launches the
// encoder kernel (encoder.hip) on the caller's stream.
//
// Two record slots alternate, so packing the next batch overlaps the copy
// and kernel of the previous one; a slot is reused once the event recorded
// after its kernel has completed. Record copies go on the encoder's own
// stream and the caller's stream waits for them, so a copy overlaps the
// work still queued ahead of the kernel.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "go_board.h"
#include "input_planes.h"
#include "leela_amd.h"
#include "lz_encoder.h"
#include "lz_position.h"
#include "mc_policy.h"

struct lz_encoder {
    int device = 0;
    int max_batch = 0;
    std::mutex mu;
    std::string err;
    struct Slot {
        lz::EncRecord *host = nullptr;  // pinned
        lz::EncRecord *dev = nullptr;
        hipEvent_t done = nullptr;    // after the kernel that reads it
        hipEvent_t copied = nullptr;  // after its copy to the device
        // Overflow strings of the boards encoded from this slot (per slot:
        // launches from both slots may run at once on different streams).
        uint64_t *ovf = nullptr;
        bool used = false;
    } slot[2];
    // The record copies run on their own stream, so they overlap whatever
    // the caller's stream is still computing; the kernel waits for them.
    hipStream_t copy_stream = nullptr;
    int next = 0;
    uint32_t *half_wins = nullptr;  // playout outcome counters, max_batch
    // Pattern playouts (allocated on first use): the policy's tables on the
    // device, one playout state per thread of the persistent grid, the
    // per-position counters of abandoned playouts, and the event after the
    // last launch that used them (the next launch waits for it, whatever
    // its stream).
    int32_t *pat_g = nullptr;
    uint16_t *pat_index[2] = {nullptr, nullptr};
    float *pat_fw = nullptr;
    float *pat_pw = nullptr;
    void *pat_states = nullptr;
    int pat_nstates = 0;
    uint32_t *pat_overflow = nullptr;
    hipEvent_t pat_done = nullptr;
    bool pat_used = false;

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
            hipMalloc(reinterpret_cast<void **>(&s.ovf),
                      sizeof(uint64_t) * lz::kEncStrOvf * lz::kEncStrWords *
                          (size_t)max_batch) != hipSuccess ||
            hipEventCreateWithFlags(&s.done, hipEventDisableTiming) !=
                hipSuccess ||
            hipEventCreateWithFlags(&s.copied, hipEventDisableTiming) !=
                hipSuccess) {
            lz_encoder_destroy(e);
            return LZ_ERR_NOMEM;
        }
    }
    if (hipStreamCreateWithFlags(&e->copy_stream, hipStreamNonBlocking) !=
        hipSuccess) {
        lz_encoder_destroy(e);
        return LZ_ERR_HIP;
    }
    if (hipMalloc(reinterpret_cast<void **>(&e->half_wins),
                  sizeof(uint32_t) * (size_t)max_batch) != hipSuccess) {
        lz_encoder_destroy(e);
        return LZ_ERR_NOMEM;
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
        if (s.copied) (void)hipEventDestroy(s.copied);
        if (s.dev) (void)hipFree(s.dev);
        if (s.ovf) (void)hipFree(s.ovf);
        if (s.host) (void)hipHostFree(s.host);
    }
    if (e->copy_stream) (void)hipStreamDestroy(e->copy_stream);
    if (e->pat_done) {
        (void)hipEventSynchronize(e->pat_done);
        (void)hipEventDestroy(e->pat_done);
    }
    for (void *p : {(void *)e->half_wins, (void *)e->pat_g,
                    (void *)e->pat_index[0], (void *)e->pat_index[1],
                    (void *)e->pat_fw, (void *)e->pat_pw, e->pat_states,
                    (void *)e->pat_overflow})
        if (p) (void)hipFree(p);
    delete e;
}

const char *lz_encoder_last_error(const lz_encoder *e) {
    return e ? e->err.c_str() : "null encoder";
}

}  // extern "C"

namespace {

// Packs n boards into the next record slot and sends them on s. n > 0,
// pos checked by the caller.
int send_records(lz_encoder *e, const lz_position *const *pos, int n,
                 hipStream_t s, lz_encoder::Slot **out) {
    if (hipSetDevice(e->device) != hipSuccess)
        return e->fail(LZ_ERR_HIP, "hipSetDevice failed");
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
                                    hipMemcpyHostToDevice, e->copy_stream);
    if (err == hipSuccess) err = hipEventRecord(sl.copied, e->copy_stream);
    if (err == hipSuccess) err = hipStreamWaitEvent(s, sl.copied, 0);
    if (err != hipSuccess)
        return e->fail(LZ_ERR_HIP, "record copy: %s", hipGetErrorString(err));
    *out = &sl;
    return LZ_OK;
}

// The slot is free again once the work enqueued after its copy is done.
int release_slot(lz_encoder *e, lz_encoder::Slot *sl, hipStream_t s) {
    const hipError_t err = hipEventRecord(sl->done, s);
    if (err != hipSuccess)
        return e->fail(LZ_ERR_HIP, "encoder event: %s", hipGetErrorString(err));
    sl->used = true;
    return LZ_OK;
}

int check_positions(lz_encoder *e, const lz_position *const *pos, int n) {
    if (!pos) return e->fail(LZ_ERR_ARG, "null positions");
    for (int i = 0; i < n; i++)
        if (!pos[i]) return e->fail(LZ_ERR_ARG, "position %d is NULL", i);
    return LZ_OK;
}

int encode(lz_encoder *e, const lz_position *const *pos, int n, int symmetry,
           uint8_t *d_policy, uint8_t *d_value, void *stream) {
    if (n < 0 || n > e->max_batch)
        return e->fail(LZ_ERR_ARG, "batch %d outside 0..%d", n, e->max_batch);
    if (symmetry < 0 || symmetry > 7)
        return e->fail(LZ_ERR_ARG, "symmetry %d", symmetry);
    if (n == 0) return LZ_OK;
    if (!d_policy && !d_value) return e->fail(LZ_ERR_ARG, "null buffer");
    int rc = check_positions(e, pos, n);
    if (rc != LZ_OK) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: default stream
    lz_encoder::Slot *sl = nullptr;
    rc = send_records(e, pos, n, s, &sl);
    if (rc != LZ_OK) return rc;
    const hipError_t err =
        lz::launch_encode(sl->dev, n, symmetry, d_policy, d_value, sl->ovf, s);
    if (err != hipSuccess)
        return e->fail(LZ_ERR_HIP, "encoder launch: %s", hipGetErrorString(err));
    return release_slot(e, sl, s);
}

int playouts(lz_encoder *e, const lz_position *const *pos, int n, int nplay,
             uint32_t seed0, float *d_winrate, void *stream) {
    if (n < 0 || n > e->max_batch)
        return e->fail(LZ_ERR_ARG, "batch %d outside 0..%d", n, e->max_batch);
    if (nplay < 0) return e->fail(LZ_ERR_ARG, "playouts %d", nplay);
    if (n == 0) return LZ_OK;
    if (!d_winrate) return e->fail(LZ_ERR_ARG, "null buffer");
    int rc = check_positions(e, pos, n);
    if (rc != LZ_OK) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: default stream
    lz_encoder::Slot *sl = nullptr;
    rc = send_records(e, pos, n, s, &sl);
    if (rc != LZ_OK) return rc;
    const hipError_t err = lz::launch_playouts(sl->dev, n, nplay, seed0,
                                               e->half_wins, d_winrate, s);
    if (err != hipSuccess)
        return e->fail(LZ_ERR_HIP, "playout launch: %s", hipGetErrorString(err));
    return release_slot(e, sl, s);
}

// Threads of the pattern playouts' persistent grid: 3,072 wavefronts (12
// per CU of the part, the kernel's occupancy), each thread with its own
// playout state (~23 KB: 4.5 GB in all). The playouts are latency-bound
// pointer chasing, so resident wavefronts are what sets the rate: a third
// of this grid ran 19x19 playouts 1.4x slower.
constexpr int kPatternThreads = 196608;

template <typename T>
hipError_t upload(T **dst, const std::vector<T> &v) {
    hipError_t err = hipMalloc(reinterpret_cast<void **>(dst),
                               sizeof(T) * v.size());
    if (err == hipSuccess)
        err = hipMemcpy(*dst, v.data(), sizeof(T) * v.size(),
                        hipMemcpyHostToDevice);
    return err;
}

// First use: the tables on the device, the playout states, the counters.
int pattern_setup(lz_encoder *e) {
    if (e->pat_states) return LZ_OK;
    std::string msg;
    const lz::mc::Tables *t = lz::mc::tables(nullptr, &msg);
    if (!t) return e->fail(LZ_ERR_STATE, "pattern tables: %s", msg.c_str());
    hipError_t err = upload(&e->pat_g, t->g);
    if (err == hipSuccess) err = upload(&e->pat_index[0], t->index[0]);
    if (err == hipSuccess) err = upload(&e->pat_index[1], t->index[1]);
    if (err == hipSuccess) err = upload(&e->pat_fw, t->feature_weight);
    if (err == hipSuccess) err = upload(&e->pat_pw, t->pattern_weight);
    if (err == hipSuccess)
        err = hipMalloc(reinterpret_cast<void **>(&e->pat_overflow),
                        sizeof(uint32_t) * (size_t)e->max_batch);
    if (err == hipSuccess)
        err = hipEventCreateWithFlags(&e->pat_done, hipEventDisableTiming);
    if (err == hipSuccess)
        err = hipMalloc(&e->pat_states,
                        lz::pattern_state_bytes() * (size_t)kPatternThreads);
    if (err != hipSuccess) {
        if (e->pat_states) (void)hipFree(e->pat_states);
        e->pat_states = nullptr;
        return e->fail(LZ_ERR_NOMEM, "pattern playout setup: %s",
                       hipGetErrorString(err));
    }
    e->pat_nstates = kPatternThreads;
    return LZ_OK;
}

int pattern_playouts(lz_encoder *e, const lz_position *const *pos, int n,
                     int nplay, uint32_t seed0, float *d_winrate,
                     void *stream) {
    if (n < 0 || n > e->max_batch)
        return e->fail(LZ_ERR_ARG, "batch %d outside 0..%d", n, e->max_batch);
    if (nplay < 0) return e->fail(LZ_ERR_ARG, "playouts %d", nplay);
    if (n == 0) return LZ_OK;
    if (!d_winrate) return e->fail(LZ_ERR_ARG, "null buffer");
    int rc = check_positions(e, pos, n);
    if (rc != LZ_OK) return rc;
    if (hipSetDevice(e->device) != hipSuccess)
        return e->fail(LZ_ERR_HIP, "hipSetDevice failed");
    rc = pattern_setup(e);
    if (rc != LZ_OK) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: default stream
    if (e->pat_used && hipStreamWaitEvent(s, e->pat_done, 0) != hipSuccess)
        return e->fail(LZ_ERR_HIP, "pattern playout wait");
    lz_encoder::Slot *sl = nullptr;
    rc = send_records(e, pos, n, s, &sl);
    if (rc != LZ_OK) return rc;
    lz::mc::TablesView tv;
    tv.g = e->pat_g;
    tv.index[0] = e->pat_index[0];
    tv.index[1] = e->pat_index[1];
    tv.feature_weight = e->pat_fw;
    tv.pattern_weight = e->pat_pw;
    hipError_t err = lz::launch_pattern_playouts(
        sl->dev, n, nplay, seed0, tv, e->pat_states, e->pat_nstates,
        e->half_wins, e->pat_overflow, d_winrate, s);
    if (err == hipSuccess) err = hipEventRecord(e->pat_done, s);
    if (err != hipSuccess)
        return e->fail(LZ_ERR_HIP, "pattern playout launch: %s",
                       hipGetErrorString(err));
    e->pat_used = true;
    return release_slot(e, sl, s);
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

int lz_playout_winrates_device(lz_encoder *e, const lz_position *const *pos,
                               int n, int playouts_per, uint32_t seed0,
                               float *d_winrates, void *stream) {
    if (!e) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(e->mu);
    return playouts(e, pos, n, playouts_per, seed0, d_winrates, stream);
}

int lz_playout_winrates_device_policy(lz_encoder *e,
                                      const lz_position *const *pos, int n,
                                      int playouts_per, uint32_t seed0,
                                      int policy, float *d_winrates,
                                      void *stream) {
    if (!e) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(e->mu);
    if (policy == LZ_PLAYOUT_LIGHT)
        return playouts(e, pos, n, playouts_per, seed0, d_winrates, stream);
    if (policy != LZ_PLAYOUT_PATTERN)
        return e->fail(LZ_ERR_ARG, "playout policy %d", policy);
    return pattern_playouts(e, pos, n, playouts_per, seed0, d_winrates,
                            stream);
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
