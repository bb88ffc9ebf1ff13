// lz_tune.cpp — autotuner of the fused kernels' launch shape.
//
// For a batch size it times every (wave grid, persistent-grid multiplier)
// candidate of each net on the device with HIP events and keeps the
// fastest for that batch bucket. The reference's counterpart is the OpenCL
// row-tile knob (cfg_rowtiles, OpenCL.cpp:605-619) and the USE_TUNER
// search-parameter hooks; here the shape is measured, not configured.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "leela_amd.h"
#include "lz_ctx.h"

namespace {

constexpr int kPos = 32 * 361;
const char *kNetName[2] = {"policy", "value"};

// Milliseconds of one forward (median of 5 after 2 warm-ups); a ragged
// batch when `ragged` is given.
int time_forward(lz_ctx *ctx, int net, int precision, int n,
                 const uint8_t *d_planes, float *d_raw, float *d_out,
                 float *ms_out, std::string *msg,
                 const lz::RaggedLayout *ragged) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        *msg = "event create failed";
        return LZ_ERR_HIP;
    }
    std::vector<float> t;
    int rc = LZ_OK;
    for (int rep = 0; rep < 7 && rc == LZ_OK; rep++) {
        hipError_t e = hipEventRecord(a, ctx->stream);
        if (e == hipSuccess)
            rc = lz::run_net(ctx->net[net], net, precision, n, d_planes, d_raw,
                             d_out, ctx->temperature, ctx->ws, ctx->stream,
                             msg, nullptr, ragged);
        if (rc == LZ_OK && e == hipSuccess) e = hipEventRecord(b, ctx->stream);
        if (rc == LZ_OK && e == hipSuccess) e = hipEventSynchronize(b);
        float ms = 0.0f;
        if (rc == LZ_OK && e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        if (rc == LZ_OK && e != hipSuccess) {
            *msg = std::string("tuning run: ") + hipGetErrorString(e);
            rc = LZ_ERR_HIP;
        }
        if (rep >= 2) t.push_back(ms);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc == LZ_OK) {
        std::sort(t.begin(), t.end());
        *ms_out = t[t.size() / 2];
    }
    return rc;
}

// Times the candidates for n boards (all 19x19, or of sides[0..n)) and
// keeps the fastest per net for the batch's bucket (its slot count for a
// ragged batch). ctx->mu held.
int autotune_locked(lz_ctx *ctx, int precision, int n, const int32_t *sides,
                    char *report, size_t report_len) {
    if (!ctx->finalized)
        return ctx->fail(LZ_ERR_STATE, "autotune before lz_finalize");
    if (!ctx->net[0].fused_ok || !ctx->net[1].fused_ok)
        return ctx->fail(LZ_ERR_UNSUPPORTED, "fused kernels not available");
    (void)hipSetDevice(ctx->device);
    std::string msg;
    int rc = lz::ensure_workspace(ctx->ws, ctx->net, n, &msg);
    if (rc != LZ_OK) return ctx->fail(rc, "%s", msg.c_str());
    size_t in_bytes = (size_t)n * kPos, out_floats = (size_t)n * 361;
    const lz::RaggedLayout *ragged = nullptr;
    int slots = n;
    if (sides) {
        rc = lz::build_ragged_layout(ctx->ragged, n, sides, ctx->stream, &msg);
        if (rc != LZ_OK) return ctx->fail(rc, "%s", msg.c_str());
        ragged = &ctx->ragged;
        slots = ctx->ragged.nslots;
        in_bytes = ctx->ragged.plane_bytes;
        out_floats = ctx->ragged.policy_floats;
    }
    // Inputs: a fixed pseudo-random 0/1 pattern (the kernels' time does not
    // depend on the values).
    std::vector<uint8_t> h(in_bytes);
    uint32_t x = 2463534242u;
    for (auto &v : h) {
        x ^= x << 13;
        x ^= x >> 17;
        x ^= x << 5;
        v = (x >> 7) & 1;
    }
    uint8_t *d_planes = nullptr;
    float *d_raw = nullptr, *d_out = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&d_planes), h.size());
    if (e == hipSuccess)
        e = hipMalloc(reinterpret_cast<void **>(&d_raw),
                      sizeof(float) * out_floats);
    if (e == hipSuccess)
        e = hipMalloc(reinterpret_cast<void **>(&d_out),
                      sizeof(float) * out_floats);
    if (e == hipSuccess)
        e = hipMemcpy(d_planes, h.data(), h.size(), hipMemcpyHostToDevice);
    std::string rep;
    const int bucket = lz::tune_bucket(slots);
    for (int net = 0; net < 2 && e == hipSuccess && rc == LZ_OK; net++) {
        lz::Net &N = ctx->net[net];
        lz::Net::Tune best;
        best.ms = -1.0f;
        for (int wm = 1; wm <= 2 && rc == LZ_OK; wm++)
            for (int gm = 0; gm <= 2 && rc == LZ_OK; gm++) {
                N.tune[bucket] = lz::Net::Tune{wm, gm, 0.0f};
                float ms = 0.0f;
                rc = time_forward(ctx, net, precision, n, d_planes, d_raw,
                                  d_out, &ms, &msg, ragged);
                if (rc != LZ_OK) break;
                char line[128];
                snprintf(line, sizeof line,
                         "%s n=%d slots=%d wm=%d grid_mult=%d: %.4f ms\n",
                         kNetName[net], n, slots, wm, gm, ms);
                rep += line;
                if (best.ms < 0.0f || ms < best.ms)
                    best = lz::Net::Tune{wm, gm, ms};
            }
        N.tune[bucket] = best;
        char line[128];
        snprintf(line, sizeof line,
                 "%s n=%d slots=%d chosen wm=%d grid_mult=%d %.4f ms\n",
                 kNetName[net], n, slots, best.wm, best.grid_mult, best.ms);
        rep += line;
    }
    (void)hipFree(d_planes);
    (void)hipFree(d_raw);
    (void)hipFree(d_out);
    if (e != hipSuccess)
        return ctx->fail(LZ_ERR_HIP, "autotune: %s", hipGetErrorString(e));
    if (rc != LZ_OK) return ctx->fail(rc, "%s", msg.c_str());
    if (report && report_len) {
        std::strncpy(report, rep.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    return LZ_OK;
}

}  // namespace

extern "C" {

int lz_autotune(lz_ctx *ctx, int precision, int n, char *report,
                size_t report_len) {
    if (!ctx || n < 1 || n > 65536) return LZ_ERR_ARG;
    if (precision != LZ_PREC_BF16 && precision != LZ_PREC_FP16)
        return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    return autotune_locked(ctx, precision, n, nullptr, report, report_len);
}

int lz_autotune_ragged(lz_ctx *ctx, int precision, int n,
                       const int32_t *sides, char *report,
                       size_t report_len) {
    if (!ctx || n < 1 || n > 65536 || !sides) return LZ_ERR_ARG;
    if (precision != LZ_PREC_BF16 && precision != LZ_PREC_FP16)
        return LZ_ERR_ARG;
    for (int i = 0; i < n; i++)
        if (lz::ragged_slot_capacity(sides[i]) == 0) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    return autotune_locked(ctx, precision, n, sides, report, report_len);
}

int lz_autotune_save(lz_ctx *ctx, const char *path) {
    if (!ctx || !path) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    FILE *f = std::fopen(path, "w");
    if (!f) return ctx->fail(LZ_ERR_ARG, "cannot write %s", path);
    std::fprintf(f, "# net bucket wm grid_mult ms\n");
    for (int net = 0; net < 2; net++)
        for (int b = 0; b < 4; b++) {
            const auto &t = ctx->net[net].tune[b];
            std::fprintf(f, "%d %d %d %d %.5f\n", net, b, t.wm, t.grid_mult,
                         t.ms);
        }
    std::fclose(f);
    return LZ_OK;
}

int lz_autotune_load(lz_ctx *ctx, const char *path) {
    if (!ctx || !path) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    FILE *f = std::fopen(path, "r");
    if (!f) return ctx->fail(LZ_ERR_ARG, "cannot read %s", path);
    char line[256];
    int loaded = 0;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        int net, b, wm, gm;
        float ms;
        if (std::sscanf(line, "%d %d %d %d %f", &net, &b, &wm, &gm, &ms) != 5 ||
            net < 0 || net > 1 || b < 0 || b > 3 || wm < 0 || wm > 2 ||
            gm < 0 || gm > 8) {
            std::fclose(f);
            return ctx->fail(LZ_ERR_ARG, "bad tuning line in %s", path);
        }
        ctx->net[net].tune[b] = lz::Net::Tune{wm, gm, ms};
        loaded++;
    }
    std::fclose(f);
    return loaded ? LZ_OK : ctx->fail(LZ_ERR_ARG, "empty tuning file");
}

}  // extern "C"
