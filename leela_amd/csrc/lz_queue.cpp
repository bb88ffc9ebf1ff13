// lz_queue.cpp — batched leaf evaluation: a multi-producer queue of single
// positions that a dispatcher thread turns into batched forwards.
//
// Reference counterpart: Network::async_scored_moves + OpenCL_Network::
// forward + forward_cb (Network.cpp:471-587, OpenCL.cpp:490-569). There each
// search thread owns an in-order queue and evaluates one position per
// forward, with at most two outstanding (OpenCL.cpp:446-454). Here any
// number of threads submit positions; the dispatcher gathers up to
// max_batch of one net (or whatever has waited max_wait_us), stages them in
// pinned memory, and replays a hipGraph captured for the smallest bucket
// (64 / 256 / 1024 / 2048) that holds the batch. Callbacks run on the
// dispatcher thread after the batch's results are back on the host, never
// inside the HIP runtime.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>
#include <vector>

#include "leela_amd.h"
#include "lz_ctx.h"

namespace lz {

namespace {
constexpr int kPos = 32 * 361;
constexpr int kBuckets[] = {64, 256, 1024, 2048};
constexpr int kNumBuckets = 4;
using Clock = std::chrono::steady_clock;
}  // namespace

struct Request {
    int net;
    lz_done_fn done;
    void *user;
    Clock::time_point t;
    std::array<uint8_t, kPos> planes;
};

class LeafQueue {
public:
    LeafQueue(lz_ctx *ctx, int precision, int max_batch, int max_wait_us)
        : ctx_(ctx), precision_(precision), max_wait_(max_wait_us) {
        max_batch_ = kBuckets[kNumBuckets - 1];
        for (int b : kBuckets)
            if (b >= max_batch) {
                max_batch_ = b;
                break;
            }
        max_take_ = std::min(max_batch, max_batch_);
    }

    ~LeafQueue() { stop(); }

    int precision() const { return precision_; }

    int start(std::string *msg) {
        hipError_t e = hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking);
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void **>(&h_planes_),
                              (size_t)max_batch_ * kPos, hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void **>(&h_raw_),
                              sizeof(float) * max_batch_ * 361,
                              hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void **>(&h_out_),
                              sizeof(float) * max_batch_ * 361,
                              hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipMalloc(reinterpret_cast<void **>(&d_planes_),
                          (size_t)max_batch_ * kPos);
        if (e == hipSuccess)
            e = hipMalloc(reinterpret_cast<void **>(&d_raw_),
                          sizeof(float) * max_batch_ * 361);
        if (e == hipSuccess)
            e = hipMalloc(reinterpret_cast<void **>(&d_out_),
                          sizeof(float) * max_batch_ * 361);
        if (e != hipSuccess) {
            *msg = std::string("queue buffers: ") + hipGetErrorString(e);
            return LZ_ERR_NOMEM;
        }
        std::memset(h_planes_, 0, (size_t)max_batch_ * kPos);
        e = hipMemcpy(d_planes_, h_planes_, (size_t)max_batch_ * kPos,
                      hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            *msg = std::string("queue buffers: ") + hipGetErrorString(e);
            return LZ_ERR_HIP;
        }
        int rc = ensure_workspace(ws_, ctx_->net, max_batch_, msg);
        if (rc != LZ_OK) return rc;
        // One graph per (net, bucket): the kernels of one forward over a
        // full bucket, on the queue's own buffers and workspace.
        for (int net = 0; net < 2; net++)
            for (int b = 0; b < kNumBuckets && kBuckets[b] <= max_batch_; b++) {
                rc = capture(net, b, msg);
                if (rc != LZ_OK) return rc;
            }
        thread_ = std::thread([this] { run(); });
        return LZ_OK;
    }

    int submit(int net, const uint8_t *planes, lz_done_fn done, void *user) {
        auto r = std::make_unique<Request>();
        r->net = net;
        r->done = done;
        r->user = user;
        r->t = Clock::now();
        std::memcpy(r->planes.data(), planes, kPos);
        {
            std::lock_guard<std::mutex> g(mu_);
            if (stopping_) return LZ_ERR_STATE;
            pending_[net].push_back(std::move(r));
            outstanding_++;
        }
        cv_.notify_one();
        return LZ_OK;
    }

    void flush() {
        {
            std::lock_guard<std::mutex> g(mu_);
            flush_ = true;
        }
        cv_.notify_one();
    }

    int join() {
        flush();
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return outstanding_ == 0; });
        return failures_ ? LZ_ERR_HIP : LZ_OK;
    }

    int depth() {
        std::lock_guard<std::mutex> g(mu_);
        return outstanding_;
    }

    void stats(int64_t *out) {
        std::lock_guard<std::mutex> g(mu_);
        out[0] = batches_;
        out[1] = positions_;
        for (int b = 0; b < kNumBuckets; b++) out[2 + b] = per_bucket_[b];
    }

    void stop() {
        {
            std::lock_guard<std::mutex> g(mu_);
            if (stopped_) return;
            stopping_ = true;
            flush_ = true;
        }
        cv_.notify_all();
        if (thread_.joinable()) thread_.join();
        for (auto &row : graphs_)
            for (auto &g : row)
                if (g) hipGraphExecDestroy(g);
        if (stream_) hipStreamDestroy(stream_);
        if (h_planes_) hipHostFree(h_planes_);
        if (h_raw_) hipHostFree(h_raw_);
        if (h_out_) hipHostFree(h_out_);
        if (d_planes_) hipFree(d_planes_);
        if (d_raw_) hipFree(d_raw_);
        if (d_out_) hipFree(d_out_);
        free_workspace(ws_);
        std::lock_guard<std::mutex> g(mu_);
        stopped_ = true;
    }

private:
    int capture(int net, int b, std::string *msg) {
        hipGraph_t graph = nullptr;
        hipError_t e =
            hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed);
        if (e != hipSuccess) {
            *msg = std::string("begin capture: ") + hipGetErrorString(e);
            return LZ_ERR_HIP;
        }
        int rc = run_net(ctx_->net[net], net, precision_, kBuckets[b],
                         d_planes_, d_raw_, d_out_, ctx_->temperature, ws_,
                         stream_, msg);
        e = hipStreamEndCapture(stream_, &graph);
        if (rc != LZ_OK) {
            if (graph) hipGraphDestroy(graph);
            return rc;
        }
        if (e == hipSuccess)
            e = hipGraphInstantiate(&graphs_[net][b], graph, nullptr, nullptr,
                                    0);
        if (graph) hipGraphDestroy(graph);
        if (e != hipSuccess) {
            *msg = std::string("graph capture: ") + hipGetErrorString(e);
            return LZ_ERR_HIP;
        }
        return LZ_OK;
    }

    // Dispatcher: take a batch when one net has max_take_ waiting, when the
    // oldest request has waited max_wait_, or on flush/stop.
    void run() {
        hipSetDevice(ctx_->device);
        std::vector<std::unique_ptr<Request>> batch;
        for (;;) {
            int net = -1;
            {
                std::unique_lock<std::mutex> lk(mu_);
                for (;;) {
                    const bool any = !pending_[0].empty() || !pending_[1].empty();
                    if (!any) {
                        flush_ = false;
                        if (stopping_) return;
                        cv_.wait(lk);
                        continue;
                    }
                    net = pick_net();
                    const auto age = Clock::now() - pending_[net].front()->t;
                    if ((int)pending_[net].size() >= max_take_ || flush_ ||
                        stopping_ ||
                        age >= std::chrono::microseconds(max_wait_))
                        break;
                    cv_.wait_for(lk, std::chrono::microseconds(max_wait_) - age);
                }
                const int n = std::min<int>(max_take_, pending_[net].size());
                batch.clear();
                for (int i = 0; i < n; i++) {
                    batch.push_back(std::move(pending_[net].front()));
                    pending_[net].pop_front();
                }
            }
            const int status = launch(net, batch);
            const size_t per = net == LZ_NET_POLICY ? 361 : 1;
            for (size_t i = 0; i < batch.size(); i++) {
                const Request &r = *batch[i];
                if (r.done)
                    r.done(r.user, status,
                           status == LZ_OK ? h_raw_ + per * i : nullptr,
                           status == LZ_OK ? h_out_ + per * i : nullptr);
            }
            {
                std::lock_guard<std::mutex> g(mu_);
                outstanding_ -= (int)batch.size();
                if (status != LZ_OK) failures_++;
            }
            done_cv_.notify_all();
        }
    }

    // The net whose oldest request has waited longest.
    int pick_net() const {
        if (pending_[0].empty()) return 1;
        if (pending_[1].empty()) return 0;
        return pending_[0].front()->t <= pending_[1].front()->t ? 0 : 1;
    }

    int launch(int net, const std::vector<std::unique_ptr<Request>> &batch) {
        const int n = (int)batch.size();
        int b = 0;
        while (kBuckets[b] < n) b++;
        for (int i = 0; i < n; i++)
            std::memcpy(h_planes_ + (size_t)i * kPos, batch[i]->planes.data(),
                        kPos);
        const size_t per = net == LZ_NET_POLICY ? 361 : 1;
        hipError_t e = hipMemcpyAsync(d_planes_, h_planes_, (size_t)n * kPos,
                                      hipMemcpyHostToDevice, stream_);
        if (e == hipSuccess) e = hipGraphLaunch(graphs_[net][b], stream_);
        if (e == hipSuccess)
            e = hipMemcpyAsync(h_raw_, d_raw_, sizeof(float) * per * n,
                               hipMemcpyDeviceToHost, stream_);
        if (e == hipSuccess)
            e = hipMemcpyAsync(h_out_, d_out_, sizeof(float) * per * n,
                               hipMemcpyDeviceToHost, stream_);
        if (e == hipSuccess) e = hipStreamSynchronize(stream_);
        std::lock_guard<std::mutex> g(mu_);
        batches_++;
        positions_ += n;
        per_bucket_[b]++;
        return e == hipSuccess ? LZ_OK : LZ_ERR_HIP;
    }

    lz_ctx *ctx_;
    int precision_;
    int max_batch_ = 0;  // largest bucket in use
    int max_take_ = 0;   // positions per batch
    int max_wait_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::deque<std::unique_ptr<Request>> pending_[2];
    int outstanding_ = 0;
    int64_t failures_ = 0, batches_ = 0, positions_ = 0;
    int64_t per_bucket_[kNumBuckets] = {0, 0, 0, 0};
    bool flush_ = false, stopping_ = false, stopped_ = false;
    std::thread thread_;
    hipStream_t stream_ = nullptr;
    hipGraphExec_t graphs_[2][kNumBuckets] = {};
    Workspace ws_;
    uint8_t *h_planes_ = nullptr, *d_planes_ = nullptr;
    float *h_raw_ = nullptr, *h_out_ = nullptr;
    float *d_raw_ = nullptr, *d_out_ = nullptr;
};

int queue_precision(const lz_ctx *ctx) {
    return ctx->queue ? ctx->queue->precision() : -1;
}

void destroy_queue(lz_ctx *ctx) {
    if (ctx->queue) {
        ctx->queue->stop();
        delete ctx->queue;
        ctx->queue = nullptr;
    }
}

}  // namespace lz

extern "C" {

int lz_queue_start(lz_ctx *ctx, int precision, int max_batch,
                   int max_wait_us) {
    if (!ctx) return LZ_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (!ctx->finalized)
        return ctx->fail(LZ_ERR_STATE, "queue before lz_finalize");
    if (ctx->queue) return ctx->fail(LZ_ERR_STATE, "queue already started");
    if (precision < LZ_PREC_FP32 || precision > LZ_PREC_FP16 ||
        max_batch < 1 || max_batch > 2048 ||
        max_wait_us < 0)
        return ctx->fail(LZ_ERR_ARG, "bad queue parameters");
    hipSetDevice(ctx->device);
    auto q = std::make_unique<lz::LeafQueue>(ctx, precision, max_batch,
                                             max_wait_us);
    std::string msg;
    int rc = q->start(&msg);
    if (rc != LZ_OK) {
        q->stop();
        return ctx->fail(rc, "%s", msg.c_str());
    }
    ctx->queue = q.release();
    return LZ_OK;
}

int lz_submit(lz_ctx *ctx, int net, const uint8_t *planes, lz_done_fn done,
              void *user) {
    if (!ctx || !planes || (net != LZ_NET_POLICY && net != LZ_NET_VALUE))
        return LZ_ERR_ARG;
    if (!ctx->queue) return LZ_ERR_STATE;
    return ctx->queue->submit(net, planes, done, user);
}

int lz_flush(lz_ctx *ctx) {
    if (!ctx) return LZ_ERR_ARG;
    if (!ctx->queue) return LZ_ERR_STATE;
    ctx->queue->flush();
    return LZ_OK;
}

int lz_join(lz_ctx *ctx) {
    if (!ctx) return LZ_ERR_ARG;
    if (!ctx->queue) return LZ_ERR_STATE;
    return ctx->queue->join();
}

int lz_queue_depth(lz_ctx *ctx) {
    if (!ctx) return LZ_ERR_ARG;
    if (!ctx->queue) return 0;
    return ctx->queue->depth();
}

int lz_queue_stats(lz_ctx *ctx, int64_t *stats) {
    if (!ctx || !stats) return LZ_ERR_ARG;
    if (!ctx->queue) return LZ_ERR_STATE;
    ctx->queue->stats(stats);
    return LZ_OK;
}

int lz_queue_stop(lz_ctx *ctx) {
    if (!ctx) return LZ_ERR_ARG;
    lz::destroy_queue(ctx);
    return LZ_OK;
}

}  // extern "C"
