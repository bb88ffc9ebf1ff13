// lz_queue.cpp — batched leaf evaluation: a multi-producer queue of single
// positions (19x19 or smaller boards) that a dispatcher thread turns into
// batched graph replays.
//
// Reference counterpart: Network::async_scored_moves + OpenCL_Network::
// forward + forward_cb (Network.cpp:471-587, OpenCL.cpp:490-569). There each
// search thread owns an in-order queue and evaluates one position per
// forward, with at most two outstanding (OpenCL.cpp:446-454), and the
// callback runs inline after queue.finish(). Here:
//
//  * Batches are built in place in one of kSets batch buffers (pinned host
//    staging, device buffers and the graphs captured on them). A producer
//    reserves the next row of its net's open batch under the lock and
//    copies its planes into that row outside it. Rows have a fixed stride
//    (32 x 361 plane bytes, 361 result floats) whatever the board side.
//  * The dispatcher closes a batch when it holds max_batch rows, when its
//    oldest row has waited max_wait_us, or on flush; once the last copy
//    into it has finished it sends the rows and the batch table (the row
//    count and, for mixed sides, the slot table of the ragged kernels),
//    replays the graph of the smallest bucket (64 / 256 / 1024 / 2048) that
//    holds the batch — the dense 19x19 graph when every board is 19x19, the
//    ragged graph otherwise — and copies the results back, all on the
//    batch's own stream, then records an event and goes on to the next
//    batch while the device works.
//  * A batch of at most layered_max dense 19x19 rows in a 16-bit form
//    skips the graph and the copies: the layered kernels, launched
//    directly, read its rows from the staging buffer (coherent host
//    memory; the stems read the rows packed one bit per plane, which the
//    producers write beside the bytes) and write the results there.
//  * A completion thread waits on the batches' events in launch order
//    (polling, then blocking), runs the callbacks (never inside the HIP
//    runtime) and returns the buffer to the pool. Up to kSets batches are
//    being filled, launched or evaluated at once.
//  * The 16-bit and FP8 kernels read the row count from device memory, so a
//    graph replayed on a part-filled bucket evaluates exactly its rows; the
//    FP32 graphs evaluate the whole bucket (each row's result is the same
//    either way).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
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
constexpr int kPos = 32 * 361;  // plane bytes of a row, any side
constexpr int kOut = 361;       // result floats of a policy row
constexpr int kP = 361;         // points of a 19x19 board
constexpr int kBuckets[] = {64, 256, 1024, 2048};
constexpr int kNumBuckets = 4;
constexpr int kSets = 4;
// Batch table: {slots, rows} then (ragged batches) the slot table.
constexpr size_t kTableHead = 16;
using Clock = std::chrono::steady_clock;

// Set on the completion thread while it runs callbacks: a submit from a
// callback must not wait for a batch buffer that only this thread frees.
thread_local bool t_in_callback = false;

size_t table_bytes(int cap) { return kTableHead + sizeof(FusedSlot) * cap; }

struct Row {
    lz_done_fn done;
    void *user;
    int side;
    bool packed;  // 19x19 row whose planes are 0 / 1, packed into h_bits
};

// A 19x19 row's planes packed one word per point (bit c = plane c), as
// the layered stem reads them; false if a plane byte is not 0 or 1 (the
// row then goes in as bytes).
bool pack_row(const uint8_t *planes, uint32_t *bits) {
    uint32_t w[kP] = {};
    uint8_t other = 0;
    for (int c = 0; c < 32; c++) {
        const uint8_t *p = planes + c * kP;
        for (int i = 0; i < kP; i++) {
            w[i] |= (uint32_t)(p[i] & 1u) << c;
            other |= (uint8_t)(p[i] & ~1u);
        }
    }
    std::memcpy(bits, w, sizeof(w));
    return other == 0;
}

struct Batch {
    enum State { kFree, kOpen, kClosed, kLaunched };
    State state = kFree;
    int net = 0;
    int rows = 0;      // reserved rows
    std::atomic<int> writers{0};  // producers still copying planes in
    bool all19 = true;
    int status = LZ_OK;
    int64_t seq = 0;   // order of opening
    Clock::time_point t0;
    std::vector<Row> meta;
    uint8_t *h_planes = nullptr, *d_planes = nullptr;
    float *h_raw = nullptr, *h_out = nullptr, *d_raw = nullptr,
          *d_out = nullptr;
    uint8_t *h_table = nullptr, *d_table = nullptr;
    uint32_t *h_bits = nullptr;  // packed 19x19 rows (coherent host)
    int *d_in_off = nullptr, *d_out_off = nullptr;
    Workspace ws;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    hipGraphExec_t graph[2][kNumBuckets][2] = {};  // [net][bucket][ragged]

    void reset() {
        state = kFree;
        rows = 0;
        writers = 0;
        all19 = true;
        status = LZ_OK;
    }
};
}  // namespace

// How long the dispatcher polls for a batch that is filling before it
// sleeps on its condition variable.
constexpr auto kDispatchPoll = std::chrono::microseconds(500);
// How long the completion thread polls a launched batch's event before it
// blocks on it.
constexpr auto kCompletePoll = std::chrono::microseconds(1000);

class LeafQueue {
public:
    LeafQueue(lz_ctx *ctx, int precision, int max_batch, int max_wait_us)
        : ctx_(ctx), precision_(precision), max_wait_(max_wait_us) {
        cap_ = kBuckets[kNumBuckets - 1];
        for (int b : kBuckets)
            if (b >= max_batch) {
                cap_ = b;
                break;
            }
        max_take_ = std::min(max_batch, cap_);
        ragged_ok_ = precision == LZ_PREC_BF16 || precision == LZ_PREC_FP16;
    }

    ~LeafQueue() { stop(); }

    int precision() const { return precision_; }

    int start(std::string *msg) {
        for (int i = 0; i < kSets; i++) {
            int rc = init_batch(batches_[i], msg);
            if (rc != LZ_OK) return rc;
        }
        dispatcher_ = std::thread([this] { dispatch(); });
        completer_ = std::thread([this] { complete(); });
        return LZ_OK;
    }

    int submit(int net, int side, const uint8_t *planes, lz_done_fn done,
               void *user) {
        if (side != 19 && !ragged_ok_) return LZ_ERR_UNSUPPORTED;
        std::unique_lock<std::mutex> lk(mu_);
        Batch *b = nullptr;
        for (;;) {
            if (stopping_) return LZ_ERR_STATE;
            b = open_[net];
            if (b) break;
            for (auto &c : batches_)
                if (c.state == Batch::kFree) {
                    b = &c;
                    break;
                }
            if (b) {
                b->state = Batch::kOpen;
                b->net = net;
                b->seq = next_seq_++;
                b->t0 = Clock::now();
                open_[net] = b;
                break;
            }
            if (t_in_callback) return LZ_ERR_STATE;
            free_cv_.wait(lk);
        }
        const int row = b->rows++;
        b->meta[row] = Row{done, user, side, false};
        if (side != 19) b->all19 = false;
        b->writers++;
        outstanding_++;
        if (b->rows == max_take_) {
            b->state = Batch::kClosed;
            open_[net] = nullptr;
        }
        lk.unlock();
        // A new batch starts the dispatcher's wait timer.
        if (row == 0) cv_.notify_one();
        std::memcpy(b->h_planes + (size_t)row * kPos, planes,
                    (size_t)32 * side * side);
        // Packed only where the batch may run the layered kernels (16-bit
        // forms, within layered_max rows); a batch with an unpacked row
        // goes in as bytes.
        if (side == 19 && row < ctx_->net[net].layered_max &&
            (precision_ == LZ_PREC_BF16 || precision_ == LZ_PREC_FP16))
            b->meta[row].packed =
                pack_row(planes, b->h_bits + (size_t)row * kP);
        // The dispatcher takes a closed batch once its last copy is in:
        // the producer that finishes it wakes the dispatcher (a batch still
        // open is taken when it fills or on the dispatcher's timer).
        if (b->writers.fetch_sub(1, std::memory_order_acq_rel) == 1) {
            lk.lock();
            const bool closed = b->state == Batch::kClosed;
            lk.unlock();
            if (closed) cv_.notify_one();
        }
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
        out[0] = launches_;
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
        free_cv_.notify_all();
        if (dispatcher_.joinable()) dispatcher_.join();
        {
            std::lock_guard<std::mutex> g(mu_);
            completer_stop_ = true;
        }
        comp_cv_.notify_all();
        if (completer_.joinable()) completer_.join();
        for (auto &b : batches_) free_batch(b);
        std::lock_guard<std::mutex> g(mu_);
        stopped_ = true;
    }

private:
    int init_batch(Batch &b, std::string *msg) {
        b.meta.resize(cap_);
        hipError_t e = hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking);
        if (e == hipSuccess)
            e = hipEventCreateWithFlags(&b.done, hipEventDisableTiming);
        // The rows and results are coherent host memory: the layered
        // kernels of small batches read and write them in place.
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void **>(&b.h_planes),
                              (size_t)cap_ * kPos, hipHostMallocCoherent);
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void **>(&b.h_raw),
                              sizeof(float) * cap_ * kOut,
                              hipHostMallocCoherent);
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void **>(&b.h_out),
                              sizeof(float) * cap_ * kOut,
                              hipHostMallocCoherent);
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void **>(&b.h_bits),
                              sizeof(uint32_t) * cap_ * kP,
                              hipHostMallocCoherent);
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void **>(&b.h_table),
                              table_bytes(cap_), hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipMalloc(reinterpret_cast<void **>(&b.d_planes),
                          (size_t)cap_ * kPos);
        if (e == hipSuccess)
            e = hipMalloc(reinterpret_cast<void **>(&b.d_raw),
                          sizeof(float) * cap_ * kOut);
        if (e == hipSuccess)
            e = hipMalloc(reinterpret_cast<void **>(&b.d_out),
                          sizeof(float) * cap_ * kOut);
        if (e == hipSuccess)
            e = hipMalloc(reinterpret_cast<void **>(&b.d_table),
                          table_bytes(cap_));
        if (e == hipSuccess)
            e = hipMalloc(reinterpret_cast<void **>(&b.d_in_off),
                          sizeof(int) * cap_);
        if (e == hipSuccess)
            e = hipMalloc(reinterpret_cast<void **>(&b.d_out_off),
                          sizeof(int) * cap_);
        if (e != hipSuccess) {
            *msg = std::string("queue buffers: ") + hipGetErrorString(e);
            return LZ_ERR_NOMEM;
        }
        // Fixed row offsets of the ragged kernels, and zero planes so that
        // the FP32 graphs' rows past a batch read defined values.
        std::vector<int> in_off(cap_), out_off(cap_);
        for (int i = 0; i < cap_; i++) {
            in_off[i] = i * kPos;
            out_off[i] = i * kOut;
        }
        std::memset(b.h_planes, 0, (size_t)cap_ * kPos);
        std::memset(b.h_table, 0, table_bytes(cap_));
        e = hipMemcpy(b.d_in_off, in_off.data(), sizeof(int) * cap_,
                      hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(b.d_out_off, out_off.data(), sizeof(int) * cap_,
                          hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(b.d_planes, b.h_planes, (size_t)cap_ * kPos,
                          hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(b.d_table, b.h_table, table_bytes(cap_),
                          hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            *msg = std::string("queue buffers: ") + hipGetErrorString(e);
            return LZ_ERR_HIP;
        }
        int rc;
        if (precision_ == LZ_PREC_FP32) {
            rc = ensure_workspace(b.ws, ctx_->net, cap_, msg);
        } else {
            e = hipMalloc(reinterpret_cast<void **>(&b.ws.h0),
                          sizeof(float) * cap_ * kOut);
            b.ws.cap = cap_;
            if (e != hipSuccess) {
                rc = LZ_ERR_NOMEM;
                *msg = std::string("queue workspace: ") + hipGetErrorString(e);
            } else {
                rc = ensure_tiles(b.ws, msg);
            }
        }
        if (rc != LZ_OK) return rc;
        // One graph per net, bucket and layout (dense 19x19 / ragged), on
        // this batch buffer's memory.
        for (int net = 0; net < 2; net++)
            for (int k = 0; k < kNumBuckets && kBuckets[k] <= cap_; k++)
                for (int rg = 0; rg < (ragged_ok_ ? 2 : 1); rg++) {
                    rc = capture(b, net, k, rg == 1, msg);
                    if (rc != LZ_OK) return rc;
                }
        return LZ_OK;
    }

    void free_batch(Batch &b) {
        if (b.stream) hipStreamSynchronize(b.stream);
        for (auto &net : b.graph)
            for (auto &bucket : net)
                for (auto &g : bucket)
                    if (g) {
                        hipGraphExecDestroy(g);
                        g = nullptr;
                    }
        if (b.done) hipEventDestroy(b.done);
        if (b.stream) hipStreamDestroy(b.stream);
        if (b.h_planes) hipHostFree(b.h_planes);
        if (b.h_raw) hipHostFree(b.h_raw);
        if (b.h_out) hipHostFree(b.h_out);
        if (b.h_table) hipHostFree(b.h_table);
        if (b.h_bits) hipHostFree(b.h_bits);
        if (b.d_planes) hipFree(b.d_planes);
        if (b.d_raw) hipFree(b.d_raw);
        if (b.d_out) hipFree(b.d_out);
        if (b.d_table) hipFree(b.d_table);
        if (b.d_in_off) hipFree(b.d_in_off);
        if (b.d_out_off) hipFree(b.d_out_off);
        free_workspace(b.ws);
        b.h_planes = b.d_planes = b.h_table = b.d_table = nullptr;
        b.h_raw = b.h_out = b.d_raw = b.d_out = nullptr;
        b.d_in_off = b.d_out_off = nullptr;
        b.h_bits = nullptr;
        b.stream = nullptr;
        b.done = nullptr;
        b.meta.clear();
        b.reset();
    }

    int capture(Batch &b, int net, int k, bool ragged, std::string *msg) {
        const int n = kBuckets[k];
        RaggedLayout lay;
        lay.n = n;
        lay.nslots = n;  // capacity; the count comes from the batch table
        lay.d_slots = reinterpret_cast<FusedSlot *>(b.d_table + kTableHead);
        lay.d_in_off = b.d_in_off;
        lay.d_out_off = b.d_out_off;
        const int *counts = reinterpret_cast<const int *>(b.d_table);
        hipGraph_t graph = nullptr;
        hipError_t e =
            hipStreamBeginCapture(b.stream, hipStreamCaptureModeRelaxed);
        if (e != hipSuccess) {
            *msg = std::string("begin capture: ") + hipGetErrorString(e);
            return LZ_ERR_HIP;
        }
        int rc = run_net(ctx_->net[net], net, precision_, n, b.d_planes,
                         b.d_raw, b.d_out, ctx_->temperature, b.ws, b.stream,
                         msg, nullptr, ragged ? &lay : nullptr, counts);
        e = hipStreamEndCapture(b.stream, &graph);
        if (rc != LZ_OK) {
            if (graph) hipGraphDestroy(graph);
            return rc;
        }
        if (e == hipSuccess)
            e = hipGraphInstantiate(&b.graph[net][k][ragged ? 1 : 0], graph,
                                    nullptr, nullptr, 0);
        if (graph) hipGraphDestroy(graph);
        if (e != hipSuccess) {
            *msg = std::string("graph capture: ") + hipGetErrorString(e);
            return LZ_ERR_HIP;
        }
        return LZ_OK;
    }

    // Closes open batches that are full of waiting (or all of them on
    // flush / stop); returns the earliest time an open batch comes due.
    Clock::time_point close_due(Clock::time_point now) {
        Clock::time_point next = Clock::time_point::max();
        for (int net = 0; net < 2; net++) {
            Batch *b = open_[net];
            if (!b || b->rows == 0) continue;
            const auto due = b->t0 + std::chrono::microseconds(max_wait_);
            if (flush_ || stopping_ || now >= due) {
                b->state = Batch::kClosed;
                open_[net] = nullptr;
            } else {
                next = std::min(next, due);
            }
        }
        return next;
    }

    void dispatch() {
        hipSetDevice(ctx_->device);
        for (;;) {
            Batch *b = nullptr;
            // While rows are coming in, the dispatcher polls (yielding its
            // core) rather than sleeping: a wake-up costs tens of
            // microseconds, as much as a small batch's launch.
            const auto poll_until = Clock::now() + kDispatchPoll;
            {
                std::unique_lock<std::mutex> lk(mu_);
                for (;;) {
                    const auto now = Clock::now();
                    const auto next = close_due(now);
                    bool pending = false;
                    for (auto &c : batches_) {
                        if (c.state != Batch::kClosed) continue;
                        pending = true;
                        if (c.writers == 0 && (!b || c.seq < b->seq)) b = &c;
                    }
                    if (b) break;
                    const bool open_rows =
                        (open_[0] && open_[0]->rows) ||
                        (open_[1] && open_[1]->rows);
                    if (!pending && !open_rows) {
                        flush_ = false;
                        if (stopping_) return;
                    } else if (now < poll_until && !stopping_) {
                        lk.unlock();
                        std::this_thread::yield();
                        lk.lock();
                        continue;
                    }
                    if (next != Clock::time_point::max())
                        cv_.wait_until(lk, next);
                    else
                        cv_.wait(lk);
                }
            }
            launch(*b);
            {
                std::lock_guard<std::mutex> g(mu_);
                b->state = Batch::kLaunched;
                inflight_.push_back(b);
            }
            comp_cv_.notify_one();
        }
    }

    // Groups rows by side (first-appearance order) into slots of as many
    // boards as ragged_slot_capacity allows; returns the slot count.
    int pack_slots(const Batch &b, FusedSlot *slots) {
        std::vector<char> used(b.rows, 0);
        int ns = 0;
        for (int i = 0; i < b.rows; i++) {
            if (used[i]) continue;
            const int s = b.meta[i].side, cap = ragged_slot_capacity(s);
            FusedSlot cur{s, 0, {0, 0, 0, 0}};
            for (int j = i; j < b.rows; j++) {
                if (used[j] || b.meta[j].side != s) continue;
                used[j] = 1;
                cur.board[cur.count++] = j;
                if (cur.count == cap) {
                    slots[ns++] = cur;
                    cur = FusedSlot{s, 0, {0, 0, 0, 0}};
                }
            }
            if (cur.count) slots[ns++] = cur;
        }
        return ns;
    }

    void launch(Batch &b) {
        const int n = b.rows;
        int k = 0;
        while (kBuckets[k] < n) k++;
        const bool ragged = !b.all19;
        int *head = reinterpret_cast<int *>(b.h_table);
        int ns = n;
        if (ragged)
            ns = pack_slots(b, reinterpret_cast<FusedSlot *>(b.h_table +
                                                             kTableHead));
        head[0] = ns;
        head[1] = n;
        const size_t tbytes =
            kTableHead + (ragged ? sizeof(FusedSlot) * ns : 0);
        const size_t per = b.net == LZ_NET_POLICY ? kOut : 1;
        const Net &net = ctx_->net[b.net];
        hipError_t e = hipSuccess;
        if (!ragged && n <= net.layered_max && b.ws.tiles) {
            // Small dense batches: the layered kernels, launched directly
            // (a graph of their ~14 nodes costs more host time to launch,
            // ~7 us a node, than the launches, which overlap the kernels
            // already running), reading the rows from and writing the
            // results into the staging buffers in host memory. No copies:
            // on either side of a small batch a copy's latency is that of
            // several of its convs. (The layered kernels take the batch
            // size as an argument, not from the batch table.)
            // The stem reads the packed rows when every row has them.
            bool packed = true;
            for (int i = 0; i < n; i++) packed = packed && b.meta[i].packed;
            std::string msg;
            if (run_net(net, b.net, precision_, n, b.h_planes, b.h_raw,
                        b.h_out, ctx_->temperature, b.ws, b.stream, &msg,
                        nullptr, nullptr, nullptr,
                        packed ? b.h_bits : nullptr) != LZ_OK)
                e = hipErrorLaunchFailure;
        } else {
            e = hipMemcpyAsync(b.d_table, b.h_table, tbytes,
                               hipMemcpyHostToDevice, b.stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(b.d_planes, b.h_planes, (size_t)n * kPos,
                                   hipMemcpyHostToDevice, b.stream);
            if (e == hipSuccess)
                e = hipGraphLaunch(b.graph[b.net][k][ragged ? 1 : 0],
                                   b.stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(b.h_raw, b.d_raw, sizeof(float) * per * n,
                                   hipMemcpyDeviceToHost, b.stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(b.h_out, b.d_out, sizeof(float) * per * n,
                                   hipMemcpyDeviceToHost, b.stream);
        }
        if (e == hipSuccess) e = hipEventRecord(b.done, b.stream);
        b.status = e == hipSuccess ? LZ_OK : LZ_ERR_HIP;
        std::lock_guard<std::mutex> g(mu_);
        launches_++;
        positions_ += n;
        per_bucket_[k]++;
    }

    // A batch's event: polled (yielding) for up to kCompletePoll, the
    // length of a small batch, then waited on; a blocking wait's wake-up
    // alone costs tens of microseconds.
    static hipError_t wait_done(hipEvent_t ev) {
        const auto until = Clock::now() + kCompletePoll;
        for (;;) {
            const hipError_t q = hipEventQuery(ev);
            if (q != hipErrorNotReady) return q;
            if (Clock::now() >= until) return hipEventSynchronize(ev);
            std::this_thread::yield();
        }
    }

    // Completion thread: batches in launch order; callbacks, then the
    // buffer goes back to the pool.
    void complete() {
        hipSetDevice(ctx_->device);
        for (;;) {
            Batch *b = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu_);
                comp_cv_.wait(lk, [&] {
                    return !inflight_.empty() || completer_stop_;
                });
                if (inflight_.empty()) return;
                b = inflight_.front();
                inflight_.pop_front();
            }
            int status = b->status;
            if (status == LZ_OK && wait_done(b->done) != hipSuccess)
                status = LZ_ERR_HIP;
            const size_t per = b->net == LZ_NET_POLICY ? kOut : 1;
            t_in_callback = true;
            for (int i = 0; i < b->rows; i++) {
                const Row &r = b->meta[i];
                if (r.done)
                    r.done(r.user, status,
                           status == LZ_OK ? b->h_raw + per * i : nullptr,
                           status == LZ_OK ? b->h_out + per * i : nullptr);
            }
            t_in_callback = false;
            {
                std::lock_guard<std::mutex> g(mu_);
                outstanding_ -= b->rows;
                if (status != LZ_OK) failures_++;
                b->reset();
            }
            free_cv_.notify_all();
            done_cv_.notify_all();
        }
    }

    lz_ctx *ctx_;
    int precision_;
    int cap_ = 0;       // rows a batch buffer holds (largest bucket used)
    int max_take_ = 0;  // rows per batch
    int max_wait_;
    bool ragged_ok_ = false;
    std::mutex mu_;
    std::condition_variable cv_, free_cv_, done_cv_, comp_cv_;
    Batch batches_[kSets];
    Batch *open_[2] = {nullptr, nullptr};
    std::deque<Batch *> inflight_;
    int64_t next_seq_ = 0;
    int outstanding_ = 0;
    int64_t failures_ = 0, launches_ = 0, positions_ = 0;
    int64_t per_bucket_[kNumBuckets] = {0, 0, 0, 0};
    bool flush_ = false, stopping_ = false, stopped_ = false;
    bool completer_stop_ = false;
    std::thread dispatcher_, completer_;
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
        max_batch < 1 || max_batch > 2048 || max_wait_us < 0)
        return ctx->fail(LZ_ERR_ARG, "bad queue parameters");
    for (int k = 0; k < 2; k++)
        if (ctx->net[k].layers[0].cin != 32)
            return ctx->fail(LZ_ERR_UNSUPPORTED,
                             "the leaf queue lays out the reference's 32 "
                             "input planes per position");
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

int lz_submit_side(lz_ctx *ctx, int net, int side, const uint8_t *planes,
                   lz_done_fn done, void *user) {
    if (!ctx || !planes || (net != LZ_NET_POLICY && net != LZ_NET_VALUE) ||
        lz::ragged_slot_capacity(side) == 0)
        return LZ_ERR_ARG;
    if (!ctx->queue) return LZ_ERR_STATE;
    return ctx->queue->submit(net, side, planes, done, user);
}

int lz_submit(lz_ctx *ctx, int net, const uint8_t *planes, lz_done_fn done,
              void *user) {
    return lz_submit_side(ctx, net, 19, planes, done, user);
}

namespace {
void count_done(void *user, int status, const float *, const float *) {
    auto *c = static_cast<std::atomic<int64_t> *>(user);
    c[status == LZ_OK ? 0 : 1]++;
}
}  // namespace

int lz_queue_benchmark(lz_ctx *ctx, int n, const int32_t *sides,
                       const uint8_t *policy_planes,
                       const uint8_t *value_planes, int threads, int rounds,
                       double *positions_per_s) {
    if (!ctx || n < 1 || !sides || !policy_planes || !value_planes ||
        threads < 1 || threads > 1024 || rounds < 1 || !positions_per_s)
        return LZ_ERR_ARG;
    if (!ctx->queue) return LZ_ERR_STATE;
    std::vector<size_t> off(n + 1, 0);
    for (int i = 0; i < n; i++) {
        if (lz::ragged_slot_capacity(sides[i]) == 0) return LZ_ERR_ARG;
        off[i + 1] = off[i] + (size_t)32 * sides[i] * sides[i];
    }
    std::atomic<int64_t> counts[2];
    counts[0] = 0;
    counts[1] = 0;
    std::atomic<int> submit_err{LZ_OK};
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back([&, t] {
            for (int r = 0; r < rounds; r++)
                for (int i = t; i < n; i += threads) {
                    int rc = lz_submit_side(ctx, LZ_NET_POLICY, sides[i],
                                            policy_planes + off[i],
                                            count_done, counts);
                    if (rc == LZ_OK)
                        rc = lz_submit_side(ctx, LZ_NET_VALUE, sides[i],
                                            value_planes + off[i],
                                            count_done, counts);
                    if (rc != LZ_OK) {
                        submit_err = rc;
                        return;
                    }
                }
        });
    for (auto &th : pool) th.join();
    int rc = ctx->queue->join();
    const double dt = std::chrono::duration<double>(
                          std::chrono::steady_clock::now() - t0)
                          .count();
    if (submit_err != LZ_OK) return submit_err;
    if (rc != LZ_OK || counts[1] != 0 ||
        counts[0] != (int64_t)2 * rounds * n)
        return LZ_ERR_HIP;
    *positions_per_s = (double)rounds * n / dt;
    return LZ_OK;
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
