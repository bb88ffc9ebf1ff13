// network.cpp — the Network facade (Network.cpp:536-832 semantics) over
// the C ABI, plus its C entry points.
#include "network.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>

#include "input_planes.h"
#include "lz_ctx.h"
#include "lz_position.h"
#include "lz_random.h"

namespace lz {
namespace {

constexpr int kP = 361;
constexpr int kPos = kPlanes * kP;

// Per-thread generator for RANDOM_ROTATION, seeded from the clock and the
// thread id as Random::get_Rng does (Random.cpp:9-22).
Rng &thread_rng() {
    thread_local Rng rng(
        (uint32_t)std::chrono::steady_clock::now().time_since_epoch().count() ^
        (uint32_t)std::hash<std::thread::id>()(std::this_thread::get_id()));
    return rng;
}

int pick_rotation(Network::Ensemble e, int rotation) {
    return e == Network::RANDOM_ROTATION ? (int)thread_rng().randfix<8>()
                                         : rotation;
}

}  // namespace

int Network::rotate_nn_idx(int vertex, int symmetry) {
    return lz::rotate_nn_idx(vertex, symmetry, 19);
}

int Network::rev_rotate_nn_idx(int vertex, int symmetry) {
    return lz::rev_rotate_nn_idx(vertex, symmetry, 19);
}

void Network::softmax(const std::vector<float> &input,
                      std::vector<float> &output, float temperature) {
    float alpha = input[0];
    for (size_t i = 1; i < output.size(); i++) alpha = std::max(alpha, input[i]);
    alpha /= temperature;
    float denom = 0.0f;
    std::vector<float> helper(output.size());
    for (size_t i = 0; i < output.size(); i++) {
        helper[i] = std::exp(input[i] / temperature - alpha);
        denom += helper[i];
    }
    for (size_t i = 0; i < output.size(); i++) output[i] = helper[i] / denom;
}

void collect_scored_moves(const go::GameState &state, const float *probs,
                          int rotation, Network::Netresult &out) {
    const go::Board &b = state.board;
    out.clear();
    for (int idx = 0; idx < kP; idx++) {
        const int x = idx % 19, y = idx / 19;
        const int v = b.vertex(x, y);
        if (b.at(v) == go::EMPTY)
            out.emplace_back(probs[Network::rev_rotate_nn_idx(idx, rotation)], v);
    }
}

void prune_ladders(const go::GameState &state, const uint8_t *ladder,
                   Network::Netresult &result) {
    for (auto &sm : result) {
        const auto xy = state.board.xy(sm.second);
        if (ladder[xy.second * 19 + xy.first]) sm.first = 0.0f;
    }
}

namespace {

// One synchronous call's rows in flight through the leaf queue. The
// caller first polls for a while (yielding its core), then sleeps: a leaf
// batch comes back within a few hundred microseconds, and a sleeping
// thread's wake-up alone costs tens of them.
struct SyncCall {
    std::mutex m;
    std::condition_variable cv;
    std::atomic<int> left{0};
    std::atomic<int> status{LZ_OK};
    float *out = nullptr;
    int width = 0;
};

struct SyncRow {
    SyncCall *call;
    int row;
};

constexpr auto kSyncPoll = std::chrono::milliseconds(2);

void sync_done(void *user, int status, const float *raw, const float *out) {
    (void)raw;
    SyncRow *r = static_cast<SyncRow *>(user);
    SyncCall *c = r->call;
    if (status == LZ_OK)
        std::copy(out, out + c->width, c->out + (size_t)r->row * c->width);
    // The count drops under the mutex, and the caller takes the mutex
    // before it returns: c is not touched once the caller may have left.
    std::lock_guard<std::mutex> g(c->m);
    if (status != LZ_OK) c->status.store(status);
    if (c->left.fetch_sub(1) == 1) c->cv.notify_all();
}

void wait_rows(SyncCall &c) {
    const auto until = std::chrono::steady_clock::now() + kSyncPoll;
    while (c.left.load() != 0 && std::chrono::steady_clock::now() < until)
        std::this_thread::yield();
    std::unique_lock<std::mutex> lk(c.m);
    c.cv.wait(lk, [&] { return c.left.load() == 0; });
}

}  // namespace

int Network::forward_rows(int net, int rows, const uint8_t *in, float *out) {
    if (forward_) return forward_(net, rows, in, out);
    // The facade encodes the reference's 32 planes per position.
    if (ctx_->net[net].layers.empty() || ctx_->net[net].layers[0].cin != 32)
        return LZ_ERR_UNSUPPORTED;
    if (queue_precision(ctx_) != precision_)
        return lz_forward(ctx_, net, precision_, rows, in, nullptr, out);
    SyncCall call;
    call.out = out;
    call.width = net == LZ_NET_POLICY ? kP : 1;
    std::vector<SyncRow> row(rows);
    int submitted = 0, rc = LZ_OK;
    call.left.store(rows);
    for (int r = 0; r < rows && rc == LZ_OK; r++) {
        row[r] = {&call, r};
        rc = lz_submit(ctx_, net, in + (size_t)r * kPos, sync_done, &row[r]);
        if (rc == LZ_OK) submitted++;
    }
    {
        std::lock_guard<std::mutex> g(call.m);
        call.left.fetch_sub(rows - submitted);  // rows never queued
    }
    wait_rows(call);
    return rc != LZ_OK ? rc : call.status.load();
}

void Network::benchmark(const go::GameState &state, int threads,
                        float *policy_ps, float *value_ps) {
    const int cpus = threads < 1 ? 1 : threads;
    auto run = [&](int amount, bool policy) {
        const int iters = (amount + cpus - 1) / cpus;
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<std::thread> pool;
        for (int i = 0; i < cpus; i++)
            pool.emplace_back([&, iters, policy] {
                go::GameState mystate = state;
                Network net(ctx_, precision_);
                for (int loop = 0; loop < iters; loop++) {
                    if (policy)
                        net.get_scored_moves(mystate, RANDOM_ROTATION);
                    else
                        net.get_value(mystate, RANDOM_ROTATION);
                }
            });
        for (auto &t : pool) t.join();
        const double s = std::chrono::duration<double>(
                             std::chrono::steady_clock::now() - t0)
                             .count();
        return (float)(amount / s);
    };
    *policy_ps = run(2000, true);
    *value_ps = run(10000, false);
}

Network::Netresult Network::get_scored_moves(const go::GameState &state,
                                             Ensemble ensemble,
                                             int rotation) {
    Netresult result;
    status_ = LZ_OK;
    if (state.board.side() != 19) return result;
    std::vector<uint8_t> planes(kPos);
    gather_features_policy(state, planes.data());
    const uint8_t *ladder = planes.data() + 25 * kP;

    const int rows = ensemble == AVERAGE_ALL ? 8 : 1;
    std::vector<int> rots(rows);
    for (int r = 0; r < rows; r++)
        rots[r] = ensemble == AVERAGE_ALL ? r : pick_rotation(ensemble, rotation);
    if (rots[0] < 0 || rots[0] > 7) {
        status_ = LZ_ERR_ARG;
        return result;
    }
    std::vector<uint8_t> in((size_t)rows * kPos);
    for (int r = 0; r < rows; r++)
        rotate_planes(planes.data(), rots[r], 19, in.data() + (size_t)r * kPos);
    std::vector<float> probs((size_t)rows * kP);
    status_ = forward_rows(LZ_NET_POLICY, rows, in.data(), probs.data());
    if (status_ != LZ_OK) return result;

    collect_scored_moves(state, probs.data(), rots[0], result);
    if (rows == 8) {
        Netresult other;
        for (int r = 1; r < 8; r++) {
            collect_scored_moves(state, probs.data() + (size_t)r * kP, rots[r],
                                 other);
            for (size_t i = 0; i < result.size(); i++)
                result[i].first += other[i].first;
        }
        for (auto &sn : result) sn.first /= 8.0f;
    }
    prune_ladders(state, ladder, result);
    return result;
}

float Network::get_value(const go::GameState &state, Ensemble ensemble) {
    status_ = LZ_OK;
    if (state.board.side() != 19) return 0.5f;
    std::vector<uint8_t> planes(kPos);
    gather_features_value(state, planes.data());
    const int rows = ensemble == AVERAGE_ALL ? 8 : 1;
    std::vector<uint8_t> in((size_t)rows * kPos);
    for (int r = 0; r < rows; r++) {
        const int rot = ensemble == AVERAGE_ALL ? r
                        : ensemble == DIRECT    ? 0
                                                : pick_rotation(ensemble, -1);
        rotate_planes(planes.data(), rot, 19, in.data() + (size_t)r * kPos);
    }
    std::vector<float> win(rows);
    status_ = forward_rows(LZ_NET_VALUE, rows, in.data(), win.data());
    if (status_ != LZ_OK) return 0.5f;
    float sum = 0.0f;
    for (float w : win) sum += w;
    return sum / rows;
}

namespace {

// What forward_cb needs after the batch (Network.cpp:472-482): a copy of
// the state, the rotation, the ladder plane and the consumer's callback.
struct AsyncJob {
    go::GameState state;
    int rotation;
    std::vector<uint8_t> ladder;
    Network::ScoringCallback cb;
};

void async_done(void *user, int status, const float *raw, const float *out) {
    (void)raw;
    std::unique_ptr<AsyncJob> job(static_cast<AsyncJob *>(user));
    Network::Netresult result;
    if (status == LZ_OK) {
        collect_scored_moves(job->state, out, job->rotation, result);
        prune_ladders(job->state, job->ladder.data(), result);
    }
    job->cb(result);
}

}  // namespace

int Network::async_scored_moves(const go::GameState &state,
                                ScoringCallback cb, Ensemble ensemble,
                                int rotation) {
    if (state.board.side() != 19) return LZ_OK;  // as Network.cpp:541-543
    if (ensemble == AVERAGE_ALL) return LZ_ERR_ARG;
    const int rot = pick_rotation(ensemble, rotation);
    if (rot < 0 || rot > 7) return LZ_ERR_ARG;
    std::vector<uint8_t> planes(kPos), in(kPos);
    gather_features_policy(state, planes.data());
    rotate_planes(planes.data(), rot, 19, in.data());
    auto job = std::make_unique<AsyncJob>();
    job->state = state;
    job->rotation = rot;
    job->ladder.assign(planes.begin() + 25 * kP, planes.begin() + 26 * kP);
    job->cb = std::move(cb);
    const int rc =
        lz_submit(ctx_, LZ_NET_POLICY, in.data(), async_done, job.get());
    if (rc == LZ_OK) job.release();
    return rc;
}

namespace {

// FastBoard::move_to_text (FastBoard.cpp:994-1018): column letter without
// 'I', then the row number from 1.
std::string move_text(const go::Board &b, int vertex) {
    if (vertex == go::kPass) return "pass";
    const auto xy = b.xy(vertex);
    std::string s(1, (char)(xy.first < 8 ? 'A' + xy.first : 'A' + xy.first + 1));
    return s + std::to_string(xy.second + 1);
}

}  // namespace

std::string Network::show_heatmap(const go::GameState &state,
                                  const Netresult &result, bool topmoves) {
    const go::Board &b = state.board;
    std::string text;
    char cell[16];
    for (int y = 18; y >= 0; y--) {
        for (int x = 0; x < 19; x++) {
            const int v = b.vertex(x, y);
            float score = 0.0f;
            for (const auto &sn : result)
                if (sn.second == v) {
                    score = sn.first;
                    break;
                }
            std::snprintf(cell, sizeof cell, "%3d ", int(score * 1000));
            text += cell;
        }
        text += "\n";
    }
    if (topmoves) {
        Netresult moves = result;
        // Descending by (probability, vertex), ties keeping their order.
        std::stable_sort(moves.rbegin(), moves.rend());
        float cum = 0.0f;
        for (size_t tried = 0; cum < 0.85f && tried < moves.size(); tried++) {
            if (moves[tried].first < 0.01f) break;
            std::snprintf(cell, sizeof cell, "%1.3f", moves[tried].first);
            text += std::string(cell) + " (" +
                    move_text(b, moves[tried].second) + ")\n";
            cum += moves[tried].first;
        }
    }
    return text;
}

std::string Network::get_backend() const {
    return ctx_ ? std::string(lz_device_name(ctx_)) : std::string("host");
}

bool Network::can_issue(int max_outstanding) const {
    return lz_queue_depth(ctx_) < max_outstanding;
}

void Network::join_outstanding_cb() { lz_join(ctx_); }

}  // namespace lz

extern "C" {

int lz_network_scored_moves(lz_ctx *ctx, int precision,
                            const lz_position *pos, int ensemble,
                            int rotation, float *probs, int32_t *vertices,
                            int *count) {
    if (!ctx || !pos || !probs || !vertices || !count || ensemble < 0 ||
        ensemble > 2)
        return LZ_ERR_ARG;
    lz::Network net(ctx, precision);
    const auto res = net.get_scored_moves(
        pos->st, static_cast<lz::Network::Ensemble>(ensemble), rotation);
    if (net.last_status() != LZ_OK) return net.last_status();
    *count = (int)res.size();
    for (size_t i = 0; i < res.size(); i++) {
        probs[i] = res[i].first;
        vertices[i] = res[i].second;
    }
    return LZ_OK;
}

int lz_network_value(lz_ctx *ctx, int precision, const lz_position *pos,
                     int ensemble, float *value) {
    if (!ctx || !pos || !value || ensemble < 0 || ensemble > 2)
        return LZ_ERR_ARG;
    lz::Network net(ctx, precision);
    *value = net.get_value(pos->st,
                           static_cast<lz::Network::Ensemble>(ensemble));
    return net.last_status();
}

int lz_network_heatmap(lz_ctx *ctx, int precision, const lz_position *pos,
                       int ensemble, int rotation, int topmoves, char *text,
                       size_t text_len) {
    if (!ctx || !pos || !text || text_len == 0 || ensemble < 0 ||
        ensemble > 2)
        return LZ_ERR_ARG;
    lz::Network net(ctx, precision);
    const auto res = net.get_scored_moves(
        pos->st, static_cast<lz::Network::Ensemble>(ensemble), rotation);
    if (net.last_status() != LZ_OK) return net.last_status();
    const std::string s =
        lz::Network::show_heatmap(pos->st, res, topmoves != 0);
    if (s.size() + 1 > text_len) return LZ_ERR_ARG;
    std::memcpy(text, s.c_str(), s.size() + 1);
    return LZ_OK;
}

int lz_network_initialize(lz_ctx *ctx) {
    if (!ctx) return LZ_ERR_ARG;
    // Network::initialize (Network.cpp:201-233): the policy net's 13 convs,
    // the value net's 12 convs and two inner products, in push order.
    struct Shape {
        int kind, filter, cin, cout;
    };
    std::vector<Shape> shapes[2];
    shapes[0].push_back({0, 5, 32, 96});
    shapes[0].push_back({0, 3, 96, 128});
    for (int i = 0; i < 10; i++) shapes[0].push_back({0, 3, 128, 128});
    shapes[0].push_back({0, 3, 128, 1});
    shapes[1].push_back({0, 5, 32, 64});
    for (int i = 0; i < 10; i++) shapes[1].push_back({0, 3, 64, 64});
    shapes[1].push_back({0, 3, 64, 1});
    shapes[1].push_back({1, 0, 361, 256});
    shapes[1].push_back({1, 0, 256, 1});
    // The reference compiles its weights in (NN128.cpp, NNValue.cpp, not
    // shipped); these are He-normal, N(0, 2 / fan_in), layer l of net k
    // from the reference RNG stream 7000 + 100 k + l, biases zero.
    for (int k = 0; k < 2; k++) {
        for (size_t l = 0; l < shapes[k].size(); l++) {
            const Shape &s = shapes[k][l];
            const int fan_in = s.kind == 0 ? s.cin * s.filter * s.filter
                                           : s.cin;
            std::vector<float> w((size_t)s.cout * fan_in), b(s.cout, 0.0f);
            lz::normal_fill(7000 + 100 * k + (uint32_t)l, w.size(),
                            (float)std::sqrt(2.0 / fan_in), w.data());
            const int rc =
                s.kind == 0
                    ? lz_push_convolve(ctx, k, s.filter, s.cin, s.cout,
                                       w.data(), b.data())
                    : lz_push_innerproduct(ctx, k, s.cin, s.cout, w.data(),
                                           b.data());
            if (rc != LZ_OK) return rc;
        }
    }
    return lz_finalize(ctx);
}

lz_ctx *lz_network_get(void) {
    static std::once_flag once;
    static lz_ctx *net = nullptr;
    std::call_once(once, [] {
        lz_ctx *c = nullptr;
        if (lz_init(&c, 0) != LZ_OK) return;
        if (lz_network_initialize(c) != LZ_OK) {
            lz_destroy(c);
            return;
        }
        net = c;
    });
    return net;
}

int lz_network_benchmark(lz_ctx *ctx, int precision, const lz_position *pos,
                         int threads, float *policy_ps, float *value_ps) {
    if (!ctx || !pos || !policy_ps || !value_ps || threads < 1)
        return LZ_ERR_ARG;
    if (pos->st.board.side() != 19) return LZ_ERR_ARG;
    lz::Network net(ctx, precision);
    net.benchmark(pos->st, threads, policy_ps, value_ps);
    return LZ_OK;
}

int lz_network_async_scored_moves(lz_ctx *ctx, const lz_position *pos,
                                  int rotation, lz_moves_fn done,
                                  void *user) {
    if (!ctx || !pos || !done || rotation < 0 || rotation > 7)
        return LZ_ERR_ARG;
    lz::Network net(ctx, LZ_PREC_FP32);  // the queue fixes the precision
    return net.async_scored_moves(
        pos->st,
        [done, user](const lz::Network::Netresult &r) {
            std::vector<float> p(r.size());
            std::vector<int32_t> v(r.size());
            for (size_t i = 0; i < r.size(); i++) {
                p[i] = r[i].first;
                v[i] = r[i].second;
            }
            done(user, (int)r.size(), p.data(), v.data());
        },
        lz::Network::DIRECT, rotation);
}

}  // extern "C"
