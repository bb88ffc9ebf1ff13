// lz_search.cpp — search-side batching: many search threads grow one game
// tree and evaluate its leaves through the network facade, whose
// synchronous calls share GPU launches through the leaf queue. This is the
// integration of UCTNode::netscore_children / UCTNode::run_value_net with
// the batched backend (UCTNode.cpp:74-121, 316-344; UCTSearch.cpp:79-149),
// in place of one blocking OpenCL forward per search thread with the
// thread_can_issue backpressure (OpenCL.cpp:446-454).
//
// The tree policy is a compact one, not a restatement of the reference's
// (UCT with RAVE, Monte Carlo playouts mixed in, net priors with cutoff
// ratios: UCTNode.cpp:543-700): selection by Q + c * P * sqrt(N) / (1 + n)
// over the policy net's priors, virtual loss on the path while a
// simulation is in flight (it counts as a loss until the backup replaces
// it, which steers the other threads to other leaves), expansion with the
// policy net and evaluation with the value net at the leaf, backup of the
// leaf's value with the sign alternating per ply. Children are the
// Netresult's points that pass the filter of UCTNode::create_children
// (UCTNode.cpp:149-158: not the ko square, not an own eye, not suicide),
// the 35 with the highest priors, or a pass when none does; the root's
// priors are the eight-symmetry average (AVERAGE_ALL). A position after two passes in a row is
// scored by FastBoard::final_mc_score.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <future>
#include <memory>
#include <thread>
#include <vector>

#include "input_planes.h"
#include "leela_amd.h"
#include "lz_ctx.h"
#include "lz_position.h"
#include "network.h"
#include "playout.h"

namespace lz {
namespace {

struct Node {
    int move = go::kPass;  // vertex that leads here (kPass for a pass)
    float prior = 1.0f;
    // Values are for the player who made `move` (the parent's side to
    // move), in [0, 1].
    std::atomic<int> visits{0};
    std::atomic<int> virtual_loss{0};
    std::atomic<double> value_sum{0.0};
    // 0 unexpanded, 1 being expanded by one thread, 2 expanded.
    std::atomic<int> state{0};
    std::vector<std::unique_ptr<Node>> children;
};

void add_value(std::atomic<double> &a, double v) {
    double cur = a.load(std::memory_order_relaxed);
    while (!a.compare_exchange_weak(cur, cur + v, std::memory_order_relaxed)) {
    }
}

// The leaf evaluator: priors over the children and a value for the side to
// move of `st`.
class Evaluator {
public:
    virtual ~Evaluator() = default;
    // Appends (vertex, prior) for the candidate moves of st; false on a
    // backend error.
    virtual bool expand(const go::GameState &st, bool at_root,
                        std::vector<std::pair<int, float>> &moves) = 0;
    virtual bool value(const go::GameState &st, uint32_t seed, float *v) = 0;
    // Both for one leaf; by default one after the other.
    virtual bool expand_value(const go::GameState &st, bool at_root,
                              uint32_t seed,
                              std::vector<std::pair<int, float>> &moves,
                              float *v) {
        return expand(st, at_root, moves) && value(st, seed, v);
    }
    // Children kept per expansion, by prior (0: all).
    virtual size_t max_children() const { return 0; }
};

bool candidate(const go::GameState &st, int v) {
    const go::Board &b = st.board;
    const int c = b.to_move();
    return b.at(v) == go::EMPTY && v != st.komove && !b.is_eye(c, v) &&
           !b.is_suicide(v, c);
}

class NetEvaluator : public Evaluator {
public:
    NetEvaluator(lz_ctx *ctx, int precision) : ctx_(ctx), precision_(precision) {}
    // A facade object per call: the threads share only the context (and
    // through it the leaf queue).
    // The root takes the eight-symmetry average (one batch of eight rows),
    // as netscore_children does at the root (UCTNode.cpp:103-106).
    bool expand(const go::GameState &st, bool at_root,
                std::vector<std::pair<int, float>> &moves) override {
        Network net(ctx_, precision_);
        const Network::Netresult r = net.get_scored_moves(
            st, at_root ? Network::AVERAGE_ALL : Network::DIRECT, 0);
        if (net.last_status() != LZ_OK) return false;
        for (const auto &sn : r)
            if (candidate(st, sn.second)) moves.emplace_back(sn.second, sn.first);
        return true;
    }
    bool value(const go::GameState &st, uint32_t, float *v) override {
        Network net(ctx_, precision_);
        *v = net.get_value(st, Network::DIRECT);
        return net.last_status() == LZ_OK;
    }
    // With the leaf queue running at this precision, a leaf's policy row
    // goes in asynchronously (lz_submit, as async_scored_moves does) while
    // this thread runs its value row through the synchronous call: one
    // queue round trip per leaf instead of two. The root keeps the
    // sequential path (its AVERAGE_ALL policy).
    bool expand_value(const go::GameState &st, bool at_root, uint32_t seed,
                      std::vector<std::pair<int, float>> &moves,
                      float *v) override {
        if (at_root || queue_precision(ctx_) != precision_)
            return Evaluator::expand_value(st, at_root, seed, moves, v);
        Pending p;
        p.st = &st;
        std::vector<uint8_t> planes((size_t)kPlanes * 361);
        gather_features_policy(st, planes.data());
        p.ladder.assign(planes.begin() + 25 * 361, planes.begin() + 26 * 361);
        std::future<int> done = p.done.get_future();
        if (lz_submit(ctx_, LZ_NET_POLICY, planes.data(), policy_done, &p) !=
            LZ_OK)
            return false;
        Network net(ctx_, precision_);
        *v = net.get_value(st, Network::DIRECT);
        const bool value_ok = net.last_status() == LZ_OK;
        // The queue runs every submitted row's callback, failed or not.
        if (done.get() != LZ_OK || !value_ok) return false;
        for (const auto &sn : p.result)
            if (candidate(st, sn.second)) moves.emplace_back(sn.second, sn.first);
        return true;
    }
    // The 35 best-scored children, as uct_select_child considers for a
    // node with net scores (UCTNode.cpp:551-552).
    size_t max_children() const override { return 35; }

private:
    struct Pending {
        const go::GameState *st = nullptr;
        std::vector<uint8_t> ladder;  // policy plane 25, board orientation
        Network::Netresult result;
        std::promise<int> done;
    };
    // On the queue's dispatcher thread: the Netresult of the row
    // (Network.cpp:820-829, 656-667), then the waiting thread is released.
    static void policy_done(void *user, int status, const float *,
                            const float *out) {
        Pending *p = static_cast<Pending *>(user);
        if (status == LZ_OK) {
            collect_scored_moves(*p->st, out, 0, p->result);
            prune_ladders(*p->st, p->ladder.data(), p->result);
        }
        p->done.set_value(status);  // p is not touched after this
    }

    lz_ctx *ctx_;
    int precision_;
};

// The same expansion and evaluation as NetEvaluator's synchronous path,
// with the nets evaluated by a caller's forward (lz_search_host): a search
// whose evaluations come from another implementation of the same nets, to
// compare trees with.
class ForwardEvaluator : public Evaluator {
public:
    explicit ForwardEvaluator(Network::ForwardFn fwd) : fwd_(std::move(fwd)) {}
    bool expand(const go::GameState &st, bool at_root,
                std::vector<std::pair<int, float>> &moves) override {
        Network net(fwd_);
        const Network::Netresult r = net.get_scored_moves(
            st, at_root ? Network::AVERAGE_ALL : Network::DIRECT, 0);
        if (net.last_status() != LZ_OK) return false;
        for (const auto &sn : r)
            if (candidate(st, sn.second)) moves.emplace_back(sn.second, sn.first);
        return true;
    }
    bool value(const go::GameState &st, uint32_t, float *v) override {
        Network net(fwd_);
        *v = net.get_value(st, Network::DIRECT);
        return net.last_status() == LZ_OK;
    }
    size_t max_children() const override { return 35; }

private:
    Network::ForwardFn fwd_;
};

// Host evaluator without a GPU: uniform priors over the candidates and the
// mean of `playouts` light playouts (Playout::mc_owner's role in the
// reference's search).
class PlayoutEvaluator : public Evaluator {
public:
    explicit PlayoutEvaluator(int playouts) : playouts_(playouts) {}
    bool expand(const go::GameState &st, bool,
                std::vector<std::pair<int, float>> &moves) override {
        const go::Board &b = st.board;
        for (int y = 0; y < b.side(); y++)
            for (int x = 0; x < b.side(); x++) {
                const int v = b.vertex(x, y);
                if (candidate(st, v)) moves.emplace_back(v, 0.0f);
            }
        for (auto &m : moves) m.second = 1.0f / (float)moves.size();
        return true;
    }
    bool value(const go::GameState &st, uint32_t seed, float *v) override {
        *v = go::mc_winrate(st, playouts_, seed);
        return true;
    }

private:
    int playouts_;
};

struct Search {
    Evaluator &eval;
    float c_puct;
    Node root;
    go::GameState root_state;
    std::atomic<long long> nodes{1}, policy_evals{0}, value_evals{0},
        collisions{0}, completed{0};
    std::atomic<int> error{LZ_OK};

    Search(Evaluator &e, float c, const go::GameState &st)
        : eval(e), c_puct(c), root_state(st) {}

    // Child of n maximising Q + U for n's side to move; one virtual loss is
    // added to it at once (two threads may still pick the same child).
    Node *select(Node &n) {
        const int parent_n = n.visits.load(std::memory_order_relaxed) +
                             n.virtual_loss.load(std::memory_order_relaxed);
        const float sq = std::sqrt((float)std::max(parent_n, 1));
        Node *best = nullptr;
        float best_score = -1e30f;
        for (auto &cp : n.children) {
            Node &c = *cp;
            const int vis = c.visits.load(std::memory_order_relaxed);
            const int vl = c.virtual_loss.load(std::memory_order_relaxed);
            const int cn = vis + vl;
            // An unvisited child counts as 0.45, a little below even (the
            // first-play value); virtual losses count as values of 0.
            const float q = cn > 0 ? (float)(c.value_sum.load(
                                         std::memory_order_relaxed) /
                                     cn)
                                   : 0.45f;
            const float u = c_puct * c.prior * sq / (1.0f + (float)cn);
            if (q + u > best_score) {
                best_score = q + u;
                best = &c;
            }
        }
        if (best) best->virtual_loss.fetch_add(1, std::memory_order_relaxed);
        return best;
    }

    // Expands `leaf` (state 1 held by this thread) and returns the value of
    // the leaf position for the player who moved into it.
    bool expand_and_evaluate(Node &leaf, const go::GameState &st,
                             uint32_t seed, float *v_in) {
        std::vector<std::pair<int, float>> moves;
        float v_to_move = 0.5f;
        if (!eval.expand_value(st, &leaf == &root, seed, moves, &v_to_move))
            return false;
        policy_evals.fetch_add(1, std::memory_order_relaxed);
        value_evals.fetch_add(1, std::memory_order_relaxed);
        if (moves.empty()) moves.emplace_back(go::kPass, 1.0f);
        std::sort(moves.begin(), moves.end(),
                  [](const std::pair<int, float> &a,
                     const std::pair<int, float> &b) {
                      return a.second > b.second;
                  });
        if (eval.max_children() && moves.size() > eval.max_children())
            moves.resize(eval.max_children());
        leaf.children.reserve(moves.size());
        for (const auto &m : moves) {
            auto c = std::make_unique<Node>();
            c->move = m.first;
            c->prior = m.second;
            leaf.children.push_back(std::move(c));
        }
        nodes.fetch_add((long long)moves.size(), std::memory_order_relaxed);
        leaf.state.store(2, std::memory_order_release);
        *v_in = 1.0f - v_to_move;
        return true;
    }

    // One simulation; false when it collided with an expansion in flight
    // (its virtual losses are already undone) or on an error.
    bool simulate(uint32_t seed) {
        go::GameState st = root_state;
        std::vector<Node *> path{&root};
        Node *n = &root;
        float v_in = 0.5f;  // value for the player who moved into `n`
        for (;;) {
            if (st.passes >= 2) {
                const float score = go::final_mc_score(st.board, st.komi);
                // The player who moved into n is the opponent of the side
                // to move; black wins with a positive score.
                const int mover = go::other(st.board.to_move());
                const float black = score > 0 ? 1.0f : score < 0 ? 0.0f : 0.5f;
                v_in = mover == go::BLACK ? black : 1.0f - black;
                break;
            }
            const int s = n->state.load(std::memory_order_acquire);
            if (s == 0) {
                int expect = 0;
                if (n->state.compare_exchange_strong(expect, 1,
                                                     std::memory_order_acq_rel)) {
                    if (!expand_and_evaluate(*n, st, seed, &v_in)) {
                        error.store(LZ_ERR_HIP);
                        n->state.store(0);
                        undo(path);
                        return false;
                    }
                    break;
                }
            }
            if (n->state.load(std::memory_order_acquire) != 2) {
                collisions.fetch_add(1, std::memory_order_relaxed);
                undo(path);
                return false;
            }
            Node *c = select(*n);
            path.push_back(c);
            st.play(c->move);
            n = c;
        }
        // Backup: v_in is for the player who moved into the leaf; one ply up
        // the perspective flips.
        float v = v_in;
        for (size_t i = path.size(); i-- > 0;) {
            Node *p = path[i];
            add_value(p->value_sum, v);
            p->visits.fetch_add(1, std::memory_order_relaxed);
            if (i > 0) p->virtual_loss.fetch_sub(1, std::memory_order_relaxed);
            v = 1.0f - v;
        }
        return true;
    }

    void undo(const std::vector<Node *> &path) {
        for (size_t i = 1; i < path.size(); i++)
            path[i]->virtual_loss.fetch_sub(1, std::memory_order_relaxed);
    }
};

void walk(const Node &n, long long *vl_left, long long *bad_sums) {
    *vl_left += n.virtual_loss.load();
    if (n.state.load() == 2 && !n.children.empty()) {
        long long sum = 0;
        for (const auto &c : n.children) {
            sum += c->visits.load();
            walk(*c, vl_left, bad_sums);
        }
        // The simulation that expanded n visited n but none of its children.
        if (sum != n.visits.load() - 1) (*bad_sums)++;
    }
}

}  // namespace
}  // namespace lz

namespace lz {
namespace {

// Runs `simulations` simulations of `threads` threads from root with ev and
// fills stats (queue counters from ctx when given).
int run_search(Evaluator &ev, lz_ctx *ctx, const lz_position *root,
               int threads, int simulations, float c_puct, uint32_t seed,
               lz_search_stats *stats) {
    Search search(ev, c_puct, root->st);

    int64_t q0[6] = {0, 0, 0, 0, 0, 0}, q1[6] = {0, 0, 0, 0, 0, 0};
    if (ctx) lz_queue_stats(ctx, q0);
    std::atomic<long long> next{0};
    const auto t0 = std::chrono::steady_clock::now();
    auto worker = [&](int tid) {
        for (;;) {
            const long long k = next.fetch_add(1, std::memory_order_relaxed);
            if (k >= simulations) return;
            // A simulation that collides is retried under a new seed.
            for (uint32_t attempt = 0;; attempt++) {
                if (search.error.load() != LZ_OK) return;
                const uint32_t s = seed + (uint32_t)k * 7919u + attempt * 104729u +
                                   (uint32_t)tid;
                if (search.simulate(s)) break;
                std::this_thread::yield();
            }
            search.completed.fetch_add(1, std::memory_order_relaxed);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; t++) pool.emplace_back(worker, t);
    worker(0);
    for (auto &th : pool) th.join();
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
            .count();
    if (ctx) lz_queue_stats(ctx, q1);
    if (search.error.load() != LZ_OK) return search.error.load();

    lz_search_stats &s = *stats;
    s = lz_search_stats{};
    s.seconds = secs;
    s.simulations = search.completed.load();
    s.policy_evals = search.policy_evals.load();
    s.value_evals = search.value_evals.load();
    s.collisions = search.collisions.load();
    s.nodes = search.nodes.load();
    s.queue_batches = q1[0] - q0[0];
    s.queue_positions = q1[1] - q0[1];
    s.root_visits = search.root.visits.load();
    s.root_value = s.root_visits
                       ? (float)(1.0 - search.root.value_sum.load() /
                                          s.root_visits)
                       : 0.5f;
    s.best_move = lz::go::kPass;
    int best = -1;
    long long child_sum = 0;
    for (const auto &c : search.root.children) {
        const int v = c->visits.load();
        child_sum += v;
        if (v > best) {
            best = v;
            s.best_move = c->move;
        }
    }
    s.best_visits = best < 0 ? 0 : best;
    s.root_children = (int32_t)search.root.children.size();
    long long vl = 0, bad = 0;
    walk(search.root, &vl, &bad);
    s.virtual_loss_left = vl;
    s.visit_sum_mismatches = bad;
    return LZ_OK;
}

}  // namespace
}  // namespace lz

extern "C" {

int lz_search(lz_ctx *ctx, int precision, const lz_position *root,
              int evaluator, int playouts, int threads, int simulations,
              float c_puct, uint32_t seed, lz_search_stats *stats) {
    if (!root || !stats || threads < 1 || threads > 1024 || simulations < 1 ||
        simulations > (1 << 22) || !(c_puct > 0.0f))
        return LZ_ERR_ARG;
    if (evaluator != LZ_EVAL_NETS && evaluator != LZ_EVAL_PLAYOUTS)
        return LZ_ERR_ARG;
    if (evaluator == LZ_EVAL_NETS && (!ctx || root->st.board.side() != 19))
        return LZ_ERR_ARG;
    if (evaluator == LZ_EVAL_PLAYOUTS && playouts < 1) return LZ_ERR_ARG;
    std::unique_ptr<lz::Evaluator> ev;
    if (evaluator == LZ_EVAL_NETS)
        ev = std::make_unique<lz::NetEvaluator>(ctx, precision);
    else
        ev = std::make_unique<lz::PlayoutEvaluator>(playouts);
    return lz::run_search(*ev, ctx, root, threads, simulations, c_puct, seed,
                          stats);
}

int lz_search_host(lz_host_forward_fn forward, void *user,
                   const lz_position *root, int threads, int simulations,
                   float c_puct, uint32_t seed, lz_search_stats *stats) {
    if (!forward || !root || !stats || threads < 1 || threads > 1024 ||
        simulations < 1 || simulations > (1 << 22) || !(c_puct > 0.0f) ||
        root->st.board.side() != 19)
        return LZ_ERR_ARG;
    lz::ForwardEvaluator ev(
        [forward, user](int net, int rows, const uint8_t *in, float *out) {
            return forward(user, net, rows, in, out);
        });
    return lz::run_search(ev, nullptr, root, threads, simulations, c_puct,
                          seed, stats);
}

}  // extern "C"
