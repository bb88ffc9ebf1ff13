// network.h — the Network facade over the C ABI: the interface UCTSearch
// calls (Network.h:23-88), with the same Netresult semantics, evaluated by
// the MI355X backend.
#pragma once

#include <atomic>
#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "go_board.h"
#include "leela_amd.h"

namespace lz {

class Network {
public:
    enum Ensemble { DIRECT, RANDOM_ROTATION, AVERAGE_ALL };
    using scored_node = std::pair<float, int>;  // (probability, vertex)
    using Netresult = std::vector<scored_node>;
    using ScoringCallback = std::function<void(const Netresult &)>;
    // A forward of `rows` rotated plane stacks into softmax probabilities
    // (policy, 361 per row) or win rates (value, one per row); LZ_OK or a
    // status. Lets the facade run on another evaluator of the same nets.
    using ForwardFn = std::function<int(int net, int rows,
                                        const uint8_t *in, float *out)>;

    static constexpr int POLICY_CHANNELS = 32;
    static constexpr int VALUE_CHANNELS = 32;
    static constexpr int MAX_CHANNELS = 128;
    static constexpr int MAX_VALUE_CHANNELS = 64;

    // ctx must be finalized; precision is an LZ_PREC_* form.
    Network(lz_ctx *ctx, int precision) : ctx_(ctx), precision_(precision) {}
    // The facade over a caller's forward (no context: the synchronous
    // calls only).
    explicit Network(ForwardFn forward)
        : ctx_(nullptr), precision_(-1), forward_(std::move(forward)) {}

    // Network::get_scored_moves (Network.cpp:624-674): probabilities of the
    // empty points in row-major order, losing-ladder escapes set to 0, not
    // renormalised; empty for boards other than 19x19. AVERAGE_ALL runs the
    // eight symmetries as one batch of eight. Returns an empty result on a
    // backend error (status in last_status()).
    Netresult get_scored_moves(const go::GameState &state, Ensemble ensemble,
                               int rotation = -1);

    // Network::get_value (Network.cpp:590-622): win probability for the
    // side to move; 0.5 for boards other than 19x19.
    float get_value(const go::GameState &state, Ensemble ensemble);

    // Network::async_scored_moves (Network.cpp:536-587) through the leaf
    // queue (lz_queue_start must have been called on the context): cb
    // receives the Netresult on the queue's dispatcher thread, as
    // forward_cb does after the OpenCL event. DIRECT or RANDOM_ROTATION.
    int async_scored_moves(const go::GameState &state, ScoringCallback cb,
                           Ensemble ensemble, int rotation = -1);

    // OpenCL::thread_can_issue / join_outstanding_cb (OpenCL.cpp:446-454,
    // 571-577) over the shared queue.
    bool can_issue(int max_outstanding) const;
    void join_outstanding_cb();

    static int rotate_nn_idx(int vertex, int symmetry);
    static int rev_rotate_nn_idx(int vertex, int symmetry);
    // Network::softmax (Network.cpp:450-469), host side.
    static void softmax(const std::vector<float> &input,
                        std::vector<float> &output, float temperature);

    int last_status() const { return status_; }

    // Network::show_heatmap (Network.cpp:834-881), returned as text: the
    // 19 rows of int(probability x 1000) as "%3d " per point, top row
    // (y = 18) first, empty for occupied points; with topmoves, then the
    // moves in descending order as "%1.3f (D4)" until they sum to 0.85 or
    // one falls below 0.01.
    static std::string show_heatmap(const go::GameState &state,
                                    const Netresult &result, bool topmoves);

    // Network::get_backend (Network.cpp:1534-1555): the device the nets
    // run on, as the OpenCL build reports its device name.
    std::string get_backend() const;

    // Network::benchmark (Network.cpp:147-199): `threads` threads, each on
    // its own copy of the state, run ceil(2000 / threads) policy
    // evaluations and then ceil(10000 / threads) value evaluations, all
    // RANDOM_ROTATION; the rates are 2000 and 10000 over the elapsed
    // seconds, as the reference prints them.
    void benchmark(const go::GameState &state, int threads, float *policy_ps,
                   float *value_ps);

private:
    // Forward of `rows` planes for the synchronous calls: through the leaf
    // queue when one runs at this precision (concurrent callers are
    // batched together, as the reference's per-thread OpenCL queues are
    // not), else one lz_forward.
    int forward_rows(int net, int rows, const uint8_t *in, float *out);

    lz_ctx *ctx_;
    int precision_;
    ForwardFn forward_;
    int status_ = LZ_OK;
};

// Builds a Netresult from the policy output of one symmetry (Network.cpp:
// 820-829): the output pixel of board point idx is rev_rotate(idx).
void collect_scored_moves(const go::GameState &state, const float *probs,
                          int rotation, Network::Netresult &out);
// Zeroes losing-ladder escapes (Network.cpp:656-667); ladder is the policy
// plane 25 in board orientation.
void prune_ladders(const go::GameState &state, const uint8_t *ladder,
                   Network::Netresult &result);

}  // namespace lz
