// mc_policy.h — the reference's Monte Carlo playout policy (pattern and
// tactical move generation, FastState::play_random_move) on a playout board
// that keeps FastBoard's playout state, for the surrogate's teacher labels.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "go_board.h"
#include "lz_random.h"

namespace lz {
namespace mc {

// Feature flags of a candidate move (MCPolicy.h:14-36); each multiplies
// the move's score by its feature weight.
enum Flag {
    kPass = 0, kNakade, kPattern, kPatternSA, kSA, kTooBigSA, kForceSA,
    kForceBigSA, kRandom, kSavingSA, kSaving1, kSaving2, kSaving3P,
    kSaving2Lib, kSaving3Lib, kSavingKill, kCapture1, kCapture2, kCapture3P,
    kSuicide, kSemeai2, kSemeai3, kNumFeatures
};

// The policy's tables: the reference's 3x3 pattern set hashed into
// pattern-weight indices per colour (Matcher, Matcher.cpp:39-91), and the
// feature and pattern weights multiplied by their SL-trained factors
// (Matcher::rescale_policy_weights, cfg_mc_softmax = 1).
struct Tables {
    static constexpr int kGSize = 8192;
    static constexpr int kVSize = 1 << 18;
    std::vector<int32_t> g;             // perfect-hash displacements
    std::vector<uint16_t> index[2];     // [colour][hash] -> pattern weight
    std::array<float, kNumFeatures> feature_weight{};
    std::vector<float> pattern_weight;
    // Pattern-weight index of a 24-bit augmented 3x3 pattern for colour c
    // to move (Matcher::matches).
    int matches(int c, uint32_t pattern) const;
};

// Loads the tables from a data file written by
// leela_amd/data/make_mc_policy.py (nullptr and a message on failure).
// With path == nullptr, the file next to the library
// (data/mc_policy.bin) or the one $LZ_MC_POLICY names; loaded once.
const Tables *tables(const char *path, std::string *err);

// A candidate move with its features (MovewFeatures, MCPolicy.h:40-120).
struct Move {
    int sq;
    int flags = 0;
    float score = 1.0f;
    Move(int v, int flag, const Tables &t) : sq(v) { add(flag, t); }
    void add(int flag, const Tables &t) {
        flags |= 1 << flag;
        score *= t.feature_weight[flag];
    }
    bool has(int flag) const { return (flags >> flag) & 1; }
};

// FastBoard's playout state: squares, packed neighbour counts, strings as
// circular lists with a parent and exact liberty and stone counts, the
// empty-point list, and the critical points found by the last move
// generation. Same updates, in the same order, as FastBoard's
// (FastBoard.cpp:130-189, 266-343, 725-875): the order of the empty list
// decides which point a random draw lands on.
class PlayBoard {
public:
    static constexpr int kMaxSq = go::kMaxSquares;
    static constexpr int kNone = kMaxSq;  // parent of empty / border points

    void reset(int size);
    // The position's stones placed in raster order (a legal position: no
    // string loses its last liberty on the way).
    void setup(const go::Board &b);

    int size() const { return size_; }
    int at(int v) const { return sq_[v]; }
    int vertex(int x, int y) const { return (y + 1) * (size_ + 2) + (x + 1); }
    int empty_count() const { return empty_cnt_; }
    int empty_at(int i) const { return empty_[i]; }

    // FastBoard::update_board_fast: returns the ko square or -1.
    int play_fast(int color, int v, bool *capture);
    // FastBoard::final_mc_score.
    float final_mc_score(float komi) const;

    // Move generators and readers (FastBoard.cpp:1179-2643).
    bool is_eye(int color, int v) const;
    bool no_eye_fill(int v, int to_move) const { return !is_eye(to_move, v); }
    bool fast_ss_suicide(int color, int v) const;
    bool is_suicide(int v, int color);
    bool self_atari(int color, int v) const;
    int enemy_atari_size(int color, int v) const;
    int string_size(int v) const { return stones_[parent_[v]]; }
    uint32_t pattern_augment(int v) const;
    void add_global_captures(int color, int komove, std::vector<Move> &m,
                             const Tables &t);
    void save_critical_neighbours(int color, int v, int komove,
                                  std::vector<Move> &m, const Tables &t);
    void add_semeai_moves(int color, int last, int komove,
                          std::vector<Move> &m, const Tables &t);
    void add_near_nakade_moves(int color, int v, int komove,
                               std::vector<Move> &m, const Tables &t);
    void add_pattern_moves(int color, int v, int komove,
                           std::vector<Move> &m, const Tables &t);

private:
    int count(int c, int v) const { return (nbr_[v] >> (4 * c)) & 7; }
    int libs_at(int v) const { return libs_[parent_[v]]; }
    void add_neighbour(int v, int color);
    void remove_neighbour(int v, int color);
    int remove_string(int v);
    void merge_strings(int ip, int aip);
    int place_eye(int color, int v);
    void take_empty(int v);
    int in_atari(int v) const;
    bool kill_or_connect(int color, int v) const;
    bool kill_neighbours(int v, int komove, std::vector<Move> &m,
                         const Tables &t);
    void try_capture(int color, int v, std::vector<Move> &m,
                     const Tables &t);
    void check_nakade(int color, int v, int komove, std::vector<Move> &m,
                      const Tables &t);
    template <int N>
    void add_string_liberties(int v, std::array<int, N> &libs,
                              size_t &cnt) const;
    template <int N>
    bool saveable_string(int parent, std::array<int, N> &libs,
                         size_t &cnt) const;

    int size_ = 0;
    int dirs_[4] = {0, 0, 0, 0};
    int extradirs_[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    std::array<uint8_t, kMaxSq> sq_{};
    std::array<uint16_t, kMaxSq> nbr_{};
    std::array<int, kMaxSq + 1> parent_{}, next_{}, libs_{}, stones_{};
    std::array<int, kMaxSq> empty_{}, empty_idx_{};
    int empty_cnt_ = 0;
    int total_[2] = {0, 0};
    std::vector<int> critical_;
};

// FastState's playout state around a PlayBoard.
struct PlayState {
    PlayBoard board;
    int to_move = go::BLACK;
    int komove = 0;
    int last_move = 0;
    bool last_was_capture = false;
    int passes = 0;
    int movenum = 0;
    float komi = 7.5f;
    std::vector<Move> moves;  // scratch of the move generator

    void setup(const go::GameState &st);
    // FastState::play_move_fast.
    int play_fast(int v);
    // FastState::play_random_move (FastState.cpp:168-255): plays and
    // returns the move drawn from the policy.
    int play_random_move(Rng &rng, const Tables &t);

private:
    bool try_move(int color, int v) const;
    int walk_empty_list(int color, Rng &rng);
    void flag_move(Move &m, int color, const Tables &t);
};

// One playout with the pattern policy from st: Playout::run with
// mc_owner's settings (four passes or move 2 * side^2); returns
// final_mc_score with the position's komi.
float pattern_playout(const go::GameState &st, Rng &rng, const Tables &t);

// Playout::mc_owner with the pattern policy: mean outcome of n playouts for
// the side to move (win 1, tie 0.5), playout j drawing from
// Rng(go::playout_seed(seed, j)) with two outputs discarded.
float pattern_winrate(const go::GameState &st, int n, uint32_t seed,
                      const Tables &t);

}  // namespace mc
}  // namespace lz
