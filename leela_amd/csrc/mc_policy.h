// mc_policy.h — host side of the reference's Monte Carlo playout policy
// (FastState::play_random_move; the policy itself is mc_core.h): its
// tables and the playouts of a game state, for the surrogate's teacher
// labels.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "go_board.h"
#include "lz_random.h"
#include "mc_core.h"

namespace lz {
namespace mc {

// The policy's tables: the reference's 3x3 pattern set hashed into
// pattern-weight indices per colour (Matcher, Matcher.cpp:39-91), and the
// feature and pattern weights multiplied by their SL-trained factors
// (Matcher::rescale_policy_weights, cfg_mc_softmax = 1).
struct Tables {
    std::vector<int32_t> g;             // perfect-hash displacements
    std::vector<uint16_t> index[2];     // [colour][hash] -> pattern weight
    std::vector<float> feature_weight;  // kNumFeatures
    std::vector<float> pattern_weight;
    TablesView view() const;
};

// Loads the tables from a data file written by
// leela_amd/data/make_mc_policy.py (nullptr and a message on failure).
// With path == nullptr, the file next to the library
// (data/mc_policy.bin) or the one $LZ_MC_POLICY names; loaded once.
const Tables *tables(const char *path, std::string *err);

// The playout state of a game state (stones in raster order).
void setup(PlayState &ps, const go::GameState &st);

// One playout with the pattern policy from st (PlayState::playout).
float pattern_playout(const go::GameState &st, Rng &rng, const Tables &t);

// Playout::mc_owner with the pattern policy: mean outcome of n playouts for
// the side to move (win 1, tie 0.5), playout j drawing from
// Rng(playout_seed(seed, j)) with two outputs discarded.
float pattern_winrate(const go::GameState &st, int n, uint32_t seed,
                      const Tables &t);

}  // namespace mc
}  // namespace lz
