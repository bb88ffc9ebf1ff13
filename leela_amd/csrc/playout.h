// playout.h — Monte Carlo playouts: the teacher (and fallback) of the
// learned leaf-value surrogate.
#pragma once

#include <cstdint>

#include "go_board.h"
#include "lz_random.h"

namespace lz {
namespace go {

// FastBoard::final_mc_score (FastBoard.cpp:420-443): black stones plus
// empty points whose four neighbours are all black (or edge), minus the
// same for white, minus komi.
float final_mc_score(const Board &b, float komi);

// One playout from `st` (played in place): Playout::run's loop
// (Playout.cpp:51-113) with mc_owner's settings — stop at 4 passes in a
// row or at move number 2 * side^2 — where each move is uniform over the
// empty points that are not the ko square, not an own eye and not suicide
// (pass when there is none). The reference draws moves from its pattern
// policy (FastState::play_random_move, FastState.cpp:168-255); this light
// policy is the surrogate's teacher here. Returns final_mc_score with the state's
// komi.
float playout(GameState &st, Rng &rng);

// Seed of playout j of a position whose playouts are seeded `seed`: the
// MurmurHash3 32-bit finalizer (fmix32) of seed ^ (j * 0x9E3779B9), so
// every playout has its own stream and playouts can run in parallel.
uint32_t playout_seed(uint32_t seed, uint32_t j);

// Playout::mc_owner (Playout.cpp:174-217): mean outcome of n playouts for
// the side to move (win 1, tie 0.5, loss 0); playout j draws from
// Rng(playout_seed(seed, j)) with its first two outputs discarded.
float mc_winrate(const GameState &st, int n, uint32_t seed);

}  // namespace go
}  // namespace lz
