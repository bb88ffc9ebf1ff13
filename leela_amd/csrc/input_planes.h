// input_planes.h — Leela's input planes and board symmetries.
#pragma once

#include <cstdint>

#include "go_board.h"

namespace lz {

constexpr int kPlanes = 32;

// Policy planes (Network::gather_features_policy, Network.cpp:883-1046):
// 0 empty, 1 side to move, 2 opponent, 3-7 own string liberties 1,2,3,4,5+,
// 8-12 the same for the opponent, 13-18 own liberties after playing here
// 1..5,6+, 19-24 the same for an opponent move here, 25 losing ladder
// escape, 26 winning ladder atari, 27 ko square, 28 last move, 29 the move
// before, 30 white stones when |komi| > 0.75, 31 third line.
// planes: uint8 [32][S*S], index y*S+x. ladder (optional): plane 25.
void gather_features_policy(const go::GameState &st, uint8_t *planes);

// Value planes (Network::gather_features_value, Network.cpp:1048-1201):
// 0-2 as above, 3-8 own liberties 1..5,6+, 9-14 opponent's, 15-20 / 21-26
// after-move liberties for us / them, 27 losing ladder, 28 winning ladder,
// 29 white stones when |komi| > 0.75, 30 third line, 31 ko square.
void gather_features_value(const go::GameState &st, uint8_t *planes);

// Network::rotate_nn_idx / rev_rotate_nn_idx (Network.cpp:1341-1379) on a
// side x side board.
int rotate_nn_idx(int idx, int symmetry, int side = 19);
int rev_rotate_nn_idx(int idx, int symmetry, int side = 19);

// in[c][h][w] = planes[c][rotate(h*S + w, symmetry)] (Network.cpp:765-773).
void rotate_planes(const uint8_t *planes, int symmetry, int side,
                   uint8_t *out);

}  // namespace lz
