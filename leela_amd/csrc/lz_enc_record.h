// lz_enc_record.h — one board as the device kernels read it (feature
// encoder, playouts).
#pragma once

#include <cstdint>

#include "go_board.h"

namespace lz {

// The grid is the padded (side+2)^2 board of the reference (FastBoard
// vertex numbering), row-major.
constexpr int kEncGrid = 448;  // >= 21 * 21, a multiple of 64
// The device encoder keeps the stones of a board's first kEncStrCap strings
// in LDS and the rest (a board has at most 361 strings) in an overflow area
// of kEncStrOvf sets of kEncStrWords 64-bit words per board in global
// memory, which the caller provides.
constexpr int kEncStrCap = 128;
constexpr int kEncStrOvf = 361 - kEncStrCap;
constexpr int kEncStrWords = 7;  // kEncGrid bits

struct alignas(64) EncRecord {
    uint8_t sq[kEncGrid];      // 0 black, 1 white, 2 empty, 3 off-board
    uint16_t head[kEncGrid];   // string representative (read at stones only)
    int32_t side;              // 5..19
    int32_t to_move;           // 0 black, 1 white
    int32_t komove;            // simple-ko vertex, <= 0 when none
    int32_t last_move;         // vertex, <= 0 when none or a pass
    int32_t prev_last;         // vertex, <= 0 when none or a pass
    int32_t white_has_komi;    // |komi| > 0.75
    int64_t out_offset;        // byte offset of this board's planes
    int32_t passes;            // passes in a row (playouts stop at 4)
    int32_t movenum;           // moves played (playouts stop at 2 side^2)
    float komi;                // scoring komi: komi + handicap
    int32_t last_was_capture;  // the last move captured stones
    uint8_t pad_[16];
};
static_assert(sizeof(EncRecord) == 1408, "record layout");

// Fills r from a game state; the board's planes go to out + offset.
void pack_record(const go::GameState &st, int64_t offset, EncRecord *r);

}  // namespace lz
