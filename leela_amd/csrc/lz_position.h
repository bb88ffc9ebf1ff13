// lz_position.h — the game state behind the C ABI's opaque lz_position.
#pragma once

#include "go_board.h"

struct lz_position {
    lz::go::GameState st;
};
