// go_board.h — Go board and game state feeding Leela's feature planes.
//
// The nets' inputs are defined by board queries of the reference engine
// (FastBoard / FullBoard / FastState): string liberties, liberties after a
// move, the saving-size and ladder readers, eye and suicide tests. This
// board answers those queries with identical results. Representation:
// a padded (side+2)^2 grid of squares; stones of a string form a circular
// list and share a representative whose entry holds the string's exact
// liberty count and stone count. Boards are small plain values (a few KB),
// so readers that try moves work on copies.
#pragma once

#include <array>
#include <cstdint>
#include <utility>

namespace lz {
namespace go {

enum Color : uint8_t { BLACK = 0, WHITE = 1, EMPTY = 2, OFFBOARD = 3 };

constexpr int kMaxSide = 19;
constexpr int kMaxWidth = kMaxSide + 2;
constexpr int kMaxSquares = kMaxWidth * kMaxWidth;
constexpr int kPass = -1;    // FastBoard::PASS
constexpr int kResign = -2;  // FastBoard::RESIGN

inline Color other(int c) { return static_cast<Color>(c ^ 1); }

class Board {
public:
    void reset(int side);

    int side() const { return side_; }
    int width() const { return side_ + 2; }
    int vertex(int x, int y) const { return (y + 1) * width() + (x + 1); }
    std::pair<int, int> xy(int v) const {
        return {v % width() - 1, v / width() - 1};
    }
    Color at(int v) const { return static_cast<Color>(sq_[v]); }
    int to_move() const { return to_move_; }
    void set_to_move(int c) { to_move_ = c; }
    int empty_count() const { return empties_; }

    // Liberties of the string through v (FastBoard::count_rliberties).
    int string_liberties(int v) const { return libs_[head_[v]]; }
    int string_size(int v) const { return stones_[head_[v]]; }
    // Empty 4-neighbours (FastBoard::count_pliberties).
    int empty_neighbours(int v) const;
    // 4-neighbours of colour c, the edge counting as both colours
    // (FastBoard::count_neighbours).
    int colour_neighbours(int c, int v) const;

    bool is_suicide(int v, int c) const;
    bool is_eye(int c, int v) const;
    bool no_eye_fill(int v) const { return !is_eye(to_move_, v); }
    bool self_atari(int c, int v) const;
    int saving_size(int c, int v) const;
    int minimum_enemy_liberties(int c, int v) const;
    // (liberties of c's string after c plays v, same for the opponent);
    // 0 when the move is suicide (FastBoard::after_liberties).
    std::pair<int, int> after_liberties(int c, int v) const;
    int after_liberties_colour(int c, int v) const;
    bool check_losing_ladder(int c, int v, int branching = 0) const;
    bool check_winning_ladder(int c, int v) const;
    // The single liberty of a string in atari, 0 if it has more.
    int atari_liberty(int v) const;

    // Game move: captures, removes a suicided string, returns the simple
    // ko square or -1 (FullBoard::update_board).
    int play(int c, int v, bool *capture);
    // Reader move (FastBoard::update_board_fast): as play(), except that a
    // stone played into an opponent eye is not removed when it captures
    // nothing. Returns the ko square for eye plays that capture one stone.
    int play_fast(int c, int v, bool *capture);

    uint64_t stones_hash() const;  // FNV-1a over the squares (tests)

    // Representative stone of the string through v, and the 4 neighbour
    // offsets of the padded grid.
    int head(int v) const { return head_[v]; }
    // The padded grid's squares (Color values) and string representatives
    // (meaningful at stones only), width() * width() entries each.
    const uint8_t *squares() const { return sq_.data(); }
    const uint16_t *heads() const { return head_.data(); }
    int dir(int k) const { return dir_[k]; }

private:
    int place(int c, int v, bool *capture, bool game_move);
    int remove_string(int v);
    void recount_liberties(int head);
    bool kill_or_connect(int c, int v) const;
    bool can_kill_neighbours(int v) const;
    // Up to N distinct liberties of the string through v appended to libs
    // (stops once N are known), as add_string_liberties<N>.
    template <int N>
    void collect_liberties(int v, std::array<int, N> &libs, int &cnt) const;

    int side_ = 0;
    int to_move_ = BLACK;
    int empties_ = 0;
    int dir_[4] = {0, 0, 0, 0};
    std::array<uint8_t, kMaxSquares> sq_{};
    std::array<uint16_t, kMaxSquares> head_{};    // representative stone
    std::array<uint16_t, kMaxSquares> next_{};    // circular string list
    std::array<uint16_t, kMaxSquares + 1> libs_{};    // by representative
    std::array<uint16_t, kMaxSquares + 1> stones_{};  // by representative
};

// Bitboards of one board over padded vertex indices (441 bits in 7 words):
// the empty points and every string's stones (indexed by representative).
// They answer after_liberties without trial moves on board copies: after c
// plays a legal, non-suicide v, the string through v is M = {v} plus the
// adjacent c strings, the opponent strings whose only liberty was v are
// captured (set C), and the string's liberties are the 4-neighbours of M
// among the empty points E - {v} + C. Borders are never empty, so plain
// shifts by 1 and by the grid width give the neighbours. Without a capture
// that is the union of the merged strings' liberty sets and v's empty
// neighbours, minus v.
class BoardBits {
public:
    static constexpr int kWords = (kMaxSquares + 63) / 64;
    struct Set {
        uint64_t w[kWords];
    };

    explicit BoardBits(const Board &b);
    // Same value as Board::after_liberties(c, v).
    std::pair<int, int> after_liberties(int c, int v) const;

private:
    int after_liberties_colour(int c, int v) const;

    const Board &b_;
    Set empty_{};
    std::array<Set, kMaxSquares> strings_;  // stones, by representative (live heads only)
    std::array<Set, kMaxSquares> libs_;     // liberties, by representative (live heads only)
};

// FastState counterpart: the board plus move history, ko and komi.
struct GameState {
    Board board;
    float komi = 7.5f;
    int komove = 0;       // simple-ko square, -1 or 0 when none
    int last_move = 0;    // 0 before the first move, kPass after a pass
    int prev_last = 0;
    int passes = 0;
    int movenum = 0;
    bool last_was_capture = false;  // the last move captured stones
    int handicap = 0;     // FastState::m_handicap: added to komi in scores

    // The komi that scores use (FastState.cpp:395-401: komi + handicap).
    float score_komi() const { return komi + (float)handicap; }

    void init(int side, float k);
    int to_move() const { return board.to_move(); }
    // FastState::play_move(color, vertex) semantics (kPass for a pass).
    void play(int vertex);
};

// Position generator: start empty (komi 7.5, black to move), seed the
// reference RNG with `seed`, discard its first two outputs, draw the move
// count m uniformly from 0..max_moves, then play m moves, each uniform
// among empty points that are not the ko square, not an own eye and not
// suicide (the candidate filter of UCTNode::create_children); pass when
// there is none, stop after two passes in a row.
void random_position(GameState &st, int side, uint32_t seed, int max_moves);

}  // namespace go
}  // namespace lz
