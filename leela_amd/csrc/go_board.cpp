// go_board.cpp — board moves and the tactical readers behind Leela's feature
// planes. Each reader follows the decision steps of the reference function
// named in its comment, so its answers match on every position.
#include "go_board.h"

#include <algorithm>
#include <vector>

#include "lz_random.h"

namespace lz {
namespace go {
namespace {

// Representative of empty and off-board squares: its liberty count is a
// large constant, so "string with few liberties" tests never fire on it
// (the reference gives those squares a sentinel parent with 16384
// liberties, FastBoard.cpp:186-188).
constexpr uint16_t kNone = kMaxSquares;
constexpr uint16_t kManyLibs = 16384;

}  // namespace

void Board::reset(int side) {
    side_ = side;
    to_move_ = BLACK;
    const int W = side + 2;
    dir_[0] = -W;
    dir_[1] = 1;
    dir_[2] = W;
    dir_[3] = -1;
    sq_.fill(OFFBOARD);
    head_.fill(kNone);
    next_.fill(kNone);
    libs_.fill(0);
    stones_.fill(0);
    libs_[kNone] = kManyLibs;
    empties_ = 0;
    for (int y = 0; y < side; y++)
        for (int x = 0; x < side; x++) {
            sq_[vertex(x, y)] = EMPTY;
            empties_++;
        }
}

int Board::empty_neighbours(int v) const {
    int n = 0;
    for (int k = 0; k < 4; k++) n += sq_[v + dir_[k]] == EMPTY;
    return n;
}

int Board::colour_neighbours(int c, int v) const {
    int n = 0;
    for (int k = 0; k < 4; k++) {
        const int s = sq_[v + dir_[k]];
        n += s == c || (s == OFFBOARD && c != EMPTY);
    }
    return n;
}

// FastBoard::is_suicide (FastBoard.cpp:191-240).
bool Board::is_suicide(int v, int c) const {
    if (empty_neighbours(v)) return false;
    bool connecting = false;
    for (int k = 0; k < 4; k++) {
        const int a = v + dir_[k];
        const int l = libs_[head_[a]];
        if (sq_[a] == c) {
            if (l > 1) return false;  // joins a string with another liberty
            connecting = true;
        } else if (l <= 1) {
            return false;  // captures a neighbour
        }
    }
    // After the stone lands every distinct neighbouring string loses v.
    bool opps_live = true, ours_die = true;
    uint16_t seen[4];
    int nseen = 0;
    int after[4];
    for (int k = 0; k < 4; k++) {
        const uint16_t h = head_[v + dir_[k]];
        bool dup = false;
        for (int i = 0; i < nseen; i++) dup |= seen[i] == h;
        if (!dup) seen[nseen++] = h;
        after[k] = libs_[h];
    }
    for (int i = 0; i < nseen; i++)
        for (int k = 0; k < 4; k++)
            if (head_[v + dir_[k]] == seen[i]) after[k] = libs_[seen[i]] - 1;
    for (int k = 0; k < 4; k++) {
        const int a = v + dir_[k];
        if (after[k] == 0 && sq_[a] != c)
            opps_live = false;
        else if (after[k] != 0 && sq_[a] == c)
            ours_die = false;
    }
    return connecting ? (opps_live && ours_die) : opps_live;
}

// FastBoard::is_eye (FastBoard.cpp:876-912): all four neighbours own (or
// edge), and the diagonals hold at most one enemy stone in the centre, none
// on the edge.
bool Board::is_eye(int c, int v) const {
    if (colour_neighbours(c, v) < 4) return false;
    const int W = width();
    const int diag[4] = {v - 1 - W, v + 1 - W, v - 1 + W, v + 1 + W};
    int enemy = 0, edge = 0;
    for (int d : diag) {
        enemy += sq_[d] == other(c);
        edge += sq_[d] == OFFBOARD;
    }
    return edge == 0 ? enemy <= 1 : enemy == 0;
}

template <int N>
void Board::collect_liberties(int v, std::array<int, N> &libs,
                              int &cnt) const {
    int s = v;
    do {
        for (int k = 0; k < 4; k++) {
            const int a = s + dir_[k];
            if (sq_[a] != EMPTY) continue;
            bool found = false;
            for (int i = 0; i < cnt; i++) found |= libs[i] == a;
            if (!found) {
                libs[cnt++] = a;
                if (cnt >= N) return;
            }
        }
        s = next_[s];
    } while (s != v);
}

// FastBoard::kill_or_connect (FastBoard.cpp:1342-1355).
bool Board::kill_or_connect(int c, int v) const {
    for (int k = 0; k < 4; k++) {
        const int a = v + dir_[k];
        const int l = libs_[head_[a]];
        if ((l <= 1 && sq_[a] == other(c)) || (l >= 3 && sq_[a] == c))
            return true;
    }
    return false;
}

// FastBoard::self_atari (FastBoard.cpp:1401-1471): the stone's string would
// have at most two liberties counting v itself, i.e. one after the move.
bool Board::self_atari(int c, int v) const {
    if (empty_neighbours(v) >= 2) return false;
    if (kill_or_connect(c, v)) return false;
    if (colour_neighbours(c, v) == 0) return true;
    std::array<int, 3> libs;
    int cnt = 0;
    libs[cnt++] = v;
    for (int k = 0; k < 4; k++) {
        const int a = v + dir_[k];
        if (sq_[a] == EMPTY) {
            bool found = false;
            for (int i = 0; i < cnt; i++) found |= libs[i] == a;
            if (!found) {
                if (cnt > 1) return false;
                libs[cnt++] = a;
            }
        } else if (sq_[a] == c && libs_[head_[a]] > 1) {
            collect_liberties<3>(a, libs, cnt);
            if (cnt > 2) return false;
        }
    }
    return true;
}

int Board::atari_liberty(int v) const {
    if (libs_[head_[v]] != 1) return 0;
    int s = v;
    do {
        for (int k = 0; k < 4; k++)
            if (sq_[s + dir_[k]] == EMPTY) return s + dir_[k];
        s = next_[s];
    } while (s != v);
    return 0;
}

// FastBoard::saving_size (FastBoard.cpp:1273-1292).
int Board::saving_size(int c, int v) const {
    for (int k = 0; k < 4; k++) {
        const int a = v + dir_[k];
        if (sq_[a] == c && libs_[head_[a]] <= 1) {
            if (!self_atari(c, atari_liberty(a))) return stones_[head_[a]];
        }
    }
    return 0;
}

// FastBoard::minimum_elib_count (FastBoard.cpp:2429-2443).
int Board::minimum_enemy_liberties(int c, int v) const {
    int m = 100;
    for (int k = 0; k < 4; k++) {
        const int a = v + dir_[k];
        if (sq_[a] == other(c) && libs_[head_[a]] < m) m = libs_[head_[a]];
    }
    return m;
}

// FastBoard::can_kill_neighbours (FastBoard.cpp:2587-2613).
bool Board::can_kill_neighbours(int v) const {
    const int kc = other(sq_[v]);
    int s = v;
    do {
        for (int k = 0; k < 4; k++) {
            const int a = s + dir_[k];
            if (sq_[a] == kc && libs_[head_[a]] <= 1) return true;
        }
        s = next_[s];
    } while (s != v);
    return false;
}

// FastBoard::after_liberties_color / after_liberties (FastBoard.cpp:
// 2511-2528).
int Board::after_liberties_colour(int c, int v) const {
    if (is_suicide(v, c)) return 0;
    Board t = *this;
    bool cap;
    t.play_fast(c, v, &cap);
    return t.string_liberties(v);
}

std::pair<int, int> Board::after_liberties(int c, int v) const {
    return {after_liberties_colour(c, v), after_liberties_colour(other(c), v)};
}

namespace {

using Set = BoardBits::Set;
constexpr int kWords = BoardBits::kWords;

inline void set_bit(Set &s, int i) { s.w[i >> 6] |= 1ull << (i & 63); }

inline void or_into(Set &d, const Set &s) {
    for (int i = 0; i < kWords; i++) d.w[i] |= s.w[i];
}

// d = s shifted towards higher (k > 0) or lower (k < 0) indices by |k| < 64.
inline void or_shifted(Set &d, const Set &s, int k) {
    if (k > 0) {
        uint64_t carry = 0;
        for (int i = 0; i < kWords; i++) {
            d.w[i] |= (s.w[i] << k) | carry;
            carry = s.w[i] >> (64 - k);
        }
    } else {
        const int m = -k;
        uint64_t carry = 0;
        for (int i = kWords - 1; i >= 0; i--) {
            d.w[i] |= (s.w[i] >> m) | carry;
            carry = s.w[i] << (64 - m);
        }
    }
}

}  // namespace

namespace {

// The 4-neighbours of s (s itself excluded only where it has no neighbour
// in s).
inline Set neighbours(const Set &s, int W) {
    Set n{};
    or_shifted(n, s, 1);
    or_shifted(n, s, -1);
    or_shifted(n, s, W);
    or_shifted(n, s, -W);
    return n;
}

inline int count_and(const Set &a, const Set &b) {
    int n = 0;
    for (int i = 0; i < kWords; i++) n += __builtin_popcountll(a.w[i] & b.w[i]);
    return n;
}

}  // namespace

BoardBits::BoardBits(const Board &b) : b_(b) {
    const int W = b.width();
    int heads[kMaxSquares];
    int nh = 0;
    for (int v = W; v < W * (W - 1); v++) {
        const int c = b.at(v);
        if (c == EMPTY) {
            set_bit(empty_, v);
        } else if (c == BLACK || c == WHITE) {
            const int h = b.head(v);
            if (h == v) {
                strings_[h] = Set{};
                heads[nh++] = h;
            }
        }
    }
    for (int v = W; v < W * (W - 1); v++) {
        const int c = b.at(v);
        if (c == BLACK || c == WHITE) set_bit(strings_[b.head(v)], v);
    }
    for (int i = 0; i < nh; i++) {
        const Set n = neighbours(strings_[heads[i]], W);
        Set &l = libs_[heads[i]];
        for (int k = 0; k < kWords; k++) l.w[k] = n.w[k] & empty_.w[k];
    }
}

int BoardBits::after_liberties_colour(int c, int v) const {
    if (b_.is_suicide(v, c)) return 0;
    int own[4], cap[4];
    int nown = 0, ncap = 0;
    for (int k = 0; k < 4; k++) {
        const int a = v + b_.dir(k);
        const int sc = b_.at(a);
        if (sc != BLACK && sc != WHITE) continue;
        const int h = b_.head(a);
        int *list = sc == c ? own : cap;
        int &cnt = sc == c ? nown : ncap;
        if (sc != c && b_.string_liberties(a) != 1) continue;
        bool dup = false;
        for (int i = 0; i < cnt; i++) dup |= list[i] == h;
        if (!dup) list[cnt++] = h;
    }
    if (ncap == 0) {
        // Liberties of the merged string: the strings' liberties and v's
        // empty neighbours, without v itself.
        if (nown == 0) return b_.empty_neighbours(v);
        Set l{};
        for (int i = 0; i < nown; i++) or_into(l, libs_[own[i]]);
        for (int k = 0; k < 4; k++) {
            const int a = v + b_.dir(k);
            if (b_.at(a) == EMPTY) set_bit(l, a);
        }
        l.w[v >> 6] &= ~(1ull << (v & 63));
        int n = 0;
        for (int i = 0; i < kWords; i++) n += __builtin_popcountll(l.w[i]);
        return n;
    }
    Set m{}, e = empty_;
    set_bit(m, v);
    e.w[v >> 6] &= ~(1ull << (v & 63));
    for (int i = 0; i < nown; i++) or_into(m, strings_[own[i]]);
    for (int i = 0; i < ncap; i++) or_into(e, strings_[cap[i]]);
    return count_and(neighbours(m, b_.width()), e);
}

std::pair<int, int> BoardBits::after_liberties(int c, int v) const {
    return {after_liberties_colour(c, v), after_liberties_colour(other(c), v)};
}

// FastBoard::check_winning_ladder (FastBoard.cpp:2530-2564): an atari on a
// neighbouring two-liberty enemy string at v whose escape reads as a
// losing ladder for the enemy.
bool Board::check_winning_ladder(int c, int v) const {
    // Whether the trial move captures: without a capture the escape's empty
    // neighbours after it are known without playing it (below).
    bool captures = false;
    for (int k = 0; k < 4; k++) {
        const int a = v + dir_[k];
        captures |= sq_[a] == other(to_move_) && libs_[head_[a]] == 1;
    }
    for (int k = 0; k < 4; k++) {
        const int a = v + dir_[k];
        if (sq_[a] != other(c) || libs_[head_[a]] != 2) continue;
        if (self_atari(c, v)) continue;
        if (!captures) {
            // The string keeps one liberty e (its other one); e's empty
            // neighbours after the move lose v if adjacent. Only an escape
            // with exactly two can start a ladder: skip the trial move
            // otherwise (same answer, no board copy).
            std::array<int, 2> lib{};
            int cnt = 0;
            collect_liberties<2>(a, lib, cnt);
            const int e = lib[0] == v ? lib[1] : lib[0];
            int en = empty_neighbours(e);
            for (int j = 0; j < 4; j++) en -= e + dir_[j] == v;
            if (en != 2) continue;
        }
        Board t = *this;
        bool cap;
        t.play_fast(t.to_move_, v, &cap);
        // A trial move that captures next to the string can leave it more
        // than one liberty: no escape point (vertex 0 has no empty
        // neighbours in the reference's neighbour counts).
        const int escape = t.atari_liberty(a);
        if (escape && t.empty_neighbours(escape) == 2) {
            t.set_to_move(other(t.to_move_));
            if (t.check_losing_ladder(t.to_move_, escape)) return true;
        }
    }
    return false;
}

// FastBoard::check_losing_ladder (FastBoard.cpp:2647-2837): does escaping
// at v with the string in atari lose to a ladder? The defender is the
// board's side to move; the attacker keeps choosing the atari that leaves
// the escape with the fewest liberties and branches (depth <= 5) on ties.
bool Board::check_losing_ladder(int c, int v, int branching) const {
    if (branching > 5) return false;
    const int elib = minimum_enemy_liberties(c, v);
    if (elib == 0 || elib == 1) return false;  // an attacker string dies
    uint16_t crit[4];
    int ncrit = 0;
    for (int k = 0; k < 4; k++) {
        const int a = v + dir_[k];
        if (sq_[a] == c && libs_[head_[a]] <= 1) {
            bool dup = false;
            for (int i = 0; i < ncrit; i++) dup |= crit[i] == head_[a];
            if (!dup) crit[ncrit++] = head_[a];
        }
    }
    if (ncrit != 1) return false;  // connects two weak strings (or none)
    if (can_kill_neighbours(crit[0])) return false;

    Board t = *this;
    bool cap;
    int atari = v;
    t.play_fast(t.to_move_, v, &cap);
    for (;;) {
        if (t.sq_[atari] == EMPTY) return true;  // the escape was suicide
        const int nl = t.string_liberties(atari);
        if (nl == 1) return true;   // still in atari
        if (nl >= 3) return false;  // escaped
        if (t.minimum_enemy_liberties(c, atari) == 1) return false;
        std::array<int, 2> lib;
        int cnt = 0;
        t.collect_liberties<2>(atari, lib, cnt);
        if (t.empty_neighbours(lib[0]) == 3 && t.empty_neighbours(lib[1]) == 3)
            return false;  // two open liberties
        const int att = other(t.to_move_);
        int l0 = t.after_liberties_colour(t.to_move_, lib[0]);
        int l1 = t.after_liberties_colour(t.to_move_, lib[1]);
        const bool bad0 = t.is_suicide(lib[0], att) || t.self_atari(att, lib[0]);
        const bool bad1 = t.is_suicide(lib[1], att) || t.self_atari(att, lib[1]);
        if (bad0 && bad1) return false;
        if (bad0) l1 = l0 + 1;
        if (bad1) l0 = l1 + 1;
        if (l0 == l1) {
            Board b = t;
            b.play_fast(att, lib[0], &cap);
            if (b.check_losing_ladder(c, lib[1], branching + 1)) return true;
            b = t;
            b.play_fast(att, lib[1], &cap);
            return b.check_losing_ladder(c, lib[0], branching + 1);
        }
        t.play_fast(att, l0 > l1 ? lib[0] : lib[1], &cap);
        atari = t.atari_liberty(atari);
        if (!atari) return false;  // unreachable: the string keeps one liberty
        t.play_fast(t.to_move_, atari, &cap);
    }
}

int Board::remove_string(int v) {
    std::array<uint16_t, kMaxSide * kMaxSide> stones;
    int n = 0;
    int s = v;
    do {
        stones[n++] = (uint16_t)s;
        s = next_[s];
    } while (s != v);
    for (int i = 0; i < n; i++) {
        sq_[stones[i]] = EMPTY;
        head_[stones[i]] = kNone;
        next_[stones[i]] = kNone;
    }
    for (int i = 0; i < n; i++) {
        const int st = stones[i];
        uint16_t seen[4];
        int nseen = 0;
        for (int k = 0; k < 4; k++) {
            const uint16_t h = head_[st + dir_[k]];
            if (h == kNone) continue;
            bool dup = false;
            for (int i = 0; i < nseen; i++) dup |= seen[i] == h;
            if (!dup) {
                seen[nseen++] = h;
                libs_[h]++;
            }
        }
    }
    empties_ += n;
    return n;
}

void Board::recount_liberties(int h) {
    std::array<uint8_t, kMaxSquares> mark{};
    int n = 0;
    int s = h;
    do {
        for (int k = 0; k < 4; k++) {
            const int a = s + dir_[k];
            if (sq_[a] == EMPTY && !mark[a]) {
                mark[a] = 1;
                n++;
            }
        }
        s = next_[s];
    } while (s != h);
    libs_[h] = (uint16_t)n;
}

int Board::place(int c, int v, bool *capture, bool game_move) {
    bool eyeplay = true;
    for (int k = 0; k < 4; k++) {
        const int s = sq_[v + dir_[k]];
        eyeplay &= s == other(c) || s == OFFBOARD;
    }
    sq_[v] = (uint8_t)c;
    head_[v] = (uint16_t)v;
    next_[v] = (uint16_t)v;
    stones_[v] = 1;
    libs_[v] = (uint16_t)empty_neighbours(v);
    empties_--;
    {
        uint16_t seen[4];
        int nseen = 0;
        for (int k = 0; k < 4; k++) {
            const uint16_t h = head_[v + dir_[k]];
            if (h == kNone) continue;
            bool dup = false;
            for (int i = 0; i < nseen; i++) dup |= seen[i] == h;
            if (!dup) {
                seen[nseen++] = h;
                libs_[h]--;
            }
        }
    }
    int captured = 0, captured_sq = -1;
    for (int k = 0; k < 4; k++) {
        const int a = v + dir_[k];
        if (sq_[a] == other(c) && libs_[head_[a]] == 0) {
            captured += remove_string(a);
            captured_sq = a;
        }
    }
    bool merged = false;
    for (int k = 0; k < 4; k++) {
        const int a = v + dir_[k];
        if (sq_[a] != c || head_[a] == head_[v]) continue;
        int keep = head_[v], gone = head_[a];
        if (stones_[gone] > stones_[keep]) std::swap(keep, gone);
        int s = gone;
        do {
            head_[s] = (uint16_t)keep;
            s = next_[s];
        } while (s != gone);
        std::swap(next_[keep], next_[gone]);
        stones_[keep] += stones_[gone];
        merged = true;
    }
    if (merged) recount_liberties(head_[v]);
    if (captured) *capture = true;
    if (libs_[head_[v]] == 0 && (game_move || !eyeplay)) remove_string(v);
    if (!game_move && eyeplay) {
        *capture = true;
        return captured == 1 ? captured_sq : -1;
    }
    return (captured == 1 && eyeplay) ? captured_sq : -1;
}

int Board::play(int c, int v, bool *capture) {
    *capture = false;
    return place(c, v, capture, true);
}

int Board::play_fast(int c, int v, bool *capture) {
    *capture = false;
    return place(c, v, capture, false);
}

uint64_t Board::stones_hash() const {
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < width() * width(); i++) {
        h ^= sq_[i];
        h *= 1099511628211ull;
    }
    return h;
}

void GameState::init(int side, float k) {
    board.reset(side);
    komi = k;
    komove = 0;
    last_move = 0;
    prev_last = 0;
    passes = 0;
    movenum = 0;
    last_was_capture = false;
    handicap = 0;
}

// FastState::play_move(color, vertex) / play_pass (FastState.cpp:345-389).
void GameState::play(int v) {
    movenum++;
    prev_last = last_move;
    if (v == kPass || v == kResign) {
        last_move = kPass;
        last_was_capture = false;
        if (passes < 4) passes++;
    } else {
        bool cap = false;
        komove = board.play(board.to_move(), v, &cap);
        last_move = v;
        last_was_capture = cap;
        passes = 0;
    }
    board.set_to_move(other(board.to_move()));
}

void random_position(GameState &st, int side, uint32_t seed, int max_moves) {
    st.init(side, 7.5f);
    Rng rng(seed);
    // The LCG seeding leaves the first outputs nearly unmixed (the first is
    // below 0.18 * 2^64 for every seed), so they are discarded.
    rng.next();
    rng.next();
    const int m = (int)rng.randuint32((uint32_t)max_moves + 1);
    std::vector<int> cand;
    cand.reserve(side * side);
    for (int i = 0; i < m; i++) {
        cand.clear();
        const Board &b = st.board;
        for (int y = 0; y < side; y++)
            for (int x = 0; x < side; x++) {
                const int v = b.vertex(x, y);
                if (b.at(v) == EMPTY && v != st.komove && b.no_eye_fill(v) &&
                    !b.is_suicide(v, b.to_move()))
                    cand.push_back(v);
            }
        if (cand.empty()) {
            st.play(kPass);
            if (st.passes >= 2) break;
        } else {
            st.play(cand[rng.randuint32((uint32_t)cand.size())]);
        }
    }
}

}  // namespace go
}  // namespace lz
