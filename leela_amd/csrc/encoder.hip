// encoder.hip — the nets' input planes computed on the GPU, one wavefront
// per board.
//
// Reference: Network::gather_features_policy / gather_features_value
// (Network.cpp:883-1201) and the FastBoard queries behind them: string
// liberties, after-move liberties (FastBoard.cpp:2511-2528), saving size
// and self-atari (:1273-1292, :1401-1471), suicide (:191-240), and the two
// ladder readers (:2530-2564, :2647-2837).
//
// The board lives in bitboards of 441 bits (7 words) over the padded grid:
// neighbours of a set are four shifts (by 1 and by the grid width), and a
// string is a flood fill of a stone through its colour. Per board:
//   1. the record (squares, string representatives) is staged in LDS; the
//      colour and empty bitboards come from wave ballots;
//   2. each string's stones are OR-ed into a per-representative bitboard,
//      its liberty count is one popcount;
//   3. lane per empty point: after-move liberties of both colours (the
//      merged string's neighbours among the empty points plus captured
//      stones) and the ladder candidates, queued as tasks;
//   4. lane per task: the ladder readers on a private bitboard board,
//      branching through a small stack in private memory;
//   5. lane per output index: the 32 plane bytes of the rotated point.
// Every output byte is written, so the output needs no clearing.
#include <hip/hip_runtime.h>

#include "lz_encoder.h"

namespace lz {
namespace {

constexpr int kWords = 7;
constexpr int kBlack = 0, kWhite = 1, kEmpty = 2, kOff = 3;
constexpr int kMaxPoints = 361;
constexpr int kMaxTasks = 5 * kMaxPoints;  // <= 4 winning + 1 losing per point
constexpr int kMaxBranch = 5;              // FastBoard.cpp:2650 (branching cap)

struct Bits {
    uint64_t w[kWords];
};

__device__ __forceinline__ Bits bits_zero() {
    Bits b;
#pragma unroll
    for (int i = 0; i < kWords; i++) b.w[i] = 0;
    return b;
}

__device__ __forceinline__ Bits bits_point(int v) {
    Bits b;
#pragma unroll
    for (int i = 0; i < kWords; i++)
        b.w[i] = (v >> 6) == i ? 1ull << (v & 63) : 0ull;
    return b;
}

__device__ __forceinline__ bool bits_test(const Bits &s, int v) {
    uint64_t w = 0;
#pragma unroll
    for (int i = 0; i < kWords; i++) w = (v >> 6) == i ? s.w[i] : w;
    return (w >> (v & 63)) & 1ull;
}

__device__ __forceinline__ void bits_set(Bits &s, int v) {
#pragma unroll
    for (int i = 0; i < kWords; i++)
        s.w[i] |= (v >> 6) == i ? 1ull << (v & 63) : 0ull;
}

__device__ __forceinline__ void bits_clear(Bits &s, int v) {
#pragma unroll
    for (int i = 0; i < kWords; i++)
        s.w[i] &= (v >> 6) == i ? ~(1ull << (v & 63)) : ~0ull;
}

__device__ __forceinline__ int bits_count(const Bits &s) {
    int n = 0;
#pragma unroll
    for (int i = 0; i < kWords; i++) n += __popcll(s.w[i]);
    return n;
}

__device__ __forceinline__ bool bits_any(const Bits &s) {
    uint64_t o = 0;
#pragma unroll
    for (int i = 0; i < kWords; i++) o |= s.w[i];
    return o != 0;
}

// Lowest member, -1 for the empty set.
__device__ __forceinline__ int bits_first(const Bits &s) {
    int r = -1;
#pragma unroll
    for (int i = kWords - 1; i >= 0; i--)
        if (s.w[i]) r = 64 * i + __builtin_ctzll(s.w[i]);
    return r;
}

__device__ __forceinline__ Bits bits_and(const Bits &a, const Bits &b) {
    Bits r;
#pragma unroll
    for (int i = 0; i < kWords; i++) r.w[i] = a.w[i] & b.w[i];
    return r;
}

__device__ __forceinline__ Bits bits_or(const Bits &a, const Bits &b) {
    Bits r;
#pragma unroll
    for (int i = 0; i < kWords; i++) r.w[i] = a.w[i] | b.w[i];
    return r;
}

__device__ __forceinline__ Bits bits_andnot(const Bits &a, const Bits &b) {
    Bits r;
#pragma unroll
    for (int i = 0; i < kWords; i++) r.w[i] = a.w[i] & ~b.w[i];
    return r;
}

// The 4-neighbours of s on a grid of width W (7 <= W <= 21). Off-board
// points may appear; callers intersect with on-board sets.
__device__ __forceinline__ Bits bits_nbrs(const Bits &s, int W) {
    Bits n;
#pragma unroll
    for (int i = 0; i < kWords; i++) {
        const uint64_t lo = i > 0 ? s.w[i - 1] : 0ull;
        const uint64_t hi = i < kWords - 1 ? s.w[i + 1] : 0ull;
        n.w[i] = (s.w[i] << 1) | (lo >> 63) | (s.w[i] >> 1) | (hi << 63) |
                 (s.w[i] << W) | (lo >> (64 - W)) | (s.w[i] >> W) |
                 (hi << (64 - W));
    }
    return n;
}

// The connected part of `mask` reachable from `seed` (seed within mask).
__device__ Bits bits_fill(Bits s, const Bits &mask, int W) {
    for (;;) {
        const Bits n = bits_nbrs(s, W);
        uint64_t grew = 0;
#pragma unroll
        for (int i = 0; i < kWords; i++) {
            const uint64_t t = s.w[i] | (n.w[i] & mask.w[i]);
            grew |= t ^ s.w[i];
            s.w[i] = t;
        }
        if (!grew) return s;
    }
}

// ---- the readers' board -------------------------------------------------

struct RBoard {
    Bits c[2];  // black, white stones
};

struct Geo {
    Bits on;    // on-board points
    int W;
    int d[4];   // FastBoard::m_dirs order: -W, +1, +W, -1
};

__device__ __forceinline__ Bits empties(const RBoard &b, const Geo &g) {
    Bits e;
#pragma unroll
    for (int i = 0; i < kWords; i++)
        e.w[i] = g.on.w[i] & ~(b.c[0].w[i] | b.c[1].w[i]);
    return e;
}

__device__ __forceinline__ int colour_at(const RBoard &b, const Geo &g, int v) {
    if (bits_test(b.c[0], v)) return kBlack;
    if (bits_test(b.c[1], v)) return kWhite;
    return bits_test(g.on, v) ? kEmpty : kOff;
}

__device__ __forceinline__ Bits string_at(const RBoard &b, const Geo &g,
                                          int col, int v) {
    return bits_fill(bits_point(v), b.c[col], g.W);
}

__device__ __forceinline__ int libs_of(const Bits &s, const Bits &E,
                                       const Geo &g) {
    return bits_count(bits_and(bits_nbrs(s, g.W), E));
}

__device__ __forceinline__ int empty_nbrs(const Bits &E, const Geo &g, int v) {
    int n = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) n += bits_test(E, v + g.d[k]);
    return n;
}

// FastBoard::update_board_fast (the readers' move, FastBoard.cpp:812-870):
// captures opponent strings left without liberties; the played string is
// removed when it has none, unless the stone went into an opponent eye.
__device__ void play_fast(RBoard &b, const Geo &g, int col, int v) {
    bool eyeplay = true;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int s = colour_at(b, g, v + g.d[k]);
        eyeplay &= s == (col ^ 1) || s == kOff;
    }
    bits_set(b.c[col], v);
    Bits E = empties(b, g);
    for (int k = 0; k < 4; k++) {
        const int a = v + g.d[k];
        if (!bits_test(b.c[col ^ 1], a)) continue;
        const Bits t = string_at(b, g, col ^ 1, a);
        if (!bits_any(bits_and(bits_nbrs(t, g.W), E))) {
            b.c[col ^ 1] = bits_andnot(b.c[col ^ 1], t);
            E = bits_or(E, t);
        }
    }
    if (!eyeplay) {
        const Bits m = string_at(b, g, col, v);
        if (!bits_any(bits_and(bits_nbrs(m, g.W), E)))
            b.c[col] = bits_andnot(b.c[col], m);
    }
}

// Liberties of the string through the stone at v (v must hold a stone).
__device__ __forceinline__ int string_libs(const RBoard &b, const Geo &g,
                                           int col, int v) {
    return libs_of(string_at(b, g, col, v), empties(b, g), g);
}

// FastBoard::minimum_elib_count (FastBoard.cpp:2429-2443).
__device__ int min_enemy_libs(const RBoard &b, const Geo &g, int col, int v) {
    int m = 100;
    const Bits E = empties(b, g);
    for (int k = 0; k < 4; k++) {
        const int a = v + g.d[k];
        if (!bits_test(b.c[col ^ 1], a)) continue;
        const int l = libs_of(string_at(b, g, col ^ 1, a), E, g);
        m = l < m ? l : m;
    }
    return m;
}

// FastBoard::is_suicide (FastBoard.cpp:191-240) for an empty v: no empty
// neighbour, every own neighbour string has v as its only liberty, and no
// opponent neighbour string is in atari.
__device__ bool is_suicide(const RBoard &b, const Geo &g, int v, int col) {
    const Bits E = empties(b, g);
    if (empty_nbrs(E, g, v)) return false;
    for (int k = 0; k < 4; k++) {
        const int a = v + g.d[k];
        if (bits_test(b.c[col], a)) {
            if (libs_of(string_at(b, g, col, a), E, g) != 1) return false;
        } else if (bits_test(b.c[col ^ 1], a)) {
            if (libs_of(string_at(b, g, col ^ 1, a), E, g) <= 1) return false;
        }
    }
    return true;
}

// FastBoard::self_atari (FastBoard.cpp:1401-1471): with v played the
// string's liberties, v included, number at most two.
__device__ bool self_atari(const RBoard &b, const Geo &g, int col, int v) {
    const Bits E = empties(b, g);
    if (empty_nbrs(E, g, v) >= 2) return false;
    int own_or_edge = 0;
    Bits u = bits_point(v);
    for (int k = 0; k < 4; k++) {
        const int a = v + g.d[k];
        const int s = colour_at(b, g, a);
        if (s == kOff) {
            own_or_edge++;
        } else if (s == kEmpty) {
            bits_set(u, a);
        } else {
            const Bits t = string_at(b, g, s, a);
            const Bits lt = bits_and(bits_nbrs(t, g.W), E);
            const int l = bits_count(lt);
            if (s == col) {
                own_or_edge++;
                if (l >= 3) return false;  // connects out
                if (l > 1) u = bits_or(u, lt);
            } else if (l <= 1) {
                return false;  // captures
            }
        }
    }
    if (own_or_edge == 0) return true;
    return bits_count(u) <= 2;
}

// FastBoard::after_liberties_color (FastBoard.cpp:2511-2528): liberties of
// the string through v after col plays it, 0 when that is suicide.
__device__ int after_libs(const RBoard &b, const Geo &g, int col, int v) {
    if (is_suicide(b, g, v, col)) return 0;
    RBoard t = b;
    play_fast(t, g, col, v);
    return string_libs(t, g, col, v);
}

// The only liberty of the string through the stone at v, 0 when it has
// another number of liberties (FastBoard::in_atari).
__device__ int atari_liberty(const RBoard &b, const Geo &g, int col, int v) {
    const Bits l = bits_and(bits_nbrs(string_at(b, g, col, v), g.W),
                            empties(b, g));
    return bits_count(l) == 1 ? bits_first(l) : 0;
}

// One pass of FastBoard::check_losing_ladder (FastBoard.cpp:2647-2837) for
// the defender `col` escaping at v on board b, at depth `branching`.
// Returns 1 (the escape fails), 0 (it works), or 2 at a tie between the
// attacker's two ataris: then b is the board before the attacker's move
// and lib0/lib1 the two candidate points.
__device__ int ladder_pass(RBoard &b, const Geo &g, int col, int v,
                           int branching, int &lib0, int &lib1) {
    if (branching > kMaxBranch) return 0;
    const int att = col ^ 1;
    const int elib = min_enemy_libs(b, g, col, v);
    if (elib == 0 || elib == 1) return 0;  // an attacking string dies
    // The own strings in atari next to v: exactly one.
    {
        const Bits E = empties(b, g);
        Bits crit = bits_zero();
        int ncrit = 0;
        for (int k = 0; k < 4; k++) {
            const int a = v + g.d[k];
            if (!bits_test(b.c[col], a)) continue;
            const Bits t = string_at(b, g, col, a);
            if (libs_of(t, E, g) > 1) continue;
            if (ncrit == 0) {
                crit = t;
                ncrit = 1;
            } else if (!bits_test(crit, a)) {
                ncrit++;
            }
        }
        if (ncrit != 1) return 0;
        // FastBoard::can_kill_neighbours (FastBoard.cpp:2587-2613).
        Bits adj = bits_and(bits_nbrs(crit, g.W), b.c[att]);
        while (bits_any(adj)) {
            const Bits t = string_at(b, g, att, bits_first(adj));
            if (libs_of(t, E, g) <= 1) return 0;
            adj = bits_andnot(adj, t);
        }
    }
    int atari = v;
    play_fast(b, g, col, v);
    for (;;) {
        if (!bits_test(b.c[col], atari)) return 1;  // the escape was suicide
        const Bits E = empties(b, g);
        const Bits s = string_at(b, g, col, atari);
        const Bits ls = bits_and(bits_nbrs(s, g.W), E);
        const int nl = bits_count(ls);
        if (nl == 1) return 1;  // still in atari
        if (nl >= 3) return 0;  // escaped
        if (min_enemy_libs(b, g, col, atari) == 1) return 0;
        const int l0p = bits_first(ls);
        Bits rest = ls;
        bits_clear(rest, l0p);
        const int l1p = bits_first(rest);
        if (empty_nbrs(E, g, l0p) == 3 && empty_nbrs(E, g, l1p) == 3)
            return 0;  // two open liberties
        int l0 = after_libs(b, g, col, l0p);
        int l1 = after_libs(b, g, col, l1p);
        const bool bad0 = is_suicide(b, g, l0p, att) || self_atari(b, g, att, l0p);
        const bool bad1 = is_suicide(b, g, l1p, att) || self_atari(b, g, att, l1p);
        if (bad0 && bad1) return 0;
        if (bad0) l1 = l0 + 1;
        if (bad1) l0 = l1 + 1;
        if (l0 == l1) {
            lib0 = l0p;
            lib1 = l1p;
            return 2;
        }
        play_fast(b, g, att, l0 > l1 ? l0p : l1p);
        atari = atari_liberty(b, g, col, atari);
        if (!atari) return 0;
        play_fast(b, g, col, atari);
    }
}

struct Pending {
    RBoard b;
    int v;
    int branching;
};

// FastBoard::check_losing_ladder with its recursion as a depth-first
// search: at a tie the result is the OR over the attacker's two ataris,
// each continuing one level deeper. At most one alternative per level is
// pending, so the stack holds kMaxBranch + 1 boards.
__device__ bool losing_ladder(RBoard b, const Geo &g, int col, int v) {
    Pending stack[kMaxBranch + 1];
    int top = 0;
    int branching = 0;
    for (;;) {
        int lib0 = 0, lib1 = 0;
        const int r = ladder_pass(b, g, col, v, branching, lib0, lib1);
        if (r == 1) return true;
        if (r == 2) {
            Pending &p = stack[top++];
            p.b = b;
            play_fast(p.b, g, col ^ 1, lib1);
            p.v = lib0;
            p.branching = branching + 1;
            play_fast(b, g, col ^ 1, lib0);
            v = lib1;
            branching++;
            continue;
        }
        if (top == 0) return false;
        --top;
        b = stack[top].b;
        v = stack[top].v;
        branching = stack[top].branching;
    }
}

// The trial of FastBoard::check_winning_ladder (FastBoard.cpp:2530-2564)
// against the opponent string through a: col ataris at v, and the escape
// (its last liberty, with two empty neighbours) reads as a losing ladder.
__device__ bool winning_trial(RBoard b, const Geo &g, int col, int v, int a) {
    play_fast(b, g, col, v);
    const int escape = atari_liberty(b, g, col ^ 1, a);
    if (!escape || empty_nbrs(empties(b, g), g, escape) != 2) return false;
    return losing_ladder(b, g, col ^ 1, escape);
}

// Network.cpp:1341-1360.
__device__ __forceinline__ int rotate_idx(int idx, int sym, int side) {
    int x = idx % side, y = idx / side;
    if (sym >= 4) {
        const int t = x;
        x = y;
        y = t;
        sym -= 4;
    }
    if (sym == 1 || sym == 3) y = side - 1 - y;
    if (sym == 2 || sym == 3) x = side - 1 - x;
    return y * side + x;
}

// First plane of each family (policy / value layouts of Network.cpp:
// 883-1046 and 1048-1201).
struct PlaneMap {
    int own_libs, opp_libs, lib_top, after_own, after_opp;
    int ladder, ladder_win, has_komi, line3, komove, last, prev;
};

__device__ __forceinline__ PlaneMap plane_map(int net) {
    if (net == 0) return {3, 8, 5, 13, 19, 25, 26, 30, 31, 27, 28, 29};
    return {3, 9, 6, 15, 21, 27, 28, 29, 30, 31, -1, -1};
}

constexpr int kTaskLose = 1 << 20;

__global__ __launch_bounds__(64) void encode_kernel(const EncRecord *recs,
                                                    int net, int symmetry,
                                                    uint8_t *out) {
    __shared__ uint8_t sq[kEncGrid];
    __shared__ uint16_t head[kEncGrid];
    __shared__ uint16_t libs[kEncGrid];        // by representative
    __shared__ uint8_t aft[2][kEncGrid];       // after-liberties, capped 255
    __shared__ uint32_t lad[kEncGrid];         // bit 0 losing, bit 1 winning
    __shared__ uint64_t strings[kEncGrid][kWords];  // stones by representative
    __shared__ uint32_t tasks[kMaxTasks];
    __shared__ int ntasks;

    const int lane = threadIdx.x;
    const EncRecord &r = recs[blockIdx.x];
    const int side = r.side, W = side + 2, G = W * W, P = side * side;
    const int tm = r.to_move, opp = tm ^ 1;

    // 1. Stage the squares and representatives; ballots give the boards.
    {
        const uint32_t *s32 = reinterpret_cast<const uint32_t *>(r.sq);
        uint32_t *d32 = reinterpret_cast<uint32_t *>(sq);
        for (int i = lane; i < kEncGrid / 4; i += 64) d32[i] = s32[i];
        const uint32_t *h32 = reinterpret_cast<const uint32_t *>(r.head);
        uint32_t *e32 = reinterpret_cast<uint32_t *>(head);
        for (int i = lane; i < kEncGrid / 2; i += 64) e32[i] = h32[i];
    }
    if (lane == 0) ntasks = 0;
    __syncthreads();
    Geo g;
    RBoard board;
    g.W = W;
    g.d[0] = -W;
    g.d[1] = 1;
    g.d[2] = W;
    g.d[3] = -1;
#pragma unroll
    for (int i = 0; i < kWords; i++) {
        const int v = 64 * i + lane;
        const int c = v < G ? sq[v] : kOff;
        board.c[0].w[i] = __ballot(c == kBlack);
        board.c[1].w[i] = __ballot(c == kWhite);
        g.on.w[i] = __ballot(c != kOff);
    }
    const Bits E = empties(board, g);

    // 2. Stones of every string, then its liberty count.
    for (int v = lane; v < G; v += 64) {
        lad[v] = 0;
        const int c = sq[v];
        if ((c == kBlack || c == kWhite) && head[v] == v) {
#pragma unroll
            for (int i = 0; i < kWords; i++) strings[v][i] = 0;
        }
    }
    __syncthreads();
    for (int v = lane; v < G; v += 64) {
        const int c = sq[v];
        if (c == kBlack || c == kWhite)
            atomicOr(reinterpret_cast<unsigned long long *>(
                         &strings[head[v]][v >> 6]),
                     1ull << (v & 63));
    }
    __syncthreads();
    for (int v = lane; v < G; v += 64) {
        const int c = sq[v];
        if ((c == kBlack || c == kWhite) && head[v] == v) {
            Bits s;
#pragma unroll
            for (int i = 0; i < kWords; i++) s.w[i] = strings[v][i];
            libs[v] = (uint16_t)libs_of(s, E, g);
        }
    }
    __syncthreads();

    // 3. Lane per empty point: after-liberties and ladder tasks.
    for (int v = lane; v < G; v += 64) {
        if (sq[v] != kEmpty) continue;
        int nh[4], ncol[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int a = v + g.d[k];
            ncol[k] = sq[a];
            nh[k] = (ncol[k] == kBlack || ncol[k] == kWhite) ? head[a] : -1;
        }
        // After-liberties (BoardBits formula): the merged string M = v plus
        // the adjacent strings of the mover; its liberties are M's
        // neighbours among the empty points without v plus the stones of
        // the captured strings (adjacent opponent strings in atari).
        int after[2];
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int col = t == 0 ? tm : opp;
            Bits m = bits_point(v);
            Bits e = E;
            bits_clear(e, v);
            for (int k = 0; k < 4; k++) {
                if (nh[k] < 0) continue;
                Bits s;
#pragma unroll
                for (int i = 0; i < kWords; i++) s.w[i] = strings[nh[k]][i];
                if (ncol[k] == col)
                    m = bits_or(m, s);
                else if (libs[nh[k]] == 1)
                    e = bits_or(e, s);
            }
            after[t] = bits_count(bits_and(bits_nbrs(m, W), e));
        }
        aft[0][v] = (uint8_t)(after[0] > 255 ? 255 : after[0]);
        aft[1][v] = (uint8_t)(after[1] > 255 ? 255 : after[1]);

        int en = 0;
        bool own_atari = false, captures = false, kill = false, connect = false;
        int own_or_edge = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            en += ncol[k] == kEmpty;
            if (ncol[k] == kOff) own_or_edge++;
            if (nh[k] < 0) continue;
            const int l = libs[nh[k]];
            if (ncol[k] == tm) {
                own_or_edge++;
                own_atari |= l == 1;
                connect |= l >= 3;
            } else {
                captures |= l == 1;
                kill |= l <= 1;
            }
        }
        // Losing-ladder escape (Network.cpp:1015-1022): two empty
        // neighbours and a saving size: an own neighbour string in atari
        // (v, its liberty, is no self-atari with two empty neighbours).
        if (en == 2 && own_atari) {
            const int i = atomicAdd(&ntasks, 1);
            tasks[i] = (uint32_t)v | kTaskLose;
        }
        // Winning ladder: ataris on adjacent two-liberty opponent strings
        // that are not self-atari; without a capture the escape must have
        // two empty neighbours after the atari to be read at all.
        bool sa;  // self_atari(tm, v), FastBoard.cpp:1401-1471
        if (en >= 2 || kill || connect) {
            sa = false;
        } else if (own_or_edge == 0) {
            sa = true;
        } else {
            Bits u = bits_point(v);
            for (int k = 0; k < 4; k++) {
                const int a = v + g.d[k];
                if (ncol[k] == kEmpty) bits_set(u, a);
                if (nh[k] >= 0 && ncol[k] == tm && libs[nh[k]] > 1) {
                    Bits s;
#pragma unroll
                    for (int i = 0; i < kWords; i++) s.w[i] = strings[nh[k]][i];
                    u = bits_or(u, bits_and(bits_nbrs(s, W), E));
                }
            }
            sa = bits_count(u) <= 2;
        }
        if (sa) continue;
        for (int k = 0; k < 4; k++) {
            if (ncol[k] != opp || libs[nh[k]] != 2) continue;
            bool dup = false;
            for (int j = 0; j < k; j++) dup |= nh[j] == nh[k];
            if (dup) continue;
            if (!captures) {
                Bits s;
#pragma unroll
                for (int i = 0; i < kWords; i++) s.w[i] = strings[nh[k]][i];
                Bits l = bits_and(bits_nbrs(s, W), E);
                bits_clear(l, v);
                const int e = bits_first(l);
                int een = 0;
                for (int j = 0; j < 4; j++)
                    een += sq[e + g.d[j]] == kEmpty && e + g.d[j] != v;
                if (een != 2) continue;
            }
            const int i = atomicAdd(&ntasks, 1);
            tasks[i] = (uint32_t)v | ((uint32_t)nh[k] << 9);
        }
    }
    __syncthreads();

    // 4. Lane per ladder task.
    const int nt = ntasks;
    for (int t = lane; t < nt; t += 64) {
        const uint32_t task = tasks[t];
        const int v = task & 511;
        if (task & kTaskLose) {
            if (losing_ladder(board, g, tm, v)) atomicOr(&lad[v], 1u);
        } else {
            const int a = (task >> 9) & 511;
            if (winning_trial(board, g, tm, v, a)) atomicOr(&lad[v], 2u);
        }
    }
    __syncthreads();

    // 5. Lane per output index: the 32 plane bytes of the rotated point.
    const PlaneMap L = plane_map(net);
    uint8_t *o = out + r.out_offset;
    for (int i = lane; i < P; i += 64) {
        const int src = rotate_idx(i, symmetry, side);
        const int x = src % side, y = src / side;
        const int v = (y + 1) * W + (x + 1);
        const int c = sq[v];
        uint32_t m = 0;
        if (x == 2 || x == side - 3 || y == 2 || y == side - 3)
            m |= 1u << L.line3;
        if (c == kEmpty) {
            m |= 1u;
            const int a0 = aft[0][v], a1 = aft[1][v];
            if (a0 >= 1) m |= 1u << (L.after_own + (a0 >= 6 ? 5 : a0 - 1));
            if (a1 >= 1) m |= 1u << (L.after_opp + (a1 >= 6 ? 5 : a1 - 1));
            if (lad[v] & 1u) m |= 1u << L.ladder;
            if (lad[v] & 2u) m |= 1u << L.ladder_win;
        } else {
            if (c == kWhite && r.white_has_komi) m |= 1u << L.has_komi;
            m |= c == tm ? 2u : 4u;
            int l = libs[head[v]];
            if (l > L.lib_top) l = L.lib_top;
            if (l >= 1) m |= 1u << ((c == tm ? L.own_libs : L.opp_libs) + l - 1);
        }
        if (r.komove > 0 && v == r.komove) m |= 1u << L.komove;
        if (L.last >= 0 && r.last_move > 0) {
            if (v == r.last_move) m |= 1u << L.last;
            if (r.prev_last > 0 && v == r.prev_last) m |= 1u << L.prev;
        }
#pragma unroll
        for (int p = 0; p < 32; p++) o[(size_t)p * P + i] = (uint8_t)((m >> p) & 1u);
    }
}

}  // namespace

hipError_t launch_encode(const EncRecord *d_rec, int n, int net, int symmetry,
                         uint8_t *d_out, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(encode_kernel, dim3(n), dim3(64), 0, stream, d_rec, net,
                       symmetry, d_out);
    return hipGetLastError();
}

}  // namespace lz
