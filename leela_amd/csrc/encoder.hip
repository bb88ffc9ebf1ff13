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
//   4. a group of 8 lanes per ladder task: the readers on their own copy
//      of the board, one bitboard word per lane, branching through a small
//      stack in private memory; string fills start from the board's own
//      strings (LDS), so a fill costs the stones the reading added, not
//      the string's length;
//   5. lane per output index: the 32 plane bytes of the rotated point, for
//      the policy net's layout, the value net's, or both from one pass.
// Every output byte is written, so the output needs no clearing.
#include <hip/hip_runtime.h>

#include "lz_encoder.h"
#include "lz_wbits.h"

namespace lz {
namespace {

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

// ---- the board in LDS ----------------------------------------------------
// One board per workgroup (one wavefront): the record's squares and string
// representatives, every string's stones by a dense string id, liberty
// counts, after-liberties, ladder flags and the ladder task list.
//
// The first kStrCap strings (ids in representative order) keep their stones
// in LDS; a board with more strings keeps the rest in its overflow area in
// global memory (generator positions after 250 moves hold at most ~120
// strings). The cap keeps the workgroup's LDS at 20 KB, eight workgroups
// (two waves per SIMD) per CU.
constexpr int kStrCap = kEncStrCap, kStrOvf = kEncStrOvf;
static_assert(kEncStrWords == kWords, "string set words");

__shared__ uint8_t s_sq[kEncGrid];
__shared__ uint16_t s_head[kEncGrid];
__shared__ uint16_t s_libs[kEncGrid];             // by representative
__shared__ uint8_t s_aft[2][kEncGrid];            // after-liberties, capped 255
__shared__ uint32_t s_lad[kEncGrid];              // bit 0 losing, bit 1 winning
__shared__ uint16_t s_sid[kEncGrid];              // string id by representative
__shared__ uint64_t s_strings[kStrCap][kWords];   // stones by string id
__shared__ uint32_t s_tasks[5 * 361];             // <= 4 winning + 1 losing per point
__shared__ int s_ntasks;
__shared__ uint64_t *s_ovf;  // this board's overflow strings (kStrOvf)

// Word w of the stones of the string whose representative is h.
__device__ __forceinline__ uint64_t str_word(int h, int w) {
    const int id = s_sid[h];
    return id < kStrCap ? s_strings[id][w]
                        : s_ovf[(id - kStrCap) * kWords + w];
}

__device__ __forceinline__ Bits base_string(int h) {
    Bits s;
#pragma unroll
    for (int i = 0; i < kWords; i++) s.w[i] = str_word(h, i);
    return s;
}

// ---- the readers: one ladder task per group of 8 lanes ------------------
// A reading works on its own copy of the board, its sets spread over the
// 8 lanes of its group (lz_wbits.h).

struct WBoard {
    uint64_t c[2];  // this lane's word of the black and white stones
    // A connected set of stones of colour hc known to the reading (the
    // escaping string) and its neighbours: fills through a stone of it, or
    // next to it, start from the whole set.
    uint64_t h, hn;
    int hc;
};

__device__ __forceinline__ uint64_t w_empties(const WBoard &b, const Grp &g) {
    return g.on & ~(b.c[0] | b.c[1]);
}

// Colours of v's four neighbours, 2 bits each in m_dirs order (one
// reduction: each neighbour lies in one lane's word).
__device__ int nbr_colours(const WBoard &b, const Grp &g, int v) {
    int x = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int p = v + g.d[k];
        if (g.gl == (p >> 6)) {
            const int sh = p & 63;
            const int c = ((b.c[0] >> sh) & 1ull)   ? kBlack
                          : ((b.c[1] >> sh) & 1ull) ? kWhite
                          : ((g.on >> sh) & 1ull)   ? kEmpty
                                                    : kOff;
            x |= c << (2 * k);
        }
    }
    return gsum(x);
}

__device__ __forceinline__ int field(int nc, int k) { return (nc >> (2 * k)) & 3; }

__device__ __forceinline__ int count_empty(int nc) {
    int n = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) n += field(nc, k) == kEmpty;
    return n;
}

// The string through the stone of colour col at v. The fill starts from
// the board's own string at v when that string is still on the reader's
// board whole: the reader only adds stones and captures whole strings, so
// then it is a connected part of the answer, and the fill ends after the
// stones added since.
// The same holds for the reading's own escaping string (b.h), which only
// grows while it is read.
__device__ uint64_t string_at(const WBoard &b, const Grp &g, int col, int v) {
    uint64_t seed = w_point(g, v);
    if (s_sq[v] == col) {
        const uint64_t base = g.gl < kWords ? str_word(s_head[v], g.gl) : 0ull;
        if (!w_any(base & ~b.c[col], g)) seed |= base;
    }
    if (col == b.hc && w_test(b.h | b.hn, g, v) && !w_any(b.h & ~b.c[col], g))
        seed |= b.h;
    return w_fill(seed, b.c[col], g);
}

__device__ __forceinline__ void set_hint(WBoard &b, const Grp &g, int col,
                                         uint64_t s) {
    b.h = s;
    b.hn = w_nbrs(s, g);
    b.hc = col;
}

__device__ __forceinline__ int libs_of(uint64_t s, uint64_t E, const Grp &g) {
    return w_count(w_nbrs(s, g) & E);
}

// FastBoard::update_board_fast (the readers' move, FastBoard.cpp:812-870):
// captures opponent strings left without liberties; the played string is
// removed when it has none, unless the stone went into an opponent eye.
__device__ void play_fast(WBoard &b, const Grp &g, int col, int v) {
    const int nc = nbr_colours(b, g, v);
    bool eyeplay = true;
#pragma unroll
    for (int k = 0; k < 4; k++)
        eyeplay &= field(nc, k) == (col ^ 1) || field(nc, k) == kOff;
    b.c[col] |= w_point(g, v);
    uint64_t E = w_empties(b, g);
    for (int k = 0; k < 4; k++) {
        if (field(nc, k) != (col ^ 1)) continue;
        const int a = v + g.d[k];
        if (!w_test(b.c[col ^ 1], g, a)) continue;  // taken with an earlier one
        const uint64_t t = string_at(b, g, col ^ 1, a);
        if (!w_any(w_nbrs(t, g) & E, g)) {
            b.c[col ^ 1] &= ~t;
            E |= t;
        }
    }
    if (!eyeplay) {
        const uint64_t m = string_at(b, g, col, v);
        if (!w_any(w_nbrs(m, g) & E, g)) b.c[col] &= ~m;
    }
}

// FastBoard::minimum_elib_count (FastBoard.cpp:2429-2443).
__device__ int min_enemy_libs(const WBoard &b, const Grp &g, int col, int v) {
    const int nc = nbr_colours(b, g, v);
    const uint64_t E = w_empties(b, g);
    int m = 100;
    for (int k = 0; k < 4; k++) {
        if (field(nc, k) != (col ^ 1)) continue;
        const int l = libs_of(string_at(b, g, col ^ 1, v + g.d[k]), E, g);
        m = l < m ? l : m;
    }
    return m;
}

// FastBoard::is_suicide (FastBoard.cpp:191-240) for an empty v: no empty
// neighbour, every own neighbour string has v as its only liberty, and no
// opponent neighbour string is in atari.
__device__ bool is_suicide(const WBoard &b, const Grp &g, int v, int col) {
    const int nc = nbr_colours(b, g, v);
    if (count_empty(nc)) return false;
    const uint64_t E = w_empties(b, g);
    for (int k = 0; k < 4; k++) {
        const int s = field(nc, k);
        if (s != kBlack && s != kWhite) continue;
        const int l = libs_of(string_at(b, g, s, v + g.d[k]), E, g);
        if (s == col ? l != 1 : l <= 1) return false;
    }
    return true;
}

// FastBoard::self_atari (FastBoard.cpp:1401-1471): with v played the
// string's liberties, v included, number at most two.
__device__ bool self_atari(const WBoard &b, const Grp &g, int col, int v) {
    const int nc = nbr_colours(b, g, v);
    if (count_empty(nc) >= 2) return false;
    const uint64_t E = w_empties(b, g);
    int own_or_edge = 0;
    uint64_t u = w_point(g, v);
    for (int k = 0; k < 4; k++) {
        const int s = field(nc, k), a = v + g.d[k];
        if (s == kOff) {
            own_or_edge++;
        } else if (s == kEmpty) {
            u |= w_point(g, a);
        } else {
            const uint64_t lt = w_nbrs(string_at(b, g, s, a), g) & E;
            const int l = w_count(lt);
            if (s == col) {
                own_or_edge++;
                if (l >= 3) return false;  // connects out
                if (l > 1) u |= lt;
            } else if (l <= 1) {
                return false;  // captures
            }
        }
    }
    if (own_or_edge == 0) return true;
    return w_count(u) <= 2;
}

// FastBoard::after_liberties_color (FastBoard.cpp:2511-2528): liberties of
// the string through v after col plays it, 0 when that is suicide.
__device__ int after_libs(const WBoard &b, const Grp &g, int col, int v) {
    if (is_suicide(b, g, v, col)) return 0;
    WBoard t = b;
    play_fast(t, g, col, v);
    return libs_of(string_at(t, g, col, v), w_empties(t, g), g);
}

// The only liberty of the string through the stone at v, 0 when it has
// another number of liberties (FastBoard::in_atari).
__device__ int atari_liberty(const WBoard &b, const Grp &g, int col, int v) {
    const uint64_t l = w_nbrs(string_at(b, g, col, v), g) & w_empties(b, g);
    return w_count(l) == 1 ? w_first(l, g) : 0;
}

// One pass of FastBoard::check_losing_ladder (FastBoard.cpp:2647-2837) for
// the defender `col` escaping at v on board b, at depth `branching`.
// Returns 1 (the escape fails), 0 (it works), or 2 at a tie between the
// attacker's two ataris: then b is the board before the attacker's move
// and lib0/lib1 the two candidate points.
__device__ int ladder_pass(WBoard &b, const Grp &g, int col, int v,
                           int branching, int &lib0, int &lib1) {
    if (branching > kMaxBranch) return 0;
    const int att = col ^ 1;
    const int elib = min_enemy_libs(b, g, col, v);
    if (elib == 0 || elib == 1) return 0;  // an attacking string dies
    // The own strings in atari next to v: exactly one.
    {
        const int nc = nbr_colours(b, g, v);
        const uint64_t E = w_empties(b, g);
        uint64_t crit = 0;
        int ncrit = 0;
        for (int k = 0; k < 4; k++) {
            if (field(nc, k) != col) continue;
            const int a = v + g.d[k];
            const uint64_t t = string_at(b, g, col, a);
            if (libs_of(t, E, g) > 1) continue;
            if (ncrit == 0) {
                crit = t;
                ncrit = 1;
            } else if (!w_test(crit, g, a)) {
                ncrit++;
            }
        }
        if (ncrit != 1) return 0;
        // FastBoard::can_kill_neighbours (FastBoard.cpp:2587-2613).
        uint64_t adj = w_nbrs(crit, g) & b.c[att];
        while (w_any(adj, g)) {
            const uint64_t t = string_at(b, g, att, w_first(adj, g));
            if (libs_of(t, E, g) <= 1) return 0;
            adj &= ~t;
        }
    }
    int atari = v;
    play_fast(b, g, col, v);
    for (;;) {
        if (!w_test(b.c[col], g, atari)) return 1;  // the escape was suicide
        const uint64_t E = w_empties(b, g);
        const uint64_t s = string_at(b, g, col, atari);
        set_hint(b, g, col, s);
        const uint64_t ls = w_nbrs(s, g) & E;
        const int nl = w_count(ls);
        if (nl == 1) return 1;  // still in atari
        if (nl >= 3) return 0;  // escaped
        if (min_enemy_libs(b, g, col, atari) == 1) return 0;
        const int l0p = w_first(ls, g);
        const int l1p = w_first(ls & ~w_point(g, l0p), g);
        if (count_empty(nbr_colours(b, g, l0p)) == 3 &&
            count_empty(nbr_colours(b, g, l1p)) == 3)
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
    WBoard b;
    int v;
    int branching;
};

// FastBoard::check_losing_ladder with its recursion as a depth-first
// search: at a tie the result is the OR over the attacker's two ataris,
// each continuing one level deeper. Pending depths grow strictly from the
// bottom of the stack, so it holds at most kMaxBranch + 1 boards.
__device__ bool losing_ladder(WBoard b, const Grp &g, int col, int v) {
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
__device__ bool winning_trial(WBoard b, const Grp &g, int col, int v, int a) {
    play_fast(b, g, col, v);
    const int escape = atari_liberty(b, g, col ^ 1, a);
    if (!escape || count_empty(nbr_colours(b, g, escape)) != 2) return false;
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
                                                    int symmetry,
                                                    uint8_t *out_policy,
                                                    uint8_t *out_value,
                                                    uint64_t *ovf) {
    const int lane = threadIdx.x;
    const EncRecord &r = recs[blockIdx.x];
    const int side = r.side, W = side + 2, G = W * W, P = side * side;
    const int tm = r.to_move, opp = tm ^ 1;

    // 1. Stage the squares and representatives; ballots give the boards.
    {
        const uint32_t *s32 = reinterpret_cast<const uint32_t *>(r.sq);
        uint32_t *d32 = reinterpret_cast<uint32_t *>(s_sq);
        for (int i = lane; i < kEncGrid / 4; i += 64) d32[i] = s32[i];
        const uint32_t *h32 = reinterpret_cast<const uint32_t *>(r.head);
        uint32_t *e32 = reinterpret_cast<uint32_t *>(s_head);
        for (int i = lane; i < kEncGrid / 2; i += 64) e32[i] = h32[i];
    }
    if (lane == 0) {
        s_ntasks = 0;
        s_ovf = ovf + (size_t)blockIdx.x * kStrOvf * kWords;
    }
    __syncthreads();
    const int d[4] = {-W, 1, W, -1};  // FastBoard::m_dirs order
    Bits stones[2], on;
#pragma unroll
    for (int i = 0; i < kWords; i++) {
        const int v = 64 * i + lane;
        const int c = v < G ? s_sq[v] : kOff;
        stones[0].w[i] = __ballot(c == kBlack);
        stones[1].w[i] = __ballot(c == kWhite);
        on.w[i] = __ballot(c != kOff);
    }
    const Bits E = bits_andnot(on, bits_or(stones[0], stones[1]));

    // 2. String ids (representatives in vertex order), the stones of every
    // string, then its liberty count.
    {
        int base = 0;
        for (int v0 = 0; v0 < G; v0 += 64) {
            const int v = v0 + lane;
            const int c = v < G ? s_sq[v] : kOff;
            const bool head = (c == kBlack || c == kWhite) && s_head[v] == v;
            const uint64_t m = __ballot(head);
            if (head) {
                const int id = base + __popcll(m & ((1ull << lane) - 1ull));
                s_sid[v] = (uint16_t)id;
#pragma unroll
                for (int i = 0; i < kWords; i++) {
                    if (id < kStrCap)
                        s_strings[id][i] = 0;
                    else
                        s_ovf[(id - kStrCap) * kWords + i] = 0;
                }
            }
            base += __popcll(m);
        }
        for (int v = lane; v < G; v += 64) s_lad[v] = 0;
    }
    __syncthreads();
    for (int v = lane; v < G; v += 64) {
        const int c = s_sq[v];
        if (c == kBlack || c == kWhite) {
            const int id = s_sid[s_head[v]];
            unsigned long long *w =
                id < kStrCap
                    ? reinterpret_cast<unsigned long long *>(
                          &s_strings[id][v >> 6])
                    : reinterpret_cast<unsigned long long *>(
                          &s_ovf[(id - kStrCap) * kWords + (v >> 6)]);
            atomicOr(w, 1ull << (v & 63));
        }
    }
    __syncthreads();
    for (int v = lane; v < G; v += 64) {
        const int c = s_sq[v];
        if ((c == kBlack || c == kWhite) && s_head[v] == v) {
            const Bits s = base_string(v);
            s_libs[v] = (uint16_t)bits_count(bits_and(bits_nbrs(s, W), E));
        }
    }
    __syncthreads();

    // 3. Lane per empty point: after-liberties and ladder tasks.
    for (int v = lane; v < G; v += 64) {
        if (s_sq[v] != kEmpty) continue;
        int nh[4], ncol[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int a = v + d[k];
            ncol[k] = s_sq[a];
            nh[k] = (ncol[k] == kBlack || ncol[k] == kWhite) ? s_head[a] : -1;
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
                for (int i = 0; i < kWords; i++) s.w[i] = str_word(nh[k], i);
                if (ncol[k] == col)
                    m = bits_or(m, s);
                else if (s_libs[nh[k]] == 1)
                    e = bits_or(e, s);
            }
            after[t] = bits_count(bits_and(bits_nbrs(m, W), e));
        }
        s_aft[0][v] = (uint8_t)(after[0] > 255 ? 255 : after[0]);
        s_aft[1][v] = (uint8_t)(after[1] > 255 ? 255 : after[1]);

        int en = 0;
        bool own_atari = false, captures = false, kill = false, connect = false;
        int own_or_edge = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            en += ncol[k] == kEmpty;
            if (ncol[k] == kOff) own_or_edge++;
            if (nh[k] < 0) continue;
            const int l = s_libs[nh[k]];
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
            const int i = atomicAdd(&s_ntasks, 1);
            s_tasks[i] = (uint32_t)v | kTaskLose;
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
                const int a = v + d[k];
                if (ncol[k] == kEmpty) bits_set(u, a);
                if (nh[k] >= 0 && ncol[k] == tm && s_libs[nh[k]] > 1) {
                    Bits s;
#pragma unroll
                    for (int i = 0; i < kWords; i++) s.w[i] = str_word(nh[k], i);
                    u = bits_or(u, bits_and(bits_nbrs(s, W), E));
                }
            }
            sa = bits_count(u) <= 2;
        }
        if (sa) continue;
        for (int k = 0; k < 4; k++) {
            if (ncol[k] != opp || s_libs[nh[k]] != 2) continue;
            bool dup = false;
            for (int j = 0; j < k; j++) dup |= nh[j] == nh[k];
            if (dup) continue;
            if (!captures) {
                Bits s;
#pragma unroll
                for (int i = 0; i < kWords; i++) s.w[i] = str_word(nh[k], i);
                Bits l = bits_and(bits_nbrs(s, W), E);
                bits_clear(l, v);
                const int e = bits_first(l);
                int een = 0;
                for (int j = 0; j < 4; j++)
                    een += s_sq[e + d[j]] == kEmpty && e + d[j] != v;
                if (een != 2) continue;
            }
            const int i = atomicAdd(&s_ntasks, 1);
            s_tasks[i] = (uint32_t)v | ((uint32_t)nh[k] << 9);
        }
    }
    __syncthreads();

    // 4. Group of 8 lanes per ladder task.
    {
        Grp g;
        g.lane = lane;
        g.gl = lane & 7;
        g.W = W;
#pragma unroll
        for (int k = 0; k < 4; k++) g.d[k] = d[k];
        WBoard wb;
        g.on = wb.c[0] = wb.c[1] = wb.h = wb.hn = 0;
        wb.hc = kEmpty;
#pragma unroll
        for (int i = 0; i < kWords; i++) {
            if (g.gl == i) {
                g.on = on.w[i];
                wb.c[0] = stones[0].w[i];
                wb.c[1] = stones[1].w[i];
            }
        }
        const int nt = s_ntasks;
        for (int t = lane >> 3; t < nt; t += 8) {
            const uint32_t task = s_tasks[t];
            const int v = task & 511;
            bool hit;
            uint32_t bit;
            if (task & kTaskLose) {
                hit = losing_ladder(wb, g, tm, v);
                bit = 1u;
            } else {
                hit = winning_trial(wb, g, tm, v, (task >> 9) & 511);
                bit = 2u;
            }
            if (hit && g.gl == 0) atomicOr(&s_lad[v], bit);
        }
    }
    __syncthreads();

    // 5. Lane per output index: the plane bytes of the rotated point, for
    // each net asked for.
    for (int i = lane; i < P; i += 64) {
        const int src = rotate_idx(i, symmetry, side);
        const int x = src % side, y = src / side;
        const int v = (y + 1) * W + (x + 1);
        const int c = s_sq[v];
        const bool line3 = x == 2 || x == side - 3 || y == 2 || y == side - 3;
        const int libs = c == kEmpty ? 0 : s_libs[s_head[v]];
        for (int net = 0; net < 2; net++) {
            uint8_t *out = net == 0 ? out_policy : out_value;
            if (!out) continue;
            const PlaneMap L = plane_map(net);
            uint32_t m = line3 ? 1u << L.line3 : 0u;
            if (c == kEmpty) {
                m |= 1u;
                const int a0 = s_aft[0][v], a1 = s_aft[1][v];
                if (a0 >= 1) m |= 1u << (L.after_own + (a0 >= 6 ? 5 : a0 - 1));
                if (a1 >= 1) m |= 1u << (L.after_opp + (a1 >= 6 ? 5 : a1 - 1));
                if (s_lad[v] & 1u) m |= 1u << L.ladder;
                if (s_lad[v] & 2u) m |= 1u << L.ladder_win;
            } else {
                if (c == kWhite && r.white_has_komi) m |= 1u << L.has_komi;
                m |= c == tm ? 2u : 4u;
                const int l = libs > L.lib_top ? L.lib_top : libs;
                if (l >= 1)
                    m |= 1u << ((c == tm ? L.own_libs : L.opp_libs) + l - 1);
            }
            if (r.komove > 0 && v == r.komove) m |= 1u << L.komove;
            if (L.last >= 0 && r.last_move > 0) {
                if (v == r.last_move) m |= 1u << L.last;
                if (r.prev_last > 0 && v == r.prev_last) m |= 1u << L.prev;
            }
            uint8_t *o = out + r.out_offset;
#pragma unroll
            for (int p = 0; p < 32; p++)
                o[(size_t)p * P + i] = (uint8_t)((m >> p) & 1u);
        }
    }
}

}  // namespace

hipError_t launch_encode(const EncRecord *d_rec, int n, int symmetry,
                         uint8_t *d_policy, uint8_t *d_value, uint64_t *d_ovf,
                         hipStream_t stream) {
    if (n <= 0 || (!d_policy && !d_value)) return hipSuccess;
    if (!d_ovf) return hipErrorInvalidValue;
    hipLaunchKernelGGL(encode_kernel, dim3(n), dim3(64), 0, stream, d_rec,
                       symmetry, d_policy, d_value, d_ovf);
    return hipGetLastError();
}

}  // namespace lz
