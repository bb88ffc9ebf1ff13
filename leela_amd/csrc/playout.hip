// playout.hip — light Monte Carlo playouts on the GPU: the surrogate's
// teacher (Playout::run / mc_owner, Playout.cpp:51-113, 174-217, with the
// light policy of playout.cpp), one group of 8 lanes per playout.
//
// Each group holds its board as bitboards spread over its lanes
// (lz_wbits.h) and plays until four passes in a row or move number
// 2 * side^2, exactly as the host playout: per move the empty points in
// vertex order are the candidates, a draw randuint32(n) picks one, and a
// point that is the ko square, an own eye or suicide is replaced by the
// last candidate and the draw repeated (the list is built in LDS, one per
// group, only once a draw is rejected); no candidate left is a pass. The move captures opponent strings left
// without liberties, and a single captured stone after a move into an
// opponent eye is the new ko square (FullBoard::update_board,
// FullBoard.cpp:130-202). The outcome (black win 2, tie 1, loss 0 in
// half points) of final_mc_score is added to the position's counter;
// playout j of position p draws from Rng(playout_seed(seed0 + p, j)), so
// the result equals the host's mc_winrate bit for bit.
#include <hip/hip_runtime.h>

#include "lz_encoder.h"
#include "lz_wbits.h"

namespace lz {
namespace {

constexpr int kPassMove = -1;
constexpr int kMaxPoints = 361;

// Random (Random.cpp:27-57): xoroshiro128+ with the LCG seed expansion.
struct DevRng {
    uint64_t s0, s1;
    __device__ explicit DevRng(uint32_t seed) {
        s0 = (uint64_t)(uint32_t)(741103597u * seed);
        s1 = 741103597ull * s0;
    }
    __device__ uint64_t next() {
        const uint64_t a = s0;
        uint64_t b = s1;
        const uint64_t r = a + b;
        b ^= a;
        s0 = ((a << 55) | (a >> 9)) ^ b ^ (b << 14);
        s1 = (b << 36) | (b >> 28);
        return r;
    }
    __device__ uint32_t randuint32(uint32_t max) {
        return (uint32_t)(((next() >> 32) * (uint64_t)max) >> 32);
    }
};

// MurmurHash3's 32-bit finalizer of seed ^ (j * 0x9E3779B9): the seed of
// playout j (as the host's playout_seed).
__device__ __forceinline__ uint32_t playout_seed_dev(uint32_t seed, uint32_t j) {
    uint32_t h = seed ^ (j * 0x9E3779B9u);
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

struct PBoard {
    uint64_t c[2];  // this lane's word of the black and white stones
};

__shared__ uint16_t s_list[8][kMaxPoints];  // per group: the move candidates

// Colours of v's 4 neighbours (m_dirs order, fields 0-3) and 4 diagonal
// points (fields 4-7), 2 bits each, in one group reduction.
__device__ int nbr8(const PBoard &b, const Grp &g, int v) {
    const int W = g.W;
    const int off[8] = {-W, 1, W, -1, -W - 1, -W + 1, W - 1, W + 1};
    int x = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int p = v + off[k];
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

__device__ __forceinline__ int fld(int f, int k) { return (f >> (2 * k)) & 3; }

// FastBoard::is_eye (FastBoard.cpp:876-912): the four neighbours own or
// edge, and the diagonals hold at most one enemy stone in the centre, none
// on the edge.
__device__ bool is_eye(int f, int col) {
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (fld(f, k) != col && fld(f, k) != kOff) return false;
    int enemy = 0, edge = 0;
#pragma unroll
    for (int k = 4; k < 8; k++) {
        enemy += fld(f, k) == (col ^ 1);
        edge += fld(f, k) == kOff;
    }
    return edge == 0 ? enemy <= 1 : enemy == 0;
}

__device__ __forceinline__ uint64_t p_empties(const PBoard &b, const Grp &g) {
    return g.on & ~(b.c[0] | b.c[1]);
}

// Whether the part of `mask` connected to s touches `target`: the fill of
// w_fill, stopped at the first contact (most strings next to a move have
// a liberty close by).
__device__ bool w_reaches(uint64_t s, uint64_t mask, uint64_t target,
                          const Grp &g) {
    for (;;) {
        const uint64_t n = w_nbrs(s, g);
        if (w_any(n & target, g)) return true;
        const uint64_t t = s | (n & mask);
        if (!w_any(t ^ s, g)) return false;
        s = t;
    }
}

// The k-th (0-based) member of s in vertex order: the lane whose prefix
// count covers k selects within its word by halving.
__device__ int w_select(uint64_t s, int k, const Grp &g) {
    const int cnt = __popcll(s);
    const int pre = gscan_excl(cnt, g.gl);
    int v = 0;
    if (k >= pre && k < pre + cnt) {
        int r = k - pre, base = 0;
        uint64_t w = s;
#pragma unroll
        for (int h = 32; h >= 1; h >>= 1) {
            const uint64_t lo = w & ((1ull << h) - 1);
            const int c = __popcll(lo);
            if (r >= c) {
                r -= c;
                base += h;
                w >>= h;
            } else {
                w = lo;
            }
        }
        v = 64 * g.gl + base;
    }
    return gsum(v);
}

// FastBoard::is_suicide (FastBoard.cpp:191-240) for an empty v: no empty
// neighbour, every own neighbour string has v as its only liberty, and no
// opponent neighbour string is in atari. (v is a liberty of each
// neighbour string, so both tests ask whether the string has another.)
__device__ bool is_suicide(const PBoard &b, const Grp &g, int v, int col, int f) {
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (fld(f, k) == kEmpty) return false;
    const uint64_t other = p_empties(b, g) & ~w_point(g, v);
    for (int k = 0; k < 4; k++) {
        const int s = fld(f, k);
        if (s != kBlack && s != kWhite) continue;
        const bool more = w_reaches(w_point(g, v + g.d[k]), b.c[s], other, g);
        if (s == col ? more : !more) return false;
    }
    return true;
}

// FullBoard::update_board for a legal, non-suicide move: returns the ko
// square (the single captured stone after a move into an opponent eye) or
// -1. The played string keeps a liberty, so it is never removed.
__device__ int play_move(PBoard &b, const Grp &g, int col, int v, int f) {
    bool eyeplay = true;
#pragma unroll
    for (int k = 0; k < 4; k++)
        eyeplay &= fld(f, k) == (col ^ 1) || fld(f, k) == kOff;
    b.c[col] |= w_point(g, v);
    uint64_t E = p_empties(b, g);
    int captured = 0, cap_sq = -1;
    for (int k = 0; k < 4; k++) {
        if (fld(f, k) != (col ^ 1)) continue;
        const int a = v + g.d[k];
        if (!w_test(b.c[col ^ 1], g, a)) continue;  // taken with an earlier one
        if (!w_reaches(w_point(g, a), b.c[col ^ 1], E, g)) {
            const uint64_t t = w_fill(w_point(g, a), b.c[col ^ 1], g);
            captured += w_count(t);
            b.c[col ^ 1] &= ~t;
            E |= t;
            cap_sq = a;
        }
    }
    return captured == 1 && eyeplay ? cap_sq : -1;
}

__global__ __launch_bounds__(64) void playout_kernel(const EncRecord *recs,
                                                     int npos, int nplay,
                                                     uint32_t seed0,
                                                     uint32_t *half_wins) {
    const int lane = threadIdx.x, grp = lane >> 3;
    const long long gid = (long long)blockIdx.x * 8 + grp;
    if (gid >= (long long)npos * nplay) return;  // whole groups leave together
    const int p = (int)(gid / nplay), j = (int)(gid % nplay);
    const EncRecord &r = recs[p];
    const int side = r.side, W = side + 2, G = W * W;

    Grp g;
    g.lane = lane;
    g.gl = lane & 7;
    g.W = W;
    g.d[0] = -W;
    g.d[1] = 1;
    g.d[2] = W;
    g.d[3] = -1;
    PBoard b;
    b.c[0] = b.c[1] = g.on = 0;
    if (g.gl < kWords) {
        for (int i = 0; i < 64; i++) {
            const int v = 64 * g.gl + i;
            const int c = v < G ? r.sq[v] : kOff;
            b.c[0] |= (uint64_t)(c == kBlack) << i;
            b.c[1] |= (uint64_t)(c == kWhite) << i;
            g.on |= (uint64_t)(c != kOff) << i;
        }
    }
    int tm = r.to_move, ko = r.komove, passes = r.passes, movenum = r.movenum;
    const int maxmoves = 2 * side * side;
    DevRng rng(playout_seed_dev(seed0 + (uint32_t)p, (uint32_t)j));
    rng.next();
    rng.next();
    uint16_t *list = s_list[grp];

    while (passes < 4 && movenum < maxmoves) {
        // The candidates are the empty points in vertex order. The first
        // draw picks from them directly (select); only after a rejection
        // is the list built in LDS (each lane writes its word's points),
        // and then rejected entries are replaced by the last one.
        uint64_t e = p_empties(b, g);
        int n = w_count(e);
        int move = kPassMove, f = 0;
        bool listed = false;
        while (n > 0) {
            const int i = (int)rng.randuint32((uint32_t)n);
            const int v = listed ? (int)list[i] : w_select(e, i, g);
            if (v != ko) {
                f = nbr8(b, g, v);
                if (!is_eye(f, tm) && !is_suicide(b, g, v, tm, f)) {
                    move = v;
                    break;
                }
            }
            if (!listed) {
                const int cnt = __popcll(e);
                int pos = gscan_excl(cnt, g.gl);
                while (e) {
                    list[pos++] = (uint16_t)(64 * g.gl + __builtin_ctzll(e));
                    e &= e - 1;
                }
                listed = true;
            }
            __builtin_amdgcn_wave_barrier();
            list[i] = list[--n];  // every lane of the group writes the same
            __builtin_amdgcn_wave_barrier();
        }
        movenum++;
        if (move == kPassMove) {
            if (passes < 4) passes++;
        } else {
            ko = play_move(b, g, tm, move, f);
            passes = 0;
        }
        tm ^= 1;
    }

    // FastBoard::final_mc_score (FastBoard.cpp:420-443): stones plus empty
    // points whose four neighbours are all of one colour (or edge).
    const uint64_t E = p_empties(b, g);
    const int bsc = w_count(b.c[0]) +
                    w_count(E & ~w_nbrs(b.c[1] | E, g));
    const int wsc = w_count(b.c[1]) +
                    w_count(E & ~w_nbrs(b.c[0] | E, g));
    const float score = (float)bsc - ((float)wsc + r.komi);
    if (g.gl == 0)
        atomicAdd(&half_wins[p], score > 0.0f ? 2u : score == 0.0f ? 1u : 0u);
}

__global__ __launch_bounds__(64) void playout_zero(uint32_t *half_wins,
                                                   int npos) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i < npos) half_wins[i] = 0;
}

// mc_owner's win rate for the side to move from the half-point counts:
// the same float operations as the host.
__global__ __launch_bounds__(64) void playout_winrate(
    const uint32_t *half_wins, const EncRecord *recs, int npos, int nplay,
    float *winrate) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= npos) return;
    float w = nplay > 0 ? (float)half_wins[i] * 0.5f / (float)nplay : 0.5f;
    if (recs[i].to_move != kBlack) w = 1.0f - w;
    winrate[i] = w;
}

}  // namespace

hipError_t launch_playouts(const EncRecord *d_rec, int npos, int nplay,
                           uint32_t seed0, uint32_t *d_half_wins,
                           float *d_winrate, hipStream_t stream) {
    if (npos <= 0) return hipSuccess;
    const int blocks = (npos + 63) / 64;
    hipLaunchKernelGGL(playout_zero, dim3(blocks), dim3(64), 0, stream,
                       d_half_wins, npos);
    const long long groups = (long long)npos * nplay;
    if (groups > 0)
        hipLaunchKernelGGL(playout_kernel, dim3((unsigned)((groups + 7) / 8)),
                           dim3(64), 0, stream, d_rec, npos, nplay, seed0,
                           d_half_wins);
    hipLaunchKernelGGL(playout_winrate, dim3(blocks), dim3(64), 0, stream,
                       d_half_wins, d_rec, npos, nplay, d_winrate);
    return hipGetLastError();
}

}  // namespace lz
