// lz_wbits.h — bitboard sets spread over groups of 8 lanes, the device
// primitive of the board kernels (feature encoder, playouts).
//
// A set of points of the padded grid (up to 7 x 64 = 448 points) lives in
// the 8 lanes of a group: lane j < 7 holds word j (points 64j..64j+63),
// lane 7 holds none. A set operation is one instruction per lane; the
// neighbours of a set need the words of lanes j-1 and j+1 (DPP row
// shifts); counts, first members, prefix sums and point tests are reduced
// within the group (DPP quad permutes, the half-row mirror, row shifts and
// group ballots). Every value a branch depends on is such a reduction, so
// the 8 lanes of a group always take the same path; different groups of a
// wavefront may diverge.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lz {
namespace {

constexpr int kWords = 7;
constexpr int kBlack = 0, kWhite = 1, kEmpty = 2, kOff = 3;

constexpr int kDppRowShl1 = 0x101;
constexpr int kDppRowShr1 = 0x111;
constexpr int kDppRowShr2 = 0x112;
constexpr int kDppRowShr4 = 0x114;
constexpr int kDppQuadXor1 = 0xB1;   // quad_perm [1,0,3,2]
constexpr int kDppQuadXor2 = 0x4E;   // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141;

template <int kCtrl>
__device__ __forceinline__ uint32_t dpp(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kCtrl, 0xF, 0xF,
                                                 true);
}

template <int kCtrl>
__device__ __forceinline__ uint64_t dpp64(uint64_t x) {
    const uint32_t lo = dpp<kCtrl>((uint32_t)x);
    const uint32_t hi = dpp<kCtrl>((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Sum and minimum over the 8 lanes of the group.
__device__ __forceinline__ int gsum(int x) {
    x += (int)dpp<kDppQuadXor1>((uint32_t)x);
    x += (int)dpp<kDppQuadXor2>((uint32_t)x);
    x += (int)dpp<kDppHalfMirror>((uint32_t)x);
    return x;
}

// Sum over the group's lanes below this one.
__device__ __forceinline__ int gscan_excl(int x, int gl) {
    int s = x, y;
    y = (int)dpp<kDppRowShr1>((uint32_t)s);
    s += gl >= 1 ? y : 0;
    y = (int)dpp<kDppRowShr2>((uint32_t)s);
    s += gl >= 2 ? y : 0;
    y = (int)dpp<kDppRowShr4>((uint32_t)s);
    s += gl >= 4 ? y : 0;
    return s - x;
}

__device__ __forceinline__ int gmin(int x) {
    int y = (int)dpp<kDppQuadXor1>((uint32_t)x);
    x = y < x ? y : x;
    y = (int)dpp<kDppQuadXor2>((uint32_t)x);
    x = y < x ? y : x;
    y = (int)dpp<kDppHalfMirror>((uint32_t)x);
    return y < x ? y : x;
}

struct Grp {
    int lane;      // lane within the wavefront
    int gl;        // this lane's word (lane within the group)
    int W;         // grid width
    int d[4];      // FastBoard::m_dirs order: -W, +1, +W, -1
    uint64_t on;   // this lane's word of the on-board points
};


__device__ __forceinline__ uint64_t w_point(const Grp &g, int v) {
    return g.gl == (v >> 6) ? 1ull << (v & 63) : 0ull;
}

// The group's 8 bits of a ballot (other groups' lanes, active or not,
// are not read).
__device__ __forceinline__ uint32_t group_ballot(bool p, int lane) {
    return (uint32_t)(__builtin_amdgcn_ballot_w64(p) >> (lane & ~7)) & 0xFFu;
}

__device__ __forceinline__ bool w_any(uint64_t s, const Grp &g) {
    return group_ballot(s != 0, g.lane) != 0;
}

__device__ __forceinline__ int w_count(uint64_t s) {
    return gsum(__popcll(s));
}

__device__ __forceinline__ bool w_test(uint64_t s, const Grp &g, int v) {
    return group_ballot(g.gl == (v >> 6) && ((s >> (v & 63)) & 1ull), g.lane) != 0;
}

// Lowest member, -1 for the empty set.
__device__ __forceinline__ int w_first(uint64_t s, const Grp &g) {
    const int big = 0x7fffffff;
    const int m = gmin(s ? 64 * g.gl + __builtin_ctzll(s) : big);
    return m == big ? -1 : m;
}

__device__ __forceinline__ uint64_t w_nbrs(uint64_t s, const Grp &g) {
    uint64_t lo = dpp64<kDppRowShr1>(s);
    const uint64_t hi = dpp64<kDppRowShl1>(s);
    if (g.gl == 0) lo = 0;
    const int W = g.W;
    const uint64_t n = (s << 1) | (lo >> 63) | (s >> 1) | (hi << 63) |
                       (s << W) | (lo >> (64 - W)) | (s >> W) |
                       (hi << (64 - W));
    return g.gl == 7 ? 0ull : n;  // lane 7 holds no points
}


__device__ uint64_t w_fill(uint64_t s, uint64_t mask, const Grp &g) {
    for (;;) {
        const uint64_t t = s | (w_nbrs(s, g) & mask);
        if (!w_any(t ^ s, g)) return t;
        s = t;
    }
}

}  // namespace
}  // namespace lz
