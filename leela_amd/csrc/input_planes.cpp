// input_planes.cpp — input planes of the policy and value nets, built from the
// board queries the reference encoders use (Network.cpp:883-1201).
#include "input_planes.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace lz {
namespace {

using go::BLACK;
using go::EMPTY;
using go::WHITE;

inline void set(uint8_t *planes, int P, int plane, int idx) {
    planes[plane * P + idx] = 1;
}

// Planes shared by both encoders: colours, liberty buckets, after-move
// liberty buckets, ladders, komi and third line. lib_top is the last
// liberty bucket (5 for the policy net's 1..5+, 6 for the value net's
// 1..6+); `p` maps the bucket families to plane numbers.
struct Layout {
    int own_libs, opp_libs, lib_top;  // first plane of each family
    int after_own, after_opp;
    int ladder, ladder_win, has_komi, line3, komove;
};

void gather(const go::GameState &st, uint8_t *planes, const Layout &L) {
    const go::Board &b = st.board;
    const int S = b.side(), P = S * S;
    std::memset(planes, 0, (size_t)kPlanes * P);
    const bool white_has_komi = !(std::fabs(st.komi) <= 0.75f);
    const int tomove = b.to_move();
    const go::BoardBits bits(b);
    for (int y = 0; y < S; y++) {
        for (int x = 0; x < S; x++) {
            const int v = b.vertex(x, y);
            const int idx = y * S + x;
            const int c = b.at(v);
            if (x == 2 || x == S - 3 || y == 2 || y == S - 3)
                set(planes, P, L.line3, idx);
            if (c != EMPTY) {
                if (c == WHITE && white_has_komi) set(planes, P, L.has_komi, idx);
                set(planes, P, c == tomove ? 1 : 2, idx);
                int libs = b.string_liberties(v);
                if (libs > L.lib_top) libs = L.lib_top;
                if (libs >= 1)
                    set(planes, P,
                        (c == tomove ? L.own_libs : L.opp_libs) + libs - 1,
                        idx);
            } else {
                set(planes, P, 0, idx);
                const auto al = bits.after_liberties(tomove, v);
                if (al.first >= 1)
                    set(planes, P, L.after_own + (al.first >= 6 ? 5 : al.first - 1),
                        idx);
                if (al.second >= 1)
                    set(planes, P,
                        L.after_opp + (al.second >= 6 ? 5 : al.second - 1), idx);
                if (b.empty_neighbours(v) == 2 && b.saving_size(tomove, v) > 0 &&
                    b.check_losing_ladder(tomove, v))
                    set(planes, P, L.ladder, idx);
                if (b.check_winning_ladder(tomove, v))
                    set(planes, P, L.ladder_win, idx);
            }
        }
    }
    if (st.komove > 0) {
        const auto xy = b.xy(st.komove);
        set(planes, P, L.komove, xy.second * S + xy.first);
    }
}

}  // namespace

void gather_features_policy(const go::GameState &st, uint8_t *planes) {
    static const Layout L = {3, 8, 5, 13, 19, 25, 26, 30, 31, 27};
    gather(st, planes, L);
    const go::Board &b = st.board;
    const int S = b.side(), P = S * S;
    if (st.last_move > 0) {
        const auto xy = b.xy(st.last_move);
        set(planes, P, 28, xy.second * S + xy.first);
        if (st.prev_last > 0) {
            const auto pxy = b.xy(st.prev_last);
            set(planes, P, 29, pxy.second * S + pxy.first);
        }
    }
}

void gather_features_value(const go::GameState &st, uint8_t *planes) {
    static const Layout L = {3, 9, 6, 15, 21, 27, 28, 29, 30, 31};
    gather(st, planes, L);
}

int rotate_nn_idx(int idx, int symmetry, int side) {
    int x = idx % side, y = idx / side;
    if (symmetry >= 4) {
        const int t = x;
        x = y;
        y = t;
        symmetry -= 4;
    }
    if (symmetry == 1 || symmetry == 3) y = side - 1 - y;
    if (symmetry == 2 || symmetry == 3) x = side - 1 - x;
    return y * side + x;
}

int rev_rotate_nn_idx(int idx, int symmetry, int side) {
    static const int inverse[8] = {0, 1, 2, 3, 4, 6, 5, 7};
    return rotate_nn_idx(idx, inverse[symmetry], side);
}

namespace {
// rotate_nn_idx as tables, per side (1..19) and symmetry, built once: the
// per-element index arithmetic dominated a whole host encode.
constexpr int kTableSides = 19;
struct RotationTables {
    std::vector<uint16_t> idx[kTableSides + 1][8];
    RotationTables() {
        for (int side = 1; side <= kTableSides; side++)
            for (int sym = 0; sym < 8; sym++) {
                auto &t = idx[side][sym];
                t.resize((size_t)side * side);
                for (int i = 0; i < side * side; i++)
                    t[i] = (uint16_t)rotate_nn_idx(i, sym, side);
            }
    }
};
}  // namespace

void rotate_planes(const uint8_t *planes, int symmetry, int side,
                   uint8_t *out) {
    const int P = side * side;
    if (symmetry == 0) {
        std::memcpy(out, planes, (size_t)kPlanes * P);
        return;
    }
    if (side < 1 || side > kTableSides) {
        for (int c = 0; c < kPlanes; c++)
            for (int i = 0; i < P; i++)
                out[c * P + i] =
                    planes[c * P + rotate_nn_idx(i, symmetry, side)];
        return;
    }
    static const RotationTables tables;
    const uint16_t *t = tables.idx[side][symmetry].data();
    for (int c = 0; c < kPlanes; c++) {
        const uint8_t *src = planes + c * P;
        uint8_t *dst = out + c * P;
        for (int i = 0; i < P; i++) dst[i] = src[t[i]];
    }
}

}  // namespace lz
