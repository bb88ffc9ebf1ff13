// mc_core.h — the reference's Monte Carlo playout policy
// (FastState::play_random_move, FastState.cpp:75-255, with the move
// generators and readers of FastBoard.cpp:191-343, 725-990, 1179-1471,
// 1486-1523, 1686-1773, 2027-2643 and Matcher.cpp:14-112) on a playout
// board that keeps FastBoard's playout state. One source for the host
// (mc_policy.cpp) and the device (mc_playout.hip): the same operations in
// the same order, so a playout is the same move sequence on both.
//
// Every container has a fixed capacity that no position can exceed (the
// bounds are derived where each is declared), so nothing allocates.
#pragma once

#include <cstdint>

#include "lz_random.h"

namespace lz {
namespace mc {

// Feature flags of a candidate move (MCPolicy.h:14-36); each multiplies
// the move's score by its feature weight.
enum Flag {
    kPass = 0, kNakade, kPattern, kPatternSA, kSA, kTooBigSA, kForceSA,
    kForceBigSA, kRandom, kSavingSA, kSaving1, kSaving2, kSaving3P,
    kSaving2Lib, kSaving3Lib, kSavingKill, kCapture1, kCapture2, kCapture3P,
    kSuicide, kSemeai2, kSemeai3, kNumFeatures
};

constexpr int kBlack = 0, kWhite = 1, kEmpty = 2, kInval = 3;
constexpr int kPassMove = -1;     // FastBoard::PASS
constexpr int kMaxSide = 19;
constexpr int kMaxSq = (kMaxSide + 2) * (kMaxSide + 2);
constexpr float kBound = 32.0f;   // cfg_bound (GTP.cpp:85)
constexpr int kRandomLoops = 4;   // cfg_random_loops (GTP.cpp:100)
constexpr int kGSize = 8192;      // Matcher's hash displacements
constexpr int kVSize = 1 << 18;   // Matcher's hash range

// The policy's tables as device- or host-resident arrays: the perfect
// hash's displacements, the pattern-weight index of each hash slot per
// colour to move, and the feature and pattern weights (each times its
// SL-trained factor; Matcher::rescale_policy_weights, cfg_mc_softmax = 1).
struct TablesView {
    const int32_t *g;
    const uint16_t *index[2];
    const float *feature_weight;
    const float *pattern_weight;
};

LZ_HD inline uint32_t rotl32(uint32_t x, int k) {
    return (x << k) | (x >> (32 - k));
}

// Matcher::PatHashG / PatHashV.
LZ_HD inline uint32_t pattern_hash(const int32_t *g, uint32_t p) {
    uint32_t q = p;
    q ^= q >> 3;
    const uint32_t d = (uint32_t)g[q % kGSize];
    p ^= rotl32(p, 11) ^ ((p + d) >> 6);
    return p % kVSize;
}

// Matcher::matches: pattern-weight index of a 24-bit augmented 3x3
// pattern for colour c to move.
LZ_HD inline int pattern_index(const TablesView &t, int c, uint32_t pattern) {
    return t.index[c][pattern_hash(t.g, pattern)];
}

// A candidate move: its vertex and feature flags packed in one word (22
// flag bits, the vertex + 1 above them; the pass is vertex -1), and its
// score (MovewFeatures, MCPolicy.h:40-120).
struct Move {
    uint32_t word;
    float score;
    LZ_HD int sq() const { return (int)(word >> 22) - 1; }
    LZ_HD bool has(int flag) const { return (word >> flag) & 1u; }
    LZ_HD void add(int flag, const TablesView &t) {
        word |= 1u << flag;
        score *= t.feature_weight[flag];
    }
};

// The candidates of one move generation. Bound: global captures <= 4 (the
// critical points), saving and killing moves <= 4 x (1 + 4), nakade <= 4,
// pattern moves <= 8, random moves and the pass <= 5, and the semeai moves
// (generated only when nothing else was): for each of the <= 4 own strings
// next to the last move, up to 3 liberties per adjacency of that string's
// stones with an enemy string, and a string has fewer than 2 x 19 x 18 =
// 684 adjacencies on 19x19 — at most 4 x 684 x 3 = 8,208. The host keeps
// that bound; a smaller list (the device's, one per playout in flight)
// records an overflow instead of writing past its end, and the playout
// that overflows is abandoned (its caller reports it). Generator positions
// need a few dozen entries (302 at most over 6.3M move generations).
constexpr int kMaxMoves = 8256;

template <int CAP>
struct MoveList {
    int n = 0;
    bool overflow = false;
    Move m[CAP];
    LZ_HD void clear() {
        n = 0;
        overflow = false;
    }
    LZ_HD bool empty() const { return n == 0; }
    LZ_HD int size() const { return n; }
    LZ_HD Move &back() { return m[n - 1]; }
    LZ_HD Move &operator[](int i) { return m[i]; }
    LZ_HD void push(int v, int flag, const TablesView &t) {
        if (n >= CAP) {
            overflow = true;
            return;
        }
        Move &x = m[n++];
        x.word = (uint32_t)(v + 1) << 22;
        x.score = 1.0f;
        x.add(flag, t);
    }
};

// FastBoard's playout state: squares, packed neighbour counts, strings as
// circular lists with a parent and exact liberty and stone counts, the
// empty-point list, and the critical points found by the last move
// generation. Same updates, in the same order, as FastBoard's
// (FastBoard.cpp:130-189, 266-343, 725-875): the order of the empty list
// decides which point a random draw lands on.
class PlayBoard {
public:
    static constexpr int kNone = kMaxSq;  // parent of empty / border points

    LZ_HD int size() const { return size_; }
    LZ_HD int at(int v) const { return sq_[v]; }
    LZ_HD int vertex(int x, int y) const {
        return (y + 1) * (size_ + 2) + (x + 1);
    }
    LZ_HD int empty_count() const { return empty_cnt_; }
    LZ_HD int empty_at(int i) const { return empty_[i]; }
    LZ_HD int string_size(int v) const { return stones_[parent_[v]]; }

    LZ_HD void reset(int size) {
        size_ = size;
        const int W = size + 2;
        dirs_[0] = -W;
        dirs_[1] = +1;
        dirs_[2] = +W;
        dirs_[3] = -1;
        const int ex[8] = {-W - 1, -W, -W + 1, -1, +1, W - 1, W, W + 1};
        for (int i = 0; i < 8; i++) extradirs_[i] = ex[i];
        for (int i = 0; i < kMaxSq; i++) {
            sq_[i] = kInval;
            nbr_[i] = 0;
            parent_[i] = kNone;
        }
        empty_cnt_ = 0;
        // Empty points column by column, with the border counting as both
        // colours in the neighbour counts.
        for (int i = 0; i < size; i++)
            for (int j = 0; j < size; j++) {
                const int v = vertex(i, j);
                sq_[v] = kEmpty;
                empty_idx_[v] = (int16_t)empty_cnt_;
                empty_[empty_cnt_++] = (int16_t)v;
                int n = 0;
                n += (i == 0 || i == size - 1) ? (1 | (1 << 4) | (1 << 8))
                                               : (2 << 8);
                n += (j == 0 || j == size - 1) ? (1 | (1 << 4) | (1 << 8))
                                               : (2 << 8);
                nbr_[v] = (uint16_t)n;
            }
        parent_[kNone] = kNone;
        libs_[kNone] = 16384;  // large: borders and empties never count low
        next_[kNone] = kNone;
        stones_[kNone] = 0;
        total_[0] = total_[1] = 0;
        ncritical_ = 0;
    }

    // The position's stones placed in raster order, read from a padded
    // (side + 2)^2 grid in this board's vertex numbering (0 black, 1 white;
    // anything else is not a stone). A legal position: no string loses its
    // last liberty on the way.
    LZ_HD void setup(int side, const uint8_t *grid) {
        reset(side);
        for (int y = 0; y < size_; y++)
            for (int x = 0; x < size_; x++) {
                const int v = vertex(x, y);
                const int c = grid[v];
                if (c != kBlack && c != kWhite) continue;
                // A game move (FullBoard::update_board) that captures
                // nothing: in a legal position every string keeps a
                // liberty.
                sq_[v] = (uint8_t)c;
                next_[v] = (int16_t)v;
                parent_[v] = (int16_t)v;
                libs_[v] = (int16_t)count(kEmpty, v);
                stones_[v] = 1;
                total_[c]++;
                add_neighbour(v, c);
                for (int k = 0; k < 4; k++) {
                    const int ai = v + dirs_[k];
                    if (sq_[ai] != c) continue;
                    const int ip = parent_[v], aip = parent_[ai];
                    if (ip != aip) {
                        if (stones_[ip] >= stones_[aip])
                            merge_strings(ip, aip);
                        else
                            merge_strings(aip, ip);
                    }
                }
                take_empty(v);
            }
    }

    // FastBoard::update_board_fast: returns the ko square or -1.
    LZ_HD int play_fast(int color, int v, bool *capture) {
        if (nbr_[v] & (4 << (4 * (1 - color)))) {  // into an opponent eye
            *capture = true;
            return place_eye(color, v);
        }
        sq_[v] = (uint8_t)color;
        next_[v] = (int16_t)v;
        parent_[v] = (int16_t)v;
        libs_[v] = (int16_t)count(kEmpty, v);
        stones_[v] = 1;
        total_[color]++;
        add_neighbour(v, color);
        for (int k = 0; k < 4; k++) {
            const int ai = v + dirs_[k];
            if (sq_[ai] > kWhite) continue;
            if (sq_[ai] == 1 - color) {
                if (libs_at(ai) <= 0) {
                    *capture = true;
                    remove_string(ai);
                }
            } else {
                const int ip = parent_[v], aip = parent_[ai];
                if (ip != aip) {
                    if (stones_[ip] >= stones_[aip])
                        merge_strings(ip, aip);
                    else
                        merge_strings(aip, ip);
                }
            }
        }
        take_empty(v);
        if (libs_at(v) == 0) remove_string(v);  // suicide
        return -1;
    }

    // FastBoard::final_mc_score.
    LZ_HD float final_mc_score(float komi) const {
        int bsc = total_[kBlack], wsc = total_[kWhite];
        for (int i = 0; i < empty_cnt_; i++) {
            const int v = empty_[i];
            if (count(kWhite, v) == 4)
                wsc++;
            else if (count(kBlack, v) == 4)
                bsc++;
        }
        return (float)bsc - ((float)wsc + komi);
    }

    LZ_HD bool is_eye(int color, int v) const {
        if (!(nbr_[v] & (4 << (4 * color)))) return false;
        const int W = size_ + 2;
        int cnt[4] = {0, 0, 0, 0};
        cnt[sq_[v - 1 - W]]++;
        cnt[sq_[v + 1 - W]]++;
        cnt[sq_[v - 1 + W]]++;
        cnt[sq_[v + 1 + W]]++;
        if (cnt[kInval] == 0) return cnt[1 - color] <= 1;
        return cnt[1 - color] == 0;
    }
    LZ_HD bool no_eye_fill(int v, int to_move) const {
        return !is_eye(to_move, v);
    }

    LZ_HD bool fast_ss_suicide(int color, int v) const {
        if (!(nbr_[v] & (4 << (4 * (1 - color))))) return false;
        const int W = size_ + 2;
        if (libs_at(v - 1) <= 1) return false;
        if (libs_at(v + 1) <= 1) return false;
        if (libs_at(v + W) <= 1) return false;
        if (libs_at(v - W) <= 1) return false;
        return true;
    }

    LZ_HD bool is_suicide(int v, int color) {
        if (count(kEmpty, v)) return false;
        bool connecting = false;
        for (int k = 0; k < 4; k++) {
            const int ai = v + dirs_[k];
            const int libs = libs_at(ai);
            if (sq_[ai] == color) {
                if (libs > 1) return false;  // joins a string that lives
                connecting = true;
            } else if (libs <= 1) {
                return false;  // captures
            }
        }
        add_neighbour(v, color);
        bool opps_live = true, ours_die = true;
        for (int k = 0; k < 4; k++) {
            const int ai = v + dirs_[k];
            const int libs = libs_at(ai);
            if (libs == 0 && sq_[ai] != color)
                opps_live = false;
            else if (libs != 0 && sq_[ai] == color)
                ours_die = false;
        }
        remove_neighbour(v, color);
        return connecting ? (opps_live && ours_die) : opps_live;
    }

    LZ_HD bool self_atari(int color, int v) const {
        if (count(kEmpty, v) >= 2) return false;
        if (kill_or_connect(color, v)) return false;
        if (count(color, v) == 0) return true;
        int libs[3];
        int cnt = 0;
        libs[cnt++] = v;
        for (int k = 0; k < 4; k++) {
            const int ai = v + dirs_[k];
            if (sq_[ai] == kEmpty) {
                bool found = false;
                for (int i = 0; i < cnt; i++)
                    if (libs[i] == ai) {
                        found = true;
                        break;
                    }
                if (!found) {
                    if (cnt > 1) return false;
                    libs[cnt++] = ai;
                }
            } else if (sq_[ai] == color) {
                if (libs_at(ai) > 1) {
                    add_string_liberties<3>(ai, libs, cnt);
                    if (cnt > 2) return false;
                }
            }
        }
        return true;
    }

    LZ_HD int enemy_atari_size(int color, int v) const {
        int size = 0;
        for (int k = 0; k < 4; k++) {
            const int ai = v + dirs_[k];
            if (sq_[ai] == 1 - color && libs_at(ai) <= 2) {
                const int s = string_size(ai);
                if (s > size) size = s;
            }
        }
        return size;
    }

    LZ_HD uint32_t pattern_augment(int v) const {
        const int W = size_ + 2;
        const int nb[8] = {v - W - 1, v - W, v - W + 1, v - 1,
                           v + 1,     v + W - 1, v + W, v + W + 1};
        uint32_t res = 0;
        for (int i = 0; i < 8; i++) res = (res << 2) | sq_[nb[i]];
        // Liberty codes of the N, W, E and S strings: min(liberties - 1, 3)
        // (-1 for a string without liberties, as the reference's
        // arithmetic gives it: all high bits set).
        const int side4[4] = {v - W, v - 1, v + 1, v + W};
        for (int i = 0; i < 4; i++) {
            int at = libs_at(side4[i]) - 1;
            if (at > 3) at = 3;
            res |= (uint32_t)at << (22 - 2 * i);
        }
        return res;
    }

    template <class ML>
    LZ_HD void add_global_captures(int color, int komove, ML &m,
                                   const TablesView &t) {
        for (int i = 0; i < ncritical_; i++)
            if (critical_[i] != komove) try_capture(color, critical_[i], m, t);
        ncritical_ = 0;
    }

    template <class ML>
    LZ_HD void save_critical_neighbours(int color, int v, int komove,
                                        ML &m, const TablesView &t) {
        for (int k = 0; k < 4; k++) {
            const int ai = v + dirs_[k];
            if (sq_[ai] != color) continue;
            const int par = parent_[ai];
            if (libs_[par] > 1) continue;
            const int atari = in_atari(ai);
            if (atari != komove) {
                const int size = stones_[par];
                m.push(atari, size == 1   ? kSaving1
                              : size == 2 ? kSaving2
                                          : kSaving3P,
                       t);
                const int libs = count(kEmpty, atari);
                if (libs == 2) m.back().add(kSaving2Lib, t);
                if (libs == 3) m.back().add(kSaving3Lib, t);
            }
            const bool sa = self_atari(color, atari);
            const bool kill = kill_neighbours(ai, komove, m, t);
            // At most one point per neighbour of the last move.
            if (!kill && (sa || atari == komove))
                critical_[ncritical_++] = (int16_t)atari;
        }
    }

    template <class ML>
    LZ_HD void add_semeai_moves(int color, int last, int komove, ML &m,
                                const TablesView &t) {
        constexpr int kMaxLibs = 3;
        int own[4], nown = 0;
        for (int k = 0; k < 4; k++) {
            const int ai = last + dirs_[k];
            if (sq_[ai] == color && libs_at(ai) <= kMaxLibs)
                own[nown++] = parent_[ai];
        }
        for (int o = 0; o < nown; o++) {
            const int og = own[o];
            if (libs_[og] <= 1) continue;  // strings in atari are handled apart
            int olibs[kMaxLibs];
            int ocnt = 0;
            if (saveable_string<kMaxLibs>(og, olibs, ocnt)) continue;
            // Enemy strings around with no more liberties than ours, in the
            // order the string's stones and directions meet them.
            int pos = og;
            do {
                for (int k = 0; k < 4; k++) {
                    const int ai = pos + dirs_[k];
                    if (sq_[ai] == 1 - sq_[og] && libs_at(ai) <= ocnt)
                        semeai_enemy(color, parent_[ai], komove, olibs, ocnt,
                                     m, t);
                }
                pos = next_[pos];
            } while (pos != og);
        }
    }

    template <class ML>
    LZ_HD void add_near_nakade_moves(int color, int v, int komove,
                                     ML &m, const TablesView &t) {
        for (int k = 0; k < 4; k++) {
            const int ai = v + dirs_[k];
            if (sq_[ai] == kEmpty) check_nakade(1 - color, ai, komove, m, t);
        }
    }

    template <class ML>
    LZ_HD void add_pattern_moves(int color, int v, int komove, ML &m,
                                 const TablesView &t) {
        for (int i = 0; i < 8; i++) {
            const int p = v + extradirs_[i];
            if (sq_[p] == kEmpty && p != komove && !fast_ss_suicide(color, p))
                m.push(p, kPattern, t);
        }
    }

private:
    LZ_HD int count(int c, int v) const { return (nbr_[v] >> (4 * c)) & 7; }
    LZ_HD int libs_at(int v) const { return libs_[parent_[v]]; }

    LZ_HD void add_neighbour(int v, int color) {
        int pars[4], cnt = 0;
        for (int k = 0; k < 4; k++) {
            const int ai = v + dirs_[k];
            nbr_[ai] = (uint16_t)(nbr_[ai] + (1 << (4 * color)) - (1 << 8));
            const int p = parent_[ai];
            bool found = false;
            for (int i = 0; i < cnt; i++) found |= pars[i] == p;
            if (!found) {
                libs_[p]--;
                pars[cnt++] = p;
            }
        }
    }

    LZ_HD void remove_neighbour(int v, int color) {
        int pars[4], cnt = 0;
        for (int k = 0; k < 4; k++) {
            const int ai = v + dirs_[k];
            nbr_[ai] = (uint16_t)(nbr_[ai] + (1 << 8) - (1 << (4 * color)));
            const int p = parent_[ai];
            bool found = false;
            for (int i = 0; i < cnt; i++) found |= pars[i] == p;
            if (!found) {
                libs_[p]++;
                pars[cnt++] = p;
            }
        }
    }

    LZ_HD int remove_string(int v) {
        const int color = sq_[v];
        int pos = v, removed = 0;
        do {
            sq_[pos] = kEmpty;
            parent_[pos] = kNone;
            total_[color]--;
            remove_neighbour(pos, color);
            empty_idx_[pos] = (int16_t)empty_cnt_;
            empty_[empty_cnt_++] = (int16_t)pos;
            removed++;
            pos = next_[pos];
        } while (pos != v);
        return removed;
    }

    LZ_HD void merge_strings(int ip, int aip) {
        stones_[ip] += stones_[aip];
        int pos = aip;
        do {
            // Liberties of the joining stones that ip's string lacks.
            for (int k = 0; k < 4; k++) {
                const int ai = pos + dirs_[k];
                if (sq_[ai] != kEmpty) continue;
                bool shared = false;
                for (int kk = 0; kk < 4; kk++)
                    shared |= parent_[ai + dirs_[kk]] == ip;
                if (!shared) libs_[ip]++;
            }
            parent_[pos] = (int16_t)ip;
            pos = next_[pos];
        } while (pos != aip);
        const int16_t tmp = next_[aip];
        next_[aip] = next_[ip];
        next_[ip] = tmp;
    }

    LZ_HD void take_empty(int v) {
        const int last = empty_[--empty_cnt_];
        empty_idx_[last] = empty_idx_[v];
        empty_[empty_idx_[v]] = (int16_t)last;
    }

    LZ_HD int place_eye(int color, int v) {
        sq_[v] = (uint8_t)color;
        next_[v] = (int16_t)v;
        parent_[v] = (int16_t)v;
        libs_[v] = 0;
        stones_[v] = 1;
        total_[color]++;
        add_neighbour(v, color);
        int captured_sq = -1, captured = 0;
        for (int k = 0; k < 4; k++) {
            const int ai = v + dirs_[k];
            if (libs_at(ai) <= 0) {
                captured += remove_string(ai);
                captured_sq = ai;
            }
        }
        take_empty(v);
        return captured == 1 ? captured_sq : -1;
    }

    LZ_HD int in_atari(int v) const {
        if (libs_at(v) > 1) return 0;
        int pos = v;
        do {
            if (count(kEmpty, pos))
                for (int k = 0; k < 4; k++) {
                    const int ai = pos + dirs_[k];
                    if (sq_[ai] == kEmpty) return ai;
                }
            pos = next_[pos];
        } while (pos != v);
        return 0;
    }

    LZ_HD bool kill_or_connect(int color, int v) const {
        for (int k = 0; k < 4; k++) {
            const int ai = v + dirs_[k];
            const int s = sq_[ai], libs = libs_at(ai);
            if ((libs <= 1 && s == 1 - color) || (libs >= 3 && s == color))
                return true;
        }
        return false;
    }

    LZ_HD static int capture_flag(int size) {
        return size == 1 ? kCapture1 : size == 2 ? kCapture2 : kCapture3P;
    }

    template <class ML>
    LZ_HD bool kill_neighbours(int v, int komove, ML &m,
                               const TablesView &t) {
        bool result = false;
        const int kcolor = 1 - sq_[v];
        int seen[4];
        int nseen = 0;
        int pos = v;
        do {
            for (int k = 0; k < 4; k++) {
                const int ai = pos + dirs_[k];
                if (sq_[ai] != kcolor) continue;
                const int par = parent_[ai];
                if (libs_[par] <= 1 && nseen < 4) {
                    bool found = false;
                    for (int i = 0; i < nseen; i++) found |= seen[i] == par;
                    if (!found) {
                        const int atari = in_atari(ai);
                        result = true;
                        if (atari != komove) {
                            m.push(atari, capture_flag(stones_[par]), t);
                            const int sav = string_size(v);
                            m.back().add(sav == 1   ? kSaving1
                                         : sav == 2 ? kSaving2
                                                    : kSaving3P,
                                         t);
                            m.back().add(kSavingKill, t);
                        }
                        seen[nseen++] = par;
                    }
                }
            }
            pos = next_[pos];
        } while (pos != v);
        return result;
    }

    template <class ML>
    LZ_HD void try_capture(int color, int v, ML &m,
                           const TablesView &t) {
        if (sq_[v] != kEmpty || !count(1 - color, v)) return;
        for (int k = 0; k < 4; k++) {
            const int ai = v + dirs_[k];
            if (sq_[ai] == 1 - color && libs_at(ai) <= 1) {
                m.push(v, capture_flag(string_size(ai)), t);
                return;
            }
        }
    }

    template <int N>
    LZ_HD void add_string_liberties(int v, int (&libs)[N], int &cnt) const {
        int pos = v;
        do {
            if (count(kEmpty, pos)) {
                for (int k = 0; k < 4; k++) {
                    const int ai = pos + dirs_[k];
                    if (sq_[ai] != kEmpty) continue;
                    bool found = false;
                    for (int i = 0; i < cnt; i++)
                        if (libs[i] == ai) {
                            found = true;
                            break;
                        }
                    if (!found) {
                        libs[cnt++] = ai;
                        if (cnt >= N) return;
                    }
                }
            }
            pos = next_[pos];
        } while (pos != v);
    }

    template <int N>
    LZ_HD static bool contains(const int (&a)[N], int cnt, int x) {
        for (int i = 0; i < cnt; i++)
            if (a[i] == x) return true;
        return false;
    }

    template <int N>
    LZ_HD bool saveable_string(int parent, int (&libs)[N], int &cnt) const {
        add_string_liberties<N>(parent, libs, cnt);
        for (int i = 0; i < cnt; i++) {
            const int lib = libs[i];
            for (int k = 0; k < 4; k++) {
                const int ai = lib + dirs_[k];
                if (sq_[ai] == kEmpty) {
                    if (!contains<N>(libs, cnt, ai))
                        return true;  // a new liberty next to one
                } else if (sq_[ai] == sq_[parent]) {
                    const int pp = parent_[ai];
                    if (pp == parent) continue;
                    if (libs_[pp] > cnt) return true;
                    if (libs_[pp] >= 2) {
                        // Connecting: do the joined liberties outnumber
                        // ours?
                        int other[N];
                        int ocnt = 0;
                        add_string_liberties<N>(pp, other, ocnt);
                        int joined = cnt;
                        for (int j = 0; j < ocnt; j++) {
                            bool dup = contains<N>(libs, cnt, other[j]);
                            for (int q = 0; q < j && !dup; q++)
                                dup = other[q] == other[j];
                            if (!dup) joined++;
                        }
                        if (joined > cnt) return true;
                    }
                }
            }
        }
        return false;
    }

    // One enemy string eg of a semeai: its liberties that are not ours, not
    // the ko square and not simple suicides become semeai moves.
    template <class ML>
    LZ_HD void semeai_enemy(int color, int eg, int komove,
                            const int (&olibs)[3], int ocnt, ML &m,
                            const TablesView &t) {
        if (libs_[eg] <= 1) return;  // captures are handled apart
        int elibs[3];
        int ecnt = 0;
        if (saveable_string<3>(eg, elibs, ecnt)) return;
        for (int i = 0; i < ecnt; i++) {
            const int lib = elibs[i];
            if (lib == komove) continue;
            if (!contains<3>(olibs, ocnt, lib) && !fast_ss_suicide(color, lib))
                m.push(lib, ocnt <= 2 ? kSemeai2 : kSemeai3, t);
        }
    }

    LZ_HD bool corner(int p) const {
        const int W = size_ + 2, x = p % W - 1, y = p / W - 1;
        return (x == 0 || x == size_ - 1) && (y == 0 || y == size_ - 1);
    }

    template <class ML>
    LZ_HD void check_nakade(int color, int v, int komove, ML &m,
                            const TablesView &t) {
        int nakade[6], empty_counts[6] = {0, 0, 0, 0, 0, 0};
        int nbr_to_coord[5] = {0, 0, 0, 0, 0};
        const int nbrs = count(kEmpty, v);
        if (nbrs == 0) return;
        int cnt = 0;
        nbr_to_coord[nbrs] = v;
        empty_counts[nbrs]++;
        nakade[cnt++] = v;
        int idx = 0;
        do {
            const int nv = nakade[idx];
            for (int k = 0; k < 4; k++) {
                const int ai = nv + dirs_[k];
                if (sq_[ai] == 1 - color) return;  // not surrounded
                if (sq_[ai] != kEmpty) continue;
                bool found = false;
                for (int j = 0; j < cnt; j++) found |= nakade[j] == ai;
                if (found) continue;
                if (cnt >= 6) return;  // too much eye space
                const int n = count(kEmpty, ai);
                nbr_to_coord[n] = ai;
                empty_counts[n]++;
                nakade[cnt++] = ai;
            }
        } while (++idx < cnt);
        // Killable eye shapes (senseis.xmp.net/?KillableEyeShapes).
        int add = -2;  // the vital point to add, -2 for none
        if (cnt <= 2) return;
        if (cnt == 3) {
            add = nbr_to_coord[2];
        } else if (cnt == 4) {
            if (empty_counts[3] == 1) {  // pyramid four
                add = nbr_to_coord[3];
            } else if (empty_counts[2] == 2 && empty_counts[1] == 2) {
                // bent four in the corner
                int crit = 0;
                bool bent4 = false;
                for (int j = 0; j < cnt; j++) {
                    const int p = nakade[j];
                    if (corner(p)) {
                        if (count(kEmpty, p) == 2) bent4 = true;
                    } else if (count(kEmpty, p) == 2) {
                        crit = p;
                    }
                }
                if (bent4) add = crit;
            }
        } else if (cnt == 5) {
            if (empty_counts[1] == 1 && empty_counts[3] == 1)  // bulky five
                add = nbr_to_coord[3];
            else if (empty_counts[4] == 1)  // crossed five
                add = nbr_to_coord[4];
        } else if (cnt == 6) {
            if (empty_counts[1] == 2 && empty_counts[4] == 1) {  // rabbity six
                add = nbr_to_coord[4];
            } else if (empty_counts[2] == 4 && empty_counts[3] == 2) {
                // rectangular six in the corner
                int crit = 0;
                bool rect6 = false;
                for (int j = 0; j < cnt; j++) {
                    const int p = nakade[j];
                    if (corner(p))
                        rect6 = true;
                    else if (count(kEmpty, p) == 3 && count(color, p) == 1 &&
                             count(1 - color, p) == 0)
                        crit = p;
                }
                if (rect6) add = crit;
            }
        }
        if (add != -2 && add != komove) m.push(add, kNakade, t);
    }

    int size_ = 0;
    int dirs_[4] = {0, 0, 0, 0};
    int extradirs_[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint8_t sq_[kMaxSq];
    uint16_t nbr_[kMaxSq];
    int16_t parent_[kMaxSq + 1], next_[kMaxSq + 1], libs_[kMaxSq + 1],
        stones_[kMaxSq + 1];
    int16_t empty_[kMaxSq], empty_idx_[kMaxSq];
    int empty_cnt_ = 0;
    int total_[2] = {0, 0};
    // Critical points of the last move generation: at most one per
    // neighbour of the last move.
    int16_t critical_[4];
    int ncritical_ = 0;
};

// FastState's playout state around a PlayBoard; CAP is the capacity of
// the move generator's candidate list.
template <int CAP>
struct PlayStateT {
    PlayBoard board;
    int to_move = kBlack;
    int komove = 0;
    int last_move = 0;
    bool last_was_capture = false;
    int passes = 0;
    int movenum = 0;
    float komi = 7.5f;
    bool overflow = false;   // a candidate list outgrew CAP: abandoned
    MoveList<CAP> moves;     // scratch of the move generator

    // FastState::play_move_fast.
    LZ_HD int play_fast(int v) {
        bool capture = false;
        if (v == kPassMove) {
            if (++passes > 4) passes = 4;
        } else {
            komove = board.play_fast(to_move, v, &capture);
            passes = 0;
        }
        last_was_capture = capture;
        last_move = v;
        to_move = 1 - to_move;
        movenum++;
        return v;
    }

    LZ_HD bool try_move(int color, int v) const {
        return v != komove && board.no_eye_fill(v, to_move) &&
               !board.fast_ss_suicide(color, v);
    }

    LZ_HD int walk_empty_list(int color, Rng &rng) {
        const int dir = (int)rng.randfix<2>();
        const int n = board.empty_count();
        const int start = (int)rng.randuint16((uint16_t)n);
        if (dir == 0) {
            for (int i = start; i < n; i++)
                if (try_move(color, board.empty_at(i))) return board.empty_at(i);
            for (int i = 0; i < start; i++)
                if (try_move(color, board.empty_at(i))) return board.empty_at(i);
        } else {
            for (int i = start; i >= 0; i--)
                if (try_move(color, board.empty_at(i))) return board.empty_at(i);
            for (int i = n - 1; i > start; i--)
                if (try_move(color, board.empty_at(i))) return board.empty_at(i);
        }
        return kPassMove;
    }

    LZ_HD void flag_move(Move &m, int color, const TablesView &t) {
        m.score *= t.pattern_weight[pattern_index(
            t, color, board.pattern_augment(m.sq()))];
        if (!board.self_atari(color, m.sq())) return;
        if (board.is_suicide(m.sq(), color)) {
            m.add(kSuicide, t);
            return;
        }
        if (board.enemy_atari_size(1 - color, m.sq()) >= 6) {
            // a self-atari of six stones leaves a living eye space of seven
            m.add(kTooBigSA, t);
        } else {
            const int dying = board.enemy_atari_size(color, m.sq());
            m.add(dying >= 5 ? kForceBigSA : dying ? kForceSA : kSA, t);
        }
        if (m.has(kPattern)) m.add(kPatternSA, t);
        if (m.has(kSaving1) || m.has(kSaving2) || m.has(kSaving3P))
            m.add(kSavingSA, t);
    }

    // FastState::play_random_move (FastState.cpp:168-255): plays and
    // returns the move drawn from the policy.
    LZ_HD int play_random_move(Rng &rng, const TablesView &t) {
        const int color = to_move;
        const int W = board.size() + 2;
        moves.clear();
        // Local and tactical moves around the opponent's last stone.
        if (last_move > 0 && last_move < W * W &&
            board.at(last_move) == 1 - color) {
            board.add_global_captures(color, komove, moves, t);
            board.save_critical_neighbours(color, last_move, komove, moves, t);
            if (moves.empty())
                board.add_semeai_moves(color, last_move, komove, moves, t);
            if (last_was_capture || (0.3f > rng.randflt()))
                board.add_near_nakade_moves(color, last_move, komove, moves,
                                            t);
            board.add_pattern_moves(color, last_move, komove, moves, t);
        }
        if (moves.overflow) return abandon();
        float cumul = 0.0f, best = 0.0f;
        for (int i = 0; i < moves.size(); i++) {
            flag_move(moves[i], color, t);
            cumul += moves[i].score;
            if (moves[i].score > best) best = moves[i].score;
        }
        // No good move yet: random moves from the empty list, and a pass.
        if (moves.empty() || best < kBound) {
            const int scored = moves.size();
            for (int loops = 0; loops < kRandomLoops; loops++) {
                const int v = walk_empty_list(color, rng);
                if (v == kPassMove) break;
                moves.push(v, kRandom, t);
            }
            moves.push(kPassMove, kPass, t);
            if (moves.overflow) return abandon();
            for (int i = scored; i < moves.size(); i++) {
                if (moves[i].sq() != kPassMove) flag_move(moves[i], color, t);
                cumul += moves[i].score;
            }
        }
        // The draw: the first move whose running sum of scores reaches
        // index (the running sums are formed again in the same order, so
        // they are the same floats).
        const float index = rng.randflt() * cumul;
        float run = 0.0f;
        for (int i = 0; i < moves.size(); i++) {
            run += moves[i].score;
            if (index <= run) return play_fast(moves[i].sq());
        }
        return play_fast(kPassMove);
    }

    // Ends the playout without a move (candidate list overflow).
    LZ_HD int abandon() {
        overflow = true;
        passes = 4;
        return kPassMove;
    }

    // Playout::run with mc_owner's settings (four passes or move
    // 2 * side^2); returns final_mc_score with the position's komi
    // (meaningless when overflow is set).
    LZ_HD float playout(Rng &rng, const TablesView &t) {
        overflow = false;
        const int S = board.size();
        const int length = 2 * S * S;
        while (passes < 4 && movenum < length) play_random_move(rng, t);
        return board.final_mc_score(komi);
    }
};

// The host's state: a candidate list no position can outgrow.
using PlayState = PlayStateT<kMaxMoves>;
// The device's state (one per thread of the persistent grid, ~23 KB): a
// list of 2,048 candidates, and an abandoned playout beyond that.
constexpr int kDevMoves = 2048;
using DevPlayState = PlayStateT<kDevMoves>;

}  // namespace mc
}  // namespace lz
