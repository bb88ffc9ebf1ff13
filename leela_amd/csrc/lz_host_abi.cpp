// lz_host_abi.cpp — C ABI of the host-side input path: game positions, the
// position generator, the feature encoders and the board queries behind
// them. No GPU work happens here.
#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "go_board.h"
#include "input_planes.h"
#include "leela_amd.h"
#include "lz_position.h"
#include "mc_policy.h"
#include "playout.h"

namespace {

bool valid_vertex(const lz::go::Board &b, int v) {
    if (v < 0 || v >= b.width() * b.width()) return false;
    const auto xy = b.xy(v);
    return xy.first >= 0 && xy.first < b.side() && xy.second >= 0 &&
           xy.second < b.side();
}

}  // namespace

extern "C" {

lz_position *lz_position_new(int side, float komi) {
    if (side < 5 || side > lz::go::kMaxSide) return nullptr;
    lz_position *p = new (std::nothrow) lz_position();
    if (p) p->st.init(side, komi);
    return p;
}

void lz_position_free(lz_position *pos) { delete pos; }

int lz_position_play(lz_position *pos, int vertex) {
    if (!pos) return LZ_ERR_ARG;
    const auto &b = pos->st.board;
    if (vertex != lz::go::kPass) {
        if (!valid_vertex(b, vertex) || b.at(vertex) != lz::go::EMPTY)
            return LZ_ERR_ARG;
    }
    pos->st.play(vertex);
    return LZ_OK;
}

int lz_position_random(lz_position *pos, uint32_t seed, int max_moves) {
    if (!pos || max_moves < 0) return LZ_ERR_ARG;
    const int side = pos->st.board.side();
    lz::go::random_position(pos->st, side, seed, max_moves);
    return LZ_OK;
}

int lz_position_info(const lz_position *pos, int32_t *info) {
    if (!pos || !info) return LZ_ERR_ARG;
    const auto &s = pos->st;
    info[0] = s.board.side();
    info[1] = s.board.to_move();
    info[2] = s.komove;
    info[3] = s.last_move;
    info[4] = s.prev_last;
    info[5] = s.movenum;
    info[6] = s.passes;
    return LZ_OK;
}

int lz_position_squares(const lz_position *pos, uint8_t *squares) {
    if (!pos || !squares) return LZ_ERR_ARG;
    const auto &b = pos->st.board;
    for (int y = 0; y < b.side(); y++)
        for (int x = 0; x < b.side(); x++)
            squares[y * b.side() + x] = b.at(b.vertex(x, y));
    return LZ_OK;
}

int lz_encode_planes(const lz_position *pos, int net, int symmetry,
                     uint8_t *planes) {
    if (!pos || !planes || symmetry < 0 || symmetry > 7) return LZ_ERR_ARG;
    if (net != LZ_NET_POLICY && net != LZ_NET_VALUE) return LZ_ERR_ARG;
    const int S = pos->st.board.side();
    std::vector<uint8_t> tmp((size_t)lz::kPlanes * S * S);
    if (net == LZ_NET_POLICY)
        lz::gather_features_policy(pos->st, tmp.data());
    else
        lz::gather_features_value(pos->st, tmp.data());
    lz::rotate_planes(tmp.data(), symmetry, S, planes);
    return LZ_OK;
}

int lz_board_queries(const lz_position *pos, int32_t *out) {
    if (!pos || !out) return LZ_ERR_ARG;
    const auto &b = pos->st.board;
    const int S = b.side(), tm = b.to_move();
    for (int y = 0; y < S; y++)
        for (int x = 0; x < S; x++) {
            const int v = b.vertex(x, y);
            int32_t *q = out + (size_t)(y * S + x) * LZ_QUERY_FIELDS;
            std::fill(q, q + LZ_QUERY_FIELDS, -1);
            q[0] = b.at(v);
            if (b.at(v) != lz::go::EMPTY) {
                q[1] = b.string_liberties(v);
                q[2] = b.string_size(v);
                continue;
            }
            const auto al = b.after_liberties(tm, v);
            q[3] = al.first;
            q[4] = al.second;
            q[5] = b.empty_neighbours(v);
            q[6] = b.saving_size(tm, v);
            // Read only where the encoders read it (Network.cpp:1015-1022).
            q[7] = (q[5] == 2 && q[6] > 0 && b.check_losing_ladder(tm, v)) ? 1 : 0;
            q[8] = b.check_winning_ladder(tm, v) ? 1 : 0;
            q[9] = b.is_suicide(v, tm) ? 1 : 0;
            q[10] = b.no_eye_fill(v) ? 1 : 0;
            q[11] = b.self_atari(tm, v) ? 1 : 0;
        }
    return LZ_OK;
}

int lz_encode_batch(const lz_position *const *pos, int n, int net,
                    int symmetry, int threads, uint8_t *planes) {
    if (n < 0 || (n > 0 && (!pos || !planes)) || symmetry < 0 || symmetry > 7 ||
        (net != LZ_NET_POLICY && net != LZ_NET_VALUE))
        return LZ_ERR_ARG;
    if (n == 0) return LZ_OK;
    for (int i = 0; i < n; i++)
        if (!pos[i] || pos[i]->st.board.side() != pos[0]->st.board.side())
            return LZ_ERR_ARG;
    const int S = pos[0]->st.board.side();
    const size_t per = (size_t)lz::kPlanes * S * S;
    if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
    threads = std::max(1, std::min(threads, n));
    auto work = [&](int t) {
        std::vector<uint8_t> tmp(symmetry ? per : 0);
        for (int i = t; i < n; i += threads) {
            uint8_t *dst = symmetry ? tmp.data() : planes + per * i;
            if (net == LZ_NET_POLICY)
                lz::gather_features_policy(pos[i]->st, dst);
            else
                lz::gather_features_value(pos[i]->st, dst);
            if (symmetry) lz::rotate_planes(tmp.data(), symmetry, S, planes + per * i);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; t++) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
    return LZ_OK;
}

int lz_make_batch(int n, int side, uint32_t seed0, int max_moves, int net,
                  int threads, uint8_t *planes) {
    if (n < 0 || !planes || side < 5 || side > lz::go::kMaxSide ||
        max_moves < 0 || (net != LZ_NET_POLICY && net != LZ_NET_VALUE))
        return LZ_ERR_ARG;
    const size_t per = (size_t)lz::kPlanes * side * side;
    if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
    threads = std::max(1, std::min(threads, std::max(n, 1)));
    auto work = [&](int t) {
        lz::go::GameState st;
        for (int i = t; i < n; i += threads) {
            lz::go::random_position(st, side, seed0 + (uint32_t)i, max_moves);
            if (net == LZ_NET_POLICY)
                lz::gather_features_policy(st, planes + per * i);
            else
                lz::gather_features_value(st, planes + per * i);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; t++) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
    return LZ_OK;
}

int lz_position_set_handicap(lz_position *pos, int handicap) {
    if (!pos || handicap < 0) return LZ_ERR_ARG;
    pos->st.handicap = handicap;
    return LZ_OK;
}

int lz_position_score(const lz_position *pos, float *score) {
    if (!pos || !score) return LZ_ERR_ARG;
    *score = lz::go::final_mc_score(pos->st.board, pos->st.score_komi());
    return LZ_OK;
}

int lz_playout_winrate(const lz_position *pos, int n, uint32_t seed,
                       float *winrate) {
    if (!pos || !winrate || n < 0) return LZ_ERR_ARG;
    *winrate = lz::go::mc_winrate(pos->st, n, seed);
    return LZ_OK;
}

int lz_mc_policy_load(const char *path, char *err, size_t err_len) {
    std::string msg;
    const lz::mc::Tables *t = lz::mc::tables(path, &msg);
    if (!t && err && err_len) {
        std::strncpy(err, msg.c_str(), err_len - 1);
        err[err_len - 1] = 0;
    }
    return t ? LZ_OK : LZ_ERR_ARG;
}

int lz_playout_winrate_policy(const lz_position *pos, int policy, int n,
                              uint32_t seed, float *winrate) {
    if (!pos || !winrate || n < 0) return LZ_ERR_ARG;
    if (policy == LZ_PLAYOUT_LIGHT)
        return lz_playout_winrate(pos, n, seed, winrate);
    if (policy != LZ_PLAYOUT_PATTERN) return LZ_ERR_ARG;
    std::string msg;
    const lz::mc::Tables *t = lz::mc::tables(nullptr, &msg);
    if (!t) return LZ_ERR_STATE;
    *winrate = lz::mc::pattern_winrate(pos->st, n, seed, *t);
    return LZ_OK;
}

int lz_playout_scores(const lz_position *pos, int policy, int n,
                      uint32_t seed, float *scores) {
    if (!pos || (n > 0 && !scores) || n < 0) return LZ_ERR_ARG;
    const lz::mc::Tables *t = nullptr;
    if (policy == LZ_PLAYOUT_PATTERN) {
        std::string msg;
        t = lz::mc::tables(nullptr, &msg);
        if (!t) return LZ_ERR_STATE;
    } else if (policy != LZ_PLAYOUT_LIGHT) {
        return LZ_ERR_ARG;
    }
    for (int j = 0; j < n; j++) {
        lz::Rng rng(lz::go::playout_seed(seed, (uint32_t)j));
        rng.next();
        rng.next();
        if (t) {
            scores[j] = lz::mc::pattern_playout(pos->st, rng, *t);
        } else {
            lz::go::GameState tmp = pos->st;
            scores[j] = lz::go::playout(tmp, rng);
        }
    }
    return LZ_OK;
}

int lz_make_value_labels(int n, int side, uint32_t seed0, int max_moves,
                         int playouts, uint32_t playout_seed, int threads,
                         uint8_t *planes, float *winrates) {
    return lz_make_value_labels_policy(LZ_PLAYOUT_LIGHT, n, side, seed0,
                                       max_moves, playouts, playout_seed,
                                       threads, planes, winrates);
}

int lz_make_value_labels_policy(int policy, int n, int side, uint32_t seed0,
                                int max_moves, int playouts,
                                uint32_t playout_seed, int threads,
                                uint8_t *planes, float *winrates) {
    if (n < 0 || side < 5 || side > lz::go::kMaxSide || max_moves < 0 ||
        playouts < 1 || !winrates ||
        (policy != LZ_PLAYOUT_LIGHT && policy != LZ_PLAYOUT_PATTERN))
        return LZ_ERR_ARG;
    const lz::mc::Tables *tab = nullptr;
    if (policy == LZ_PLAYOUT_PATTERN) {
        std::string msg;
        tab = lz::mc::tables(nullptr, &msg);
        if (!tab) return LZ_ERR_STATE;
    }
    const size_t per = (size_t)lz::kPlanes * side * side;
    if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
    threads = std::max(1, std::min(threads, std::max(n, 1)));
    auto work = [&](int t) {
        lz::go::GameState st;
        for (int i = t; i < n; i += threads) {
            lz::go::random_position(st, side, seed0 + (uint32_t)i, max_moves);
            if (planes) lz::gather_features_value(st, planes + per * i);
            const uint32_t s = playout_seed + (uint32_t)i;
            winrates[i] = tab ? lz::mc::pattern_winrate(st, playouts, s, *tab)
                              : lz::go::mc_winrate(st, playouts, s);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; t++) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
    return LZ_OK;
}

}  // extern "C"
