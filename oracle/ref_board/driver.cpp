// driver.cpp — TEST INFRASTRUCTURE ONLY. C entry points over the
// reference's own board engine (FastBoard / FullBoard / FastState, compiled
// from /root/reference by oracle/ref_board/Makefile into oracle/_ref/) that
// the golden-vector script and the board tests call. Nothing shipped
// links this.
#include <cstdint>
#include <vector>

#include "FastBoard.h"
#include "FastState.h"
#include "GTP.h"
#include "Matcher.h"
#include "Random.h"
#include "Zobrist.h"

namespace {

void init_once() {
    static bool done = false;
    if (!done) {
        Random rng(5489);  // Leela.cpp:279-280
        Zobrist::init_zobrist(rng);
        done = true;
    }
}

// The playout policy's settings at their GTP defaults
// (GTP::setup_default_parameters, GTP.cpp:64-104), then the pattern tables
// (Leela.cpp:284).
void init_policy_once() {
    static bool done = false;
    if (!done) {
        init_once();
        cfg_mc_softmax = 1.0f;
        cfg_random_loops = 4;
        cfg_bound = 32.0f;
        Matcher::get_Matcher();
        done = true;
    }
}

void fill(FastState &st, int32_t *info, uint8_t *squares, int32_t *q) {
    FastBoard &b = st.board;
    const int S = b.get_boardsize(), tm = st.get_to_move();
    info[0] = S;
    info[1] = tm;
    info[2] = st.get_komove();
    info[3] = st.get_last_move();
    info[4] = st.get_prevlast_move();
    info[5] = st.get_movenum();
    info[6] = st.get_passes();
    for (int y = 0; y < S; y++)
        for (int x = 0; x < S; x++) {
            const int v = b.get_vertex(x, y);
            const int i = y * S + x;
            squares[i] = (uint8_t)b.get_square(v);
            int32_t *r = q + i * 12;
            for (int k = 0; k < 12; k++) r[k] = -1;
            r[0] = b.get_square(v);
            if (b.get_square(v) != FastBoard::EMPTY) {
                r[1] = b.count_rliberties(v);
                r[2] = b.string_size(v);
                continue;
            }
            auto al = b.after_liberties(tm, v);
            r[3] = al.first;
            r[4] = al.second;
            r[5] = b.count_pliberties(v);
            r[6] = b.saving_size(tm, v);
            // Guarded as the encoders call it (Network.cpp:1015-1022).
            r[7] = (r[5] == 2 && r[6] > 0) ? (b.check_losing_ladder(tm, v) ? 1 : 0) : 0;
            r[8] = b.check_winning_ladder(tm, v) ? 1 : 0;
            r[9] = b.is_suicide(v, tm) ? 1 : 0;
            r[10] = b.no_eye_fill(v) ? 1 : 0;
            r[11] = b.self_atari(tm, v) ? 1 : 0;
        }
}

}  // namespace

extern "C" {

// The build's synthetic-position generator run on the reference engine.
int ref_random_position(int side, uint32_t seed, int max_moves,
                        int32_t *info, uint8_t *squares, int32_t *queries,
                        int32_t *moves, int32_t *nmoves) {
    init_once();
    FastState st;
    st.init_game(side, 7.5f);
    Random rng(seed);
    rng.randuint32();  // the generator discards the first two outputs
    rng.randuint32();
    const int m = (int)rng.randuint32((uint32_t)max_moves + 1);
    int played = 0;
    std::vector<int> cand;
    for (int i = 0; i < m; i++) {
        cand.clear();
        for (int y = 0; y < side; y++)
            for (int x = 0; x < side; x++) {
                const int v = st.board.get_vertex(x, y);
                if (st.board.get_square(v) == FastBoard::EMPTY &&
                    v != st.get_komove() && st.board.no_eye_fill(v) &&
                    !st.board.is_suicide(v, st.board.get_to_move()))
                    cand.push_back(v);
            }
        int mv = FastBoard::PASS;
        if (!cand.empty()) mv = cand[rng.randuint32((uint32_t)cand.size())];
        st.play_move(mv);
        moves[played++] = mv;
        if (mv == FastBoard::PASS && st.get_passes() >= 2) break;
    }
    *nmoves = played;
    fill(st, info, squares, queries);
    return 0;
}

int ref_replay(int side, float komi, int n, const int32_t *moves,
               int32_t *info, uint8_t *squares, int32_t *queries) {
    init_once();
    FastState st;
    st.init_game(side, komi);
    for (int i = 0; i < n; i++) st.play_move(moves[i]);
    fill(st, info, squares, queries);
    return 0;
}

// FastBoard::final_mc_score of the position after `moves`.
float ref_final_score(int side, float komi, int n, const int32_t *moves) {
    init_once();
    FastState st;
    st.init_game(side, komi);
    for (int i = 0; i < n; i++) st.play_move(moves[i]);
    return st.board.final_mc_score(komi);
}

// Playouts of the reference's policy (FastState::play_random_move) from the
// position after `moves`, the loop of Playout::run with postpassout and no
// resignation (Playout.cpp:55-80). The position is rebuilt with its stones
// placed in raster order as game moves that capture nothing
// (FullBoard::update_board), the state fields copied from the replay: the
// same empty-point order as the library's playout board gets from a
// position. Playout j seeds the thread's generator with seeds[j] and
// discards two outputs; scores[j] = calculate_mc_score at its end.
int ref_pattern_playouts(int side, float komi, int n_moves,
                         const int32_t *moves, int n, const uint32_t *seeds,
                         float *scores) {
    init_policy_once();
    FastState st;
    st.init_game(side, komi);
    for (int i = 0; i < n_moves; i++) st.play_move(moves[i]);
    FastState base;
    base.init_game(side, komi);
    for (int y = 0; y < side; y++)
        for (int x = 0; x < side; x++) {
            const int v = st.board.get_vertex(x, y);
            const int c = st.board.get_square(v);
            if (c != FastBoard::BLACK && c != FastBoard::WHITE) continue;
            bool capture = false;
            base.board.update_board(c, v, capture);
        }
    base.board.set_to_move(st.board.get_to_move());
    base.m_komove = st.m_komove;
    base.m_lastmove = st.m_lastmove;
    base.m_onebutlastmove = st.m_onebutlastmove;
    base.m_last_was_capture = st.m_last_was_capture;
    base.m_passes = st.m_passes;
    base.m_movenum = st.m_movenum;
    const int length = 2 * side * side;
    for (int j = 0; j < n; j++) {
        Random *rng = Random::get_Rng();
        rng->seedrandom(seeds[j]);
        rng->randuint32();
        rng->randuint32();
        FastState tmp = base;
        while (tmp.get_passes() < 4 && (int)tmp.get_movenum() < length)
            tmp.play_random_move(tmp.get_to_move());
        scores[j] = tmp.calculate_mc_score();
    }
    return 0;
}

void ref_rng_u32(uint32_t seed, int n, uint32_t max, uint32_t *out) {
    Random rng(seed);
    for (int i = 0; i < n; i++)
        out[i] = max ? rng.randuint32(max) : rng.randuint32();
}

void ref_rng_flt(uint32_t seed, int n, float *out) {
    Random rng(seed);
    for (int i = 0; i < n; i++) out[i] = rng.randflt();
}

}  // extern "C"
