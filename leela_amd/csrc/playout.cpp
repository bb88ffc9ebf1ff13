// playout.cpp — light Monte Carlo playouts (see playout.h).
#include "playout.h"

#include <array>

namespace lz {
namespace go {

float final_mc_score(const Board &b, float komi) {
    int bsc = 0, wsc = 0;
    const int S = b.side();
    for (int y = 0; y < S; y++)
        for (int x = 0; x < S; x++) {
            const int v = b.vertex(x, y);
            const int c = b.at(v);
            if (c == BLACK) {
                bsc++;
            } else if (c == WHITE) {
                wsc++;
            } else if (b.colour_neighbours(WHITE, v) == 4) {
                wsc++;
            } else if (b.colour_neighbours(BLACK, v) == 4) {
                bsc++;
            }
        }
    return (float)bsc - ((float)wsc + komi);
}

float playout(GameState &st, Rng &rng) {
    Board &b = st.board;
    const int S = b.side();
    const int maxmoves = 2 * S * S;
    std::array<int, kMaxSide * kMaxSide> empt;
    while (st.passes < 4 && st.movenum < maxmoves) {
        int n = 0;
        for (int y = 0; y < S; y++)
            for (int x = 0; x < S; x++) {
                const int v = b.vertex(x, y);
                if (b.at(v) == EMPTY) empt[n++] = v;
            }
        // Uniform over legal candidates by rejection: draw among the
        // remaining empties, drop rejects.
        int move = kPass;
        while (n > 0) {
            const int i = (int)rng.randuint32((uint32_t)n);
            const int v = empt[i];
            if (v != st.komove && b.no_eye_fill(v) &&
                !b.is_suicide(v, b.to_move())) {
                move = v;
                break;
            }
            empt[i] = empt[--n];
        }
        st.play(move);
    }
    return final_mc_score(b, st.score_komi());
}

uint32_t playout_seed(uint32_t seed, uint32_t j) {
    return lz::playout_seed(seed, j);
}

float mc_winrate(const GameState &st, int n, uint32_t seed) {
    float bwins = 0.0f;
    for (int i = 0; i < n; i++) {
        Rng rng(playout_seed(seed, (uint32_t)i));
        rng.next();
        rng.next();
        GameState tmp = st;
        const float score = playout(tmp, rng);
        if (score == 0.0f)
            bwins += 0.5f;
        else if (score > 0.0f)
            bwins += 1.0f;
    }
    float w = n > 0 ? bwins / (float)n : 0.5f;
    if (st.to_move() != BLACK) w = 1.0f - w;
    return w;
}

}  // namespace go
}  // namespace lz
