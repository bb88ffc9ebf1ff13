// lz_enc_record.cpp — packing a game state into a device board record.
#include "lz_enc_record.h"

#include <cmath>
#include <cstring>

namespace lz {

void pack_record(const go::GameState &st, int64_t offset, EncRecord *r) {
    const go::Board &b = st.board;
    const int W = b.width(), G = W * W;
    std::memcpy(r->sq, b.squares(), G);
    std::memcpy(r->head, b.heads(), sizeof(uint16_t) * G);
    std::memset(r->sq + G, go::OFFBOARD, kEncGrid - G);
    std::memset(r->head + G, 0, sizeof(uint16_t) * (kEncGrid - G));
    r->side = b.side();
    r->to_move = b.to_move();
    r->komove = st.komove;
    r->last_move = st.last_move;
    r->prev_last = st.prev_last;
    r->white_has_komi = !(std::fabs(st.komi) <= 0.75f);
    r->out_offset = offset;
    r->passes = st.passes;
    r->movenum = st.movenum;
    r->komi = st.score_komi();
    r->last_was_capture = st.last_was_capture ? 1 : 0;
}

}  // namespace lz
