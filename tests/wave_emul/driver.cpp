// Host driver of the device encoder kernel under the wavefront emulation
// (hip/hip_runtime.h beside this file): replays boards with the library's
// board engine, packs them into records exactly as lz_encode_device does,
// and runs encoder.hip's kernel source on the CPU. Test infrastructure.
#include <cstdint>
#include <cstring>
#include <vector>

#include "go_board.h"
#include "input_planes.h"
#include "lz_encoder.h"
#include "playout.h"

using lz::go::GameState;

namespace {

void run(const std::vector<GameState> &boards, int net, int sym,
         uint8_t *out, uint8_t *out2 = nullptr) {
    std::vector<lz::EncRecord> rec(boards.size());
    int64_t off = 0;
    for (size_t i = 0; i < boards.size(); i++) {
        lz::pack_record(boards[i], off, &rec[i]);
        const int S = boards[i].board.side();
        off += (int64_t)lz::kPlanes * S * S;
    }
    std::vector<uint64_t> ovf((size_t)lz::kEncStrOvf * lz::kEncStrWords *
                              boards.size());
    if (out2)  // both layouts from one launch
        lz::launch_encode(rec.data(), (int)rec.size(), sym, out, out2,
                          ovf.data(), nullptr);
    else
        lz::launch_encode(rec.data(), (int)rec.size(), sym,
                          net == 0 ? out : nullptr, net == 1 ? out : nullptr,
                          ovf.data(), nullptr);
}

}  // namespace

extern "C" {

// Planes of the board after `moves` (passes as -1) on a side x side board.
int emul_encode_moves(int side, int n, const int32_t *moves, int net, int sym,
                      uint8_t *out) {
    std::vector<GameState> b(1);
    b[0].init(side, 7.5f);
    for (int i = 0; i < n; i++) b[0].play(moves[i]);
    run(b, net, sym, out);
    return 0;
}

// Generator positions seeds seed0.. (sides[i], max_moves[i]) in one
// ragged batch; out gets the kernel's planes, ref the host encoder's.
int emul_encode_generated(int n, const int32_t *sides, const uint32_t *seeds,
                          const int32_t *max_moves, int net, int sym,
                          uint8_t *out, uint8_t *ref) {
    std::vector<GameState> b(n);
    int64_t off = 0;
    std::vector<uint8_t> tmp;
    for (int i = 0; i < n; i++) {
        lz::go::random_position(b[i], sides[i], seeds[i], max_moves[i]);
        const int S = sides[i];
        tmp.assign((size_t)lz::kPlanes * S * S, 0);
        if (net == 0)
            lz::gather_features_policy(b[i], tmp.data());
        else
            lz::gather_features_value(b[i], tmp.data());
        lz::rotate_planes(tmp.data(), sym, S, ref + off);
        off += (int64_t)lz::kPlanes * S * S;
    }
    run(b, net, sym, out);
    return 0;
}

// Playout win rates of generator positions (sides[i], seeds[i],
// max_moves[i]), nplay playouts each, position i seeded seed0 + i: out
// from the kernel, ref from the host's mc_winrate.
int emul_playouts(int n, const int32_t *sides, const uint32_t *seeds,
                  const int32_t *max_moves, int nplay, uint32_t seed0,
                  float *out, float *ref) {
    std::vector<GameState> b(n);
    std::vector<lz::EncRecord> rec(n);
    for (int i = 0; i < n; i++) {
        lz::go::random_position(b[i], sides[i], seeds[i], max_moves[i]);
        lz::pack_record(b[i], 0, &rec[i]);
        ref[i] = lz::go::mc_winrate(b[i], nplay, seed0 + (uint32_t)i);
    }
    std::vector<uint32_t> half(n);
    lz::launch_playouts(rec.data(), n, nplay, seed0, half.data(), out, nullptr);
    return 0;
}

// Both layouts of the board after `moves` from one launch.
int emul_encode_moves_both(int side, int n, const int32_t *moves, int sym,
                           uint8_t *policy, uint8_t *value) {
    std::vector<GameState> b(1);
    b[0].init(side, 7.5f);
    for (int i = 0; i < n; i++) b[0].play(moves[i]);
    run(b, 0, sym, policy, value);
    return 0;
}

}  // extern "C"
