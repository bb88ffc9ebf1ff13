// lz_fused.h — arguments of the persistent fused layer-chain kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace lz {

constexpr int kMaxFusedConvs = 16;

// One workgroup pass of a ragged batch: up to four boards of one side,
// stacked in the LDS tile (board ids index in_off / out_off / h0 rows).
struct FusedSlot {
    int side;
    int count;
    int board[4];
};

// Boards of `side` one slot holds (0 for unsupported sides): 1 for 19x19,
// 2 for 13x13, 4 for 9x9; limited by the 384-pixel column tiling and the
// 456-pixel LDS tile.
int ragged_slot_capacity(int side);

// Device pointers of one net in the fused kernels' layouts. w[l] holds conv
// l packed (in the operand type) in MFMA fragment order: [k-step][m-tile][lane 0..63][8 values],
// k-steps tap-major with 32 input channels each (see pack_conv_fragments).
struct FusedArgs {
    const void *w[kMaxFusedConvs];
    const float *b[kMaxFusedConvs];
    // FP8 form: per-output-channel weight scales of conv l (l >= 1), the
    // packed e4m3 weights being W / scale (see pack_conv_fp8).
    const float *ws[kMaxFusedConvs] = {};
    const float *fc1_wt = nullptr;  // value head: [361][hidden]
    const float *fc1_b = nullptr;
    const float *fc2_w = nullptr;
    const float *fc2_b = nullptr;
    float *h0 = nullptr;            // value head plane [n][361], per call
    // Ragged batches (per call; slots == nullptr for n 19x19 boards):
    const FusedSlot *slots = nullptr;
    int nslots = 0;
    const int *in_off = nullptr;    // board -> byte offset of its planes
    const int *out_off = nullptr;   // board -> offset of its policy outputs
    // Device-side {slots, boards} of the batch, read when the kernels start
    // (nullptr: the launch's own counts). A graph captured for a bucket's
    // capacity then evaluates exactly the batch it is replayed on.
    const int *counts = nullptr;
    int hidden = 0;
    int nconv = 0;
    int c = 0;   // trunk width
    int m1 = 0;  // outputs of the 5x5 stem
};

// MFMA operand type of the fused kernels: the conv weights are packed in
// it and the LDS activation tile holds it.
// kFP8: FP16 stem, e4m3 operands for every later conv (19x19 only).
enum class FusedType { kBF16 = 0, kFP16 = 1, kFP8 = 2 };

// which: 0 policy (13 convs: 5x5 32->96, 3x3 96->128, 10x 128->128,
// 128->1), 1 value (12 convs: 5x5 32->64, 10x 64->64, 64->1, + 2 FC).
// grid <= 0 launches one workgroup per board.
// Launch shape of a fused forward: wm waves along output channels (1 or 2;
// 0 = the net's default) and the persistent grid (workgroups; <= 0 or more
// than the slots = one workgroup per slot).
struct FusedLaunch {
    int wm = 0;
    int grid = 0;
};

int fused_default_wm(int which);

// Resident workgroups per CU of the fused kernel of net `which` with wave
// grid wm x 4 (LDS and register limited).
int fused_blocks_per_cu(int which, int wm);

hipError_t fused_forward(int which, FusedType type, const uint8_t *planes,
                         const FusedArgs &args, float *raw, float *out, int n,
                         float temperature, const FusedLaunch &cfg,
                         hipStream_t stream);

// Layered form of fused_forward for at most kLayeredMaxBoards 19x19 boards
// (BF16 / FP16, not ragged): one launch per conv, each conv of a board
// split over several workgroups, activations between convs in `tiles`
// (layered_tile_bytes(), zeroed once before first use: the tiles' halos
// must stay zero). Bitwise the same outputs as fused_forward, at a lower
// latency for small batches.
constexpr int kLayeredMaxBoards = 64;
size_t layered_tile_bytes();
// bits (optional): the same boards' planes packed one 32-bit word per
// point, bit c = plane c ([n][361]), which the stem reads instead of the
// 32 x 361 plane bytes (planes of 0 / 1 only).
hipError_t fused_forward_layered(int which, FusedType type,
                                 const uint8_t *planes, const FusedArgs &args,
                                 float *raw, float *out, int n,
                                 float temperature, void *tiles,
                                 hipStream_t stream,
                                 const uint32_t *bits = nullptr);

}  // namespace lz
