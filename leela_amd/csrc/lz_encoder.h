// lz_encoder.h — device feature encoder: per-board records in, the nets'
// uint8 input planes out (Network::gather_features_policy/value,
// Network.cpp:883-1201, computed on the GPU).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lz_enc_record.h"
#include "mc_core.h"

namespace lz {

// Planes of n boards (uint8 [32][S*S] each at out + out_offset) for the
// policy net's layout into d_policy and the value net's into d_value
// (either may be NULL), rotated by `symmetry` (0..7, Network.cpp:765-773).
// d_ovf: n * kEncStrOvf * kEncStrWords words of device memory, the
// overflow strings of boards with more strings than the kernel keeps in
// LDS (lz_enc_record.h).
hipError_t launch_encode(const EncRecord *d_rec, int n, int symmetry,
                         uint8_t *d_policy, uint8_t *d_value, uint64_t *d_ovf,
                         hipStream_t stream);

// Light playouts (playout.cpp's policy) of n boards, nplay each, playout j
// of board p seeded playout_seed(seed0 + p, j): d_winrate[p] = mc_winrate
// of the board for its side to move. d_half_wins (n entries) is scratch.
hipError_t launch_playouts(const EncRecord *d_rec, int npos, int nplay,
                           uint32_t seed0, uint32_t *d_half_wins,
                           float *d_winrate, hipStream_t stream);

// Pattern playouts (mc_core.h's policy, the reference's) of n boards,
// nplay each, playout j of board p seeded playout_seed(seed0 + p, j):
// d_winrate[p] = mc_winrate for the side to move, NaN if one of its
// playouts outgrew the device's candidate list. d_states holds nstates
// playout states (a multiple of 64, pattern_state_bytes() each): the
// persistent grid's threads. d_half_wins and d_overflow (n entries each)
// are scratch.
size_t pattern_state_bytes();
hipError_t launch_pattern_playouts(const EncRecord *d_rec, int npos,
                                   int nplay, uint32_t seed0,
                                   const mc::TablesView &tables,
                                   void *d_states, int nstates,
                                   uint32_t *d_half_wins,
                                   uint32_t *d_overflow, float *d_winrate,
                                   hipStream_t stream);

}  // namespace lz
