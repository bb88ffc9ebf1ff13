// lz_encoder.h — device feature encoder: per-board records in, the nets'
// uint8 input planes out (Network::gather_features_policy/value,
// Network.cpp:883-1201, computed on the GPU).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lz_enc_record.h"

namespace lz {

// Planes of n boards (uint8 [32][S*S] each at out + out_offset), rotated
// by `symmetry` (0..7, Network.cpp:765-773). net: 0 policy, 1 value.
hipError_t launch_encode(const EncRecord *d_rec, int n, int net, int symmetry,
                         uint8_t *d_out, hipStream_t stream);

}  // namespace lz
