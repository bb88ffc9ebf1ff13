// lz_internal.h — shared declarations of the HIP backend (not part of the
// C ABI). Kernels are compiled for gfx950 only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lz {

// Largest board side the kernels accept (the reference nets are 19x19 only,
// Network.cpp:541-543,591-594,627-629).
constexpr int kMaxSide = 19;

// ---- FP32 plain form (conv_fp32.hip) --------------------------------------
// Output-channel block of one FP32 conv workgroup; packed weights are laid
// out [cout/64][cin*k*k][64] with zero rows past cout.
constexpr int kF32MBlock = 64;

// K rows of a packed FP32 conv: the input channels are staged in chunks of
// 8 (5x5) or 16 (3x3) channels, so cin is rounded up to whole chunks and
// the extra rows are zero.
__host__ __device__ inline int f32_packed_k(int filter, int cin) {
    const int ck = filter == 5 ? 8 : 16;
    return (cin + ck - 1) / ck * ck * filter * filter;
}

// Input planes per board the first conv may take (the reference's nets
// take 32) and the widest trunk conv of the FP32 form.
constexpr int kMaxInputPlanes = 64;
constexpr int kMaxTrunkChannels = 512;

// Conv activations (LZ_ACT_*).
constexpr int kActElu = 0, kActRelu = 1, kActNone = 2;

// What the FP32 conv epilogue applies after the bias: the activation, then
// (if bn_mean is set) bn_inv_std[o] * (v - bn_mean[o]).
struct ConvEpilogue {
    int act = kActElu;
    const float *bn_mean = nullptr;
    const float *bn_inv_std = nullptr;
};

hipError_t conv_f32(int filter, bool u8_input, const void *in,
                    const float *wpk, const float *bias, float *out, int n,
                    int cin, int cout, int S, const ConvEpilogue &epi,
                    hipStream_t stream);

// Last 1-channel conv + activation + softmax(T): raw/out [n][S*S].
hipError_t policy_head_f32(const float *in, const float *w, const float *b,
                           float *raw, float *out, int n, int cin, int S,
                           float temperature, int act, hipStream_t stream);

// Last 1-channel conv + activation of the value net into h0 [n][S*S].
hipError_t value_headconv_f32(const float *in, const float *w, const float *b,
                              float *h0, int n, int cin, int S, int act,
                              hipStream_t stream);

// Value inner-product head over a batch: FC (P -> hidden) + ELU, FC
// (hidden -> 1), (1 + tanh) / 2. fc1_wt is the first FC transposed:
// [P][hidden]. Shared by every precision (the FCs stay FP32). n_dev (if not
// null) is a device-side board count that caps n when the kernel runs (a
// graph captured for n boards replayed on fewer).
hipError_t value_fc_head(const float *h0, const float *fc1_wt,
                         const float *fc1_b, const float *fc2_w,
                         const float *fc2_b, float *raw, float *out, int n,
                         int P, int hidden, hipStream_t stream,
                         const int *n_dev = nullptr);

}  // namespace lz
