// lz_surrogate.h — device kernels of the learned playout surrogate (not part
// of the C ABI). The network:
//
//   a0 = elu(conv3x3(planes) + b0)                       32 -> 64
//   block k = 1..6: u = elu(conv(h) + b), h' = elu(conv(u) + b + h)
//   s  = elu(wh . h6 + bh)    (1x1, 64 -> 1)
//   f  = elu(F1 s + c1)       (361 -> 64)
//   t  = tanh(F2 . f + c2),   win rate (1 + t) / 2
//
// Activations and conv gradients live in HBM as BF16 HWC maps [n][361][64];
// parameters, gradients and Adam moments are FP32 (flat vector: per conv
// W[64][cin][3][3] then b[64]; then wh[64], bh, F1[64][361], c1[64], F2[64],
// c2). Convs run on MFMA with BF16 operands and FP32 accumulation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lz {
namespace sg {

constexpr int kSide = 19;
constexpr int kP = kSide * kSide;
constexpr int kCh = 64;
constexpr int kNConv = 13;
constexpr int kHid = 64;
// Head parameters in flat order: wh[64], bh, F1[64][361], c1[64], F2[64], c2.
constexpr int kHeadParams = kCh + 1 + kHid * kP + kHid + kHid + 1;

inline int conv_cin(int layer) { return layer == 0 ? 32 : kCh; }
inline int conv_weights(int layer) { return kCh * conv_cin(layer) * 9; }

// Packed MFMA A-operand weights of one conv, forward or (64 -> 64 only)
// backward-data (transposed, taps flipped): 9 * cin * 64 BF16 values.
hipError_t pack_conv(const float *w, uint16_t *wpk, int cin, bool bwd,
                     hipStream_t stream);

// One conv over n boards. Forward (bwd = false): out = elu(conv(in) + bias
// + res). Backward-data (bwd = true, cin 64): out = (conv(in) + res) masked
// where mask > 0. res and mask may be null. in is u8 planes [n][32][361]
// when cin == 32, else BF16 [n][361][64].
hipError_t conv(const void *in, int cin, const uint16_t *wpk,
                const float *bias, const uint16_t *res, const uint16_t *mask,
                uint16_t *out, int n, bool bwd, hipStream_t stream);

// Weight and bias gradient of one conv: gW (OIHW) and gb are written (not
// accumulated). e: BF16 [n][361][64] gradient at the conv's pre-activation;
// in: the conv's input as for conv(). part must hold
// wgrad_part_floats(cin, n) floats.
size_t wgrad_part_floats(int cin, int n);
hipError_t wgrad(const void *in, int cin, const uint16_t *e, float *part,
                 float *gW, float *gb, int n, hipStream_t stream);

// Head forward: s [n][361], f [n][64], t [n] kept for the backward pass;
// win [n] = (1 + t) / 2 (may be null).
hipError_t head_fwd(const uint16_t *h, const float *hp, float *s, float *f,
                    float *t, float *win, int n, hipStream_t stream);

// Loss mean_n (t - (2y - 1))^2, head gradients (into ghp, written), the
// BF16 gradient e [n][361][64] at the last conv's pre-activation, and the
// loss into *loss (device). part must hold head_part_floats(n) floats.
size_t head_part_floats(int n);
hipError_t head_bwd(const uint16_t *h, const float *hp, const float *s,
                    const float *f, const float *t, const float *y,
                    uint16_t *e, float *part, float *ghp, float *loss, int n,
                    hipStream_t stream);

// Adam (Kingma & Ba 2015, Algorithm 1) over count parameters; c1 = 1 -
// beta1^step, c2 = 1 - beta2^step.
hipError_t adam(float *p, const float *g, float *m, float *v, int count,
                float lr, float beta1, float beta2, float eps, float c1,
                float c2, hipStream_t stream);

}  // namespace sg
}  // namespace lz
