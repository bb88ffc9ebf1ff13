// conv_fp32.hip — the FP32 "plain" form of Leela's policy and value nets.
//
// One workgroup computes one board x 64 output channels of a same-padded
// conv as an implicit GEMM on FP32 MFMA (v_mfma_f32_16x16x4_f32):
//   D[cout x pixels] = W[cout x cin*k*k] * im2col(X)[cin*k*k x pixels]
// which is the product the reference forms explicitly with im2col + sgemm
// (Network.cpp:344-393, Im2Col.h:9-50). The input board is staged in LDS
// one channel chunk at a time with a zero halo, so the im2col matrix is
// never materialised: a B-operand element is one LDS read at
// chunk_offset(c, ky, kx) + pixel_offset(y, x). Channels past cin in the
// last chunk stage as zeros (their packed weight rows are zero too), so any
// input channel count works. Bias, the activation (ELU, Network.cpp:
// 382-392; or ReLU / none for Caffe nets) and an optional batch
// normalisation of the activated output (Network.cpp:427-448) are fused
// into the epilogue.
#include "lz_internal.h"

namespace lz {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;           // 4 waves
constexpr int kNTiles = 24;             // 16-pixel tiles: 384 >= 361
constexpr int kTilesPerWave = kNTiles / 4;
constexpr int kMTiles = kF32MBlock / 16;

__device__ __forceinline__ float elu(float v) {
    return v > 0.0f ? v : (expf(v) - 1.0f);
}

__device__ __forceinline__ float activate(float v, int act) {
    if (act == kActElu) return elu(v);
    if (act == kActRelu) return v > 0.0f ? v : 0.0f;
    return v;
}

template <int F>
struct ConvF32Cfg {
    static constexpr int CK = (F == 5) ? 8 : 16;  // input channels per chunk
    static constexpr int KC = CK * F * F;         // K rows per chunk
    static constexpr int SP = kMaxSide + 2 * (F / 2);
};

template <int F, typename InT>
__global__ __launch_bounds__(kThreads) void conv_f32_kernel(
    const InT *__restrict__ in, const float *__restrict__ wpk,
    const float *__restrict__ bias, float *__restrict__ out, int cin,
    int cout, int S, ConvEpilogue epi) {
    using C = ConvF32Cfg<F>;
    constexpr int PAD = F / 2;
    __shared__ float s_in[C::CK * C::SP * C::SP];
    __shared__ f32x4 s_w4[C::KC * kF32MBlock / 4];
    __shared__ int s_koff[C::KC];
    const float *s_w = reinterpret_cast<const float *>(s_w4);

    const int board = blockIdx.x;
    const int mb = blockIdx.y;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int P = S * S;
    const int SPs = S + 2 * PAD;  // padded side of this board
    const int plane = SPs * SPs;

    // Offsets of the (c, ky, kx) rows of one chunk inside s_in.
    for (int i = tid; i < C::KC; i += kThreads) {
        const int c = i / (F * F), r = i % (F * F);
        s_koff[i] = c * plane + (r / F) * SPs + (r % F);
    }
    // Offset of this lane's pixel in each of its wave's N tiles. Columns
    // past the board read in-bounds garbage and are never stored.
    int poff[kTilesPerWave];
#pragma unroll
    for (int t = 0; t < kTilesPerWave; t++) {
        const int j = (wave * kTilesPerWave + t) * 16 + (lane & 15);
        poff[t] = j < P ? (j / S) * SPs + (j % S) : 0;
    }

    f32x4 acc[kMTiles][kTilesPerWave];
#pragma unroll
    for (int m = 0; m < kMTiles; m++)
#pragma unroll
        for (int t = 0; t < kTilesPerWave; t++) acc[m][t] = f32x4{0, 0, 0, 0};

    const InT *src = in + (size_t)board * cin * P;
    const int K = f32_packed_k(F, cin);  // cin rounded up to whole chunks
    const f32x4 *wsrc4 = reinterpret_cast<const f32x4 *>(
        wpk + (size_t)mb * K * kF32MBlock);
    const int kq = lane >> 4;
    const int ml = lane & 15;

    for (int c0 = 0; c0 < cin; c0 += C::CK) {
        __syncthreads();
        for (int i = tid; i < C::CK * plane; i += kThreads) {
            const int c = i / plane, r = i % plane;
            const int y = r / SPs - PAD, x = r % SPs - PAD;
            float v = 0.0f;
            if (y >= 0 && y < S && x >= 0 && x < S && c0 + c < cin)
                v = (float)src[(size_t)(c0 + c) * P + y * S + x];
            s_in[i] = v;
        }
        const f32x4 *wchunk = wsrc4 + (size_t)c0 * F * F * kF32MBlock / 4;
        for (int i = tid; i < C::KC * kF32MBlock / 4; i += kThreads)
            s_w4[i] = wchunk[i];
        __syncthreads();

        for (int ks = 0; ks < C::KC; ks += 4) {
            const int k = ks + kq;
            const int ko = s_koff[k];
            float a[kMTiles], b[kTilesPerWave];
#pragma unroll
            for (int m = 0; m < kMTiles; m++)
                a[m] = s_w[k * kF32MBlock + m * 16 + ml];
#pragma unroll
            for (int t = 0; t < kTilesPerWave; t++) b[t] = s_in[ko + poff[t]];
#pragma unroll
            for (int m = 0; m < kMTiles; m++)
#pragma unroll
                for (int t = 0; t < kTilesPerWave; t++)
                    acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                        a[m], b[t], acc[m][t], 0, 0, 0);
        }
    }

    // D layout of 16x16x4: lane holds rows 4*(lane/16)+r of column lane%16.
    float *dst = out + (size_t)board * cout * P;
#pragma unroll
    for (int m = 0; m < kMTiles; m++) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int o = mb * kF32MBlock + m * 16 + 4 * (lane >> 4) + r;
            if (o >= cout) continue;
            const float bo = bias[o];
            const float bm = epi.bn_mean ? epi.bn_mean[o] : 0.0f;
            const float ba = epi.bn_mean ? epi.bn_inv_std[o] : 1.0f;
#pragma unroll
            for (int t = 0; t < kTilesPerWave; t++) {
                const int j = (wave * kTilesPerWave + t) * 16 + (lane & 15);
                if (j >= P) continue;
                float v = activate(acc[m][t][r] + bo, epi.act);
                // Network.cpp:445: scale_stddiv * (in - mean).
                if (epi.bn_mean) v = ba * (v - bm);
                dst[(size_t)o * P + j] = v;
            }
        }
    }
}

// ---- 1-channel head convs -------------------------------------------------

constexpr int kHeadCK = 16;
constexpr int kHeadSP = kMaxSide + 2;

// Computes the activated 3x3 (cin -> 1) conv of one board into
// s_out[0..P). Each thread owns pixels tid and tid + 256.
__device__ void head_conv3(const float *__restrict__ src,
                           const float *__restrict__ w, float b, int cin,
                           int S, int act, float *s_in, float *s_w,
                           float *s_out) {
    const int tid = threadIdx.x;
    const int P = S * S, SPs = S + 2, plane = SPs * SPs;
    float acc0 = 0.0f, acc1 = 0.0f;
    const int p0 = tid, p1 = tid + kThreads;
    const int o0 = p0 < P ? (p0 / S) * SPs + p0 % S : 0;
    const int o1 = p1 < P ? (p1 / S) * SPs + p1 % S : 0;
    for (int c0 = 0; c0 < cin; c0 += kHeadCK) {
        __syncthreads();
        for (int i = tid; i < kHeadCK * plane; i += kThreads) {
            const int c = i / plane, r = i % plane;
            const int y = r / SPs - 1, x = r % SPs - 1;
            float v = 0.0f;
            if (y >= 0 && y < S && x >= 0 && x < S && c0 + c < cin)
                v = src[(size_t)(c0 + c) * P + y * S + x];
            s_in[i] = v;
        }
        for (int i = tid; i < kHeadCK * 9; i += kThreads)
            s_w[i] = c0 + i / 9 < cin ? w[c0 * 9 + i] : 0.0f;
        __syncthreads();
        for (int c = 0; c < kHeadCK; c++) {
            const float *pl = s_in + c * plane;
#pragma unroll
            for (int ky = 0; ky < 3; ky++)
#pragma unroll
                for (int kx = 0; kx < 3; kx++) {
                    const float wv = s_w[c * 9 + ky * 3 + kx];
                    acc0 += wv * pl[o0 + ky * SPs + kx];
                    acc1 += wv * pl[o1 + ky * SPs + kx];
                }
        }
    }
    if (p0 < P) s_out[p0] = activate(acc0 + b, act);
    if (p1 < P) s_out[p1] = activate(acc1 + b, act);
    __syncthreads();
}

__device__ float block_reduce(float v, bool is_max, float *s_red) {
    for (int off = 32; off > 0; off >>= 1) {
        const float o = __shfl_xor(v, off, 64);
        v = is_max ? fmaxf(v, o) : v + o;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) s_red[wave] = v;
    __syncthreads();
    float r = s_red[0];
    for (int i = 1; i < kThreads / 64; i++)
        r = is_max ? fmaxf(r, s_red[i]) : r + s_red[i];
    return r;
}

__global__ __launch_bounds__(kThreads) void policy_head_f32_kernel(
    const float *__restrict__ in, const float *__restrict__ w,
    const float *__restrict__ b, float *__restrict__ raw,
    float *__restrict__ out, int cin, int S, float temperature, int act) {
    __shared__ float s_in[kHeadCK * kHeadSP * kHeadSP];
    __shared__ float s_w[kHeadCK * 9];
    __shared__ float s_logit[kMaxSide * kMaxSide];
    __shared__ float s_red[kThreads / 64];
    const int board = blockIdx.x, tid = threadIdx.x, P = S * S;
    head_conv3(in + (size_t)board * cin * P, w, b[0], cin, S, act, s_in, s_w,
               s_logit);
    // Network::softmax (Network.cpp:450-469): exp(x/T - max/T) / sum.
    float mx = -INFINITY;
    for (int p = tid; p < P; p += kThreads) mx = fmaxf(mx, s_logit[p]);
    mx = block_reduce(mx, true, s_red);
    const float alpha = mx / temperature;
    float e[2] = {0.0f, 0.0f};
    float sum = 0.0f;
    for (int i = 0, p = tid; p < P; p += kThreads, i++) {
        e[i] = expf(s_logit[p] / temperature - alpha);
        sum += e[i];
    }
    sum = block_reduce(sum, false, s_red);
    for (int i = 0, p = tid; p < P; p += kThreads, i++) {
        if (raw) raw[(size_t)board * P + p] = s_logit[p];
        if (out) out[(size_t)board * P + p] = e[i] / sum;
    }
}

// Last 1-channel conv + activation of the value net, written to h0
// [n][S*S] for value_fc_head.
__global__ __launch_bounds__(kThreads) void value_headconv_f32_kernel(
    const float *__restrict__ in, const float *__restrict__ w,
    const float *__restrict__ b, float *__restrict__ h0, int cin, int S,
    int act) {
    __shared__ float s_in[kHeadCK * kHeadSP * kHeadSP];
    __shared__ float s_w[kHeadCK * 9];
    __shared__ float s_h0[kMaxSide * kMaxSide];
    const int board = blockIdx.x, tid = threadIdx.x, P = S * S;
    head_conv3(in + (size_t)board * cin * P, w, b[0], cin, S, act, s_in, s_w,
               s_h0);
    for (int p = tid; p < P; p += kThreads) h0[(size_t)board * P + p] = s_h0[p];
}

// Batched inner-product head of the value net (Network.cpp:733-736):
// hidden = ELU(W1 h0 + b1), v = w2 . hidden + b2, out = (1 + tanh v) / 2.
// A workgroup takes kFcBoards boards so every W1 element loaded from L2 is
// used kFcBoards times; thread o accumulates output o for all of them. Four
// boards per workgroup keeps a batch of 1024 on 256 workgroups (one per CU).
constexpr int kFcBoards = 4;

__global__ __launch_bounds__(kThreads) void value_fc_kernel(
    const float *__restrict__ h0, const float *__restrict__ w1t,
    const float *__restrict__ b1, const float *__restrict__ w2,
    const float *__restrict__ b2, float *__restrict__ raw,
    float *__restrict__ out, int n, int P, int H,
    const int *__restrict__ n_dev) {
    __shared__ float s_h[kFcBoards][kMaxSide * kMaxSide];
    __shared__ float s_red[kFcBoards][kThreads / 64];
    // A replayed graph sizes its grid for the bucket; the batch's board
    // count is then read from device memory.
    if (n_dev) n = min(n, *n_dev);
    const int b0 = blockIdx.x * kFcBoards, tid = threadIdx.x;
    if (b0 >= n) return;
    const int nb = min(kFcBoards, n - b0);
    for (int i = tid; i < kFcBoards * P; i += kThreads) {
        const int bb = i / P, p = i % P;
        s_h[bb][p] = bb < nb ? h0[(size_t)(b0 + bb) * P + p] : 0.0f;
    }
    __syncthreads();
    float part[kFcBoards];
#pragma unroll
    for (int bb = 0; bb < kFcBoards; bb++) part[bb] = 0.0f;
    for (int o = tid; o < H; o += kThreads) {
        float acc[kFcBoards];
#pragma unroll
        for (int bb = 0; bb < kFcBoards; bb++) acc[bb] = 0.0f;
#pragma unroll 8
        for (int i = 0; i < P; i++) {
            const float wv = w1t[(size_t)i * H + o];
#pragma unroll
            for (int bb = 0; bb < kFcBoards; bb++) acc[bb] += wv * s_h[bb][i];
        }
        const float bo = b1[o], wo = w2[o];
#pragma unroll
        for (int bb = 0; bb < kFcBoards; bb++) part[bb] += wo * elu(acc[bb] + bo);
    }
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int bb = 0; bb < kFcBoards; bb++) {
        float v = part[bb];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) s_red[bb][wave] = v;
    }
    __syncthreads();
    if (tid < nb) {
        float v = b2[0];
        for (int w = 0; w < kThreads / 64; w++) v += s_red[tid][w];
        if (raw) raw[b0 + tid] = v;
        if (out) out[b0 + tid] = (1.0f + tanhf(v)) / 2.0f;
    }
}

template <int F, typename InT>
hipError_t launch_conv(const void *in, const float *wpk, const float *bias,
                       float *out, int n, int cin, int cout, int S,
                       const ConvEpilogue &epi, hipStream_t stream) {
    static_assert(ConvF32Cfg<F>::CK == (F == 5 ? 8 : 16),
                  "chunk size must match f32_packed_k");
    dim3 grid(n, (cout + kF32MBlock - 1) / kF32MBlock);
    hipLaunchKernelGGL((conv_f32_kernel<F, InT>), grid, dim3(kThreads), 0,
                       stream, static_cast<const InT *>(in), wpk, bias, out,
                       cin, cout, S, epi);
    return hipGetLastError();
}

bool act_ok(int act) {
    return act == kActElu || act == kActRelu || act == kActNone;
}

}  // namespace

hipError_t conv_f32(int filter, bool u8_input, const void *in,
                    const float *wpk, const float *bias, float *out, int n,
                    int cin, int cout, int S, const ConvEpilogue &epi,
                    hipStream_t stream) {
    if (S < 1 || S > kMaxSide || n <= 0 || cin < 1 || cout < 1 ||
        !act_ok(epi.act) || (epi.bn_mean && !epi.bn_inv_std))
        return hipErrorInvalidValue;
    if (filter == 3)
        return u8_input ? launch_conv<3, uint8_t>(in, wpk, bias, out, n, cin,
                                                  cout, S, epi, stream)
                        : launch_conv<3, float>(in, wpk, bias, out, n, cin,
                                                cout, S, epi, stream);
    if (filter == 5)
        return u8_input ? launch_conv<5, uint8_t>(in, wpk, bias, out, n, cin,
                                                  cout, S, epi, stream)
                        : launch_conv<5, float>(in, wpk, bias, out, n, cin,
                                                cout, S, epi, stream);
    return hipErrorInvalidValue;
}

hipError_t policy_head_f32(const float *in, const float *w, const float *b,
                           float *raw, float *out, int n, int cin, int S,
                           float temperature, int act, hipStream_t stream) {
    if (S < 1 || S > kMaxSide || n <= 0 || cin < 1 || !act_ok(act))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(policy_head_f32_kernel, dim3(n), dim3(kThreads), 0,
                       stream, in, w, b, raw, out, cin, S, temperature, act);
    return hipGetLastError();
}

hipError_t value_headconv_f32(const float *in, const float *w,
                              const float *b, float *h0, int n, int cin,
                              int S, int act, hipStream_t stream) {
    if (S < 1 || S > kMaxSide || n <= 0 || cin < 1 || !act_ok(act))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(value_headconv_f32_kernel, dim3(n), dim3(kThreads), 0,
                       stream, in, w, b, h0, cin, S, act);
    return hipGetLastError();
}

hipError_t value_fc_head(const float *h0, const float *fc1_wt,
                         const float *fc1_b, const float *fc2_w,
                         const float *fc2_b, float *raw, float *out, int n,
                         int P, int hidden, hipStream_t stream,
                         const int *n_dev) {
    if (n <= 0) return hipSuccess;
    if (P < 1 || P > kMaxSide * kMaxSide || hidden < 1)
        return hipErrorInvalidValue;
    const int grid = (n + kFcBoards - 1) / kFcBoards;
    hipLaunchKernelGGL(value_fc_kernel, dim3(grid), dim3(kThreads), 0, stream,
                       h0, fc1_wt, fc1_b, fc2_w, fc2_b, raw, out, n, P,
                       hidden, n_dev);
    return hipGetLastError();
}

}  // namespace lz
