// fused_bf16.hip — persistent fused layer-chain kernel: the whole policy
// (or value) net of one board per workgroup pass, BF16 on MFMA with FP32
// accumulation, activations resident in LDS.
//
// Work decomposition
//   * A workgroup (4 waves) owns one board at a time and loops over boards
//     (persistent grid). The board's activations live in LDS as BF16 in an
//     HWC tile with a zero halo: [SP][SP][PS], SP = 19 + 2*pad, PS = C + 8
//     (the 8-element pad staggers pixels across LDS banks).
//   * Each conv is an implicit GEMM D[cout x 361] = W[cout x K] B[K x 361],
//     K = k*k*cin ordered tap-major, 32 channels per MFMA k-step
//     (v_mfma_f32_16x16x32_bf16). A wave owns all output channels of six
//     16-pixel column tiles (4 waves x 6 x 16 = 384 >= 361), so every
//     B fragment is read from LDS exactly once per layer: one 16-byte
//     ds_read per lane at pixel_base + tap_offset + channel_block.
//   * Weights are pre-packed on the host in MFMA fragment order, [k-step]
//     [m-tile][lane][8], so the A operand is one coalesced 16-byte global
//     load per lane, prefetched one k-step ahead. All waves of a workgroup
//     walk the same k-steps together, so the 4 reads of each fragment hit
//     in L1 and each CU pulls every layer's weights from L2 once per board.
//   * Epilogue: bias + ELU in FP32 (Network.cpp:382-392), rounded to BF16,
//     written back over the input tile after a barrier. The last
//     1-channel conv keeps FP32 logits for softmax(T) (Network.cpp:450-469)
//     or feeds the FP32 value head (ip13 + ELU, ip14, (1 + tanh) / 2,
//     Network.cpp:733-736).
#include "lz_fused.h"

namespace lz {
namespace {

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kWaves = 4;
constexpr int kNT = 6;  // 16-pixel tiles per wave
constexpr int kP = 361;
constexpr int kInPS = 40;  // pixel stride of the 32-plane input tile

__device__ __forceinline__ float elu(float v) {
    return v > 0.0f ? v : (expf(v) - 1.0f);
}

// Offset (in bf16 elements) of pixel j's window origin in a tile of side SP
// and pixel stride PS. Padding columns (j >= 361) alias pixel 0.
template <int SP, int PS>
__device__ __forceinline__ int pix_base(int j) {
    if (j >= kP) j = 0;
    return ((j / 19) * SP + (j % 19)) * PS;
}

// One same-padded conv of the tile `act` (side SP, stride PS, CIN input
// channels) into acc[MT][kNT]: the wave's column tiles, all MT*16 rows.
template <int F, int CIN, int MT, int SP, int PS>
__device__ __forceinline__ void conv_tile(const bf16 *act,
                                          const bf16x8 *__restrict__ wpk,
                                          f32x4 (&acc)[MT][kNT],
                                          const int (&pb)[kNT], int lane) {
    constexpr int CB = CIN / 32;
    constexpr int STEPS = F * F * CB;
    const int g8 = (lane >> 4) * 8;
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int t = 0; t < kNT; t++) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};

    // One A register per m-tile: refilled with the next k-step's fragment
    // as soon as the current step's MFMAs on it have issued.
    bf16x8 a[MT];
#pragma unroll
    for (int m = 0; m < MT; m++) a[m] = wpk[m * 64 + lane];

    for (int s = 0; s < STEPS; s++) {
        const int tap = s / CB, cb = s % CB;
        const int off = ((tap / F) * SP + (tap % F)) * PS + cb * 32 + g8;
        bf16x8 b[kNT];
#pragma unroll
        for (int t = 0; t < kNT; t++)
            b[t] = *reinterpret_cast<const bf16x8 *>(act + pb[t] + off);
        const int nxt = s + 1 < STEPS ? s + 1 : s;
#pragma unroll
        for (int m = 0; m < MT; m++) {
#pragma unroll
            for (int t = 0; t < kNT; t++)
                acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    a[m], b[t], acc[m][t], 0, 0, 0);
            a[m] = wpk[(nxt * MT + m) * 64 + lane];
        }
    }
}

// Bias + ELU of acc, written as BF16 into the output tile (side 21, stride
// PSO). D layout of 16x16 MFMA: lane holds rows 4*(lane/16)+r of column
// lane%16, i.e. four consecutive output channels of one pixel.
template <int MT, int PSO>
__device__ __forceinline__ void store_tile(bf16 *act, const f32x4 (&acc)[MT][kNT],
                                           const float *__restrict__ bias,
                                           int cout, int wave, int lane) {
    const int g4 = (lane >> 4) * 4;
#pragma unroll
    for (int m = 0; m < MT; m++) {
        float bv[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int o = m * 16 + g4 + r;
            bv[r] = o < cout ? bias[o] : 0.0f;
        }
#pragma unroll
        for (int t = 0; t < kNT; t++) {
            const int j = (wave * kNT + t) * 16 + (lane & 15);
            if (j >= kP) continue;
            bf16x4 v;
#pragma unroll
            for (int r = 0; r < 4; r++) v[r] = (bf16)elu(acc[m][t][r] + bv[r]);
            const int y = j / 19 + 1, x = j % 19 + 1;
            *reinterpret_cast<bf16x4 *>(act + (y * 21 + x) * PSO + m * 16 + g4) = v;
        }
    }
}

__device__ __forceinline__ const bf16x8 *wfrag(const FusedArgs &a, int l) {
    return static_cast<const bf16x8 *>(a.w[l]);
}

__device__ float block_sum(float v, float *s_red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = 0.0f;
    for (int i = 0; i < kWaves; i++) r += s_red[i];
    return r;
}

__device__ float block_max(float v, float *s_red) {
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = s_red[0];
    for (int i = 1; i < kWaves; i++) r = fmaxf(r, s_red[i]);
    return r;
}

// C: trunk width (128 policy, 64 value); M1: outputs of the 5x5 stem;
// NMID: 3x3 convs between stem and head (the first takes M1 inputs).
template <int C, int M1, int NMID, bool VALUE>
__global__ __launch_bounds__(kThreads, 1) void fused_net_bf16(
    const uint8_t *__restrict__ planes, FusedArgs args, float *__restrict__ raw,
    float *__restrict__ out, int n, float temperature) {
    constexpr int PS = C + 8;
    constexpr int SP = 21;
    constexpr int TILE = SP * SP * PS;         // trunk tile (bf16 elements)
    constexpr int TILE0 = 23 * 23 * kInPS;     // stem input tile
    constexpr int TILE_MAX = TILE > TILE0 ? TILE : TILE0;
    constexpr int MT = C / 16;
    constexpr int MT1 = (M1 + 15) / 16;
    static_assert(C % 32 == 0 && M1 % 32 == 0, "channel blocking");

    __shared__ __attribute__((aligned(16))) bf16 s_act[TILE_MAX];
    __shared__ float s_vec[384];
    __shared__ float s_red[kWaves];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    int pb_in[kNT], pb[kNT];
#pragma unroll
    for (int t = 0; t < kNT; t++) {
        const int j = (wave * kNT + t) * 16 + (lane & 15);
        pb_in[t] = pix_base<23, kInPS>(j);
        pb[t] = pix_base<SP, PS>(j);
    }

    for (int board = blockIdx.x; board < n; board += gridDim.x) {
        // Stage the 32 binary planes as BF16 with a 2-wide zero halo.
        const uint8_t *src = planes + (size_t)board * 32 * kP;
        for (int i = tid; i < 23 * 23 * 4; i += kThreads) {
            const int pp = i >> 2, cg = (i & 3) * 8;
            const int y = pp / 23 - 2, x = pp % 23 - 2;
            bf16x8 v;
            const bool in = y >= 0 && y < 19 && x >= 0 && x < 19;
#pragma unroll
            for (int e = 0; e < 8; e++)
                v[e] = (bf16)(in ? (float)src[(cg + e) * kP + y * 19 + x] : 0.0f);
            *reinterpret_cast<bf16x8 *>(s_act + pp * kInPS + cg) = v;
        }
        __syncthreads();

        // Stem: 5x5, 32 -> M1.
        {
            f32x4 acc[MT1][kNT];
            conv_tile<5, 32, MT1, 23, kInPS>(s_act, wfrag(args, 0), acc, pb_in, lane);
            __syncthreads();
            // Switch to the trunk geometry: clear tile and halo, then write.
            for (int i = tid * 8; i < TILE; i += kThreads * 8)
                *reinterpret_cast<bf16x8 *>(s_act + i) = bf16x8{};
            __syncthreads();
            store_tile<MT1, PS>(s_act, acc, args.b[0], M1, wave, lane);
            __syncthreads();
        }
        // Middle 3x3 convs; the first reads M1 channels.
        {
            f32x4 acc[MT][kNT];
            conv_tile<3, M1, MT, SP, PS>(s_act, wfrag(args, 1), acc, pb, lane);
            __syncthreads();
            store_tile<MT, PS>(s_act, acc, args.b[1], C, wave, lane);
            __syncthreads();
        }
        for (int l = 2; l <= NMID; l++) {
            f32x4 acc[MT][kNT];
            conv_tile<3, C, MT, SP, PS>(s_act, wfrag(args, l), acc, pb, lane);
            __syncthreads();
            store_tile<MT, PS>(s_act, acc, args.b[l], C, wave, lane);
            __syncthreads();
        }
        // Head: 3x3, C -> 1, kept in FP32.
        {
            f32x4 acc[1][kNT];
            conv_tile<3, C, 1, SP, PS>(s_act, wfrag(args, NMID + 1), acc, pb, lane);
            const float hb = args.b[NMID + 1][0];
            if (lane < 16) {
#pragma unroll
                for (int t = 0; t < kNT; t++) {
                    const int j = (wave * kNT + t) * 16 + lane;
                    if (j < kP) s_vec[j] = elu(acc[0][t][0] + hb);
                }
            }
            __syncthreads();
        }
        if (!VALUE) {
            float mx = -INFINITY;
            for (int p = tid; p < kP; p += kThreads) mx = fmaxf(mx, s_vec[p]);
            mx = block_max(mx, s_red);
            const float alpha = mx / temperature;
            float e[2] = {0.f, 0.f}, sum = 0.f;
            for (int i = 0, p = tid; p < kP; p += kThreads, i++) {
                e[i] = expf(s_vec[p] / temperature - alpha);
                sum += e[i];
            }
            sum = block_sum(sum, s_red);
            for (int i = 0, p = tid; p < kP; p += kThreads, i++) {
                if (raw) raw[(size_t)board * kP + p] = s_vec[p];
                if (out) out[(size_t)board * kP + p] = e[i] / sum;
            }
        } else {
            float part = 0.0f;
            for (int o = tid; o < args.hidden; o += kThreads) {
                float a = 0.0f;
                for (int i = 0; i < kP; i++)
                    a += args.fc1_wt[(size_t)i * args.hidden + o] * s_vec[i];
                part += args.fc2_w[o] * elu(a + args.fc1_b[o]);
            }
            const float v = block_sum(part, s_red) + args.fc2_b[0];
            if (tid == 0) {
                if (raw) raw[board] = v;
                if (out) out[board] = (1.0f + tanhf(v)) / 2.0f;
            }
        }
        __syncthreads();
    }
}

}  // namespace

int fused_blocks_per_cu(int which) {
    int blocks = 0;
    hipError_t e =
        which == 0
            ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &blocks, fused_net_bf16<128, 96, 11, false>, kThreads, 0)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &blocks, fused_net_bf16<64, 64, 10, true>, kThreads, 0);
    return (e == hipSuccess && blocks > 0) ? blocks : 1;
}

hipError_t fused_forward_bf16(int which, const uint8_t *planes,
                              const FusedArgs &args, float *raw, float *out,
                              int n, float temperature, int grid,
                              hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    if (grid <= 0 || grid > n) grid = n;
    if (which == 0) {
        if (args.nconv != 13 || args.c != 128 || args.m1 != 96)
            return hipErrorInvalidValue;
        hipLaunchKernelGGL((fused_net_bf16<128, 96, 11, false>), dim3(grid),
                           dim3(kThreads), 0, stream, planes, args, raw, out,
                           n, temperature);
    } else {
        if (args.nconv != 12 || args.c != 64 || args.m1 != 64)
            return hipErrorInvalidValue;
        hipLaunchKernelGGL((fused_net_bf16<64, 64, 10, true>), dim3(grid),
                           dim3(kThreads), 0, stream, planes, args, raw, out,
                           n, temperature);
    }
    return hipGetLastError();
}

}  // namespace lz
