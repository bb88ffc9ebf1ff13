// fused_net.hip — persistent fused layer-chain kernel: all convs of the
// policy (or value) net for one slot of boards per workgroup pass, 16-bit
// operands (BF16 or FP16) on MFMA with FP32 accumulation, activations
// resident in LDS.
//
// Work decomposition
//   * A workgroup owns one slot at a time and loops over slots (persistent
//     grid). A slot is one 19x19 board, or up to four smaller boards of one
//     side (two 13x13, four 9x9: ragged batches). The slot's activations
//     live in LDS as 16-bit values in an HWC tile with zero halos:
//     boards are stacked vertically with one zero separator row between
//     them, row pitch S + 2 pixels plus a row pad, pixel stride PS = C + 16
//     elements (the LDS bank layout that these choose is given with
//     row_pad below).
//   * Each conv is an implicit GEMM D[cout x pixels] = W[cout x K] B[K x
//     pixels], K = k*k*cin ordered tap-major, 32 channels per MFMA k-step
//     (v_mfma_f32_16x16x32_{bf16,f16}). The waves form a WM (output
//     channels) x 4 (pixels) grid (policy 2 x 4, value 1 x 4); a wave owns
//     1/WM of the output channels of six 16-pixel column tiles
//     (4 x 6 x 16 = 384 >= 361 pixels; two 13x13 boards are 338, four 9x9
//     are 324). Its B fragments are one 16-byte ds_read per lane at
//     pixel_base + tap_offset + channel_block, prefetched one k-step ahead.
//   * Weights are pre-packed on the host in MFMA fragment order, [k-step]
//     [m-tile][lane][8], so an A operand is one coalesced 16-byte global load
//     per lane, prefetched one k-step ahead. The pixel-waves of a channel
//     group walk the same k-steps together, so repeated fragment reads hit
//     in L1 and each CU pulls a layer's weights from L2 once per slot.
//   * Bias (the accumulators' starting value) + ELU in FP32 (Network.cpp:
//     382-392), rounded to the operand type in the epilogue and written
//     back over the input tile after a barrier. The
//     last 1-channel conv keeps FP32: softmax(T) per board over its logits
//     (Network.cpp:450-469) for the policy net; for the value net the ELU'd
//     plane goes to HBM and the batched inner-product head (value_fc_head)
//     finishes it.
#include "lz_fused.h"
#include "lz_internal.h"

namespace lz {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Op;
template <>
struct Op<__bf16> {
    typedef __bf16 v8 __attribute__((ext_vector_type(8)));
    typedef __bf16 v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ f32x4 mfma(v8 a, v8 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
template <>
struct Op<_Float16> {
    typedef _Float16 v8 __attribute__((ext_vector_type(8)));
    typedef _Float16 v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ f32x4 mfma(v8 a, v8 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

constexpr int kWN = 4;                // waves along pixels
constexpr int kNT = 6;                // 16-pixel tiles per wave
constexpr int kP = 361;
constexpr int kInPS = 48;             // pixel stride of the 32-plane input
constexpr int kTilePixels = 456;      // slot capacity in tile pixels

// Boards of `side` one ragged slot holds: as many as fit the 384 pixels of
// the column tiling and kTilePixels tile pixels (1 for 19x19, 2 for 13x13,
// 4 for 9x9), at most 4.
constexpr int slot_capacity(int side) {
    if (side < 5 || side > 19) return 0;
    int k = 1;
    while (k < 4 && (k + 1) * side * side <= 384 &&
           (1 + (k + 1) * (side + 1)) * (side + 2) <= kTilePixels)
        k++;
    return k;
}

// LDS bank layout of the tiles. A 16-byte ds_read (the B fragments of the
// convs) is served per half-wavefront, from 128 4-byte banks: the 32 lanes
// of a half-wave are conflict-free when their 16-byte units are distinct mod
// 32 (fitted to rocprofv3's SQ_LDS_BANK_CONFLICT over a set of lane
// patterns on gfx950). Lanes 0-31 of a fragment read 16 pixels x two
// 16-byte channel groups (units u*pixel + 0 and + 1, u = PS / 8). With u = 2
// mod 4 (PS = C + 16: 18 units for 128 channels, 10 for 64; 6 for the
// 32-plane input at 48) 16 consecutive pixels take the 16 even residues and
// the second channel group the odd ones. A 16-pixel column tile also wraps
// from the end of a board row to the next row, skipping the `halo` zero
// columns; the pad at the end of each row makes the row pitch, in units,
// congruent mod 32 to u times the S pixels of the row, so the pixels after
// a wrap continue the same residue sequence. row_pad is that pad, in
// elements (u * halo + pad = 0 mod 32 units).
constexpr int row_pad(int ps, int halo) {
    return 8 * ((((-halo * (ps / 8)) % 32) + 32) % 32);
}

// LDS elements of a slot's tile: the stem's input (pitch S + 4, 32
// channels at kInPS) or the trunk (pitch S + 2, ps), whichever is larger,
// over the sides a slot can hold.
constexpr int tile_elems(int ps, bool ragged) {
    int best = 0;
    for (int S = ragged ? 5 : 19; S <= 19; S++) {
        const int k = ragged ? slot_capacity(S) : 1;
        const int trunk = (1 + k * (S + 1)) * ((S + 2) * ps + row_pad(ps, 2));
        const int stem =
            (2 + k * (S + 2)) * ((S + 4) * kInPS + row_pad(kInPS, 4));
        best = trunk > best ? trunk : best;
        best = stem > best ? stem : best;
    }
    return best;
}

__device__ __forceinline__ float elu(float v) {
    return v > 0.0f ? v : (expf(v) - 1.0f);
}

// ELU of a trunk output that is about to be rounded to a 16-bit operand:
// exp(v) = 2^(v log2 e) on the hardware base-2 exponential (1 ulp), so
// exp(v) - 1 carries the same ~1e-7 absolute error as with the library
// expf, far below the 16-bit rounding that follows, without expf's
// range reduction, which dominated the epilogue's VALU time.
__device__ __forceinline__ float elu_operand(float v) {
    return v > 0.0f ? v
                    : __builtin_amdgcn_exp2f(v * 1.44269504088896341f) - 1.0f;
}

// One same-padded conv of the tile `act` (row pitch rp elements, pixel
// stride PS, CIN input channels) into acc[MTW][NT]: m-tiles mt0..mt0+MTW-1 (of MT in the
// packed weights) for the wave's column tiles; pb[t] is the lane's window
// origin in column tile t.
template <typename T, int F, int CIN, int MT, int MTW, int PS,
          int NT = kNT>
__device__ __forceinline__ void conv_tile(const T *act,
                                          const typename Op<T>::v8 *__restrict__ wpk,
                                          f32x4 (&acc)[MTW][NT],
                                          const int (&pb)[NT], int rp,
                                          int mt0, int lane,
                                          const float *__restrict__ bias =
                                              nullptr) {
    typedef typename Op<T>::v8 v8;
    constexpr int CB = CIN / 32;
    constexpr int STEPS = F * F * CB;
    const int g8 = (lane >> 4) * 8;
    // The accumulators start at the bias (when given): the epilogue then
    // has one add less per output.
#pragma unroll
    for (int m = 0; m < MTW; m++) {
        f32x4 init = {0.f, 0.f, 0.f, 0.f};
        if (bias) {
            const int o0 = (mt0 + m) * 16 + (lane >> 4) * 4;
#pragma unroll
            for (int r = 0; r < 4; r++) init[r] = bias[o0 + r];
        }
#pragma unroll
        for (int t = 0; t < NT; t++) acc[m][t] = init;
    }

    // (The empty asm keeps the compiler from hoisting the per-step
    // fragment addresses of a fully unrolled conv, the 1-channel head's,
    // out of the slot loop, where they would stay live, and spill, for the
    // whole kernel.)
    int lanev = lane;
    asm volatile("" : "+v"(lanev));
    const v8 *wl = wpk + mt0 * 64 + lanev;
    // Operands of k-step s, clamped to the last step (the loads issued past
    // the end re-read it and go unused).
    auto b_off = [&](int s) {
        s = s < STEPS ? s : STEPS - 1;
        const int tap = s / CB, cb = s % CB;
        return (tap / F) * rp + (tap % F) * PS + cb * 32 + g8;
    };
    auto a_at = [&](int s, int m) {
        s = s < STEPS ? s : STEPS - 1;
        return wl[(s * MT + m) * 64];
    };
    // B fragments (LDS) are loaded one k-step ahead into the other of two
    // register sets; the step loop is unrolled by two so that the sets
    // alternate without copies.
    v8 a[MTW], b0[NT], b1[NT];
#pragma unroll
    for (int m = 0; m < MTW; m++) a[m] = a_at(0, m);
#pragma unroll
    for (int t = 0; t < NT; t++)
        b0[t] = *reinterpret_cast<const v8 *>(act + pb[t] + g8);

    // Step s: acc += a * b while the B fragments of step s + 1 load into
    // bn. The MFMAs run m-tile by m-tile at raised wave priority; each A
    // fragment reloads with step s + 1's right after its last MFMA (a full
    // step of MFMAs ahead of its next use: the loads come from L2), and
    // the NT LDS reads are spread over the first half of the MFMAs.
    auto step = [&](const v8(&b)[NT], v8(&bn)[NT], int s) {
        const int off = b_off(s + 1);
#pragma unroll
        for (int t = 0; t < NT; t++)
            bn[t] = *reinterpret_cast<const v8 *>(act + pb[t] + off);
        __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int m = 0; m < MTW; m++) {
#pragma unroll
            for (int t = 0; t < NT; t++)
                acc[m][t] = Op<T>::mfma(a[m], b[t], acc[m][t]);
            a[m] = a_at(s + 1, m);
        }
        constexpr int kMfma = MTW * NT;
        constexpr int kDsEvery = (kMfma / 2) / NT > 0 ? (kMfma / 2) / NT : 1;
        int ds = 0;
#pragma unroll
        for (int m = 0; m < MTW; m++) {
#pragma unroll
            for (int t = 0; t < NT; t++) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
                if ((m * NT + t) % kDsEvery == kDsEvery - 1 && ds < NT) {
                    __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
                    ds++;
                }
            }
            __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // VMEM read
        }
        __builtin_amdgcn_s_setprio(0);
    };
    for (int s = 0; s < STEPS; s += 2) {
        step(b0, b1, s);
        if (s + 1 < STEPS) step(b1, b0, s + 1);
    }
}

// conv_tile for the layered kernels, whose waves issue few MFMAs per
// k-step (one wave per SIMD, 10 MFMAs per step in a trunk band): fully
// unrolled, the A fragments prefetched PF k-steps ahead so that more of
// their L2 latency is covered, the B fragments one step ahead. Every
// accumulator sees the same MFMAs in the same order as in conv_tile.
// prefetch() issues the bias loads and the first PF steps' A fragments,
// so that a kernel can overlap them with the staging of its input tile;
// run() then does the convolution.
template <typename T, int F, int CIN, int MT, int MTW, int PF>
struct DeepConv {
    typedef typename Op<T>::v8 v8;
    static constexpr int CB = CIN / 32;
    static constexpr int STEPS = F * F * CB;
    const v8 *wl;
    v8 a[PF][MTW];
    f32x4 init[MTW];

    __device__ __forceinline__ void prefetch(const v8 *__restrict__ wpk,
                                             int mt0, int lane,
                                             const float *__restrict__ bias) {
#pragma unroll
        for (int m = 0; m < MTW; m++) {
            init[m] = f32x4{0.f, 0.f, 0.f, 0.f};
            if (bias) {
                const int o0 = (mt0 + m) * 16 + (lane >> 4) * 4;
#pragma unroll
                for (int r = 0; r < 4; r++) init[m][r] = bias[o0 + r];
            }
        }
        wl = wpk + mt0 * 64 + lane;
#pragma unroll
        for (int p = 0; p < PF; p++)
            if (p < STEPS) {
#pragma unroll
                for (int m = 0; m < MTW; m++) a[p][m] = wl[(p * MT + m) * 64];
            }
    }

    template <int PS, int NT>
    __device__ __forceinline__ void run(const T *act, f32x4 (&acc)[MTW][NT],
                                        const int (&pb)[NT], int rp,
                                        int lane) {
        const int g8 = (lane >> 4) * 8;
#pragma unroll
        for (int m = 0; m < MTW; m++)
#pragma unroll
            for (int t = 0; t < NT; t++) acc[m][t] = init[m];
        auto b_off = [&](int s) {
            const int tap = s / CB, cb = s % CB;
            return (tap / F) * rp + (tap % F) * PS + cb * 32 + g8;
        };
        v8 b[2][NT];
#pragma unroll
        for (int t = 0; t < NT; t++)
            b[0][t] = *reinterpret_cast<const v8 *>(act + pb[t] + b_off(0));
#pragma unroll
        for (int s = 0; s < STEPS; s++) {
            if (s + 1 < STEPS) {
#pragma unroll
                for (int t = 0; t < NT; t++)
                    b[(s + 1) & 1][t] = *reinterpret_cast<const v8 *>(
                        act + pb[t] + b_off(s + 1));
            }
#pragma unroll
            for (int m = 0; m < MTW; m++) {
#pragma unroll
                for (int t = 0; t < NT; t++)
                    acc[m][t] =
                        Op<T>::mfma(a[s % PF][m], b[s & 1][t], acc[m][t]);
                if (s + PF < STEPS)
                    a[s % PF][m] = wl[((s + PF) * MT + m) * 64];
            }
        }
    }
};

// ELU of acc (whose bias the conv folded into the accumulator start),
// written into the trunk tile at the lane's pixel offsets po[t] (-1:
// padding column). D layout of 16x16 MFMA: lane holds rows 4*(lane/16)+r
// of column lane%16, i.e. four consecutive output channels of one pixel.
template <typename T, int MTW, int NT = kNT>
__device__ __forceinline__ void store_tile(T *act,
                                           const f32x4 (&acc)[MTW][NT],
                                           int mt0, const int (&po)[NT],
                                           int lane) {
    typedef typename Op<T>::v4 v4;
    const int g4 = (lane >> 4) * 4;
#pragma unroll
    for (int m = 0; m < MTW; m++) {
        const int o0 = (mt0 + m) * 16 + g4;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            if (po[t] < 0) continue;
            v4 v;
#pragma unroll
            for (int r = 0; r < 4; r++) v[r] = (T)elu_operand(acc[m][t][r]);
            *reinterpret_cast<v4 *>(act + po[t] + o0) = v;
        }
    }
}

// A board's policy outputs from its head plane v (SS <= 384 pixels), by
// one wavefront: raw = v and out = softmax(v / T) (Network.cpp:450-469).
// Wave-level reductions only, so the workgroup's other waves go on to the
// next slot meanwhile (its staging and convs do not touch v).
__device__ void wave_softmax(const float *v, int SS, float temperature,
                             float *raw, float *out, int lane) {
    float mx = -INFINITY;
    for (int p = lane; p < SS; p += 64) mx = fmaxf(mx, v[p]);
    for (int off = 32; off > 0; off >>= 1)
        mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    const float alpha = mx / temperature;
    float e[6], sum = 0.f;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const int p = lane + 64 * i;
        e[i] = p < SS ? expf(v[p] / temperature - alpha) : 0.f;
        sum += e[i];
    }
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const int p = lane + 64 * i;
        if (p < SS) {
            if (raw) raw[p] = v[p];
            if (out) out[p] = e[i] / sum;
        }
    }
}

// ---- FP8 e4m3 trunk ---------------------------------------------------------
// Fragments of v_mfma_scale_f32_16x16x128_f8f6f4 (format 0 = e4m3, unit
// block scales): A lane holds 32 k-values of row lane%16, B lane 32 k-values
// of column lane%16, lane group lane/16 the k-range 32*(lane/16)..+31 of
// the 128-wide k-step (weights and activations use the same order).
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v4i __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma_f8(v8i a, v8i b, f32x4 c) {
    return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, c, 0, 0, 0,
                                                            127, 0, 127);
}

constexpr float kE4M3Max = 448.0f;

// conv_tile on an e4m3 tile (pixel stride PS8 bytes): k = tap * CIN + c in
// steps of 128; a lane's 32 k-values are 32 channels of one tap (CIN is a
// multiple of 32). Past K = F*F*CIN the packed weights are zero and the
// lane reads the last tap (finite values).
template <int F, int CIN, int MT, int MTW, int PS8>
__device__ __forceinline__ void conv_tile_f8(const uint8_t *act,
                                             const v8i *__restrict__ wpk,
                                             f32x4 (&acc)[MTW][kNT],
                                             const int (&pb)[kNT], int sp,
                                             int mt0, int lane) {
    constexpr int STEPS = (F * F * CIN + 127) / 128;
    const int g32 = (lane >> 4) * 32;
#pragma unroll
    for (int m = 0; m < MTW; m++)
#pragma unroll
        for (int t = 0; t < kNT; t++) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const v8i *wl = wpk + mt0 * 64 + lane;
    // A fragments are prefetched one k-step ahead; the step loop stays
    // rolled so that the compiler does not hoist every step's loads.
    v8i a[MTW];
#pragma unroll
    for (int m = 0; m < MTW; m++) a[m] = wl[m * 64];
#pragma unroll 1
    for (int s = 0; s < STEPS; s++) {
        const int k0 = s * 128 + g32;
        int tap = k0 / CIN;
        const int c0 = k0 - tap * CIN;
        if (tap >= F * F) tap = F * F - 1;
        const int off = ((tap / F) * sp + (tap % F)) * PS8 + c0;
        const int nxt = s + 1 < STEPS ? s + 1 : s;
        v8i an[MTW];
#pragma unroll
        for (int m = 0; m < MTW; m++) an[m] = wl[(nxt * MT + m) * 64];
#pragma unroll
        for (int t = 0; t < kNT; t++) {
            const v8i b = *reinterpret_cast<const v8i *>(act + pb[t] + off);
#pragma unroll
            for (int m = 0; m < MTW; m++) acc[m][t] = mfma_f8(a[m], b, acc[m][t]);
        }
#pragma unroll
        for (int m = 0; m < MTW; m++) a[m] = an[m];
    }
}

// Epilogue of an e4m3 layer: v = ELU(acc * wscale[o] * s_in + bias[o]) in
// FP32; the board's activation scale s_out = max|v| / 448 (a block
// reduction: every thread must call this); e4m3(v / s_out) is written as
// four bytes (four consecutive channels of one pixel) per accumulator.
// Returns s_out. wscale == nullptr means unit weight scales.
template <int MTW, int kWaves>
__device__ __forceinline__ float store_tile_f8(uint8_t *act,
                                               f32x4 (&acc)[MTW][kNT],
                                               const float *__restrict__ bias,
                                               const float *__restrict__ wscale,
                                               float s_in, int mt0,
                                               const int (&po)[kNT], int lane,
                                               float *s_red) {
    const int g4 = (lane >> 4) * 4;
    float mx = 0.0f;
#pragma unroll
    for (int m = 0; m < MTW; m++) {
        const int o0 = (mt0 + m) * 16 + g4;
        float bv[4], sc[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            bv[r] = bias[o0 + r];
            sc[r] = (wscale ? wscale[o0 + r] : 1.0f) * s_in;
        }
#pragma unroll
        for (int t = 0; t < kNT; t++) {
            if (po[t] < 0) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float v = elu_operand(acc[m][t][r] * sc[r] + bv[r]);
                acc[m][t][r] = v;
                mx = fmaxf(mx, fabsf(v));
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = mx;
    __syncthreads();
    mx = s_red[0];
    for (int i = 1; i < kWaves; i++) mx = fmaxf(mx, s_red[i]);
    const float s_out = mx > 0.0f ? mx / kE4M3Max : 1.0f;
    const float inv = 1.0f / s_out;
#pragma unroll
    for (int m = 0; m < MTW; m++) {
        const int o0 = (mt0 + m) * 16 + g4;
#pragma unroll
        for (int t = 0; t < kNT; t++) {
            if (po[t] < 0) continue;
            float q[4];
#pragma unroll
            for (int r = 0; r < 4; r++)
                q[r] = fminf(fmaxf(acc[m][t][r] * inv, -kE4M3Max), kE4M3Max);
            int packed = __builtin_amdgcn_cvt_pk_fp8_f32(q[0], q[1], 0, false);
            packed = __builtin_amdgcn_cvt_pk_fp8_f32(q[2], q[3], packed, true);
            *reinterpret_cast<int *>(act + po[t] + o0) = packed;
        }
    }
    return s_out;
}

// The FP8 form: the 5x5 stem on FP16 operands (its 32 binary input planes
// are exact in any format), every later conv (trunk, and the 1-channel head
// conv) on e4m3 operands with per-output-channel weight scales
// (args.ws[l]) and a per-board activation scale chosen after each layer
// from the layer's output range. 19x19 boards only (no ragged slots).
// Outputs as fused_net.
template <int C, int M1, int NMID, bool VALUE, int kWM>
__global__ __launch_bounds__(kWM * kWN * 64,
                             !VALUE && kWM == 1 ? 1 : 2) void fused_net_f8(
    const uint8_t *__restrict__ planes, FusedArgs args,
    float *__restrict__ raw, float *__restrict__ out, int nslots,
    float temperature) {
    typedef _Float16 T;
    typedef typename Op<T>::v8 v8;
    constexpr int PS8 = C + 16;
    constexpr int STEM = 23 * 23 * kInPS * (int)sizeof(T);
    constexpr int TRUNK = 21 * 21 * PS8;
    constexpr int BYTES = STEM > TRUNK ? STEM : TRUNK;
    constexpr int kWaves = kWM * kWN;
    constexpr int kThreads = kWaves * 64;
    constexpr int MT = C / 16, MTW = MT / kWM;
    constexpr int MT1 = M1 / 16, MTW1 = MT1 / kWM;
    static_assert(C % 32 == 0 && M1 % 32 == 0 && MT % kWM == 0 &&
                      MT1 % kWM == 0, "channel blocking");

    __shared__ __attribute__((aligned(16))) uint8_t s_raw[BYTES];
    __shared__ float s_vec[384];
    __shared__ float s_red[kWaves];
    T *s_stem = reinterpret_cast<T *>(s_raw);
    uint8_t *s_t8 = s_raw;
    if (args.counts) nslots = min(nslots, args.counts[0]);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / kWN, wn = wave % kWN;
    const int nt0 = wn * kNT;
    constexpr int S = 19, SS = 361, pin = S + 4, pt = S + 2;
    int pb_in[kNT], pb[kNT], po[kNT];
#pragma unroll
    for (int t = 0; t < kNT; t++) {
        const int j = (nt0 + t) * 16 + (lane & 15);
        const int jj = j < SS ? j : 0;
        const int y = jj / S, x = jj % S;
        pb_in[t] = (y * pin + x) * kInPS;
        pb[t] = (y * pt + x) * PS8;
        po[t] = j < SS ? pb[t] + (pt + 1) * PS8 : -1;
    }

    for (int board = blockIdx.x; board < nslots; board += gridDim.x) {
        const uint8_t *src = planes + (size_t)board * 32 * kP;
        for (int i = tid; i < pin * pin * 4; i += kThreads) {
            const int pp = i >> 2, cg = (i & 3) * 8;
            const int yy = pp / pin - 2, xx = pp % pin - 2;
            const bool in = yy >= 0 && yy < S && xx >= 0 && xx < S;
            v8 v;
#pragma unroll
            for (int e = 0; e < 8; e++)
                v[e] = (T)(in ? (float)src[(cg + e) * SS + yy * S + xx] : 0.0f);
            *reinterpret_cast<v8 *>(s_stem + pp * kInPS + cg) = v;
        }
        __syncthreads();

        float s_act;
        {
            f32x4 acc[MTW1][kNT];
            conv_tile<T, 5, 32, MT1, MTW1, kInPS>(
                s_stem, static_cast<const v8 *>(args.w[0]), acc, pb_in,
                pin * kInPS, wm * MTW1, lane);
            __syncthreads();
            for (int i = tid * 16; i < TRUNK; i += kThreads * 16)
                *reinterpret_cast<v4i *>(s_t8 + i) = v4i{0, 0, 0, 0};
            s_act = store_tile_f8<MTW1, kWaves>(s_t8, acc, args.b[0], nullptr,
                                                1.0f, wm * MTW1, po, lane,
                                                s_red);
            __syncthreads();
        }
        {
            f32x4 acc[MTW][kNT];
            conv_tile_f8<3, M1, MT, MTW, PS8>(
                s_t8, static_cast<const v8i *>(args.w[1]), acc, pb, pt,
                wm * MTW, lane);
            __syncthreads();
            s_act = store_tile_f8<MTW, kWaves>(s_t8, acc, args.b[1], args.ws[1],
                                               s_act, wm * MTW, po, lane, s_red);
            __syncthreads();
        }
        for (int l = 2; l <= NMID; l++) {
            f32x4 acc[MTW][kNT];
            conv_tile_f8<3, C, MT, MTW, PS8>(
                s_t8, static_cast<const v8i *>(args.w[l]), acc, pb, pt,
                wm * MTW, lane);
            __syncthreads();
            s_act = store_tile_f8<MTW, kWaves>(s_t8, acc, args.b[l], args.ws[l],
                                               s_act, wm * MTW, po, lane, s_red);
            __syncthreads();
        }
        if (wm == 0) {
            f32x4 acc[1][kNT];
            conv_tile_f8<3, C, 1, 1, PS8>(
                s_t8, static_cast<const v8i *>(args.w[NMID + 1]), acc, pb, pt,
                0, lane);
            const float hb = args.b[NMID + 1][0];
            const float hs = args.ws[NMID + 1][0] * s_act;
            if (lane < 16) {
#pragma unroll
                for (int t = 0; t < kNT; t++) {
                    const int j = (nt0 + t) * 16 + lane;
                    if (j < SS) s_vec[j] = elu(acc[0][t][0] * hs + hb);
                }
            }
        }
        __syncthreads();
        // Outputs by wave 0 alone (no barrier): the next slot's staging and
        // convs do not touch s_vec, and its head conv comes after several
        // barriers that wave 0 reaches only once done here.
        if (wave == 0) {
            if (VALUE) {
                float *h = args.h0 + (size_t)board * kP;
                for (int p = lane; p < kP; p += 64) h[p] = s_vec[p];
            } else {
                const size_t o = (size_t)board * kP;
                wave_softmax(s_vec, SS, temperature, raw ? raw + o : nullptr,
                             out ? out + o : nullptr, lane);
            }
        }
    }
}

// T: operand type; C: trunk width (128 policy, 64 value); M1: outputs of
// the 5x5 stem; NMID: 3x3 convs between stem and head (the first takes M1
// inputs). RAGGED: slots from args.slots (boards of any side up to 19,
// planes and policy outputs at per-board offsets), else slot b = 19x19
// board b with planes at 32*361*b and outputs at 361*b.
// Policy: writes raw = ELU'd logits and out = softmax(raw / T) per board.
// Value: writes each board's ELU'd head plane to h0 [n][361] (a smaller
// board's y*S+x pixel at y*19+x, zeros elsewhere) for value_fc_head.
// kWM waves split the output channels, kWN = 4 the pixels; OCC is the
// waves per SIMD the register budget is sized for.
template <typename T, int C, int M1, int NMID, bool VALUE, bool RAGGED,
          int kWM, int OCC>
__global__ __launch_bounds__(kWM * kWN * 64, OCC) void fused_net(
    const uint8_t *__restrict__ planes, FusedArgs args,
    float *__restrict__ raw, float *__restrict__ out, int nslots,
    float temperature) {
    typedef typename Op<T>::v8 v8;
    constexpr int PS = C + 16;
    static_assert((PS / 8) % 4 == 2 && (kInPS / 8) % 4 == 2,
                  "pixel strides of 2 mod 4 16-byte units");
    constexpr int RPAD = row_pad(PS, 2), RPAD_IN = row_pad(kInPS, 4);
    constexpr int TILE = tile_elems(PS, RAGGED);
    constexpr int kWaves = kWM * kWN;
    constexpr int kThreads = kWaves * 64;
    constexpr int MT = C / 16, MTW = MT / kWM;
    constexpr int MT1 = M1 / 16, MTW1 = MT1 / kWM;
    static_assert(C % 32 == 0 && M1 % 32 == 0 && MT % kWM == 0 &&
                      MT1 % kWM == 0, "channel blocking");

    __shared__ __attribute__((aligned(16))) T s_act[TILE];
    __shared__ float s_vec[384];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / kWN, wn = wave % kWN;
    const int nt0 = wn * kNT;
    auto W = [&](int l) { return static_cast<const v8 *>(args.w[l]); };
    if (args.counts) nslots = min(nslots, args.counts[0]);

    for (int slot = blockIdx.x; slot < nslots; slot += gridDim.x) {
        // Slot geometry: k boards of side S stacked in the tile.
        int S = 19, k = 1;
        int board[4] = {slot, 0, 0, 0};
        if (RAGGED) {
            const FusedSlot sl = args.slots[slot];
            S = sl.side;
            k = sl.count;
#pragma unroll
            for (int i = 0; i < 4; i++) board[i] = sl.board[i];
        }
        const int SS = S * S, npix = k * SS;
        const int pin = S + 4, pt = S + 2;  // stem / trunk row pixels
        const int rp_in = pin * kInPS + RPAD_IN, rp = pt * PS + RPAD;
        int pb_in[kNT], pb[kNT], po[kNT], pix[kNT];
#pragma unroll
        for (int t = 0; t < kNT; t++) {
            const int j = (nt0 + t) * 16 + (lane & 15);
            pix[t] = j < npix ? j : -1;
            // A lane past the last pixel reads the pixel a multiple of 16
            // before it (npix >= 25): the tile keeps its bank pattern.
            const int jj = j < npix ? j : j - ((j - npix) / 16 + 1) * 16;
            const int kb = jj / SS, r = jj % SS, y = r / S, x = r % S;
            pb_in[t] = (kb * (S + 2) + y) * rp_in + x * kInPS;
            pb[t] = (kb * (S + 1) + y) * rp + x * PS;
            po[t] = j < npix ? pb[t] + rp + PS : -1;
        }

        // Stage the 32 binary planes with a 2-wide zero halo.
        const int rows_in = 2 + k * (S + 2);
        for (int i = tid; i < rows_in * pin * 4; i += kThreads) {
            const int pp = i >> 2, cg = (i & 3) * 8;
            const int row = pp / pin - 2, xx = pp % pin - 2;
            const int kb = row >= 0 ? row / (S + 2) : 0;
            const int yy = row >= 0 ? row % (S + 2) : -1;
            const bool in = row >= 0 && kb < k && yy < S && xx >= 0 && xx < S;
            const uint8_t *src =
                planes + (RAGGED ? (size_t)args.in_off[board[kb]]
                                 : (size_t)board[0] * 32 * kP);
            v8 v;
#pragma unroll
            for (int e = 0; e < 8; e++)
                v[e] = (T)(in ? (float)src[(cg + e) * SS + yy * S + xx] : 0.0f);
            *reinterpret_cast<v8 *>(s_act + (row + 2) * rp_in +
                                    (xx + 2) * kInPS + cg) = v;
        }
        __syncthreads();

        // Stem: 5x5, 32 -> M1.
        {
            f32x4 acc[MTW1][kNT];
            conv_tile<T, 5, 32, MT1, MTW1, kInPS>(s_act, W(0), acc, pb_in,
                                                  rp_in, wm * MTW1, lane,
                                                  args.b[0]);
            __syncthreads();
            // Switch to the trunk geometry: clear tile and halos, then write.
            const int trunk = (1 + k * (S + 1)) * rp;
            for (int i = tid * 8; i < trunk; i += kThreads * 8)
                *reinterpret_cast<v8 *>(s_act + i) = v8{};
            __syncthreads();
            store_tile<T, MTW1>(s_act, acc, wm * MTW1, po, lane);
            __syncthreads();
        }
        // Middle 3x3 convs; the first reads M1 channels.
        {
            f32x4 acc[MTW][kNT];
            conv_tile<T, 3, M1, MT, MTW, PS>(s_act, W(1), acc, pb, rp,
                                             wm * MTW, lane, args.b[1]);
            __syncthreads();
            store_tile<T, MTW>(s_act, acc, wm * MTW, po, lane);
            __syncthreads();
        }
        for (int l = 2; l <= NMID; l++) {
            f32x4 acc[MTW][kNT];
            conv_tile<T, 3, C, MT, MTW, PS>(s_act, W(l), acc, pb, rp,
                                            wm * MTW, lane, args.b[l]);
            __syncthreads();
            store_tile<T, MTW>(s_act, acc, wm * MTW, po, lane);
            __syncthreads();
        }
        // Head: 3x3, C -> 1, in FP32; the channel-group-0 waves cover all
        // pixels.
        if (wm == 0) {
            f32x4 acc[1][kNT];
            conv_tile<T, 3, C, 1, 1, PS>(s_act, W(NMID + 1), acc, pb, rp, 0,
                                         lane);
            const float hb = args.b[NMID + 1][0];
            if (lane < 16) {
#pragma unroll
                for (int t = 0; t < kNT; t++)
                    if (pix[t] >= 0) s_vec[pix[t]] = elu(acc[0][t][0] + hb);
            }
        }
        __syncthreads();
        // Outputs by wave 0 alone (no barrier): the next slot's staging and
        // convs do not touch s_vec, and its head conv comes after several
        // barriers that wave 0 reaches only once done here.
        if (wave == 0) {
            for (int kb = 0; kb < k; kb++) {
                const float *v = s_vec + kb * SS;
                if (VALUE) {
                    float *h = args.h0 + (size_t)board[kb] * kP;
                    for (int p = lane; p < kP; p += 64) {
                        const int y = p / 19, x = p % 19;
                        h[p] = (y < S && x < S) ? v[y * S + x] : 0.0f;
                    }
                } else {
                    const size_t o = RAGGED ? (size_t)args.out_off[board[kb]]
                                            : (size_t)board[kb] * kP;
                    wave_softmax(v, SS, temperature, raw ? raw + o : nullptr,
                                 out ? out + o : nullptr, lane);
                }
            }
        }
    }
}

// Policy: one 8-wave workgroup per CU (136 KB LDS tile, 148 KB in the
// ragged build), 2 waves per SIMD. Value: 4-wave workgroups, two per CU
// (75 / 80 KB tile each), 2 waves per SIMD.
// Wave grids the autotuner chooses from. Policy: 2 x 4 (8 waves, 96
// accumulators per lane, 2 waves/SIMD; default) or 1 x 4 (4 waves, 192
// accumulators, 1 wave/SIMD). Value: 1 x 4 (4 waves, two workgroups per CU;
// default) or 2 x 4 (8 waves).
template <typename T, bool RAGGED, int WM>
struct PolicyKernel {
    static constexpr auto fn =
        fused_net<T, 128, 96, 11, false, RAGGED, WM, WM == 2 ? 2 : 1>;
};
template <typename T, bool RAGGED, int WM>
struct ValueKernel {
    static constexpr auto fn = fused_net<T, 64, 64, 10, true, RAGGED, WM, 2>;
};

template <typename T, bool RAGGED, int WM>
hipError_t launch_wm(int which, const uint8_t *planes, const FusedArgs &args,
                     float *raw, float *out, int nslots, int nboards,
                     float temperature, int grid, hipStream_t stream) {
    const int threads = WM * kWN * 64;
    if (which == 0) {
        if (args.nconv != 13 || args.c != 128 || args.m1 != 96)
            return hipErrorInvalidValue;
        hipLaunchKernelGGL((PolicyKernel<T, RAGGED, WM>::fn), dim3(grid),
                           dim3(threads), 0, stream, planes, args, raw, out,
                           nslots, temperature);
        return hipGetLastError();
    }
    if (args.nconv != 12 || args.c != 64 || args.m1 != 64 || !args.h0)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL((ValueKernel<T, RAGGED, WM>::fn), dim3(grid),
                       dim3(threads), 0, stream, planes, args, raw, out,
                       nslots, temperature);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return value_fc_head(args.h0, args.fc1_wt, args.fc1_b, args.fc2_w,
                         args.fc2_b, raw, out, nboards, kP, args.hidden,
                         stream, args.counts ? args.counts + 1 : nullptr);
}

template <int WM>
hipError_t launch_f8_wm(int which, const uint8_t *planes, const FusedArgs &args,
                        float *raw, float *out, int n, float temperature,
                        int grid, hipStream_t stream) {
    const int threads = WM * kWN * 64;
    if (which == 0) {
        if (args.nconv != 13 || args.c != 128 || args.m1 != 96)
            return hipErrorInvalidValue;
        hipLaunchKernelGGL((fused_net_f8<128, 96, 11, false, WM>), dim3(grid),
                           dim3(threads), 0, stream, planes, args, raw, out, n,
                           temperature);
        return hipGetLastError();
    }
    if (args.nconv != 12 || args.c != 64 || args.m1 != 64 || !args.h0)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL((fused_net_f8<64, 64, 10, true, WM>), dim3(grid),
                       dim3(threads), 0, stream, planes, args, raw, out, n,
                       temperature);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return value_fc_head(args.h0, args.fc1_wt, args.fc1_b, args.fc2_w,
                         args.fc2_b, raw, out, n, kP, args.hidden, stream,
                         args.counts ? args.counts + 1 : nullptr);
}

template <typename T, bool RAGGED>
hipError_t launch(int which, int wm, const uint8_t *planes,
                  const FusedArgs &args, float *raw, float *out, int nslots,
                  int nboards, float temperature, int grid,
                  hipStream_t stream) {
    if (grid <= 0 || grid > nslots) grid = nslots;
    return wm == 1 ? launch_wm<T, RAGGED, 1>(which, planes, args, raw, out,
                                             nslots, nboards, temperature,
                                             grid, stream)
                   : launch_wm<T, RAGGED, 2>(which, planes, args, raw, out,
                                             nslots, nboards, temperature,
                                             grid, stream);
}

// ---- Layered path: small batches of 19x19 boards ---------------------------
// The persistent kernel runs a board's whole layer chain on one CU (~0.2 ms
// for the policy net), which is the latency of a small leaf batch whatever
// its size. The layered path spreads each conv of a board over several
// workgroups (groups of kLMT output m-tiles), one launch per conv, with the
// activations between convs in global memory in the LDS tile layout
// (zero halos written once, never overwritten). Every output is formed by
// the same MFMA sequence from the same operands, the same bias start,
// activation and rounding as in fused_net, so the two paths agree bit for
// bit.
constexpr int kLS = 19;
constexpr int kLMT = 2;  // output m-tiles per workgroup
// A prefetch distance of the layered convs, in k-steps. Measured at 1 / 16
// / 64 boards (policy, FP16): 2 steps 97 / 113 / 155 us per forward; 1, 3,
// 4 and 6 steps 118-122 / 141 / 186 (the value net 93 / 100 / 117 for
// every distance).
constexpr int kLayeredPF = 2;
template <int PS>
constexpr int lrp() { return (kLS + 2) * PS + row_pad(PS, 2); }
constexpr int lrp_in() { return (kLS + 4) * kInPS + row_pad(kInPS, 4); }
// A board's trunk tile, rounded up to whole KiB (512 16-bit elements) so
// that it moves as KiB chunks.
template <int PS>
constexpr int ltile() {
    return ((1 + (kLS + 1)) * lrp<PS>() + 511) / 512 * 512;
}
constexpr int ltile_in() { return (2 + (kLS + 2)) * lrp_in(); }
constexpr int kLPlaneBytes = 32 * kP;  // a board's input planes

typedef __attribute__((address_space(3))) void lds_void_t;
typedef const __attribute__((address_space(1))) void gbl_void_t;

// Copies `chunks` KiB from global memory to LDS (both 16-byte aligned) with
// direct global-to-LDS loads, 16 bytes per lane (no registers, all in
// flight at once); the waves take chunks round robin. The caller waits
// (vmcnt(0)) and synchronises the workgroup.
__device__ __forceinline__ void copy_to_lds(const void *src, void *dst,
                                            int chunks, int wave, int lane) {
    const char *g = static_cast<const char *>(src) + lane * 16;
    char *l = static_cast<char *>(dst);
    for (int c = wave; c < chunks; c += kWN)
        __builtin_amdgcn_global_load_lds((gbl_void_t *)(g + c * 1024),
                                         (lds_void_t *)(l + c * 1024), 16, 0,
                                         0);
}

__device__ __forceinline__ void wait_lds_copies() {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// fused_net's per-lane pixel offsets for one 19x19 board (k = 1).
template <int PS>
__device__ __forceinline__ void lgeom(int nt0, int lane, int (&pb_in)[kNT],
                                      int (&pb)[kNT], int (&po)[kNT],
                                      int (&pix)[kNT]) {
    constexpr int S = kLS, SS = S * S, npix = SS;
    constexpr int rp_in = lrp_in(), rp = lrp<PS>();
#pragma unroll
    for (int t = 0; t < kNT; t++) {
        const int j = (nt0 + t) * 16 + (lane & 15);
        pix[t] = j < npix ? j : -1;
        const int jj = j < npix ? j : j - ((j - npix) / 16 + 1) * 16;
        const int y = jj / S, x = jj % S;
        pb_in[t] = y * rp_in + x * kInPS;
        pb[t] = y * rp + x * PS;
        po[t] = j < npix ? pb[t] + rp + PS : -1;
    }
}

__device__ __forceinline__ int layered_boards(int n, const int *counts) {
    return counts ? min(n, counts[1]) : n;
}

// Stem: 5x5, 32 planes -> M1, written into the trunk tile of the board.
// bits (optional): the planes packed one word per point, read instead.
template <typename T, int PS, int M1>
__global__ __launch_bounds__(kWN * 64) void layered_stem(
    const uint8_t *__restrict__ planes, const uint32_t *__restrict__ bits,
    const void *w, const float *__restrict__ bias, T *__restrict__ out_tiles,
    int n, const int *counts) {
    typedef typename Op<T>::v8 v8;
    constexpr int MT1 = M1 / 16, G = MT1 / kLMT;
    constexpr int S = kLS, pin = S + 4, rp_in = lrp_in();
    static_assert(MT1 % kLMT == 0, "stem channel groups");
    const int board = blockIdx.x / G, cg = blockIdx.x % G;
    if (board >= layered_boards(n, counts)) return;
    __shared__ __attribute__((aligned(16))) T s_act[ltile_in()];
    __shared__ __attribute__((aligned(16))) uint8_t s_pl[kLPlaneBytes];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // The board's planes into LDS: whole KiB chunks directly, the 288-byte
    // tail through registers (no read past the board).
    const uint8_t *src = planes + (size_t)board * kLPlaneBytes;
    const uint32_t *wb = bits ? bits + (size_t)board * kP : nullptr;
    DeepConv<T, 5, 32, MT1, kLMT, kLayeredPF> dc;
    constexpr int kFull = kLPlaneBytes / 1024, kTail = kLPlaneBytes % 1024;
    static_assert(kTail % 16 == 0 && kLPlaneBytes % 16 == 0, "plane copy");
    if (!wb) {
        copy_to_lds(src, s_pl, kFull, wave, lane);
        if (tid < kTail / 16)
            reinterpret_cast<uint4 *>(s_pl + kFull * 1024)[tid] =
                reinterpret_cast<const uint4 *>(src + kFull * 1024)[tid];
        dc.prefetch(static_cast<const v8 *>(w), cg * kLMT, lane, bias);
    }
    int pb_in[kNT], pb[kNT], po[kNT], pix[kNT];
    lgeom<PS>(wave * kNT, lane, pb_in, pb, po, pix);
    wait_lds_copies();
    __syncthreads();
    // The 32 binary planes with a 2-wide zero halo (as fused_net stages
    // them).
    constexpr int rows_in = 2 + (S + 2);
    if (wb) {
        // Packed rows (host memory): each thread's words are all loaded
        // before any is used, one round trip instead of one per pixel.
        constexpr int kPix = rows_in * pin, kIter = (kPix + 255) / 256;
        static_assert(kWN * 64 == 256, "stem staging");
        uint32_t word[kIter];
#pragma unroll
        for (int it = 0; it < kIter; it++) {
            const int pp = tid + it * 256;
            const int row = pp / pin - 2, xx = pp % pin - 2;
            const bool in =
                pp < kPix && row >= 0 && row < S && xx >= 0 && xx < S;
            word[it] = in ? wb[row * S + xx] : 0u;
        }
        // (after the words: waiting for them does not wait for these)
        dc.prefetch(static_cast<const v8 *>(w), cg * kLMT, lane, bias);
#pragma unroll
        for (int it = 0; it < kIter; it++) {
            const int pp = tid + it * 256;
            if (pp >= kPix) continue;
            const int row = pp / pin - 2, xx = pp % pin - 2;
#pragma unroll
            for (int cg8 = 0; cg8 < 32; cg8 += 8) {
                v8 v;
#pragma unroll
                for (int e = 0; e < 8; e++)
                    v[e] = (T)(float)((word[it] >> (cg8 + e)) & 1u);
                *reinterpret_cast<v8 *>(s_act + (row + 2) * rp_in +
                                        (xx + 2) * kInPS + cg8) = v;
            }
        }
    } else {
        for (int i = tid; i < rows_in * pin * 4; i += kWN * 64) {
            const int pp = i >> 2, cg8 = (i & 3) * 8;
            const int row = pp / pin - 2, xx = pp % pin - 2;
            const bool in = row >= 0 && row < S && xx >= 0 && xx < S;
            v8 v;
#pragma unroll
            for (int e = 0; e < 8; e++)
                v[e] = (T)(in ? (float)s_pl[(cg8 + e) * kP + row * S + xx]
                              : 0.0f);
            *reinterpret_cast<v8 *>(s_act + (row + 2) * rp_in +
                                    (xx + 2) * kInPS + cg8) = v;
        }
    }
    __syncthreads();
    f32x4 acc[kLMT][kNT];
    dc.template run<kInPS, kNT>(s_act, acc, pb_in, rp_in, lane);
    store_tile<T, kLMT>(out_tiles + (size_t)board * ltile<PS>(), acc,
                        cg * kLMT, po, lane);
}

// Row bands: a 3x3 conv of a board is split into kBands workgroups of
// kBandRows image rows (the last band the rest), each staging only its rows
// and their two halo rows of the input tile and computing every output
// channel of its pixels: four waves over the m-tiles, kBandNT column tiles
// of 16 of the band's pixels each.
constexpr int kBandRows = 4;
constexpr int kBands = (kLS + kBandRows - 1) / kBandRows;
constexpr int kBandNT = (kBandRows * kLS + 15) / 16;

// Band b's pixel offsets: window origins in the staged rows (pb), output
// offsets in the board's global tile (po, -1 past the band), pixel index
// on the board (pix, -1 past the band).
template <int PS>
__device__ __forceinline__ void band_geom(int b, int lane,
                                          int (&pb)[kBandNT],
                                          int (&po)[kBandNT],
                                          int (&pix)[kBandNT]) {
    constexpr int S = kLS, rp = lrp<PS>();
    const int y0 = b * kBandRows;
    const int rows = min(kBandRows, S - y0), npix = rows * S;
#pragma unroll
    for (int t = 0; t < kBandNT; t++) {
        const int j = t * 16 + (lane & 15);
        // A lane past the band reads a pixel a multiple of 16 before it.
        const int jj = j < npix ? j : j - ((j - npix) / 16 + 1) * 16;
        const int yl = jj / S, x = jj % S;
        pb[t] = yl * rp + x * PS;
        po[t] = j < npix ? (y0 + yl + 1) * rp + (x + 1) * PS : -1;
        pix[t] = j < npix ? (y0 + yl) * S + x : -1;
    }
}

// Stages band b's rows y0 - 1 .. y0 + rows of the board's tile (global
// rows y0 .. y0 + rows + 1) into LDS: whole KiB chunks directly, the tail
// through registers.
template <typename T, int PS>
__device__ __forceinline__ void band_stage(const T *tile, int b, T *s_act,
                                           int tid, int wave, int lane) {
    constexpr int rp = lrp<PS>();
    const int y0 = b * kBandRows;
    const int rows = min(kBandRows, kLS - y0) + 2;
    const char *src = reinterpret_cast<const char *>(tile + (size_t)y0 * rp);
    const int bytes = rows * rp * (int)sizeof(T);
    const int full = bytes / 1024, tail = bytes % 1024;
    copy_to_lds(src, s_act, full, wave, lane);
    if (tid * 16 < tail)
        reinterpret_cast<uint4 *>(reinterpret_cast<char *>(s_act) +
                                  full * 1024)[tid] =
            reinterpret_cast<const uint4 *>(src + full * 1024)[tid];
}

// One 3x3 conv CIN -> C of the trunk for one band of one board.
template <typename T, int PS, int CIN, int C>
__global__ __launch_bounds__(kWN * 64) void layered_conv(
    const T *__restrict__ in_tiles, const void *w,
    const float *__restrict__ bias, T *__restrict__ out_tiles, int n,
    const int *counts) {
    typedef typename Op<T>::v8 v8;
    constexpr int MT = C / 16, MTW = MT / kWN, TILE = ltile<PS>();
    static_assert(MT % kWN == 0 && TILE % 512 == 0, "trunk tiling");
    static_assert((lrp<PS>() * (int)sizeof(T)) % 16 == 0 &&
                      (kBandRows + 2) * lrp<PS>() * (int)sizeof(T) % 1024 <=
                          kWN * 64 * 16,
                  "band staging");
    const int board = blockIdx.x / kBands, b = blockIdx.x % kBands;
    if (board >= layered_boards(n, counts)) return;
    __shared__ __attribute__((aligned(16))) T s_act[(kBandRows + 2) * lrp<PS>()];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    band_stage<T, PS>(in_tiles + (size_t)board * TILE, b, s_act, tid, wave,
                      lane);
    DeepConv<T, 3, CIN, MT, MTW, kLayeredPF> dc;
    dc.prefetch(static_cast<const v8 *>(w), wave * MTW, lane, bias);
    int pb[kBandNT], po[kBandNT], pix[kBandNT];
    band_geom<PS>(b, lane, pb, po, pix);
    wait_lds_copies();
    __syncthreads();
    f32x4 acc[MTW][kBandNT];
    dc.template run<PS, kBandNT>(s_act, acc, pb, lrp<PS>(), lane);
    store_tile<T, MTW, kBandNT>(out_tiles + (size_t)board * TILE, acc,
                                wave * MTW, po, lane);
}

// The 1-channel head conv (FP32 accumulation, ELU) of one band of one
// board, by wave 0, into plane [n][361].
template <typename T, int PS, int C>
__global__ __launch_bounds__(kWN * 64) void layered_head(
    const T *__restrict__ in_tiles, const void *w,
    const float *__restrict__ hbias, float *__restrict__ plane, int n,
    const int *counts) {
    typedef typename Op<T>::v8 v8;
    constexpr int TILE = ltile<PS>();
    const int board = blockIdx.x / kBands, b = blockIdx.x % kBands;
    if (board >= layered_boards(n, counts)) return;
    __shared__ __attribute__((aligned(16))) T s_act[(kBandRows + 2) * lrp<PS>()];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    band_stage<T, PS>(in_tiles + (size_t)board * TILE, b, s_act, tid, wave,
                      lane);
    DeepConv<T, 3, C, 1, 1, kLayeredPF> dc;
    if (wave == 0)
        dc.prefetch(static_cast<const v8 *>(w), 0, lane, nullptr);
    int pb[kBandNT], po[kBandNT], pix[kBandNT];
    band_geom<PS>(b, lane, pb, po, pix);
    wait_lds_copies();
    __syncthreads();
    if (wave != 0) return;
    f32x4 acc[1][kBandNT];
    dc.template run<PS, kBandNT>(s_act, acc, pb, lrp<PS>(), lane);
    const float hb = hbias[0];
    if (lane < 16) {
#pragma unroll
        for (int t = 0; t < kBandNT; t++)
            if (pix[t] >= 0)
                plane[(size_t)board * kP + pix[t]] = elu(acc[0][t][0] + hb);
    }
}

// The policy head of one board, whole: the 1-channel conv (FP32
// accumulation, ELU) over the board's tile staged in full, by the four
// waves as in fused_net's head (six column tiles each), then softmax(T) by
// wave 0 (Network.cpp:450-469), raw logits and probabilities straight to
// the outputs. One launch in place of a banded head and a softmax.
template <typename T, int PS, int C>
__global__ __launch_bounds__(kWN * 64) void layered_policy_out(
    const T *__restrict__ in_tiles, const void *w,
    const float *__restrict__ hbias, float *__restrict__ raw,
    float *__restrict__ out, int n, const int *counts, float temperature) {
    typedef typename Op<T>::v8 v8;
    constexpr int TILE = ltile<PS>();
    const int board = blockIdx.x;
    if (board >= layered_boards(n, counts)) return;
    __shared__ __attribute__((aligned(16))) T s_act[TILE];
    __shared__ float s_vec[kP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    copy_to_lds(in_tiles + (size_t)board * TILE, s_act,
                TILE * (int)sizeof(T) / 1024, wave, lane);
    DeepConv<T, 3, C, 1, 1, kLayeredPF> dc;
    dc.prefetch(static_cast<const v8 *>(w), 0, lane, nullptr);
    int pb_in[kNT], pb[kNT], po[kNT], pix[kNT];
    lgeom<PS>(wave * kNT, lane, pb_in, pb, po, pix);
    wait_lds_copies();
    __syncthreads();
    f32x4 acc[1][kNT];
    dc.template run<PS, kNT>(s_act, acc, pb, lrp<PS>(), lane);
    const float hb = hbias[0];
    if (lane < 16) {
#pragma unroll
        for (int t = 0; t < kNT; t++)
            if (pix[t] >= 0) s_vec[pix[t]] = elu(acc[0][t][0] + hb);
    }
    __syncthreads();
    if (wave == 0) {
        const size_t o = (size_t)board * kP;
        wave_softmax(s_vec, kP, temperature, raw ? raw + o : nullptr,
                     out ? out + o : nullptr, lane);
    }
}

template <typename T>
hipError_t layered_run(int which, const uint8_t *planes, const FusedArgs &a,
                       float *raw, float *out, int n, float temperature,
                       void *tiles, hipStream_t s, const uint32_t *bits) {
    const dim3 blk(kWN * 64);
    // Two policy tiles per board, then two value tiles: each net keeps its
    // own layout's zero halos.
    T *pol = static_cast<T *>(tiles);
    T *val = pol + 2 * (size_t)kLayeredMaxBoards * ltile<144>();
    T *buf[2];
    if (which == 0) {
        buf[0] = pol;
        buf[1] = pol + (size_t)kLayeredMaxBoards * ltile<144>();
    } else {
        buf[0] = val;
        buf[1] = val + (size_t)kLayeredMaxBoards * ltile<80>();
    }
    const int *cnt = a.counts;
    if (which == 0) {
        if (a.nconv != 13 || a.c != 128 || a.m1 != 96)
            return hipErrorInvalidValue;
        hipLaunchKernelGGL((layered_stem<T, 144, 96>), dim3(n * 3), blk, 0, s,
                           planes, bits, a.w[0], a.b[0], buf[0], n, cnt);
        hipLaunchKernelGGL((layered_conv<T, 144, 96, 128>), dim3(n * kBands),
                           blk, 0, s, buf[0], a.w[1], a.b[1], buf[1], n, cnt);
        for (int l = 2; l <= 11; l++)
            hipLaunchKernelGGL((layered_conv<T, 144, 128, 128>),
                               dim3(n * kBands), blk, 0, s, buf[(l + 1) & 1],
                               a.w[l], a.b[l], buf[l & 1], n, cnt);
        hipLaunchKernelGGL((layered_policy_out<T, 144, 128>), dim3(n), blk, 0,
                           s, buf[1], a.w[12], a.b[12], raw, out, n, cnt,
                           temperature);
        return hipGetLastError();
    }
    if (a.nconv != 12 || a.c != 64 || a.m1 != 64 || !a.h0)
        return hipErrorInvalidValue;  // (h0: the value head planes)
    hipLaunchKernelGGL((layered_stem<T, 80, 64>), dim3(n * 2), blk, 0, s,
                       planes, bits, a.w[0], a.b[0], buf[0], n, cnt);
    for (int l = 1; l <= 10; l++)
        hipLaunchKernelGGL((layered_conv<T, 80, 64, 64>), dim3(n * kBands),
                           blk, 0, s, buf[(l + 1) & 1], a.w[l], a.b[l],
                           buf[l & 1], n, cnt);
    hipLaunchKernelGGL((layered_head<T, 80, 64>), dim3(n * kBands), blk, 0, s,
                       buf[0], a.w[11], a.b[11], a.h0, n, cnt);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return value_fc_head(a.h0, a.fc1_wt, a.fc1_b, a.fc2_w, a.fc2_b, raw, out,
                         n, kP, a.hidden, s, cnt ? cnt + 1 : nullptr);
}

}  // namespace

int fused_default_wm(int which) { return which == 0 ? 2 : 1; }

int fused_blocks_per_cu(int which, int wm) {
    // The ragged instantiation has the larger LDS tile of the two.
    int blocks = 0;
    hipError_t e;
    const int threads = wm * kWN * 64;
    if (which == 0)
        e = wm == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &blocks, PolicyKernel<__bf16, true, 1>::fn, threads, 0)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &blocks, PolicyKernel<__bf16, true, 2>::fn, threads, 0);
    else
        e = wm == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &blocks, ValueKernel<__bf16, true, 1>::fn, threads, 0)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &blocks, ValueKernel<__bf16, true, 2>::fn, threads, 0);
    return (e == hipSuccess && blocks > 0) ? blocks : 1;
}

int ragged_slot_capacity(int side) { return slot_capacity(side); }

size_t layered_tile_bytes() {
    return 2 * (size_t)kLayeredMaxBoards * (ltile<144>() + ltile<80>()) *
           sizeof(uint16_t);
}

hipError_t fused_forward_layered(int which, FusedType type,
                                 const uint8_t *planes, const FusedArgs &args,
                                 float *raw, float *out, int n,
                                 float temperature, void *tiles,
                                 hipStream_t stream, const uint32_t *bits) {
    if (n <= 0) return hipSuccess;
    if (n > kLayeredMaxBoards || !tiles || args.slots ||
        type == FusedType::kFP8)
        return hipErrorInvalidValue;
    return type == FusedType::kBF16
               ? layered_run<__bf16>(which, planes, args, raw, out, n,
                                     temperature, tiles, stream, bits)
               : layered_run<_Float16>(which, planes, args, raw, out, n,
                                       temperature, tiles, stream, bits);
}

hipError_t fused_forward(int which, FusedType type, const uint8_t *planes,
                         const FusedArgs &args, float *raw, float *out, int n,
                         float temperature, const FusedLaunch &cfg,
                         hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const bool ragged = args.slots != nullptr;
    const int nslots = ragged ? args.nslots : n;
    const int wm = cfg.wm == 1 || cfg.wm == 2 ? cfg.wm : fused_default_wm(which);
    int grid = cfg.grid;
    if (type == FusedType::kFP8) {
        if (ragged) return hipErrorInvalidValue;
        if (grid <= 0 || grid > n) grid = n;
        return wm == 1 ? launch_f8_wm<1>(which, planes, args, raw, out, n,
                                         temperature, grid, stream)
                       : launch_f8_wm<2>(which, planes, args, raw, out, n,
                                         temperature, grid, stream);
    }
    if (type == FusedType::kBF16)
        return ragged ? launch<__bf16, true>(which, wm, planes, args, raw, out,
                                             nslots, n, temperature, grid,
                                             stream)
                      : launch<__bf16, false>(which, wm, planes, args, raw,
                                              out, nslots, n, temperature,
                                              grid, stream);
    return ragged ? launch<_Float16, true>(which, wm, planes, args, raw, out,
                                           nslots, n, temperature, grid, stream)
                  : launch<_Float16, false>(which, wm, planes, args, raw, out,
                                            nslots, n, temperature, grid,
                                            stream);
}

}  // namespace lz
