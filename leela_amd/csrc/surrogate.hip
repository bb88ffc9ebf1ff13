// surrogate.hip — training and inference kernels of the learned playout
// surrogate (network in lz_surrogate.h).
//
// Convolution (forward and backward-data) — sg_conv
//   One workgroup per board, 4 waves. The input map is staged into LDS as a
//   21x21 BF16 HWC tile with a zero halo (pixel stride cin + 8 staggers
//   pixels across banks). Implicit GEMM D[64 x 368] = A[64 x 9cin] B, K
//   tap-major, 32 channels per v_mfma_f32_16x16x32_bf16 k-step. A (weights)
//   is pre-packed in fragment order [k-step][m-tile][lane][8]: one 16-byte
//   global load per lane (L1/L2 resident). Wave w owns pixel tiles w, w+4,
//   ... (6 or 5 of the 23) for all 64 output channels. The FP32 epilogue
//   adds bias and residual, applies ELU (forward) or multiplies by ELU's
//   derivative taken from the stored forward activation (backward), and
//   stores BF16 quads.
//   Backward-data is the same kernel: dX = conv(dY; W') with
//   W'[c][o][tap] = W[o][c][8 - tap] (taps flipped), packed by pack_conv.
//
// Weight gradient — sg_wgrad
//   dW[o][c][tap] = sum over boards and pixels of e[p][o] x[p + tap][c]:
//   a GEMM whose reduction runs over pixels, so both operands are staged
//   transposed (channel rows, pixel columns) in LDS. Pixels are indexed in
//   a 19x21 frame (two zero columns per row, 399 -> 416 = 13 k-steps), and
//   the input is staged already shifted by the workgroup's tap, so every
//   fragment is one aligned 16-byte ds_read. Grid = 9 taps x board chunks;
//   each workgroup writes its FP32 partial, and a second kernel sums the
//   chunks in a fixed order (deterministic) into OIHW.
//
// Head (ELU activations), loss, Adam: FP32, small kernels (the head is
// 0.2% of the FLOPs).
#include "lz_surrogate.h"

namespace lz {
namespace sg {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

constexpr int S = kSide, P = kP;
constexpr int HS = S + 2, HP = HS * HS;   // halo frame 21 x 21
constexpr int NT = (P + 15) / 16;         // 23 pixel tiles
constexpr int KF = 416;                   // 19 x 21 frame, 13 k-steps of 32
constexpr int KFP = KF + 8;               // LDS row pitch (bank stagger)
constexpr int kThreads = 256;
constexpr int kMaxChunks = 64;

__device__ inline f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// The activation of every layer: ELU, as in the reference nets
// (Network.cpp:387-392). With ReLU, training runs collapsed to a constant
// output (every unit of a layer switched off); ELU keeps a gradient
// everywhere.
__device__ inline float elu(float v) { return v > 0.f ? v : expf(v) - 1.f; }
// d ELU / d v expressed through the output a = ELU(v).
__device__ inline float elu_grad(float a) { return a > 0.f ? 1.f : a + 1.f; }

__device__ inline float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

int chunks_for(int n) { return n < kMaxChunks ? n : kMaxChunks; }

// ---- weight packing --------------------------------------------------------
template <int CIN, bool BWD>
__global__ __launch_bounds__(kThreads) void sg_pack(const float *__restrict__ w,
                                                    __bf16 *__restrict__ wpk) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= 9 * CIN * kCh) return;
    const int e = i & 7, lane = (i >> 3) & 63, mt = (i >> 9) & 3, ks = i >> 11;
    const int k = ks * 32 + (lane >> 4) * 8 + e;
    const int row = mt * 16 + (lane & 15);
    float v;
    if (!BWD) {
        const int tap = k / CIN, c = k % CIN;
        v = w[(row * CIN + c) * 9 + tap];
    } else {
        // Row = forward input channel; k = (tap', forward output channel).
        const int tap = k / kCh, o = k % kCh;
        v = w[(o * kCh + row) * 9 + (8 - tap)];
    }
    wpk[i] = (__bf16)v;
}

// ---- convolution -------------------------------------------------------------
template <int CIN, bool BWD>
__global__ __launch_bounds__(kThreads, 2) void sg_conv(
    const void *__restrict__ in, const bf16x8 *__restrict__ wpk,
    const float *__restrict__ bias, const __bf16 *__restrict__ res,
    const __bf16 *__restrict__ mask, __bf16 *__restrict__ out) {
    constexpr int PS = CIN + 8;
    constexpr int C8 = CIN / 8;
    __shared__ __attribute__((aligned(16))) __bf16 tile[HP * PS];
    const int b = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;

    const bf16x8 zero8 = {};
    for (int i = tid; i < HP * C8; i += kThreads) {
        const int hp = i / C8, c8 = i % C8;
        const int hy = hp / HS, hx = hp % HS;
        if (hy == 0 || hy == HS - 1 || hx == 0 || hx == HS - 1)
            *(bf16x8 *)&tile[hp * PS + c8 * 8] = zero8;
    }
    if (CIN == 32) {
        const uint8_t *pl = (const uint8_t *)in + (size_t)b * 32 * P;
        for (int i = tid; i < 32 * P; i += kThreads) {
            const int c = i / P, p = i % P;
            const int y = p / S, x = p % S;
            tile[((y + 1) * HS + x + 1) * PS + c] = (__bf16)(float)pl[i];
        }
    } else {
        const bf16x8 *src = (const bf16x8 *)in + (size_t)b * P * C8;
        for (int i = tid; i < P * C8; i += kThreads) {
            const int p = i / C8, c8 = i % C8;
            const int y = p / S, x = p % S;
            *(bf16x8 *)&tile[((y + 1) * HS + x + 1) * PS + c8 * 8] = src[i];
        }
    }
    __syncthreads();

    int hp0[6];
#pragma unroll
    for (int jj = 0; jj < 6; jj++) {
        const int px = (w + 4 * jj) * 16 + (lane & 15);
        hp0[jj] = px < P ? (px / S + 1) * HS + px % S + 1 : -1;
    }
    f32x4 acc[6][4];
#pragma unroll
    for (int jj = 0; jj < 6; jj++)
#pragma unroll
        for (int mt = 0; mt < 4; mt++) acc[jj][mt] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int kq = (lane >> 4) * 8;
    constexpr int KS = 9 * CIN / 32;
    for (int ks = 0; ks < KS; ks++) {
        const int tap = (ks * 32) / CIN, c0 = (ks * 32) % CIN;
        const int off = (tap / 3 - 1) * HS + (tap % 3 - 1);
        bf16x8 a[4];
#pragma unroll
        for (int mt = 0; mt < 4; mt++) a[mt] = wpk[(ks * 4 + mt) * 64 + lane];
#pragma unroll
        for (int jj = 0; jj < 6; jj++) {
            if (w + 4 * jj >= NT) continue;
            const int hp = hp0[jj] >= 0 ? hp0[jj] + off : 0;
            const bf16x8 bv = *(const bf16x8 *)&tile[hp * PS + c0 + kq];
#pragma unroll
            for (int mt = 0; mt < 4; mt++) acc[jj][mt] = mfma(a[mt], bv, acc[jj][mt]);
        }
    }

    const int orow = (lane >> 4) * 4;
#pragma unroll
    for (int jj = 0; jj < 6; jj++) {
        const int px = (w + 4 * jj) * 16 + (lane & 15);
        if (w + 4 * jj >= NT || px >= P) continue;
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            const int o = mt * 16 + orow;
            const size_t at = ((size_t)b * P + px) * kCh + o;
            float v[4];
#pragma unroll
            for (int r = 0; r < 4; r++) v[r] = acc[jj][mt][r];
            if (!BWD) {
#pragma unroll
                for (int r = 0; r < 4; r++) v[r] += bias[o + r];
            }
            if (res) {
                const bf16x4 rv = *(const bf16x4 *)(res + at);
#pragma unroll
                for (int r = 0; r < 4; r++) v[r] += (float)rv[r];
            }
            if (!BWD) {
#pragma unroll
                for (int r = 0; r < 4; r++) v[r] = elu(v[r]);
            } else if (mask) {
                const bf16x4 mv = *(const bf16x4 *)(mask + at);
#pragma unroll
                for (int r = 0; r < 4; r++)
                    v[r] *= elu_grad((float)mv[r]);
            }
            bf16x4 ov;
#pragma unroll
            for (int r = 0; r < 4; r++) ov[r] = (__bf16)v[r];
            *(bf16x4 *)(out + at) = ov;
        }
    }
}

// ---- weight gradient -----------------------------------------------------------
template <int CIN>
__global__ __launch_bounds__(kThreads, 1) void sg_wgrad(
    const void *__restrict__ in, const __bf16 *__restrict__ e,
    float *__restrict__ part, int n, int bpc) {
    __shared__ __attribute__((aligned(16))) __bf16 eT[kCh * KFP];
    __shared__ __attribute__((aligned(16))) __bf16 xT[CIN * KFP];
    constexpr int NTC = CIN / 16;
    const int tap = blockIdx.x % 9, chunk = blockIdx.x / 9;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int b0 = chunk * bpc;
    const int b1 = b0 + bpc < n ? b0 + bpc : n;

    f32x4 acc[NTC];
#pragma unroll
    for (int nt = 0; nt < NTC; nt++) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    float dbacc = 0.f;

    for (int b = b0; b < b1; b++) {
        // e^T[o][p'], p' = y * 21 + x in the 19 x 21 frame.
        for (int i = tid; i < KF * 8; i += kThreads) {
            const int pp = i >> 3, o8 = i & 7;
            const int y = pp / HS, x = pp % HS;
            bf16x8 v = {};
            if (y < S && x < S)
                v = *(const bf16x8 *)(e + ((size_t)b * P + y * S + x) * kCh + o8 * 8);
#pragma unroll
            for (int k = 0; k < 8; k++) eT[(o8 * 8 + k) * KFP + pp] = v[k];
        }
        // x^T[c][p'] = input at (y + dy, x + dx), zero off the board.
        if (CIN == 32) {
            const uint8_t *pl = (const uint8_t *)in + (size_t)b * 32 * P;
            for (int i = tid; i < 32 * KF; i += kThreads) {
                const int c = i / KF, pp = i % KF;
                const int y = pp / HS + dy, x = pp % HS + dx;
                float v = 0.f;
                if (pp / HS < S && pp % HS < S && y >= 0 && y < S && x >= 0 && x < S)
                    v = (float)pl[c * P + y * S + x];
                xT[c * KFP + pp] = (__bf16)v;
            }
        } else {
            for (int i = tid; i < KF * 8; i += kThreads) {
                const int pp = i >> 3, c8 = i & 7;
                const int y = pp / HS + dy, x = pp % HS + dx;
                bf16x8 v = {};
                if (pp / HS < S && pp % HS < S && y >= 0 && y < S && x >= 0 && x < S)
                    v = *(const bf16x8 *)((const __bf16 *)in +
                                          ((size_t)b * P + y * S + x) * kCh + c8 * 8);
#pragma unroll
                for (int k = 0; k < 8; k++) xT[(c8 * 8 + k) * KFP + pp] = v[k];
            }
        }
        __syncthreads();
        for (int ks = 0; ks < KF / 32; ks++) {
            const int k0 = ks * 32 + (lane >> 4) * 8;
            const bf16x8 a = *(const bf16x8 *)&eT[(w * 16 + (lane & 15)) * KFP + k0];
#pragma unroll
            for (int nt = 0; nt < NTC; nt++) {
                const bf16x8 bv = *(const bf16x8 *)&xT[(nt * 16 + (lane & 15)) * KFP + k0];
                acc[nt] = mfma(a, bv, acc[nt]);
            }
        }
        if (tap == 4 && tid < kCh) {
            float s = 0.f;
            for (int pp = 0; pp < KF; pp++) s += (float)eT[tid * KFP + pp];
            dbacc += s;
        }
        __syncthreads();
    }
    // Partial layout: [chunk][tap][o][c], then [chunk][o] bias sums after
    // all chunks' weight partials.
    float *dst = part + ((size_t)chunk * 9 + tap) * kCh * CIN;
#pragma unroll
    for (int nt = 0; nt < NTC; nt++)
#pragma unroll
        for (int r = 0; r < 4; r++)
            dst[(w * 16 + (lane >> 4) * 4 + r) * CIN + nt * 16 + (lane & 15)] = acc[nt][r];
    if (tap == 4 && tid < kCh) {
        const int nch = gridDim.x / 9;
        part[(size_t)nch * 9 * kCh * CIN + chunk * kCh + tid] = dbacc;
    }
}

template <int CIN>
__global__ __launch_bounds__(kThreads) void sg_wgrad_reduce(
    const float *__restrict__ part, int nch, float *__restrict__ gW,
    float *__restrict__ gb) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    constexpr int N = 9 * kCh * CIN;
    if (i < N) {
        float s = 0.f;
        for (int ch = 0; ch < nch; ch++) s += part[(size_t)ch * N + i];
        const int tap = i / (kCh * CIN), o = (i / CIN) % kCh, c = i % CIN;
        gW[(o * CIN + c) * 9 + tap] = s;
    }
    if (i < kCh) {
        const float *db = part + (size_t)nch * N;
        float s = 0.f;
        for (int ch = 0; ch < nch; ch++) s += db[ch * kCh + i];
        gb[i] = s;
    }
}

// ---- head ------------------------------------------------------------------
// hp: wh[64], bh, F1[64][361], c1[64], F2[64], c2.
__global__ __launch_bounds__(kThreads) void sg_head_fwd(
    const __bf16 *__restrict__ h, const float *__restrict__ hp,
    float *__restrict__ s_out, float *__restrict__ f_out,
    float *__restrict__ t_out, float *__restrict__ win) {
    __shared__ float s[P];
    __shared__ float f[kHid];
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const float *wh = hp, *F1 = hp + kCh + 1;
    const float *c1 = F1 + kHid * P, *F2 = c1 + kHid;
    const float bh = hp[kCh], c2 = F2[kHid];
    for (int p = tid; p < P; p += kThreads) {
        const bf16x8 *hv = (const bf16x8 *)(h + ((size_t)b * P + p) * kCh);
        float d = bh;
#pragma unroll
        for (int c8 = 0; c8 < 8; c8++) {
            const bf16x8 v = hv[c8];
#pragma unroll
            for (int k = 0; k < 8; k++) d += wh[c8 * 8 + k] * (float)v[k];
        }
        d = elu(d);
        s[p] = d;
        s_out[(size_t)b * P + p] = d;
    }
    __syncthreads();
    for (int jj = 0; jj < kHid / 4; jj++) {
        const int j = w * (kHid / 4) + jj;
        float d = 0.f;
        for (int p = lane; p < P; p += 64) d += F1[j * P + p] * s[p];
        d = wave_sum(d);
        if (lane == 0) {
            d = elu(d + c1[j]);
            f[j] = d;
            f_out[(size_t)b * kHid + j] = d;
        }
    }
    __syncthreads();
    if (w == 0) {
        const float d = wave_sum(F2[lane] * f[lane]);
        if (lane == 0) {
            const float t = tanhf(d + c2);
            t_out[b] = t;
            if (win) win[b] = 0.5f * (1.f + t);
        }
    }
}

constexpr int kF1 = kHid * P;                     // 23,104
constexpr int kF1PerThread = (kF1 + kThreads - 1) / kThreads;  // 91
constexpr int kHeadPart = kHeadParams + 1;        // + loss

__global__ __launch_bounds__(kThreads, 1) void sg_head_bwd(
    const __bf16 *__restrict__ h, const float *__restrict__ hp,
    const float *__restrict__ s_in, const float *__restrict__ f_in,
    const float *__restrict__ t_in, const float *__restrict__ y,
    __bf16 *__restrict__ e, float *__restrict__ part, int n, int bpc) {
    __shared__ float s[P];
    __shared__ float ds[P];
    __shared__ float df[kHid];
    __shared__ float red[kThreads];
    const int chunk = blockIdx.x, tid = threadIdx.x;
    const float *wh = hp, *F1 = hp + kCh + 1;
    const float *F2 = F1 + kF1 + kHid;
    const float inv_n = 1.f / (float)n;
    const int b0 = chunk * bpc;
    const int b1 = b0 + bpc < n ? b0 + bpc : n;

    float aF1[kF1PerThread];
#pragma unroll
    for (int k = 0; k < kF1PerThread; k++) aF1[k] = 0.f;
    float aF2 = 0.f, ac1 = 0.f, ac2 = 0.f, aloss = 0.f, awh = 0.f, abh = 0.f;
    const int c = tid & 63, g = tid >> 6;
    const float whc = wh[c];

    for (int b = b0; b < b1; b++) {
        const float t = t_in[b];
        const float ystar = 2.f * y[b] - 1.f;
        const float dz = 2.f * (t - ystar) * inv_n * (1.f - t * t);
        if (tid == 0) {
            aloss += (t - ystar) * (t - ystar);
            ac2 += dz;
        }
        if (tid < kHid) {
            const float fj = f_in[(size_t)b * kHid + tid];
            const float d = dz * F2[tid] * elu_grad(fj);
            df[tid] = d;
            aF2 += dz * fj;
            ac1 += d;
        }
        for (int p = tid; p < P; p += kThreads) s[p] = s_in[(size_t)b * P + p];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kF1PerThread; k++) {
            const int i = tid + k * kThreads;
            if (i < kF1) aF1[k] += df[i / P] * s[i % P];
        }
        for (int p = tid; p < P; p += kThreads) {
            float d = 0.f;
            for (int j = 0; j < kHid; j++) d += F1[j * P + p] * df[j];
            ds[p] = d * elu_grad(s[p]);
        }
        __syncthreads();
        for (int p = g; p < P; p += 4) {
            const size_t at = ((size_t)b * P + p) * kCh + c;
            const float hv = (float)h[at];
            const float d = ds[p];
            awh += d * hv;
            if (c == 0) abh += d;
            e[at] = (__bf16)(d * whc * elu_grad(hv));
        }
        __syncthreads();
    }
    float *dst = part + (size_t)chunk * kHeadPart;
    // wh and bh: sum the four pixel groups.
    red[tid] = awh;
    __syncthreads();
    if (tid < kCh) dst[tid] = red[tid] + red[tid + 64] + red[tid + 128] + red[tid + 192];
    __syncthreads();
    red[tid] = abh;
    __syncthreads();
    if (tid == 0) dst[kCh] = red[0] + red[64] + red[128] + red[192];
#pragma unroll
    for (int k = 0; k < kF1PerThread; k++) {
        const int i = tid + k * kThreads;
        if (i < kF1) dst[kCh + 1 + i] = aF1[k];
    }
    if (tid < kHid) {
        dst[kCh + 1 + kF1 + tid] = ac1;
        dst[kCh + 1 + kF1 + kHid + tid] = aF2;
    }
    if (tid == 0) {
        dst[kHeadParams - 1] = ac2;
        dst[kHeadParams] = aloss;
    }
}

__global__ __launch_bounds__(kThreads) void sg_head_reduce(
    const float *__restrict__ part, int nch, float inv_n,
    float *__restrict__ ghp, float *__restrict__ loss) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= kHeadPart) return;
    float s = 0.f;
    for (int ch = 0; ch < nch; ch++) s += part[(size_t)ch * kHeadPart + i];
    if (i < kHeadParams)
        ghp[i] = s;
    else
        *loss = s * inv_n;
}

__global__ __launch_bounds__(kThreads) void sg_adam(
    float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
    float *__restrict__ v, int count, float lr, float b1, float b2, float eps,
    float c1, float c2) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

int blocks(long long count) { return (int)((count + kThreads - 1) / kThreads); }

}  // namespace

hipError_t pack_conv(const float *w, uint16_t *wpk, int cin, bool bwd,
                     hipStream_t stream) {
    const int g = blocks(9LL * cin * kCh);
    __bf16 *o = (__bf16 *)wpk;
    if (cin == 32 && !bwd)
        sg_pack<32, false><<<g, kThreads, 0, stream>>>(w, o);
    else if (cin == kCh && !bwd)
        sg_pack<kCh, false><<<g, kThreads, 0, stream>>>(w, o);
    else if (cin == kCh && bwd)
        sg_pack<kCh, true><<<g, kThreads, 0, stream>>>(w, o);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t conv(const void *in, int cin, const uint16_t *wpk,
                const float *bias, const uint16_t *res, const uint16_t *mask,
                uint16_t *out, int n, bool bwd, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const bf16x8 *w = (const bf16x8 *)wpk;
    const __bf16 *r = (const __bf16 *)res, *m = (const __bf16 *)mask;
    __bf16 *o = (__bf16 *)out;
    if (cin == 32 && !bwd)
        sg_conv<32, false><<<n, kThreads, 0, stream>>>(in, w, bias, r, m, o);
    else if (cin == kCh && !bwd)
        sg_conv<kCh, false><<<n, kThreads, 0, stream>>>(in, w, bias, r, m, o);
    else if (cin == kCh && bwd)
        sg_conv<kCh, true><<<n, kThreads, 0, stream>>>(in, w, bias, r, m, o);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

size_t wgrad_part_floats(int cin, int n) {
    const int nch = chunks_for(n);
    return (size_t)nch * 9 * kCh * cin + (size_t)nch * kCh;
}

hipError_t wgrad(const void *in, int cin, const uint16_t *e, float *part,
                 float *gW, float *gb, int n, hipStream_t stream) {
    if (n <= 0) return hipErrorInvalidValue;
    const int nch = chunks_for(n);
    const int bpc = (n + nch - 1) / nch;
    const int used = (n + bpc - 1) / bpc;   // chunks that own boards
    const __bf16 *ev = (const __bf16 *)e;
    if (cin == 32) {
        sg_wgrad<32><<<9 * used, kThreads, 0, stream>>>(in, ev, part, n, bpc);
        sg_wgrad_reduce<32><<<blocks(9 * kCh * 32), kThreads, 0, stream>>>(
            part, used, gW, gb);
    } else if (cin == kCh) {
        sg_wgrad<kCh><<<9 * used, kThreads, 0, stream>>>(in, ev, part, n, bpc);
        sg_wgrad_reduce<kCh><<<blocks(9 * kCh * kCh), kThreads, 0, stream>>>(
            part, used, gW, gb);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t head_fwd(const uint16_t *h, const float *hp, float *s, float *f,
                    float *t, float *win, int n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    sg_head_fwd<<<n, kThreads, 0, stream>>>((const __bf16 *)h, hp, s, f, t,
                                            win);
    return hipGetLastError();
}

size_t head_part_floats(int n) { return (size_t)chunks_for(n) * kHeadPart; }

hipError_t head_bwd(const uint16_t *h, const float *hp, const float *s,
                    const float *f, const float *t, const float *y,
                    uint16_t *e, float *part, float *ghp, float *loss, int n,
                    hipStream_t stream) {
    if (n <= 0) return hipErrorInvalidValue;
    const int nch = chunks_for(n);
    const int bpc = (n + nch - 1) / nch;
    const int used = (n + bpc - 1) / bpc;
    sg_head_bwd<<<used, kThreads, 0, stream>>>((const __bf16 *)h, hp, s, f, t,
                                               y, (__bf16 *)e, part, n, bpc);
    sg_head_reduce<<<blocks(kHeadPart), kThreads, 0, stream>>>(
        part, used, 1.f / (float)n, ghp, loss);
    return hipGetLastError();
}

hipError_t adam(float *p, const float *g, float *m, float *v, int count,
                float lr, float beta1, float beta2, float eps, float c1,
                float c2, hipStream_t stream) {
    if (count <= 0) return hipSuccess;
    sg_adam<<<blocks(count), kThreads, 0, stream>>>(p, g, m, v, count, lr,
                                                    beta1, beta2, eps, c1, c2);
    return hipGetLastError();
}

}  // namespace sg
}  // namespace lz
