// lz_net.cpp — validation, weight packing/upload and the per-precision
// layer schedule of the policy and value nets.
#include "lz_net.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "leela_amd.h"
#include "lz_internal.h"
#include "lz_random.h"

namespace lz {
namespace {

template <typename T>
hipError_t dalloc(T **p, size_t count) {
    return hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * (count ? count : 1));
}

template <typename T>
hipError_t upload(T **p, const std::vector<T> &h) {
    hipError_t e = dalloc(p, h.size());
    if (e != hipSuccess) return e;
    return hipMemcpy(*p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
}

void dfree(void *p) {
    if (p) hipFree(p);
}

}  // namespace

int validate_net(Net &net, int which, std::string *msg) {
    auto &L = net.layers;
    int i = 0;
    while (i < (int)L.size() && L[i].kind == Layer::kConv) i++;
    const int nconv = i;
    if (nconv < 2) {
        *msg = "needs at least one trunk conv and a 1-channel head conv";
        return LZ_ERR_UNSUPPORTED;
    }
    net.n_trunk = nconv - 1;
    net.head = nconv - 1;
    net.side = 19;
    net.max_ch = L[0].cin;
    net.general = false;
    // The reference's nets take its 32 input planes (POLICY_CHANNELS =
    // VALUE_CHANNELS = 32, Network.h:48-49), which every path that builds
    // planes (the leaf queue, the facade, the encoders) lays out; a Caffe
    // net takes its own count through lz_forward / lz_forward_device.
    if (L[0].cin > kMaxInputPlanes) {
        *msg = "the first conv takes at most " +
               std::to_string(kMaxInputPlanes) + " input planes";
        return LZ_ERR_UNSUPPORTED;
    }
    for (int k = 0; k < nconv; k++) {
        if (k > 0 && L[k].cin != L[k - 1].cout) {
            *msg = "conv " + std::to_string(k) + " input channels do not "
                   "match the previous layer's outputs";
            return LZ_ERR_ARG;
        }
        if (k < nconv - 1 && L[k].cout > kMaxTrunkChannels) {
            *msg = "trunk convs wider than " +
                   std::to_string(kMaxTrunkChannels) +
                   " channels are not built";
            return LZ_ERR_UNSUPPORTED;
        }
        if (!L[k].bn_mean.empty() && k == nconv - 1) {
            *msg = "batch normalisation of the 1-channel head conv is not "
                   "built";
            return LZ_ERR_UNSUPPORTED;
        }
        if (L[k].act != kActElu || !L[k].bn_mean.empty()) net.general = true;
        if (L[k].cout > net.max_ch) net.max_ch = L[k].cout;
    }
    if (L[net.head].cout != 1 || L[net.head].filter != 3) {
        *msg = "the last conv must be 3x3 with one output channel";
        return LZ_ERR_UNSUPPORTED;
    }
    const int P = net.side * net.side;
    if (which == LZ_NET_POLICY) {
        if ((int)L.size() != nconv) {
            *msg = "policy net takes no inner-product layers";
            return LZ_ERR_UNSUPPORTED;
        }
    } else {
        if ((int)L.size() != nconv + 2 || L[nconv].kind != Layer::kFC ||
            L[nconv + 1].kind != Layer::kFC) {
            *msg = "value net must end in exactly two inner products";
            return LZ_ERR_UNSUPPORTED;
        }
        net.fc1 = nconv;
        net.fc2 = nconv + 1;
        if (L[net.fc1].cin != P || L[net.fc2].cin != L[net.fc1].cout ||
            L[net.fc2].cout != 1) {
            *msg = "value inner products must be (S*S -> H) and (H -> 1)";
            return LZ_ERR_ARG;
        }
    }
    return LZ_OK;
}

int upload_net(Net &net, int which, std::string *msg) {
    (void)which;
    hipError_t e = hipSuccess;
    for (int k = 0; k < (int)net.layers.size() && e == hipSuccess; k++) {
        Layer &L = net.layers[k];
        std::vector<float> packed;
        if (L.kind == Layer::kConv && k < net.head) {
            // [cout/64][Kp][64]: a workgroup's 64 output channels for one K
            // row are contiguous, so the LDS copy is a straight stream; K
            // rows past cin * k * k (a partial last chunk) stay zero.
            const int K = L.cin * L.filter * L.filter;
            const int Kp = f32_packed_k(L.filter, L.cin);
            const int mbs = (L.cout + kF32MBlock - 1) / kF32MBlock;
            packed.assign((size_t)mbs * Kp * kF32MBlock, 0.0f);
            for (int o = 0; o < L.cout; o++)
                for (int kk = 0; kk < K; kk++)
                    packed[((size_t)(o / kF32MBlock) * Kp + kk) * kF32MBlock +
                           o % kF32MBlock] = L.w[(size_t)o * K + kk];
        } else if (L.kind == Layer::kFC && k == net.fc1) {
            packed.assign(L.w.size(), 0.0f);
            for (int o = 0; o < L.cout; o++)
                for (int i = 0; i < L.cin; i++)
                    packed[(size_t)i * L.cout + o] = L.w[(size_t)o * L.cin + i];
        } else {
            packed = L.w;
        }
        e = upload(&L.d_w32, packed);
        if (e == hipSuccess) e = upload(&L.d_b, L.b);
        if (e == hipSuccess && !L.bn_mean.empty()) {
            e = upload(&L.d_bn_mean, L.bn_mean);
            if (e == hipSuccess) e = upload(&L.d_bn_inv_std, L.bn_inv_std);
        }
    }
    if (e != hipSuccess) {
        *msg = std::string("weight upload: ") + hipGetErrorString(e);
        return LZ_ERR_HIP;
    }
    // The fused kernels are specialised for the reference's two shape sets
    // (Network.cpp:82-137): policy 5x5 32->96, 3x3 96->128, 10x 128->128,
    // 128->1; value 5x5 32->64, 10x 64->64, 64->1, FC 361->H, FC H->1.
    const auto &L = net.layers;
    const int nconv = net.head + 1;
    bool ok = !net.general && nconv <= kMaxFusedConvs && L[0].filter == 5 &&
              L[0].cin == 32;
    const int c = nconv > 1 ? L[1].cout : 0;
    for (int k = 1; ok && k < net.head; k++)
        ok = L[k].filter == 3 && L[k].cout == c && (k == 1 || L[k].cin == c);
    if (which == LZ_NET_POLICY)
        ok = ok && nconv == 13 && c == 128 && L[0].cout == 96;
    else
        ok = ok && nconv == 12 && c == 64 && L[0].cout == 64;
    net.fused_ok = ok;
    if (!ok) return LZ_OK;
    FusedArgs &A = net.fused;
    for (int k = 0; k < nconv && e == hipSuccess; k++) {
        e = upload(&net.layers[k].d_wbf16,
                   pack_conv_fragments(L[k], FusedType::kBF16));
        if (e == hipSuccess)
            e = upload(&net.layers[k].d_wf16,
                       pack_conv_fragments(L[k], FusedType::kFP16));
        A.w[k] = net.layers[k].d_wbf16;
        A.b[k] = net.layers[k].d_b;
    }
    if (e != hipSuccess) {
        *msg = std::string("16-bit weight upload: ") + hipGetErrorString(e);
        return LZ_ERR_HIP;
    }
    A.nconv = nconv;
    A.c = c;
    A.m1 = L[0].cout;
    if (which == LZ_NET_VALUE) {
        A.fc1_wt = L[net.fc1].d_w32;
        A.fc1_b = L[net.fc1].d_b;
        A.fc2_w = L[net.fc2].d_w32;
        A.fc2_b = L[net.fc2].d_b;
        A.hidden = L[net.fc1].cout;
    }
    net.fused16 = A;
    for (int k = 0; k < nconv; k++) net.fused16.w[k] = net.layers[k].d_wf16;
    net.fused8 = net.fused16;
    for (int k = 1; k < nconv && e == hipSuccess; k++) {
        std::vector<float> scales;
        e = upload(&net.layers[k].d_wf8, pack_conv_fp8(L[k], &scales));
        if (e == hipSuccess) e = upload(&net.layers[k].d_ws8, scales);
        net.fused8.w[k] = net.layers[k].d_wf8;
        net.fused8.ws[k] = net.layers[k].d_ws8;
    }
    if (e != hipSuccess) {
        *msg = std::string("FP8 weight upload: ") + hipGetErrorString(e);
        return LZ_ERR_HIP;
    }
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                dev);
    net.cus = cus > 0 ? cus : 1;
    net.occ[1] = fused_blocks_per_cu(which, 1);
    net.occ[2] = fused_blocks_per_cu(which, 2);
    return LZ_OK;
}

int tune_bucket(int n) {
    return n <= 64 ? 0 : n <= 256 ? 1 : n <= 1024 ? 2 : 3;
}

FusedLaunch fused_launch_for(const Net &net, int which, int nslots) {
    const Net::Tune &t = net.tune[tune_bucket(nslots)];
    FusedLaunch cfg;
    cfg.wm = t.wm == 1 || t.wm == 2 ? t.wm : fused_default_wm(which);
    cfg.grid = t.grid_mult <= 0 ? 0 : net.cus * net.occ[cfg.wm] * t.grid_mult;
    return cfg;
}

uint16_t float_to_bf16_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u)  // inf / nan: truncate, keep nan
        return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

uint16_t float_to_f16_rne(float f) {
    // The host compiler's _Float16 conversion rounds to nearest even and
    // handles subnormals, overflow and NaN.
    const _Float16 h = (_Float16)f;
    uint16_t u;
    std::memcpy(&u, &h, 2);
    return u;
}

uint8_t float_to_e4m3_rne(float f) {
    if (std::isnan(f)) return 0x7f;
    const uint8_t sign = std::signbit(f) ? 0x80 : 0x00;
    const double a = std::fabs((double)f);
    if (a >= 448.0) return sign | 0x7e;
    if (a < 0.015625) {  // below 2^-6: subnormals, multiples of 2^-9
        // nearbyint rounds half to even in the default rounding mode; 8
        // becomes the smallest normal's code 0x08.
        return sign | (uint8_t)std::nearbyint(a * 512.0);
    }
    int e = 0;
    const double m = std::frexp(a, &e);   // a = m 2^e, m in [0.5, 1)
    int E = e - 1;
    int q = (int)std::nearbyint((2.0 * m - 1.0) * 8.0);
    if (q == 8) {
        q = 0;
        E += 1;
    }
    int code = ((E + 7) << 3) | q;
    if (code > 0x7e) code = 0x7e;
    return sign | (uint8_t)code;
}

std::vector<uint8_t> pack_conv_fp8(const Layer &L, std::vector<float> *scales) {
    const int F = L.filter, cin = L.cin, cout = L.cout;
    const int K = F * F * cin, steps = (K + 127) / 128, MT = (cout + 15) / 16;
    scales->assign((size_t)MT * 16, 1.0f);
    for (int o = 0; o < cout; o++) {
        float mx = 0.0f;
        for (int k = 0; k < K; k++)
            mx = std::max(mx, std::fabs(L.w[(size_t)o * K + k]));
        if (mx > 0.0f) (*scales)[o] = mx / 448.0f;
    }
    std::vector<uint8_t> out((size_t)steps * MT * 64 * 32, 0);
    size_t idx = 0;
    for (int s = 0; s < steps; s++)
        for (int m = 0; m < MT; m++)
            for (int lane = 0; lane < 64; lane++)
                for (int j = 0; j < 32; j++, idx++) {
                    const int k = s * 128 + (lane >> 4) * 32 + j;
                    const int o = m * 16 + (lane & 15);
                    if (k >= K || o >= cout) continue;
                    const int tap = k / cin, c = k % cin;
                    const float w = L.w[(((size_t)o * cin + c) * F + tap / F) * F +
                                        tap % F];
                    out[idx] = float_to_e4m3_rne(w / (*scales)[o]);
                }
    return out;
}

std::vector<uint16_t> pack_conv_fragments(const Layer &L, FusedType type) {
    const int F = L.filter, cin = L.cin, cout = L.cout;
    const int CB = cin / 32, steps = F * F * CB, MT = (cout + 15) / 16;
    std::vector<uint16_t> out((size_t)steps * MT * 64 * 8, 0);
    size_t idx = 0;
    for (int s = 0; s < steps; s++) {
        const int tap = s / CB, cb = s % CB, dy = tap / F, dx = tap % F;
        for (int m = 0; m < MT; m++)
            for (int lane = 0; lane < 64; lane++)
                for (int e = 0; e < 8; e++, idx++) {
                    const int o = m * 16 + (lane & 15);
                    const int ch = cb * 32 + (lane >> 4) * 8 + e;
                    if (o < cout) {
                        const float w =
                            L.w[(((size_t)o * cin + ch) * F + dy) * F + dx];
                        out[idx] = type == FusedType::kBF16
                                       ? float_to_bf16_rne(w)
                                       : float_to_f16_rne(w);
                    }
                }
    }
    return out;
}

void free_net(Net &net) {
    for (auto &L : net.layers) {
        dfree(L.d_w32);
        dfree(L.d_b);
        dfree(L.d_wbf16);
        dfree(L.d_wf16);
        dfree(L.d_wf8);
        dfree(L.d_ws8);
        dfree(L.d_bn_mean);
        dfree(L.d_bn_inv_std);
        L.d_bn_mean = L.d_bn_inv_std = nullptr;
        L.d_w32 = L.d_b = nullptr;
        L.d_wbf16 = L.d_wf16 = nullptr;
        L.d_wf8 = nullptr;
        L.d_ws8 = nullptr;
    }
}

int ensure_workspace(Workspace &ws, const Net nets[2], int n,
                     std::string *msg) {
    if (n <= ws.cap) return LZ_OK;
    free_workspace(ws);
    int ch = 32;
    int cin0 = 0;
    for (int k = 0; k < 2; k++) {
        if (nets[k].max_ch > ch) ch = nets[k].max_ch;
        if (!nets[k].layers.empty() && nets[k].layers[0].cin > cin0)
            cin0 = nets[k].layers[0].cin;
    }
    const size_t P = (size_t)kMaxSide * kMaxSide;
    hipError_t e = dalloc(&ws.planes, (size_t)n * cin0 * P);
    if (e == hipSuccess) e = dalloc(&ws.act[0], (size_t)n * ch * P);
    if (e == hipSuccess) e = dalloc(&ws.act[1], (size_t)n * ch * P);
    if (e == hipSuccess) e = dalloc(&ws.raw, (size_t)n * P);
    if (e == hipSuccess) e = dalloc(&ws.out, (size_t)n * P);
    if (e == hipSuccess) e = dalloc(&ws.h0, (size_t)n * P);
    if (e != hipSuccess) {
        free_workspace(ws);
        *msg = std::string("workspace allocation: ") + hipGetErrorString(e);
        return LZ_ERR_NOMEM;
    }
    ws.cap = n;
    return ensure_tiles(ws, msg);
}

int ensure_tiles(Workspace &ws, std::string *msg) {
    if (ws.tiles) return LZ_OK;
    const size_t bytes = layered_tile_bytes();
    hipError_t e = hipMalloc(&ws.tiles, bytes);
    if (e == hipSuccess) e = hipMemset(ws.tiles, 0, bytes);
    if (e != hipSuccess) {
        if (ws.tiles) (void)hipFree(ws.tiles);
        ws.tiles = nullptr;
        *msg = std::string("layered tiles: ") + hipGetErrorString(e);
        return LZ_ERR_NOMEM;
    }
    return LZ_OK;
}

void free_workspace(Workspace &ws) {
    dfree(ws.planes);
    dfree(ws.act[0]);
    dfree(ws.act[1]);
    dfree(ws.raw);
    dfree(ws.out);
    dfree(ws.h0);
    if (ws.tiles) (void)hipFree(ws.tiles);
    ws = Workspace();
}

hipError_t LaunchTimer::mark(hipStream_t s) {
    hipEvent_t ev;
    hipError_t e = hipEventCreate(&ev);
    if (e != hipSuccess) return e;
    events.push_back(ev);
    return hipEventRecord(ev, s);
}

LaunchTimer::~LaunchTimer() {
    for (auto ev : events) hipEventDestroy(ev);
}

static int run_fp32(const Net &net, int which, int n, const uint8_t *planes,
                    float *raw, float *out, float temperature, Workspace &ws,
                    hipStream_t s, std::string *msg, LaunchTimer *timer) {
    const int S = net.side;
    const void *x = planes;
    bool u8 = true;
    hipError_t e = timer ? timer->mark(s) : hipSuccess;
    for (int k = 0; k < net.n_trunk && e == hipSuccess; k++) {
        const Layer &L = net.layers[k];
        float *y = ws.act[k & 1];
        ConvEpilogue epi;
        epi.act = L.act;
        epi.bn_mean = L.d_bn_mean;
        epi.bn_inv_std = L.d_bn_inv_std;
        e = conv_f32(L.filter, u8, x, L.d_w32, L.d_b, y, n, L.cin, L.cout, S,
                     epi, s);
        if (e == hipSuccess && timer) e = timer->mark(s);
        x = y;
        u8 = false;
    }
    const Layer &H = net.layers[net.head];
    if (e == hipSuccess) {
        if (which == LZ_NET_POLICY) {
            e = policy_head_f32(static_cast<const float *>(x), H.d_w32, H.d_b,
                                raw, out, n, H.cin, S, temperature, H.act, s);
        } else {
            const Layer &F1 = net.layers[net.fc1];
            const Layer &F2 = net.layers[net.fc2];
            e = value_headconv_f32(static_cast<const float *>(x), H.d_w32,
                                   H.d_b, ws.h0, n, H.cin, S, H.act, s);
            if (e == hipSuccess) e = value_fc_head(ws.h0, F1.d_w32, F1.d_b,
                                                   F2.d_w32, F2.d_b, raw, out,
                                                   n, S * S, F1.cout, s);
        }
        if (e == hipSuccess && timer) e = timer->mark(s);
    }
    if (e != hipSuccess) {
        *msg = std::string("FP32 forward launch: ") + hipGetErrorString(e);
        return LZ_ERR_HIP;
    }
    return LZ_OK;
}

int run_net(const Net &net, int which, int precision, int n,
            const uint8_t *d_planes, float *d_raw, float *d_out,
            float temperature, Workspace &ws, hipStream_t stream,
            std::string *msg, LaunchTimer *timer,
            const RaggedLayout *ragged, const int *counts,
            const uint32_t *bits) {
    if (ragged && precision != LZ_PREC_BF16 && precision != LZ_PREC_FP16) {
        *msg = "ragged batches run on the 16-bit fused kernels only";
        return LZ_ERR_UNSUPPORTED;
    }
    if (precision == LZ_PREC_FP32)
        return run_fp32(net, which, n, d_planes, d_raw, d_out, temperature,
                        ws, stream, msg, timer);
    if (precision == LZ_PREC_FP8) {
        if (!net.fused_ok) {
            *msg = "FP8 form implements only the reference layer shapes";
            return LZ_ERR_UNSUPPORTED;
        }
        hipError_t e = timer ? timer->mark(stream) : hipSuccess;
        FusedArgs args = net.fused8;
        args.h0 = ws.h0;
        args.counts = counts;
        // Default wave grid, one workgroup per board (the FP8 tile is
        // smaller than the 16-bit one, so occupancy differs from theirs).
        FusedLaunch cfg;
        if (e == hipSuccess)
            e = fused_forward(which, FusedType::kFP8, d_planes, args, d_raw,
                              d_out, n, temperature, cfg, stream);
        if (e == hipSuccess && timer) e = timer->mark(stream);
        if (e != hipSuccess) {
            *msg = std::string("FP8 forward launch: ") + hipGetErrorString(e);
            return LZ_ERR_HIP;
        }
        return LZ_OK;
    }
    if (precision == LZ_PREC_BF16 || precision == LZ_PREC_FP16) {
        if (!net.fused_ok) {
            *msg = "16-bit fused path implements only the reference layer "
                   "shapes";
            return LZ_ERR_UNSUPPORTED;
        }
        const bool bf = precision == LZ_PREC_BF16;
        hipError_t e = timer ? timer->mark(stream) : hipSuccess;
        FusedArgs args = bf ? net.fused : net.fused16;
        args.h0 = ws.h0;
        args.counts = counts;
        if (ragged) {
            args.slots = ragged->d_slots;
            args.nslots = ragged->nslots;
            args.in_off = ragged->d_in_off;
            args.out_off = ragged->d_out_off;
        }
        const FusedLaunch cfg =
            fused_launch_for(net, which, ragged ? ragged->nslots : n);
        const FusedType ft = bf ? FusedType::kBF16 : FusedType::kFP16;
        if (e == hipSuccess) {
            // (The layered stem copies planes 16 bytes at a time.)
            if (!ragged && ws.tiles && n <= net.layered_max &&
                reinterpret_cast<uintptr_t>(d_planes) % 16 == 0)
                e = fused_forward_layered(which, ft, d_planes, args, d_raw,
                                          d_out, n, temperature, ws.tiles,
                                          stream, bits);
            else
                e = fused_forward(which, ft, d_planes, args, d_raw, d_out, n,
                                  temperature, cfg, stream);
        }
        if (e == hipSuccess && timer) e = timer->mark(stream);
        if (e != hipSuccess) {
            *msg = std::string("16-bit forward launch: ") +
                   hipGetErrorString(e);
            return LZ_ERR_HIP;
        }
        return LZ_OK;
    }
    *msg = "precision not built yet";
    return LZ_ERR_UNSUPPORTED;
}

// Box-Muller over the reference RNG: each pair of open-interval uniforms
// (u1, u2) gives sqrt(-2 ln u1) * (cos 2 pi u2, sin 2 pi u2).
void normal_fill(uint32_t seed, size_t n, float stddev, float *out) {
    Rng rng(seed);
    const double two_pi = 6.283185307179586476925286766559;
    size_t i = 0;
    while (i < n) {
        const double u1 = rng.uniform_open();
        const double u2 = rng.uniform_open();
        const double r = std::sqrt(-2.0 * std::log(u1));
        out[i++] = (float)(stddev * r * std::cos(two_pi * u2));
        if (i < n) out[i++] = (float)(stddev * r * std::sin(two_pi * u2));
    }
}

}  // namespace lz

namespace lz {

int build_ragged_layout(RaggedLayout &lay, int n, const int32_t *sides,
                        hipStream_t stream, std::string *msg) {
    std::vector<int> in_off(n), out_off(n);
    std::vector<FusedSlot> slots;
    size_t pb = 0, pf = 0;
    for (int i = 0; i < n; i++) {
        if (ragged_slot_capacity(sides[i]) == 0) {
            *msg = "board side " + std::to_string(sides[i]) +
                   " outside 5..19";
            return LZ_ERR_ARG;
        }
        in_off[i] = (int)pb;
        out_off[i] = (int)pf;
        pb += (size_t)32 * sides[i] * sides[i];
        pf += (size_t)sides[i] * sides[i];
    }
    // Group by side in first-appearance order; fill slots to capacity.
    std::vector<int> order;
    std::vector<char> done(n, 0);
    for (int i = 0; i < n; i++) {
        if (done[i]) continue;
        const int s = sides[i], cap = ragged_slot_capacity(s);
        FusedSlot cur{s, 0, {0, 0, 0, 0}};
        for (int j = i; j < n; j++) {
            if (done[j] || sides[j] != s) continue;
            done[j] = 1;
            cur.board[cur.count++] = j;
            if (cur.count == cap) {
                slots.push_back(cur);
                cur = FusedSlot{s, 0, {0, 0, 0, 0}};
            }
        }
        if (cur.count) slots.push_back(cur);
    }
    hipError_t e = hipSuccess;
    if (n > lay.cap_boards || (int)slots.size() > lay.cap_slots) {
        free_ragged_layout(lay);
        e = dalloc(&lay.d_slots, std::max<size_t>(slots.size(), 1));
        if (e == hipSuccess) e = dalloc(&lay.d_in_off, (size_t)std::max(n, 1));
        if (e == hipSuccess) e = dalloc(&lay.d_out_off, (size_t)std::max(n, 1));
        lay.cap_boards = n;
        lay.cap_slots = (int)slots.size();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(lay.d_slots, slots.data(),
                           sizeof(FusedSlot) * slots.size(),
                           hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(lay.d_in_off, in_off.data(), sizeof(int) * n,
                           hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(lay.d_out_off, out_off.data(), sizeof(int) * n,
                           hipMemcpyHostToDevice, stream);
    // The host vectors go out of scope: finish the copies first.
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        *msg = std::string("ragged layout upload: ") + hipGetErrorString(e);
        return LZ_ERR_HIP;
    }
    lay.n = n;
    lay.nslots = (int)slots.size();
    lay.plane_bytes = pb;
    lay.policy_floats = pf;
    return LZ_OK;
}

void free_ragged_layout(RaggedLayout &lay) {
    dfree(lay.d_slots);
    dfree(lay.d_in_off);
    dfree(lay.d_out_off);
    lay = RaggedLayout();
}

}  // namespace lz
