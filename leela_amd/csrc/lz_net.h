// lz_net.h — layer tables, device weights and workspaces behind lz_ctx.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "lz_fused.h"

namespace lz {

struct Layer {
    enum Kind { kConv = 0, kFC = 1 };
    int kind = kConv;
    int filter = 0;
    int cin = 0;
    int cout = 0;
    std::vector<float> w;  // conv OIHW, FC [out][in]
    std::vector<float> b;
    int act = 0;           // conv activation (kActElu / kActRelu / kActNone)
    // Batch normalisation of the activated conv output (empty: none):
    // mean[c] / scale and 1 / sqrt(variance[c] / scale + 1e-5), formed in
    // FP32 as Network.cpp:438-441 forms them.
    std::vector<float> bn_mean, bn_inv_std;
    float *d_bn_mean = nullptr;
    float *d_bn_inv_std = nullptr;
    // Device copies.
    float *d_w32 = nullptr;  // FP32: trunk packed [cout/64][K][64]; head conv
                             // OIHW; FC1 transposed [in][out]; FC2 as is
    float *d_b = nullptr;
    uint16_t *d_wbf16 = nullptr;  // BF16 MFMA fragments (fused kernels)
    uint16_t *d_wf16 = nullptr;   // FP16 MFMA fragments (fused kernels)
    uint8_t *d_wf8 = nullptr;     // e4m3 MFMA fragments (FP8 form)
    float *d_ws8 = nullptr;       // their per-output-channel scales
};

// A net as the reference pushes it (Network.cpp:206-233): a chain of convs
// each followed by ELU whose last conv has one output channel, then either
// nothing (policy: softmax head) or two inner products (value: FC+ELU, FC,
// then (1 + tanh) / 2). A Caffe net's convs may instead take ReLU or no
// activation and a batch normalisation after it (`general`: FP32 only).
struct Net {
    std::vector<Layer> layers;
    bool general = false;   // an activation other than ELU, or a BN layer
    int side = 19;
    int n_trunk = 0;    // convs before the 1-channel head conv
    int head = -1;      // index of the 1-channel conv
    int fc1 = -1, fc2 = -1;
    int max_ch = 0;     // widest trunk activation
    bool fused_ok = false;  // shapes the fused kernels implement
    FusedArgs fused;        // device pointers for the BF16 fused kernel
    FusedArgs fused16;      // device pointers for the FP16 fused kernel
    FusedArgs fused8;       // FP8 form: FP16 stem, e4m3 later convs
    // Batches of at most this many 19x19 boards run the layered form of
    // the 16-bit kernels (fused_forward_layered); 0 disables it.
    int layered_max = kLayeredMaxBoards;
    int cus = 0;            // compute units of the device
    int occ[3] = {0, 0, 0}; // resident workgroups per CU for wm = 1, 2
    // Launch shape per batch bucket (<=64, <=256, <=1024, more), set by
    // the autotuner: wm 0 = the net's default wave grid; grid_mult 0 = one
    // workgroup per slot, else CUs x occupancy x grid_mult (persistent).
    struct Tune {
        int wm = 0;
        int grid_mult = 1;
        float ms = 0.0f;  // measured forward time of the choice
    } tune[4];
};

int tune_bucket(int n);
FusedLaunch fused_launch_for(const Net &net, int which, int nslots);

struct Workspace {
    int cap = 0;             // positions the buffers hold
    uint8_t *planes = nullptr;
    float *act[2] = {nullptr, nullptr};
    float *raw = nullptr;
    float *out = nullptr;
    float *h0 = nullptr;     // value head plane [cap][361]
    void *tiles = nullptr;   // layered form's activation tiles (zeroed)
};

// Allocates and zeroes ws.tiles if absent (layered_tile_bytes()).
int ensure_tiles(Workspace &ws, std::string *msg);

// Device copy of a ragged batch's layout: slots and per-board offsets.
struct RaggedLayout {
    int n = 0, nslots = 0;
    size_t plane_bytes = 0;   // total planes bytes of the batch
    size_t policy_floats = 0; // total policy outputs of the batch
    FusedSlot *d_slots = nullptr;
    int *d_in_off = nullptr;
    int *d_out_off = nullptr;
    int cap_boards = 0, cap_slots = 0;
};

// Packs boards of sides[0..n) into slots (grouped by side, stable) and
// uploads the layout on `stream`. Sides must be 5..19.
int build_ragged_layout(RaggedLayout &lay, int n, const int32_t *sides,
                        hipStream_t stream, std::string *msg);
void free_ragged_layout(RaggedLayout &lay);

int validate_net(Net &net, int which, std::string *msg);
int upload_net(Net &net, int which, std::string *msg);
void free_net(Net &net);
int ensure_workspace(Workspace &ws, const Net nets[2], int n,
                     std::string *msg);
void free_workspace(Workspace &ws);
// Records a HIP event on the launch stream before the first and after
// every kernel launch of a forward.
struct LaunchTimer {
    std::vector<hipEvent_t> events;
    hipError_t mark(hipStream_t s);
    ~LaunchTimer();
};

int run_net(const Net &net, int which, int precision, int n,
            const uint8_t *d_planes, float *d_raw, float *d_out,
            float temperature, Workspace &ws, hipStream_t stream,
            std::string *msg, LaunchTimer *timer = nullptr,
            const RaggedLayout *ragged = nullptr,
            const int *counts = nullptr, const uint32_t *bits = nullptr);
// counts: device-side {slots, boards} the fused forms read when they start
// (graph replays of a bucket on a smaller batch); n (and the ragged
// layout's nslots) are then capacities. The FP32 form ignores it and
// evaluates all n rows.

// Host packing of a conv into MFMA fragment order (16-bit values, round to
// nearest even): [k-step][m-tile][lane][8], k-step s = tap * (cin/32) +
// channel block, lane = row + 16 * g holding input channels
// 32*cb + 8*g .. +7.
std::vector<uint16_t> pack_conv_fragments(const Layer &L, FusedType type);

// OCP e4m3 ("E4M3FN": bias 7, max 448, no infinities) with round to
// nearest even; magnitudes beyond 448 saturate to 448.
uint8_t float_to_e4m3_rne(float f);

// e4m3 fragments of v_mfma_scale_f32_16x16x128_f8f6f4: [k-step of 128]
// [m-tile][lane][32 bytes], k = tap * cin + c, lane's bytes k = 128 s +
// 32 (lane / 16) + j of output channel 16 m + lane % 16; zero past K and
// past cout. Each output channel o is scaled by scales[o] = max |W[o]| /
// 448 (1 for an all-zero channel) before rounding; scales has 16 * MT
// entries.
std::vector<uint8_t> pack_conv_fp8(const Layer &L, std::vector<float> *scales);
uint16_t float_to_bf16_rne(float f);
uint16_t float_to_f16_rne(float f);

void normal_fill(uint32_t seed, size_t n, float stddev, float *out);

}  // namespace lz
