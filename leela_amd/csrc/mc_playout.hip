// mc_playout.hip — the reference's pattern playout policy on the GPU: one
// playout per thread, running the same source as the host (mc_core.h), so
// each playout is the same move sequence and final score as the host's for
// the same position and seed (Playout::run / mc_owner, Playout.cpp:51-113,
// 174-217, with FastState::play_random_move, FastState.cpp:168-255).
//
// The policy's state is pointer-chasing over a FastBoard-like board (string
// lists, the empty list, the candidate list) that differs per playout, so
// the kernel gives each thread its own state in global memory (a persistent
// grid, states reused playout after playout) and relies on many resident
// wavefronts to hide the latency; the lanes of a wavefront diverge freely.
// This file is compiled with -ffp-contract=off: the scores' products and
// running sums must round exactly as the host's do.
#include <hip/hip_runtime.h>

#include "lz_encoder.h"
#include "mc_core.h"

namespace lz {
namespace {

using mc::DevPlayState;

__global__ __launch_bounds__(64) void pat_zero(uint32_t *half_wins,
                                               uint32_t *overflow, int npos) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i < npos) {
        half_wins[i] = 0;
        overflow[i] = 0;
    }
}

// Thread t plays playouts t, t + T, ... (T threads) of the npos x nplay
// grid; playout j of position p is seeded playout_seed(seed0 + p, j) with
// two outputs discarded. Outcomes are summed as half points per position.
__global__ __launch_bounds__(64) void pat_playout_kernel(
    const EncRecord *__restrict__ recs, int npos, int nplay, uint32_t seed0,
    mc::TablesView tables, DevPlayState *states, uint32_t *half_wins,
    uint32_t *overflow) {
    const long long T = (long long)gridDim.x * 64;
    const long long tid = (long long)blockIdx.x * 64 + threadIdx.x;
    const long long total = (long long)npos * nplay;
    DevPlayState &ps = states[tid];
    for (long long i = tid; i < total; i += T) {
        const int p = (int)(i / nplay), j = (int)(i % nplay);
        const EncRecord &r = recs[p];
        ps.board.setup(r.side, r.sq);
        ps.to_move = r.to_move;
        ps.komove = r.komove;
        ps.last_move = r.last_move;
        ps.last_was_capture = r.last_was_capture != 0;
        ps.passes = r.passes;
        ps.movenum = r.movenum;
        ps.komi = r.komi;
        Rng rng(playout_seed(seed0 + (uint32_t)p, (uint32_t)j));
        rng.next();
        rng.next();
        const float score = ps.playout(rng, tables);
        if (ps.overflow) {
            atomicAdd(&overflow[p], 1u);
        } else {
            const uint32_t h = score > 0.0f ? 2u : score == 0.0f ? 1u : 0u;
            if (h) atomicAdd(&half_wins[p], h);
        }
    }
}

// mc_owner's win rate for the side to move from the half-point counts,
// with the host's float operations; NaN for a position with an abandoned
// playout.
__global__ __launch_bounds__(64) void pat_winrate(
    const uint32_t *half_wins, const uint32_t *overflow,
    const EncRecord *recs, int npos, int nplay, float *winrate) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= npos) return;
    float w = nplay > 0 ? (float)half_wins[i] * 0.5f / (float)nplay : 0.5f;
    if (recs[i].to_move != mc::kBlack) w = 1.0f - w;
    winrate[i] = overflow[i] ? __builtin_nanf("") : w;
}

}  // namespace

size_t pattern_state_bytes() { return sizeof(DevPlayState); }

hipError_t launch_pattern_playouts(const EncRecord *d_rec, int npos,
                                   int nplay, uint32_t seed0,
                                   const mc::TablesView &tables,
                                   void *d_states, int nstates,
                                   uint32_t *d_half_wins,
                                   uint32_t *d_overflow, float *d_winrate,
                                   hipStream_t stream) {
    if (npos <= 0) return hipSuccess;
    if (nstates < 64 || nstates % 64 != 0) return hipErrorInvalidValue;
    const int blocks = (npos + 63) / 64;
    hipLaunchKernelGGL(pat_zero, dim3(blocks), dim3(64), 0, stream,
                       d_half_wins, d_overflow, npos);
    const long long total = (long long)npos * nplay;
    if (total > 0) {
        // Every thread of the grid owns a state; no more threads than
        // playouts (whole wavefronts).
        long long threads = (total + 63) / 64 * 64;
        if (threads > nstates) threads = nstates;
        hipLaunchKernelGGL(pat_playout_kernel, dim3((unsigned)(threads / 64)),
                           dim3(64), 0, stream, d_rec, npos, nplay, seed0,
                           tables, static_cast<DevPlayState *>(d_states),
                           d_half_wins, d_overflow);
    }
    hipLaunchKernelGGL(pat_winrate, dim3(blocks), dim3(64), 0, stream,
                       d_half_wins, d_overflow, d_rec, npos, nplay, d_winrate);
    return hipGetLastError();
}

}  // namespace lz
