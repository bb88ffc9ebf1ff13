// Host emulation of the few HIP device features the encoder kernel uses,
// for running its source on the CPU under test. Each workgroup runs as 64
// lanes (one wavefront), each lane a coroutine (ucontext) on one host
// thread, switched at synchronisation points only: __shared__ data is
// shared by the lanes, __syncthreads is a barrier of all 64, __ballot
// gathers the 64 predicates behind such a barrier, and DPP moves and
// __builtin_amdgcn_ballot_w64 exchange values within groups of 8 lanes
// (the only way the kernel uses them) behind a barrier of the group, so
// groups may take different paths as on the hardware; a DPP read from
// outside the group returns 0.
// Test infrastructure only; the product builds the kernel with hipcc.
#pragma once

#include <ucontext.h>

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static

typedef int hipError_t;
typedef void *hipStream_t;
constexpr hipError_t hipSuccess = 0;
constexpr hipError_t hipErrorInvalidValue = 1;
inline hipError_t hipGetLastError() { return hipSuccess; }

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

namespace wave_emul {

struct Idx {
    unsigned x = 0, y = 0, z = 0;
};

constexpr int kLanes = 64;
constexpr size_t kStack = 1 << 18;

struct Barrier {
    int n = 0, arrived = 0;
    unsigned gen = 0;
};

inline ucontext_t sched_ctx;
inline ucontext_t lane_ctx[kLanes];
inline bool lane_done[kLanes];
inline int cur = 0;
inline Idx thread_idx, block_idx;
inline Barrier wave_bar, group_bar[kLanes / 8];
inline bool pred[kLanes];
inline int gslot[kLanes];
inline std::vector<char> stacks;
inline Barrier *wait_bar[kLanes];  // the barrier a lane waits at, if any
inline unsigned wait_gen[kLanes];

inline void arrive(Barrier &b) {
    const unsigned gen = b.gen;
    if (++b.arrived == b.n) {
        b.arrived = 0;
        b.gen++;
        return;
    }
    const int me = cur;
    wait_bar[me] = &b;
    wait_gen[me] = gen;
    while (b.gen == gen) swapcontext(&lane_ctx[me], &sched_ctx);
    wait_bar[me] = nullptr;
}

}  // namespace wave_emul

#define threadIdx (wave_emul::thread_idx)
#define blockIdx (wave_emul::block_idx)

inline void __syncthreads() { wave_emul::arrive(wave_emul::wave_bar); }

inline unsigned long long __ballot(int p) {
    wave_emul::pred[threadIdx.x] = p != 0;
    __syncthreads();
    unsigned long long m = 0;
    for (int i = 0; i < 64; i++) m |= (unsigned long long)wave_emul::pred[i] << i;
    __syncthreads();
    return m;
}

inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }

// The kernel's group-scoped ballots: the caller's group of 8 lanes, bits
// in place, every other bit 0 (the kernel reads only its own group's
// byte of such a ballot).
inline unsigned long long __builtin_amdgcn_ballot_w64(bool p) {
    const int lane = (int)threadIdx.x, base = lane & ~7;
    wave_emul::Barrier &b = wave_emul::group_bar[lane >> 3];
    wave_emul::pred[lane] = p;
    wave_emul::arrive(b);
    unsigned long long m = 0;
    for (int i = base; i < base + 8; i++)
        m |= (unsigned long long)wave_emul::pred[i] << i;
    wave_emul::arrive(b);
    return m;
}

// DPP controls the kernels use: row_shl:1, row_shr:1/2/4, quad_perm
// [1,0,3,2] and [2,3,0,1], row_half_mirror.
inline int __builtin_amdgcn_update_dpp(int, int src, int ctrl, int, int, bool) {
    const int lane = (int)threadIdx.x, gl = lane & 7;
    wave_emul::Barrier &b = wave_emul::group_bar[lane >> 3];
    wave_emul::gslot[lane] = src;
    wave_emul::arrive(b);
    int from = -1;
    switch (ctrl) {
        case 0x101: from = gl == 7 ? -1 : lane + 1; break;
        case 0x111: from = gl < 1 ? -1 : lane - 1; break;
        case 0x112: from = gl < 2 ? -1 : lane - 2; break;
        case 0x114: from = gl < 4 ? -1 : lane - 4; break;
        case 0xB1: from = lane ^ 1; break;
        case 0x4E: from = lane ^ 2; break;
        case 0x141: from = (lane & ~7) | (7 - gl); break;
        default: std::abort();
    }
    const int r = from < 0 ? 0 : wave_emul::gslot[from];
    wave_emul::arrive(b);
    return r;
}

// The kernels order a group's LDS writes before the group's next reads
// with a wave barrier (on the hardware the wavefront executes in order);
// here the group's lanes meet.
inline void __builtin_amdgcn_wave_barrier() {
    wave_emul::arrive(wave_emul::group_bar[threadIdx.x >> 3]);
}

inline unsigned long long atomicOr(unsigned long long *p, unsigned long long v) {
    const unsigned long long o = *p;
    *p = o | v;
    return o;
}
inline unsigned atomicOr(unsigned *p, unsigned v) {
    const unsigned o = *p;
    *p = o | v;
    return o;
}
inline int atomicAdd(int *p, int v) {
    const int o = *p;
    *p = o + v;
    return o;
}
inline unsigned atomicAdd(unsigned *p, unsigned v) {
    const unsigned o = *p;
    *p = o + v;
    return o;
}

namespace wave_emul {

template <typename F>
struct Entry {
    static inline F *fn = nullptr;
    static void run(int lane) {
        cur = lane;
        thread_idx.x = (unsigned)lane;
        (*fn)();
        lane_done[lane] = true;
    }
};

}  // namespace wave_emul

// Lanes run in turn; a lane gives up the host thread only inside a
// barrier, so every read-modify-write between barriers is atomic.
template <typename K, typename... A>
void wave_emul_launch(K kernel, dim3 grid, dim3 block, A... args) {
    using namespace wave_emul;
    if (block.x != kLanes) std::abort();
    stacks.resize(kStack * kLanes);
    for (unsigned b = 0; b < grid.x; b++) {
        block_idx.x = b;
        wave_bar = Barrier{kLanes};
        for (auto &g : group_bar) g = Barrier{8};
        auto body = [&] { kernel(args...); };
        using E = Entry<decltype(body)>;
        E::fn = &body;
        for (int l = 0; l < kLanes; l++) {
            lane_done[l] = false;
            wait_bar[l] = nullptr;
            getcontext(&lane_ctx[l]);
            lane_ctx[l].uc_stack.ss_sp = stacks.data() + kStack * l;
            lane_ctx[l].uc_stack.ss_size = kStack;
            lane_ctx[l].uc_link = &sched_ctx;
            makecontext(&lane_ctx[l], (void (*)())E::run, 1, l);
        }
        for (;;) {
            bool all = true, ran = false;
            for (int l = 0; l < kLanes; l++) {
                if (lane_done[l]) continue;
                all = false;
                if (wait_bar[l] && wait_bar[l]->gen == wait_gen[l]) continue;
                cur = l;
                thread_idx.x = (unsigned)l;
                ran = true;
                swapcontext(&sched_ctx, &lane_ctx[l]);
            }
            if (all) break;
            if (!ran) std::abort();  // every live lane waits: a divergent barrier
        }
    }
}

#define hipLaunchKernelGGL(k, g, b, sh, st, ...) \
    wave_emul_launch(k, g, b, __VA_ARGS__)
