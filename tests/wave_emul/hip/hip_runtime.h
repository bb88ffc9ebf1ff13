// Host emulation of the few HIP device features the encoder kernel uses,
// for running its source on the CPU under test: each workgroup runs as 64
// host threads (one wavefront), __shared__ data is shared by them,
// __syncthreads is a barrier and __ballot gathers the 64 predicates.
// Test infrastructure only; the product builds the kernel with hipcc.
#pragma once

#include <atomic>
#include <barrier>
#include <cstddef>
#include <cstdint>
#include <thread>
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
inline hipError_t hipGetLastError() { return hipSuccess; }

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

namespace wave_emul {
struct Idx {
    unsigned x = 0, y = 0, z = 0;
};
inline thread_local Idx thread_idx, block_idx;
inline std::barrier<> *bar = nullptr;
inline bool pred[64];
}  // namespace wave_emul

#define threadIdx (wave_emul::thread_idx)
#define blockIdx (wave_emul::block_idx)

inline void __syncthreads() { wave_emul::bar->arrive_and_wait(); }

inline unsigned long long __ballot(int p) {
    wave_emul::pred[threadIdx.x] = p != 0;
    wave_emul::bar->arrive_and_wait();
    unsigned long long m = 0;
    for (int i = 0; i < 64; i++) m |= (unsigned long long)wave_emul::pred[i] << i;
    wave_emul::bar->arrive_and_wait();
    return m;
}

inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }

inline unsigned long long atomicOr(unsigned long long *p, unsigned long long v) {
    return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST);
}
inline unsigned atomicOr(unsigned *p, unsigned v) {
    return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST);
}
inline int atomicAdd(int *p, int v) {
    return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}

template <typename K, typename... A>
void wave_emul_launch(K kernel, dim3 grid, dim3 block, A... args) {
    for (unsigned b = 0; b < grid.x; b++) {
        std::barrier<> br((std::ptrdiff_t)block.x);
        wave_emul::bar = &br;
        std::vector<std::thread> th;
        for (unsigned t = 0; t < block.x; t++)
            th.emplace_back([&, t] {
                wave_emul::thread_idx.x = t;
                wave_emul::block_idx.x = b;
                kernel(args...);
            });
        for (auto &x : th) x.join();
    }
}

#define hipLaunchKernelGGL(k, g, b, sh, st, ...) \
    wave_emul_launch(k, g, b, __VA_ARGS__)
