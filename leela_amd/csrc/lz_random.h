// lz_random.h — the reference's random number generator, restated for the
// host side of this library: xoroshiro128+ with an LCG seed expansion
// (Random.cpp:27-57, Random.h:81-110). Streams are bit-identical to
// Random(seed) in the reference, which the position generator and the
// synthetic weight sets rely on for reproducibility.
#pragma once

#include <cstdint>
#include <limits>

// Host and device: the playout policy (mc_core.h) runs on both.
#if defined(__HIPCC__)
#define LZ_HD __host__ __device__
#else
#define LZ_HD
#endif

namespace lz {

class Rng {
public:
    LZ_HD explicit Rng(uint32_t seed) { seedrandom(seed); }

    // Random::seedrandom (Random.cpp:51-57). The first product is formed in
    // 32-bit unsigned arithmetic (int literal times uint32) and then
    // widened; the second multiplies the 64-bit state word.
    LZ_HD void seedrandom(uint32_t seed) {
        s_[0] = (uint64_t)(uint32_t)(741103597u * seed);
        s_[1] = 741103597ull * s_[0];
    }

    // Random::random (Random.cpp:27-37).
    LZ_HD uint64_t next() {
        const uint64_t s0 = s_[0];
        uint64_t s1 = s_[1];
        const uint64_t r = s0 + s1;
        s1 ^= s0;
        s_[0] = rotl(s0, 55) ^ s1 ^ (s1 << 14);
        s_[1] = rotl(s1, 36);
        return r;
    }

    // Random::randuint16 / randuint32(max) / randuint32() (Random.cpp:39-49).
    LZ_HD uint16_t randuint16(uint16_t max) {
        return (uint16_t)(((next() >> 48) * max) >> 16);
    }
    LZ_HD uint32_t randuint32(uint32_t max) {
        return (uint32_t)(((next() >> 32) * (uint64_t)max) >> 32);
    }
    LZ_HD uint32_t randuint32() { return (uint32_t)(next() >> 32); }

    // Random::randfix<MAX> (Random.h:81-89, 107-110).
    template <int MAX>
    LZ_HD uint32_t randfix() {
        static_assert(MAX > 0, "randfix range");
        if (MAX == 2) return next() > (std::numeric_limits<uint64_t>::max() / 2);
        return (uint32_t)(next() % MAX);
    }

    // Random::randflt (Random.cpp:59-65).
    LZ_HD float randflt() {
        const float umax = 1.0f / (float)(1u << 24);
        return (float)(uint32_t)(next() >> 40) * umax;
    }

    // Uniform double in (0, 1) from the top 53 bits (never 0 or 1); used by
    // the Box-Muller sampler of the synthetic weights.
    LZ_HD double uniform_open() {
        return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    }

private:
    LZ_HD static uint64_t rotl(uint64_t x, int k) {
        return (x << k) | (x >> (64 - k));
    }
    uint64_t s_[2];
};

// Seed of playout j of a position whose playouts are seeded `seed`
// (MurmurHash3's 32-bit finalizer of seed ^ j * 0x9E3779B9): independent
// streams, so the playouts of a position can run in any order or at once.
LZ_HD inline uint32_t playout_seed(uint32_t seed, uint32_t j) {
    uint32_t h = seed ^ (j * 0x9E3779B9u);
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

}  // namespace lz
