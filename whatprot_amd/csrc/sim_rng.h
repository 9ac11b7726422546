// Random numbers of the seeded simulators, shared by the host (simulate.cpp)
// and the GPU (simulate_gpu.hip) so both draw the same values: xoshiro256**
// (Blackman and Vigna) seeded by four splitmix64 outputs; uniform doubles
// from the top 53 bits; normals by Marsaglia's polar method. Stream g of a run
// with seed s starts from splitmix64 state s + (g + 1) * 0xD1B54A32D192ED03.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace wp {

struct Rng {
    uint64_t s[4];
    __host__ __device__ explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; i++) {
            x += 0x9e3779b97f4a7c15ull;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            s[i] = z ^ (z >> 31);
        }
    }
    __host__ __device__ static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    __host__ __device__ uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    __host__ __device__ double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    __host__ __device__ bool bernoulli(double p) { return uniform() < p; }
    // uniform integer in [lo, hi]
    __host__ __device__ int range(int lo, int hi) { return lo + (int)(next() % (uint64_t)(hi - lo + 1)); }
    __host__ __device__ double normal() {  // Marsaglia polar method
#pragma clang fp contract(off)
        for (;;) {
            double u = 2.0 * uniform() - 1.0, v = 2.0 * uniform() - 1.0;
            double s2 = u * u + v * v;
            if (s2 > 0.0 && s2 < 1.0) return u * sqrt(-2.0 * log(s2) / s2);
        }
    }
};

__host__ __device__ inline Rng stream_rng(uint64_t seed, uint64_t g) {
    return Rng(seed + (g + 1) * 0xD1B54A32D192ED03ull);
}

}  // namespace wp
