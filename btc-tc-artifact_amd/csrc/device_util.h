// Wave64 helpers for gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace btcx {

constexpr int kWave = 64;

// Length of a kernel's index range: a value fixed at launch, or one read
// from device memory when the kernel runs (Cnt::dev, plus `add`), which a
// captured launch needs when a preceding kernel computes it.
struct Cnt {
    int64_t v = 0;
    const uint32_t* p = nullptr;
    Cnt(int64_t x) : v(x) {}
    static Cnt dev(const uint32_t* q, int64_t add = 0) {
        Cnt c(add);
        c.p = q;
        return c;
    }
    __device__ __forceinline__ int64_t get() const { return p ? (int64_t)*p + v : v; }
};

// Make a value the compiler treats as per-lane into a wave-uniform one, so
// index loads derived from it go through the scalar unit.
__device__ __forceinline__ uint32_t uniform(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Workgroup sum of v (blockDim.x = NT, a multiple of 64), valid in thread
// 0; every thread calls it (lds: NT / 64 words, free again on return).
template <int NT>
__device__ __forceinline__ uint64_t block_sum_u64(uint64_t v, uint64_t* lds) {
    v = wave_sum_u64(v);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    uint64_t s = 0;
    if (threadIdx.x == 0)
#pragma unroll
        for (int w = 0; w < NT / 64; ++w) s += lds[w];
    __syncthreads();
    return s;
}

__device__ __forceinline__ uint32_t group_min_u32(uint32_t v, int width) {
    for (int o = width >> 1; o > 0; o >>= 1) {
        uint32_t w = __shfl_xor(v, o, width);
        v = w < v ? w : v;
    }
    return v;
}

__device__ __forceinline__ int popc64(uint64_t x) { return __popcll(x); }

}  // namespace btcx
