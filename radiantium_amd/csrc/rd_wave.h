// rd_wave.h — device helpers shared by the wavefront kernels: path flags,
// wavefront-aggregated queue appends and counters, and the wave's
// path-index -> (pixel, sample) mapping.
#pragma once
#include "rd_kernels.h"
#include "rd_math.h"

namespace rd {

enum : int {
    kFlagSpecular = 1,
    kFlagNeePending = 2,
    kFlagShadowIssued = 4,
    kFlagMisIssued = 8,
    kFlagDone = 16,        // the path ended this bounce (retired after resolve)
    kFlagVertex = 32,      // adjoint replay: a material vertex recorded its derivative terms
};

// ---------------------------------------------------------------- wave helpers
// Appends one slot per lane with `pred` set; one atomic per wavefront.
// Must be reached by every lane of the wavefront (converged).
__device__ __forceinline__ int wave_append(int* counter, bool pred) {
    uint64_t mask = __ballot(pred);
    if (mask == 0) return -1;
    int lane = __lane_id();
    int leader = __ffsll((unsigned long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader);
    uint64_t below = lane == 0 ? 0ull : (mask & ((~0ull) >> (64 - lane)));
    return pred ? base + __popcll(below) : -1;
}

// As wave_append, with one atomic per block of kBlock threads (kBlock / 64
// wavefronts): slots are handed out wave by wave in block order.  Every
// thread of the block must reach it (block-uniform control flow); `sh` is
// kBlock / 64 + 1 ints of shared memory, reused only after the call returns.
template <int kBlock>
__device__ __forceinline__ int block_append(int* counter, bool pred, int* sh) {
    constexpr int kWaves = kBlock / 64;
    const uint64_t mask = __ballot(pred);
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = __popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        int total = 0;
        for (int w = 0; w < kWaves; w++) {
            const int c = sh[w];
            sh[w] = total;   // the wave's offset within the block's slots
            total += c;
        }
        sh[kWaves] = total ? atomicAdd(counter, total) : 0;
    }
    __syncthreads();
    const uint64_t below = lane == 0 ? 0ull : (mask & ((~0ull) >> (64 - lane)));
    const int slot = sh[kWaves] + sh[wave] + __popcll(below);
    __syncthreads();
    return pred ? slot : -1;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ void add_counter(unsigned long long* c, uint64_t v) {
    uint64_t s = wave_sum_u64(v);
    if (__lane_id() == 0 && s) atomicAdd(c, (unsigned long long)s);
}

// Per query class (0 extension, 1 MIS, 2 shadow): rays, node visits and
// primitive tests, for the algorithmic-bytes accounting of the trace kernels.
__device__ __forceinline__ void add_class_counters(int* counts, int cls, uint64_t rays, uint64_t nv, uint64_t nt) {
    add_counter((unsigned long long*)(counts + kCntClassRays) + cls, rays);
    add_counter((unsigned long long*)(counts + kCntClassNodes) + cls, nv);
    add_counter((unsigned long long*)(counts + kCntClassPrims) + cls, nt);
}

__device__ __forceinline__ f3 ld3(const float4* a, int i) { float4 v = a[i]; return mk(v.x, v.y, v.z); }
__device__ __forceinline__ void st3(float4* a, int i, f3 v, float w = 0.0f) { a[i] = make_float4(v.x, v.y, v.z, w); }

}  // namespace rd
