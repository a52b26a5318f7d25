// rd_sort.hip — coherence order of an extension-ray queue.
//
// Rays are keyed by their direction octant (3 bits, which fixes the
// near/far child choice at every node, Bvh.cs:211-222) above a 27-bit Morton
// code of the origin inside the scene's root box, then radix-sorted; the
// trace kernel walks the queue through the sorted permutation and writes
// every result back at the ray's own queue position, so the order is a
// schedule only: hits and visit counts are unchanged.
#include <rocprim/device/device_radix_sort.hpp>

#include "rd_kernels.h"

namespace rd {

__device__ __forceinline__ uint32_t spread3(uint32_t v) {   // 9 bits -> every third bit
    v &= 0x1ff;
    v = (v | (v << 16)) & 0x030000ff;
    v = (v | (v << 8)) & 0x0300f00f;
    v = (v | (v << 4)) & 0x030c30c3;
    v = (v | (v << 2)) & 0x09249249;
    return v;
}

__global__ __launch_bounds__(256) void k_ray_keys(const float4* __restrict__ nodes, RayQueue q, int n, int bits,
                                                  uint32_t* __restrict__ keys, int* __restrict__ vals) {
    float4 lo = nodes[0], hi = nodes[1];
    float mnx = lo.x, mny = lo.y, mnz = lo.z, mxx = lo.w, mxy = hi.x, mxz = hi.y;
    float sx = mxx > mnx ? 511.0f / (mxx - mnx) : 0.0f;
    float sy = mxy > mny ? 511.0f / (mxy - mny) : 0.0f;
    float sz = mxz > mnz ? 511.0f / (mxz - mnz) : 0.0f;
    for (int r = blockIdx.x * 256 + threadIdx.x; r < n; r += gridDim.x * 256) {
        float4 o = q.org[r], d = q.dir[r];
        uint32_t oct = (d.x < 0 ? 1u : 0u) | (d.y < 0 ? 2u : 0u) | (d.z < 0 ? 4u : 0u);
        auto cell = [](float v, float m, float s) {
            float c = (v - m) * s;
            c = c > 0.0f ? (c < 511.0f ? c : 511.0f) : 0.0f;   // NaN -> 0
            return (uint32_t)c;
        };
        uint32_t m = spread3(cell(o.x, mnx, sx)) | (spread3(cell(o.y, mny, sy)) << 1) |
                     (spread3(cell(o.z, mnz, sz)) << 2);
        // `bits` key bits: the octant above the top bits - 3 of the Morton code
        keys[r] = (oct << (bits - 3)) | (m >> (30 - bits));
        vals[r] = r;
    }
}

size_t ray_sort_temp_bytes(int max_n) {
    size_t bytes = 0;
    if (rocprim::radix_sort_pairs((void*)nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                  (const int*)nullptr, (int*)nullptr, max_n, 0, 30, (hipStream_t)0) != hipSuccess)
        return 0;
    return bytes;
}

hipError_t launch_ray_sort(const float4* nodes, RayQueue q, int n, int bits, uint32_t* keys, uint32_t* keys_alt,
                           int* vals, int* perm, void* temp, size_t temp_bytes, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (bits < 3) bits = 3;
    if (bits > 30) bits = 30;
    int blocks = (n + 255) / 256;
    if (blocks > persistent_blocks()) blocks = persistent_blocks();
    hipLaunchKernelGGL(k_ray_keys, dim3(blocks), dim3(256), 0, s, nodes, q, n, bits, keys, vals);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t bytes = temp_bytes;
    return rocprim::radix_sort_pairs(temp, bytes, keys, keys_alt, vals, perm, n, 0, (unsigned)bits, s);
}

}  // namespace rd
