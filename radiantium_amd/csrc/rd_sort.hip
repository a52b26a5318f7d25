// rd_sort.hip — coherence order of an extension-ray queue.
//
// Rays are keyed by their direction octant (3 bits, which fixes the
// near/far child choice at every node, Bvh.cs:211-222) above a 27-bit Morton
// code of the origin inside the scene's root box, then sorted by a
// least-significant-digit radix sort written for this chip (8-bit digits, one
// histogram / scan / scatter triple per digit); the trace kernel walks the
// queue through the sorted permutation and writes every result back at the
// ray's own queue position, so the order is a schedule only: hits and visit
// counts are unchanged.
//
// Per digit pass, over tiles of kTile keys (256 threads x 16):
//   k_sort_hist     tile histograms, digit-major: hist[d * ntiles + t]
//   k_sort_scan     one block per digit: exclusive scan of its row over the
//                   tiles, the row's total into totals[d]
//   k_sort_scatter  ranks every key of a tile among the tile's keys of its
//                   digit (in index order, so the sort is stable), stages the
//                   tile in LDS ordered by digit, and writes each digit's run
//                   to base(d) + hist[d * ntiles + t], base(d) = the totals of
//                   the smaller digits: consecutive lanes write consecutive
//                   addresses.
// No decoupled look-back: a pass's blocks never wait on each other, which on
// a chip whose L2 is split over eight dies costs a device-scope round trip
// per tile.
#include <algorithm>

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
                                                  uint32_t* __restrict__ keys) {
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
    }
}

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kTile = kSortThreads * kSortItems;

__device__ __forceinline__ uint64_t lanes_below() {
    int lane = __lane_id();
    return lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
}

// Lanes of the wavefront (among `valid`) holding the same 8-bit digit.
__device__ __forceinline__ uint64_t digit_peers(uint32_t d, bool valid) {
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
        uint64_t ones = __ballot(valid && ((d >> b) & 1u));
        m &= ((d >> b) & 1u) ? ones : ~ones;
    }
    return m;
}

__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const uint32_t* __restrict__ keys, int n, int shift,
                                                            int ntiles, int* __restrict__ hist) {
    __shared__ int h[256];
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        h[threadIdx.x] = 0;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kSortItems; i++) {
            const int idx = t * kTile + i * kSortThreads + threadIdx.x;
            const bool valid = idx < n;
            const uint32_t d = valid ? (keys[idx] >> shift) & 255u : 0u;
            const uint64_t peers = digit_peers(d, valid);
            // the lowest lane of each digit group adds the group's size
            if (valid && (peers & lanes_below()) == 0) atomicAdd(&h[d], __popcll(peers));
        }
        __syncthreads();
        hist[threadIdx.x * ntiles + t] = h[threadIdx.x];
        __syncthreads();
    }
}

// Block d: exclusive scan of hist[d * ntiles ...] in place, total to totals[d].
__global__ __launch_bounds__(kSortThreads) void k_sort_scan(int* __restrict__ hist, int ntiles,
                                                            int* __restrict__ totals) {
    __shared__ int wsum[kSortThreads / 64];
    __shared__ int carry;
    int* row = hist + (size_t)blockIdx.x * ntiles;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = __lane_id();
    for (int base = 0; base < ntiles; base += kSortThreads) {
        const int i = base + threadIdx.x;
        const int v = i < ntiles ? row[i] : 0;
        // inclusive wavefront scan
        int x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        int before = carry;
        for (int w = 0; w < wave; w++) before += wsum[w];
        if (i < ntiles) row[i] = before + x - v;
        __syncthreads();
        if (threadIdx.x == kSortThreads - 1) carry = before + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// One digit pass: keys_in -> keys_out (null: not written), vals_in (null:
// the identity) -> vals_out.
__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(const uint32_t* __restrict__ keys_in,
                                                               const int* __restrict__ vals_in, int n, int shift,
                                                               int ntiles, const int* __restrict__ hist,
                                                               const int* __restrict__ totals,
                                                               uint32_t* __restrict__ keys_out,
                                                               int* __restrict__ vals_out) {
    __shared__ uint32_t skey[kTile];
    __shared__ int sval[kTile];
    __shared__ int base[256];       // first output position of digit d (all tiles)
    __shared__ int gpos[256];       // this tile's first output position of digit d
    __shared__ int cnt[256];        // keys of digit d ranked so far in this tile
    __shared__ int wcnt[kSortThreads / 64][256];
    __shared__ int wsum[kSortThreads / 64];
    const int tid = threadIdx.x, wave = tid >> 6, lane = __lane_id();
    // base = exclusive scan of totals
    {
        const int v = totals[tid];
        int x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        int before = 0;
        for (int w = 0; w < wave; w++) before += wsum[w];
        base[tid] = before + x - v;
        __syncthreads();
    }
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        gpos[tid] = base[tid] + hist[tid * ntiles + t];
        cnt[tid] = 0;
        uint32_t key[kSortItems];
        int val[kSortItems], rank[kSortItems];
#pragma unroll
        for (int i = 0; i < kSortItems; i++) {
            const int idx = t * kTile + i * kSortThreads + tid;
            const bool valid = idx < n;
            key[i] = valid ? keys_in[idx] : 0u;
            val[i] = valid ? (vals_in ? vals_in[idx] : idx) : 0;
            const uint32_t d = (key[i] >> shift) & 255u;
            const uint64_t peers = digit_peers(d, valid);
#pragma unroll
            for (int w = 0; w < kSortThreads / 64; w++) wcnt[w][tid] = 0;
            __syncthreads();
            if (valid && (peers & lanes_below()) == 0) wcnt[wave][d] = __popcll(peers);
            __syncthreads();
            if (valid) {
                int r = cnt[d] + __popcll(peers & lanes_below());
                for (int w = 0; w < wave; w++) r += wcnt[w][d];
                rank[i] = r;
            } else {
                rank[i] = -1;
            }
            __syncthreads();
            int add = 0;
#pragma unroll
            for (int w = 0; w < kSortThreads / 64; w++) add += wcnt[w][tid];
            cnt[tid] += add;
        }
        __syncthreads();
        // tile-local start of each digit: exclusive scan of cnt (reuses wsum)
        int start;
        {
            const int v = cnt[tid];
            int x = v;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                int y = __shfl_up(x, o);
                if (lane >= o) x += y;
            }
            if (lane == 63) wsum[wave] = x;
            __syncthreads();
            int before = 0;
            for (int w = 0; w < wave; w++) before += wsum[w];
            start = before + x - v;
        }
        __syncthreads();
        cnt[tid] = start;   // cnt now holds the digit's start inside the staged tile
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kSortItems; i++) {
            if (rank[i] < 0) continue;
            const int slot = cnt[(key[i] >> shift) & 255u] + rank[i];
            skey[slot] = key[i];
            sval[slot] = val[i];
        }
        __syncthreads();
        const int tn = min(kTile, n - t * kTile);
        for (int sidx = tid; sidx < tn; sidx += kSortThreads) {
            const uint32_t k = skey[sidx];
            const uint32_t d = (k >> shift) & 255u;
            const int pos = gpos[d] + (sidx - cnt[d]);
            if (keys_out) keys_out[pos] = k;
            vals_out[pos] = sval[sidx];
        }
        __syncthreads();
    }
}

static int sort_tiles(int n) { return (n + kTile - 1) / kTile; }

size_t ray_sort_temp_bytes(int max_n) {
    return ((size_t)256 * (size_t)sort_tiles(max_n > 0 ? max_n : 1) + 256) * sizeof(int);
}

hipError_t launch_ray_sort(const float4* nodes, RayQueue q, int n, int bits, uint32_t* keys, uint32_t* keys_alt,
                           int* vals, int* perm, void* temp, size_t temp_bytes, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (bits < 3) bits = 3;
    if (bits > 30) bits = 30;
    const int ntiles = sort_tiles(n);
    if (temp_bytes < ray_sort_temp_bytes(n)) return hipErrorInvalidValue;
    int* hist = (int*)temp;
    int* totals = hist + (size_t)256 * ntiles;
    int blocks = (n + 255) / 256;
    if (blocks > persistent_blocks()) blocks = persistent_blocks();
    hipLaunchKernelGGL(k_ray_keys, dim3(blocks), dim3(256), 0, s, nodes, q, n, bits, keys);
    const int tblocks = std::min(ntiles, persistent_blocks());
    const int npass = (bits + 7) / 8;
    const uint32_t* kin = keys;
    const int* vin = nullptr;   // the identity
    for (int p = 0; p < npass; p++) {
        const bool last = p == npass - 1;
        uint32_t* kout = last ? nullptr : (kin == keys ? keys_alt : keys);
        int* vout = ((npass - 1 - p) % 2 == 0) ? perm : vals;   // the last pass writes perm
        hipLaunchKernelGGL(k_sort_hist, dim3(tblocks), dim3(kSortThreads), 0, s, kin, n, 8 * p, ntiles, hist);
        hipLaunchKernelGGL(k_sort_scan, dim3(256), dim3(kSortThreads), 0, s, hist, ntiles, totals);
        hipLaunchKernelGGL(k_sort_scatter, dim3(tblocks), dim3(kSortThreads), 0, s, kin, vin, n, 8 * p, ntiles,
                           (const int*)hist, (const int*)totals, kout, vout);
        kin = kout;
        vin = vout;
    }
    return hipGetLastError();
}

}  // namespace rd
