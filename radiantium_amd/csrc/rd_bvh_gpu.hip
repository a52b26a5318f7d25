// rd_bvh_gpu.hip — the SAH BVH of Bvh.RecursiveBuild + FlattenBVHTree
// (Accelerators/Bvh.cs:234-442), built on the device, node for node the tree
// rd_bvh_build.cpp builds on the host.
//
// Level by level: every node of a level is one block (k_bvh_level: 256
// threads while a level has few nodes, then one wavefront per node),
// which reduces its range's bounds and centroid bounds, evaluates the 12
// SAH buckets, and partitions its range the way MathExt.Partition
// (MathExt.cs:170-201) does, then appends its two children to the next
// level.  The host reads the next level's size once per level.  Then the
// flatten: subtree sizes bottom-up, pre-order offsets top-down, records.
//
// Equal results, not just an equal tree shape:
//   - bounds reduce with the reference's selects (BoundingBox3F.Union:
//     `min < x ? min : x`), which on ties keep the later element; a lane
//     folds a contiguous run of the range in order and lanes combine in lane
//     order, so a tie between +0 and -0 resolves as the host's scan does;
//   - bucket indices, costs and areas use the host's float expressions;
//   - the two-pointer partition swaps the i-th false element from the left
//     with the i-th true element from the right while the first lies left of
//     the second, and returns the count of true elements, except that when no
//     false element precedes the range's last element it returns the last
//     index (its loop stops there without testing it);
//   - pre-order numbering needs only the tree, so the order in which a
//     level's wavefronts append children does not show.
#include <algorithm>
#include <cfloat>
#include <string>
#include <vector>

#include "rd_bvh_build.h"
#include "rd_kernels.h"

namespace rd {

namespace {

constexpr int kBuckets = 12;
constexpr int kBigLevel = 512;   // levels of at most this many nodes use 256-thread blocks

struct BNode {
    float mn[3], mx[3];
    int start, count;   // primitive range
    int left, right;    // children (build ids), -1: leaf
    int axis, depth;
    int size, off;      // subtree nodes, pre-order index
};

__device__ __forceinline__ float sel_min(float a, float b) { return a < b ? a : b; }   // a earlier
__device__ __forceinline__ float sel_max(float a, float b) { return a > b ? a : b; }

// Lane-ordered reduction over the wavefront: the lower lane's value is the
// earlier one.  Every lane ends with the result.
__device__ __forceinline__ float wave_min_ordered(float v) {
    const int lane = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const float p = __shfl_xor(v, o);
        v = (lane & o) ? sel_min(p, v) : sel_min(v, p);
    }
    return v;
}
__device__ __forceinline__ float wave_max_ordered(float v) {
    const int lane = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const float p = __shfl_xor(v, o);
        v = (lane & o) ? sel_max(p, v) : sel_max(v, p);
    }
    return v;
}
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ float box_area(const float* mn, const float* mx) {   // BoundingBox3F.SurfaceArea
    const float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    return 2 * (dx * dy + dx * dz + dy * dz);
}

__device__ __forceinline__ void swap_rec(float4* lo, float4* hi, int a, int b) {
    const float4 la = lo[a], ha = hi[a], lb = lo[b], hb = hi[b];
    lo[a] = lb; hi[a] = hb;
    lo[b] = la; hi[b] = ha;
}

// Block-wide ordered reductions over kT threads (kT / 64 wavefronts): each
// wavefront reduces in lane order, then every thread folds the wavefronts'
// results in wavefront order.  lds: kT / 64 floats of scratch per value.
template <int kT>
__device__ __forceinline__ float block_min_ordered(float v, float* lds) {
    v = wave_min_ordered(v);
    if constexpr (kT == 64) return v;
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if (__lane_id() == 0) lds[w] = v;
    __syncthreads();
    float r = lds[0];
    for (int k = 1; k < kT / 64; k++) r = sel_min(r, lds[k]);
    return r;
}
template <int kT>
__device__ __forceinline__ float block_max_ordered(float v, float* lds) {
    v = wave_max_ordered(v);
    if constexpr (kT == 64) return v;
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if (__lane_id() == 0) lds[w] = v;
    __syncthreads();
    float r = lds[0];
    for (int k = 1; k < kT / 64; k++) r = sel_max(r, lds[k]);
    return r;
}
template <int kT>
__device__ __forceinline__ int block_sum(int v, int* lds) {
    v = wave_sum(v);
    if constexpr (kT == 64) return v;
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if (__lane_id() == 0) lds[w] = v;
    __syncthreads();
    int r = 0;
    for (int k = 0; k < kT / 64; k++) r += lds[k];
    return r;
}

// One node per block of kT threads (64 for levels of many small nodes, 256
// for the top levels, whose few nodes hold most primitives).  lo[i].w holds
// the primitive's input index.  err: 1 = leaf over 65535 primitives, 2 =
// empty child.
template <int kT>
__global__ __launch_bounds__(kT) void k_bvh_level(float4* __restrict__ lo, float4* __restrict__ hi,
                                                  int* __restrict__ fpos, int* __restrict__ tpos,
                                                  const int* __restrict__ level, BNode* __restrict__ bn,
                                                  int* __restrict__ counters, int* __restrict__ next, int max_prims,
                                                  int* __restrict__ err) {
    __shared__ float sred[kT / 64];
    __shared__ int sint[kT / 64];
    __shared__ int swt[kT / 64], swf[kT / 64];
    const int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
    const int id = level[blockIdx.x];
    const int start = bn[id].start, count = bn[id].count, end = start + count;
    // a contiguous run per thread, threads in order
    const int per = (count + kT - 1) / kT;
    const int r0 = min(start + tid * per, end), r1 = min(r0 + per, end);
    float bmn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, bmx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    float cmn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, cmx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int i = r0; i < r1; i++) {
        const float4 l = lo[i], h = hi[i];
        const float c[3] = {0.5f * l.x + 0.5f * h.x, 0.5f * l.y + 0.5f * h.y, 0.5f * l.z + 0.5f * h.z};
        const float lv[3] = {l.x, l.y, l.z}, hv[3] = {h.x, h.y, h.z};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            bmn[k] = sel_min(bmn[k], lv[k]);
            bmx[k] = sel_max(bmx[k], hv[k]);
            cmn[k] = sel_min(cmn[k], c[k]);
            cmx[k] = sel_max(cmx[k], c[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        bmn[k] = block_min_ordered<kT>(bmn[k], sred);
        bmx[k] = block_max_ordered<kT>(bmx[k], sred);
        cmn[k] = block_min_ordered<kT>(cmn[k], sred);
        cmx[k] = block_max_ordered<kT>(cmx[k], sred);
    }
    if (tid == 0) {
        for (int k = 0; k < 3; k++) { bn[id].mn[k] = bmn[k]; bn[id].mx[k] = bmx[k]; }
    }
    auto leaf = [&]() {
        if (tid == 0) {
            bn[id].left = bn[id].right = -1;
            bn[id].axis = 0;
            if (count > 65535) atomicOr(err, 1);
        }
    };
    if (count == 1) { leaf(); return; }
    // BoundingBox3F.MaximumExtent
    const float d0 = cmx[0] - cmn[0], d1 = cmx[1] - cmn[1], d2 = cmx[2] - cmn[2];
    const int dim = (d0 > d1 && d0 > d2) ? 0 : (d1 > d2 ? 1 : 2);
    const float cdmin = dim == 0 ? cmn[0] : (dim == 1 ? cmn[1] : cmn[2]);
    const float cdmax = dim == 0 ? cmx[0] : (dim == 1 ? cmx[1] : cmx[2]);
    if (cdmax == cdmin) { leaf(); return; }
    auto cen_dim = [&](const float4& l, const float4& h) {
        return dim == 0 ? 0.5f * l.x + 0.5f * h.x : (dim == 1 ? 0.5f * l.y + 0.5f * h.y : 0.5f * l.z + 0.5f * h.z);
    };
    int mid;
    if (count <= 2) {
        mid = start + 1;
        if (tid == 0) {
            const float lv = cen_dim(lo[start], hi[start]), rv = cen_dim(lo[start + 1], hi[start + 1]);
            if (!(lv == rv) && !(lv < rv)) swap_rec(lo, hi, start, start + 1);   // SwapIfGreater
        }
    } else {
        const float ext = cdmax - cdmin;
        auto bucket_of = [&](const float4& l, const float4& h) {
            return (int)(kBuckets * ((cen_dim(l, h) - cdmin) / ext));
        };
        int cnt[kBuckets];
        float bbmn[kBuckets][3], bbmx[kBuckets][3];
#pragma unroll
        for (int b = 0; b < kBuckets; b++) {
            cnt[b] = 0;
#pragma unroll
            for (int k = 0; k < 3; k++) { bbmn[b][k] = FLT_MAX; bbmx[b][k] = -FLT_MAX; }
        }
        for (int i = r0; i < r1; i++) {
            const float4 l = lo[i], h = hi[i];
            int bk = bucket_of(l, h);
            if (bk == kBuckets) bk = kBuckets - 1;
            const float lv[3] = {l.x, l.y, l.z}, hv[3] = {h.x, h.y, h.z};
#pragma unroll
            for (int b = 0; b < kBuckets; b++)
                if (bk == b) {
                    cnt[b]++;
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        bbmn[b][k] = sel_min(bbmn[b][k], lv[k]);
                        bbmx[b][k] = sel_max(bbmx[b][k], hv[k]);
                    }
                }
        }
#pragma unroll
        for (int b = 0; b < kBuckets; b++) {
            cnt[b] = block_sum<kT>(cnt[b], sint);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                bbmn[b][k] = block_min_ordered<kT>(bbmn[b][k], sred);
                bbmx[b][k] = block_max_ordered<kT>(bbmx[b][k], sred);
            }
        }
        // costs as the host computes them (every thread holds the same values)
        float cost[kBuckets - 1];
        const float area = box_area(bmn, bmx);
#pragma unroll
        for (int i = 0; i < kBuckets - 1; i++) {
            float b0n[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, b0x[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            float b1n[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, b1x[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            int c0 = 0, c1 = 0;
#pragma unroll
            for (int j = 0; j < kBuckets; j++) {
                if (j <= i) {
#pragma unroll
                    for (int k = 0; k < 3; k++) { b0n[k] = sel_min(b0n[k], bbmn[j][k]); b0x[k] = sel_max(b0x[k], bbmx[j][k]); }
                    c0 += cnt[j];
                } else {
#pragma unroll
                    for (int k = 0; k < 3; k++) { b1n[k] = sel_min(b1n[k], bbmn[j][k]); b1x[k] = sel_max(b1x[k], bbmx[j][k]); }
                    c1 += cnt[j];
                }
            }
            cost[i] = 1 + (c0 * box_area(b0n, b0x) + c1 * box_area(b1n, b1x)) / area;
        }
        float min_cost = cost[0];
        int split = 0;
#pragma unroll
        for (int i = 1; i < kBuckets - 1; i++)
            if (cost[i] < min_cost) { min_cost = cost[i]; split = i; }
        const float leaf_cost = (float)count;
        if (!(count > max_prims || min_cost < leaf_cost)) { leaf(); return; }
        // MathExt.Partition(list, start, end - 1, bucket <= split)
        int ntrue = 0;
        for (int i = start + tid; i < end; i += kT) ntrue += bucket_of(lo[i], hi[i]) <= split ? 1 : 0;
        ntrue = block_sum<kT>(ntrue, sint);
        const int nfalse = count - ntrue;
        int nf = 0, nt = 0;   // false / true elements before the current chunk
        const uint64_t below = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
        for (int base = start; base < end; base += kT) {
            const int i = base + tid;
            const bool valid = i < end;
            const bool t = valid && bucket_of(lo[i], hi[i]) <= split;
            const uint64_t tm = __ballot(t), fm = __ballot(valid && !t);
            int wt = 0, wf = 0, ct = __popcll(tm), cf = __popcll(fm);   // earlier wavefronts of the chunk
            if constexpr (kT > 64) {
                __syncthreads();
                if (lane == 0) { swt[wave] = ct; swf[wave] = cf; }
                __syncthreads();
                ct = 0; cf = 0;
                for (int k = 0; k < kT / 64; k++) {
                    if (k < wave) { wt += swt[k]; wf += swf[k]; }
                    ct += swt[k];
                    cf += swf[k];
                }
            }
            if (valid) {
                if (t) tpos[start + (ntrue - 1 - (nt + wt + __popcll(tm & below)))] = i;   // rank from the right
                else fpos[start + nf + wf + __popcll(fm & below)] = i;                      // rank from the left
            }
            nt += ct;
            nf += cf;
        }
        __syncthreads();
        const int m = min(nfalse, ntrue);
        const int first_false = nfalse > 0 ? fpos[start] : -1;
        for (int k = tid; k < m; k += kT) {
            const int p = fpos[start + k], q = tpos[start + k];
            if (p < q) swap_rec(lo, hi, p, q);
        }
        mid = (nfalse == 0 || first_false == end - 1) ? end - 1 : start + ntrue;
    }
    if (tid == 0) {
        if (mid <= start || mid >= end) {
            atomicOr(err, 2);
            bn[id].left = bn[id].right = -1;
            return;
        }
        const int c = atomicAdd(counters + 0, 2);   // build ids
        BNode L{}, R{};
        L.start = start; L.count = mid - start; L.depth = bn[id].depth + 1; L.left = L.right = -1;
        R.start = mid; R.count = end - mid; R.depth = L.depth; R.left = R.right = -1;
        bn[c] = L;
        bn[c + 1] = R;
        bn[id].left = c;
        bn[id].right = c + 1;
        bn[id].axis = dim;
        const int k = atomicAdd(counters + 1, 2);   // next level
        next[k] = c;
        next[k + 1] = c + 1;
    }
}

__global__ __launch_bounds__(256) void k_bvh_sizes(const int* __restrict__ level, int n, BNode* __restrict__ bn) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int id = level[i];
    bn[id].size = bn[id].left < 0 ? 1 : 1 + bn[bn[id].left].size + bn[bn[id].right].size;
}

__global__ __launch_bounds__(256) void k_bvh_offsets(const int* __restrict__ level, int n, BNode* __restrict__ bn) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int id = level[i];
    if (bn[id].left < 0) return;
    const int o = bn[id].off;
    bn[bn[id].left].off = o + 1;
    bn[bn[id].right].off = o + 1 + bn[bn[id].left].size;
}

__global__ __launch_bounds__(256) void k_bvh_write(const BNode* __restrict__ bn, int nnodes, NodeRec* __restrict__ out,
                                                   const float4* __restrict__ lo, int nprims,
                                                   int32_t* __restrict__ order) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < nnodes) {
        const BNode b = bn[i];
        NodeRec r;
        for (int k = 0; k < 3; k++) { r.mn[k] = b.mn[k]; r.mx[k] = b.mx[k]; }
        if (b.left < 0) {
            r.offset = b.start;
            r.count = (uint16_t)b.count;
            r.axis = 0;
        } else {
            r.offset = bn[b.right].off;
            r.count = 0;
            r.axis = (uint8_t)b.axis;
        }
        r._pad = 0;
        out[b.off] = r;
    }
    if (i < nprims) order[i] = __float_as_int(lo[i].w);
}

__global__ __launch_bounds__(256) void k_bvh_init(const float* __restrict__ bmin, const float* __restrict__ bmax, int n,
                                                  float4* __restrict__ lo, float4* __restrict__ hi) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    lo[i] = make_float4(bmin[3 * i], bmin[3 * i + 1], bmin[3 * i + 2], __int_as_float(i));
    hi[i] = make_float4(bmax[3 * i], bmax[3 * i + 1], bmax[3 * i + 2], 0.0f);
}

__global__ void k_bvh_root(BNode* bn, int n, int* counters, int* level0) {
    BNode r{};
    r.start = 0; r.count = n; r.left = r.right = -1; r.depth = 0; r.off = 0;
    bn[0] = r;
    counters[0] = 1;   // build ids used
    counters[1] = 0;   // next level's size
    level0[0] = 0;
}

__global__ void k_zero_int(int* p) { *p = 0; }

}  // namespace

// Device build from device bounds (bmin / bmax: [n][3] floats).  nodes: at
// least 2n - 1 records; order: n indices.  Returns the node count, or -1 with
// *err set.
int gpu_build_bvh_device(int n, const float* d_bmin, const float* d_bmax, int max_prims, NodeRec* d_nodes,
                         int32_t* d_order, hipStream_t s, int* max_depth, std::string* err) {
    *max_depth = 0;
    if (n <= 0) return 0;
    if (max_prims < 1) max_prims = 1;
    const size_t cap = 2 * (size_t)n;
    float4 *lo = nullptr, *hi = nullptr;
    int *fpos = nullptr, *tpos = nullptr, *levels = nullptr, *counters = nullptr;
    BNode* bn = nullptr;
    int* hcnt = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(lo); (void)hipFree(hi); (void)hipFree(fpos); (void)hipFree(tpos);
        (void)hipFree(levels); (void)hipFree(counters); (void)hipFree(bn);
        if (hcnt) (void)hipHostFree(hcnt);
    };
    if (hipMalloc(&lo, sizeof(float4) * n) != hipSuccess || hipMalloc(&hi, sizeof(float4) * n) != hipSuccess ||
        hipMalloc(&fpos, sizeof(int) * n) != hipSuccess || hipMalloc(&tpos, sizeof(int) * n) != hipSuccess ||
        hipMalloc(&levels, sizeof(int) * cap) != hipSuccess || hipMalloc(&counters, sizeof(int) * 4) != hipSuccess ||
        hipMalloc(&bn, sizeof(BNode) * cap) != hipSuccess ||
        hipHostMalloc(&hcnt, sizeof(int) * 4, hipHostMallocDefault) != hipSuccess) {
        cleanup();
        *err = "device memory for the BVH build";
        return -1;
    }
    const int g = (n + 255) / 256;
    hipLaunchKernelGGL(k_bvh_init, dim3(g), dim3(256), 0, s, d_bmin, d_bmax, n, lo, hi);
    hipLaunchKernelGGL(k_bvh_root, dim3(1), dim3(1), 0, s, bn, n, counters, levels);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, counters + 2);   // error flags
    // levels[] holds the levels back to back: [lvl_off[d], lvl_off[d] + lvl_n[d])
    std::vector<int> lvl_off{0}, lvl_n{1};
    for (int d = 0;; d++) {
        const int cur = lvl_off[d], ncur = lvl_n[d];
        // few, large nodes: 256 threads each; many small ones: a wavefront each
        if (ncur <= kBigLevel)
            hipLaunchKernelGGL(k_bvh_level<256>, dim3(ncur), dim3(256), 0, s, lo, hi, fpos, tpos, levels + cur, bn,
                               counters, levels + cur + ncur, max_prims, counters + 2);
        else
            hipLaunchKernelGGL(k_bvh_level<64>, dim3(ncur), dim3(64), 0, s, lo, hi, fpos, tpos, levels + cur, bn,
                               counters, levels + cur + ncur, max_prims, counters + 2);
        if (hipMemcpyAsync(hcnt, counters, sizeof(int) * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            cleanup();
            *err = "BVH build level failed";
            return -1;
        }
        if (hcnt[2] & 1) { cleanup(); *err = "BVH leaf holds more than 65535 primitives"; return -1; }
        if (hcnt[2] & 2) { cleanup(); *err = "BVH split produced an empty child"; return -1; }
        const int nnext = hcnt[1];
        if (nnext == 0) break;
        *max_depth = d + 1;
        lvl_off.push_back(cur + ncur);
        lvl_n.push_back(nnext);
        if ((size_t)(cur + ncur + nnext) > cap || hcnt[0] > (int)cap) {
            cleanup();
            *err = "BVH build overflow";
            return -1;
        }
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, counters + 1);
    }
    const int nnodes = hcnt[0];
    for (int d = (int)lvl_n.size() - 1; d >= 0; d--)
        hipLaunchKernelGGL(k_bvh_sizes, dim3((lvl_n[d] + 255) / 256), dim3(256), 0, s, levels + lvl_off[d], lvl_n[d], bn);
    for (size_t d = 0; d < lvl_n.size(); d++)
        hipLaunchKernelGGL(k_bvh_offsets, dim3((lvl_n[d] + 255) / 256), dim3(256), 0, s, levels + lvl_off[d], lvl_n[d], bn);
    const int w = std::max(nnodes, n);
    hipLaunchKernelGGL(k_bvh_write, dim3((w + 255) / 256), dim3(256), 0, s, bn, nnodes, d_nodes, lo, n, d_order);
    hipError_t e = hipStreamSynchronize(s);
    cleanup();
    if (e != hipSuccess) { *err = std::string("BVH build: ") + hipGetErrorString(e); return -1; }
    return nnodes;
}

// Host bounds in, host result out (the shape build_bvh returns).
bool gpu_build_bvh(int n, const float* bmin, const float* bmax, int max_prims, hipStream_t s, BvhResult* out,
                   std::string* err) {
    out->nodes.clear();
    out->order.clear();
    out->leaves = 0;
    out->max_depth = 0;
    if (n <= 0) return true;
    float *db = nullptr, *dx = nullptr;
    NodeRec* dn = nullptr;
    int32_t* dord = nullptr;
    auto cleanup = [&]() { (void)hipFree(db); (void)hipFree(dx); (void)hipFree(dn); (void)hipFree(dord); };
    if (hipMalloc(&db, sizeof(float) * 3 * (size_t)n) != hipSuccess ||
        hipMalloc(&dx, sizeof(float) * 3 * (size_t)n) != hipSuccess ||
        hipMalloc(&dn, sizeof(NodeRec) * 2 * (size_t)n) != hipSuccess ||
        hipMalloc(&dord, sizeof(int32_t) * (size_t)n) != hipSuccess) {
        cleanup();
        *err = "device memory for the BVH build";
        return false;
    }
    if (hipMemcpyAsync(db, bmin, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dx, bmax, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, s) != hipSuccess) {
        cleanup();
        *err = "BVH bounds upload";
        return false;
    }
    int depth = 0;
    const int nn = gpu_build_bvh_device(n, db, dx, max_prims, dn, dord, s, &depth, err);
    if (nn < 0) { cleanup(); return false; }
    out->nodes.resize(nn);
    out->order.resize(n);
    if (hipMemcpyAsync(out->nodes.data(), dn, sizeof(NodeRec) * nn, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(out->order.data(), dord, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        cleanup();
        *err = "BVH download";
        return false;
    }
    cleanup();
    for (const NodeRec& r : out->nodes) out->leaves += r.count > 0 ? 1 : 0;
    out->max_depth = depth;
    return true;
}

}  // namespace rd
