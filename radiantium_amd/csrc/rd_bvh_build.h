// rd_bvh_build.h — host SAH BVH builder (product side).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

namespace rd {

// One flattened node, the reference's 32-byte LinearBVHNode layout.
struct NodeRec {
    float mn[3], mx[3];
    int32_t offset;
    uint16_t count;
    uint8_t axis;
    uint8_t _pad;
};
static_assert(sizeof(NodeRec) == 32, "LinearBVHNode is 32 bytes");

struct BvhResult {
    std::vector<NodeRec> nodes;  // depth-first, left child adjacent
    std::vector<int32_t> order;  // BVH order -> input primitive index
    int leaves = 0;
    int max_depth = 0;           // deepest node, root = 0
};

// Builds the binary SAH BVH of Accelerators/Bvh.cs (12 buckets, leaf when
// count <= max_prims and SAH cost >= count, two-primitive nodes always split
// by centroid order, equal-centroid ranges become one leaf) over n
// primitive bounds (bmin/bmax: [n][3]).  Returns false with *err on failure.
bool build_bvh(int n, const float* bmin, const float* bmax, int max_prims, BvhResult* out, std::string* err);

// The same tree built on the device (rd_bvh_gpu.hip): host bounds in, host
// records out, on stream s.
bool gpu_build_bvh(int n, const float* bmin, const float* bmax, int max_prims, hipStream_t s, BvhResult* out,
                   std::string* err);
// Device bounds in, device records out; returns the node count or -1.
int gpu_build_bvh_device(int n, const float* d_bmin, const float* d_bmax, int max_prims, NodeRec* d_nodes,
                         int32_t* d_order, hipStream_t s, int* max_depth, std::string* err);

}  // namespace rd
