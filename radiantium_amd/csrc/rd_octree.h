// rd_octree.h — host builder of the reference's loose octree accelerator
// (Radiantium.Offline/Accelerators/Octree.cs), product side.
#pragma once
#include <stdint.h>
#include <string>
#include <vector>

namespace rd {

// One flattened node in the reference's 32-byte LinearNode layout
// (Octree.cs:28-35): the bound, then NodeCount (interior) / LeafIndex
// (leaf) and NextHead (-1 for a leaf, the first child's index otherwise).
struct OctNode {
    float mn[3], mx[3];
    int32_t count_or_leaf;
    int32_t next_head;
};
static_assert(sizeof(OctNode) == 32, "LinearNode is 32 bytes");

struct OctResult {
    std::vector<OctNode> nodes;        // root at 0; an interior node's children contiguous
    std::vector<int32_t> leaf_first;   // per leaf: first entry of leaf_prims
    std::vector<int32_t> leaf_count;
    std::vector<int32_t> leaf_prims;   // primitive indices, leaf after leaf (a primitive may repeat)
    int max_depth = 0;                 // deepest node built (root = 0)
    // entries the traversal's LIFO can hold at once, at worst: expanding a
    // node pushes its NodeCount children, and each ancestor still holds its
    // other children below them (every box hit), so the maximum over
    // interior nodes of NodeCount plus the ancestors' NodeCount - 1
    // (1 for a root leaf, 0 for an empty tree)
    int stack_needed = 0;
};

// Octree ctor (Octree.cs:48-93): RecursionBuild over the n primitive bounds
// (bmin/bmax: [n][3]) from the union of them all, then FlattenNode.
// max_depth / max_count / out_bound: the accel's "max_depth", "max_count",
// "out_bound".  An empty scene gives an empty tree.
bool build_octree(int n, const float* bmin, const float* bmax, int max_depth, int max_count, float out_bound,
                  OctResult* out, std::string* err);

}  // namespace rd
