// rd_octree.cpp — the reference's loose octree (Accelerators/Octree.cs),
// built on the host.  The build decides which leaves hold a primitive and
// in which order the traversal meets them, so it follows RecursionBuild
// (:197-281) and FlattenNode (:283-315) step for step: children in corner
// order (bit 0 x, bit 1 y, bit 2 z, BoundingBox3F.GetCorner), a primitive
// into the first child whose inner box strictly contains its bound's centre
// and whose loosened box strictly contains its bound, else into every child
// whose inner box it strictly overlaps (a bound lying flat on a child's
// face overlaps none and is dropped, as in the reference), each child's box
// then shrunk to the union of its primitives' bounds.
#include "rd_octree.h"

#include <algorithm>
#include <new>
#include <utility>

namespace rd {

namespace {

struct Box {
    float mn[3], mx[3];
};

Box empty_box() {   // BoundingBox3F(): Min = float.MaxValue, Max = float.MinValue
    Box b;
    for (int k = 0; k < 3; k++) { b.mn[k] = 3.40282347e+38f; b.mx[k] = -3.40282347e+38f; }
    return b;
}
// Vector3.Min / Max per component
inline float vmin(float a, float b) { return a < b ? a : b; }
inline float vmax(float a, float b) { return a > b ? a : b; }
void box_union(Box& a, const Box& b) {
    for (int k = 0; k < 3; k++) { a.mn[k] = vmin(a.mn[k], b.mn[k]); a.mx[k] = vmax(a.mx[k], b.mx[k]); }
}
bool contains_point(const Box& b, const float* p) {
    return p[0] > b.mn[0] && p[0] < b.mx[0] && p[1] > b.mn[1] && p[1] < b.mx[1] && p[2] > b.mn[2] && p[2] < b.mx[2];
}
bool contains_box(const Box& b, const Box& o) {
    return o.mn[0] > b.mn[0] && o.mx[0] < b.mx[0] && o.mn[1] > b.mn[1] && o.mx[1] < b.mx[1] && o.mn[2] > b.mn[2] &&
           o.mx[2] < b.mx[2];
}
bool overlaps(const Box& b, const Box& o) {
    return o.mn[0] < b.mx[0] && o.mx[0] > b.mn[0] && o.mn[1] < b.mx[1] && o.mx[1] > b.mn[1] && o.mn[2] < b.mx[2] &&
           o.mx[2] > b.mn[2];
}

struct BNode {
    Box box;
    std::vector<int> prims;   // leaf
    int child[8];             // interior: build-node index or -1
    bool leaf = false;
};

// A primitive that fits no loose cell is copied into every cell it
// overlaps, at every level, so a scene of large primitives with a small
// max_count and a deep max_depth multiplies its entries by up to 8 per
// level.  The build stops with an error past this many primitive entries
// (4 GiB of indices) instead of exhausting host memory.
constexpr int64_t kMaxEntries = (int64_t)1 << 30;

struct TooLarge {};

struct Builder {
    const std::vector<Box>& bounds;
    int max_depth, max_count;
    float out_bound;
    std::vector<BNode> nodes;
    int now_depth = 0;
    int64_t entries = 0;

    int build(const std::vector<int>& idx, const Box& inner_box, const Box& outer_box, int depth) {
        const int cnt = (int)idx.size();
        if (cnt == 0) return -1;
        now_depth = std::max(now_depth, depth);
        if (cnt <= max_count || depth >= max_depth) {
            BNode n;
            n.box = outer_box;
            n.prims = idx;
            n.leaf = true;
            nodes.push_back(std::move(n));
            return (int)nodes.size() - 1;
        }
        float center[3];
        for (int k = 0; k < 3; k++) center[k] = (inner_box.mx[k] + inner_box.mn[k]) * 0.5f;   // Center
        Box inner[8], outer[8];
        std::vector<int> child_data[8];
        const float extend = (out_bound - 1) * 0.5f;
        for (int i = 0; i < 8; i++) {
            const float corner[3] = {(i & 1) == 0 ? inner_box.mn[0] : inner_box.mx[0],
                                     (i & 2) == 0 ? inner_box.mn[1] : inner_box.mx[1],
                                     (i & 4) == 0 ? inner_box.mn[2] : inner_box.mx[2]};
            for (int k = 0; k < 3; k++) {
                inner[i].mn[k] = vmin(center[k], corner[k]);   // BoundingBox3F(center).ExpendBy(corner)
                inner[i].mx[k] = vmax(center[k], corner[k]);
                outer[i].mn[k] = inner[i].mn[k] - extend;
                outer[i].mx[k] = inner[i].mx[k] + extend;
            }
        }
        for (int p : idx) {
            const Box& b = bounds[p];
            float c[3];
            for (int k = 0; k < 3; k++) c[k] = (b.mx[k] + b.mn[k]) * 0.5f;
            bool inserted = false;
            for (int j = 0; j < 8; j++) {
                if (contains_point(inner[j], c) && contains_box(outer[j], b)) {
                    child_data[j].push_back(p);
                    inserted = true;
                    break;
                }
            }
            if (!inserted)
                for (int j = 0; j < 8; j++)
                    if (overlaps(inner[j], b)) child_data[j].push_back(p);
        }
        for (int j = 0; j < 8; j++) entries += (int64_t)child_data[j].size();
        if (entries > kMaxEntries) throw TooLarge{};
        for (int i = 0; i < 8; i++) {
            Box real = empty_box();
            for (int p : child_data[i]) box_union(real, bounds[p]);
            outer[i] = real;
        }
        int children[8];
        for (int i = 0; i < 8; i++) children[i] = build(child_data[i], inner[i], outer[i], depth + 1);
        BNode n;
        n.box = outer_box;
        n.leaf = false;
        for (int i = 0; i < 8; i++) n.child[i] = children[i];
        nodes.push_back(std::move(n));
        return (int)nodes.size() - 1;
    }

    void flatten(int bn, int index, int* all, OctResult* out) {
        const BNode& n = nodes[bn];
        OctNode& o = out->nodes[index];
        for (int k = 0; k < 3; k++) { o.mn[k] = n.box.mn[k]; o.mx[k] = n.box.mx[k]; }
        if (n.leaf) {
            out->leaf_first.push_back((int)out->leaf_prims.size());
            out->leaf_count.push_back((int)n.prims.size());
            out->leaf_prims.insert(out->leaf_prims.end(), n.prims.begin(), n.prims.end());
            o.count_or_leaf = (int)out->leaf_first.size() - 1;
            o.next_head = -1;
            return;
        }
        int not_null = 0;
        for (int i = 0; i < 8; i++) not_null += n.child[i] >= 0 ? 1 : 0;
        o.count_or_leaf = not_null;
        o.next_head = *all;
        const int head = *all;
        *all += not_null;
        for (int i = 0, j = 0; i < 8; i++) {
            if (n.child[i] >= 0) {
                flatten(n.child[i], head + j, all, out);
                j++;
            }
        }
    }
};

}  // namespace

bool build_octree(int n, const float* bmin, const float* bmax, int max_depth, int max_count, float out_bound,
                  OctResult* out, std::string* err) {
    *out = OctResult{};
    if (n < 0) { *err = "negative primitive count"; return false; }
    std::vector<Box> bounds(n);
    Box root = empty_box();
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) { bounds[i].mn[k] = bmin[3 * i + k]; bounds[i].mx[k] = bmax[3 * i + k]; }
        box_union(root, bounds[i]);
        idx[i] = i;
    }
    Builder b{bounds, max_depth, max_count, out_bound, {}, 0, 0};
    int all = 1;
    try {
        const int r = b.build(idx, root, root, 0);
        if (r < 0) return true;   // an empty tree
        out->nodes.resize(b.nodes.size());
        b.flatten(r, 0, &all, out);
    } catch (const TooLarge&) {
        *out = OctResult{};
        *err = "octree over 2^30 primitive entries (primitives copied into every cell they overlap)";
        return false;
    } catch (const std::bad_alloc&) {
        *out = OctResult{};
        *err = "out of host memory building the octree";
        return false;
    }
    if (all != (int)b.nodes.size()) { *err = "octree flatten count mismatch"; return false; }
    out->max_depth = b.now_depth;
    // worst-case LIFO size, depth first with the pending siblings below
    std::vector<std::pair<int, int>> todo = {{0, 0}};   // (node, entries pending below it)
    out->stack_needed = 1;
    while (!todo.empty()) {
        const auto [i, below] = todo.back();
        todo.pop_back();
        const OctNode& o = out->nodes[i];
        if (o.next_head < 0) continue;
        out->stack_needed = std::max(out->stack_needed, below + o.count_or_leaf);
        for (int k = 0; k < o.count_or_leaf; k++) todo.push_back({o.next_head + k, below + o.count_or_leaf - 1});
    }
    return true;
}

}  // namespace rd
