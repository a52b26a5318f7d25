// rd_bvh_build.cpp — host SAH BVH builder.
//
// Produces the same tree as Bvh.RecursiveBuild + FlattenBVHTree
// (Accelerators/Bvh.cs:234-442) with an explicit task stack instead of
// recursion and structure-of-arrays primitive records.  Nodes are emitted in
// pre-order as they are created, so the left child is always the next node
// and only the right child's index is patched into its parent; leaves append
// their primitives in the same left-to-right order as the reference's
// orderedPrims.  An interior node's box is the union over its primitive
// range, which equals Union(left, right) because min/max are exact.
#include "rd_bvh_build.h"

#include <algorithm>
#include <cfloat>

namespace rd {
namespace {

struct Box {
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    // Vector3.Min / Vector3.Max component selects (BoundingBox3F.Union)
    inline void grow(const float* lo, const float* hi) {
        for (int k = 0; k < 3; k++) {
            mn[k] = mn[k] < lo[k] ? mn[k] : lo[k];
            mx[k] = mx[k] > hi[k] ? mx[k] : hi[k];
        }
    }
    inline void grow(const Box& b) { grow(b.mn, b.mx); }
    // BoundingBox3F.SurfaceArea (BoundingBox3F.cs:24-31)
    inline float area() const {
        float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
        return 2 * (dx * dy + dx * dz + dy * dz);
    }
};

struct Task {
    int start, end, depth, parent;  // parent >= 0: patch parent's offset with this node
};

}  // namespace

bool build_bvh(int n, const float* bmin, const float* bmax, int max_prims, BvhResult* out, std::string* err) {
    out->nodes.clear();
    out->order.clear();
    out->leaves = 0;
    out->max_depth = 0;
    if (n <= 0) return true;
    if (max_prims < 1) max_prims = 1;

    // primitive records (BVHPrimitiveInfo, Bvh.cs:56-67), kept in SoA
    std::vector<int32_t> num(n);
    std::vector<float> lo(3 * (size_t)n), hi(3 * (size_t)n), cen(3 * (size_t)n);
    for (int i = 0; i < n; i++) {
        num[i] = i;
        for (int k = 0; k < 3; k++) {
            lo[3 * i + k] = bmin[3 * i + k];
            hi[3 * i + k] = bmax[3 * i + k];
            cen[3 * i + k] = 0.5f * bmin[3 * i + k] + 0.5f * bmax[3 * i + k];
        }
    }
    auto swap_rec = [&](int a, int b) {
        std::swap(num[a], num[b]);
        for (int k = 0; k < 3; k++) {
            std::swap(lo[3 * a + k], lo[3 * b + k]);
            std::swap(hi[3 * a + k], hi[3 * b + k]);
            std::swap(cen[3 * a + k], cen[3 * b + k]);
        }
    };

    out->nodes.reserve(2 * (size_t)n);
    out->order.reserve(n);
    std::vector<Task> stack;
    stack.push_back({0, n, 0, -1});
    constexpr int kBuckets = 12;

    while (!stack.empty()) {
        Task t = stack.back();
        stack.pop_back();
        int idx = (int)out->nodes.size();
        out->nodes.push_back(NodeRec{});
        if (t.parent >= 0) out->nodes[t.parent].offset = idx;
        if (t.depth > out->max_depth) out->max_depth = t.depth;

        Box bounds, cb;
        for (int i = t.start; i < t.end; i++) {
            bounds.grow(&lo[3 * i], &hi[3 * i]);
            cb.grow(&cen[3 * i], &cen[3 * i]);
        }
        int count = t.end - t.start;
        auto make_leaf = [&]() -> bool {
            if (count > 65535) { *err = "BVH leaf holds more than 65535 primitives"; return false; }
            NodeRec& r = out->nodes[idx];
            for (int k = 0; k < 3; k++) { r.mn[k] = bounds.mn[k]; r.mx[k] = bounds.mx[k]; }
            r.offset = (int32_t)out->order.size();
            r.count = (uint16_t)count;
            r.axis = 0;
            for (int i = t.start; i < t.end; i++) out->order.push_back(num[i]);
            out->leaves++;
            return true;
        };
        if (count == 1) { if (!make_leaf()) return false; continue; }
        // BoundingBox3F.MaximumExtent (BoundingBox3F.cs:216-225)
        float d0 = cb.mx[0] - cb.mn[0], d1 = cb.mx[1] - cb.mn[1], d2 = cb.mx[2] - cb.mn[2];
        int dim = (d0 > d1 && d0 > d2) ? 0 : (d1 > d2 ? 1 : 2);
        if (cb.mx[dim] == cb.mn[dim]) { if (!make_leaf()) return false; continue; }
        int mid;
        if (count <= 2) {
            mid = (t.start + t.end) / 2;
            float lv = cen[3 * t.start + dim], rv = cen[3 * (t.start + 1) + dim];
            if (!(lv == rv) && !(lv < rv)) swap_rec(t.start, t.start + 1);   // SwapIfGreater
        } else {
            const float cmin = cb.mn[dim];
            const float ext = cb.mx[dim] - cb.mn[dim];
            // BoundingBox3F.Offset along dim, then (int)(12 * offset)
            auto bucket_of = [&](int i) { return (int)(kBuckets * ((cen[3 * i + dim] - cmin) / ext)); };
            int cnt[kBuckets] = {0};
            Box bb[kBuckets];
            for (int i = t.start; i < t.end; i++) {
                int b = bucket_of(i);
                if (b == kBuckets) b = kBuckets - 1;
                cnt[b]++;
                bb[b].grow(&lo[3 * i], &hi[3 * i]);
            }
            float cost[kBuckets - 1];
            const float area = bounds.area();
            for (int i = 0; i < kBuckets - 1; i++) {
                Box b0, b1;
                int c0 = 0, c1 = 0;
                for (int j = 0; j <= i; j++) { b0.grow(bb[j]); c0 += cnt[j]; }
                for (int j = i + 1; j < kBuckets; j++) { b1.grow(bb[j]); c1 += cnt[j]; }
                cost[i] = 1 + (c0 * b0.area() + c1 * b1.area()) / area;
            }
            float min_cost = cost[0];
            int split = 0;
            for (int i = 1; i < kBuckets - 1; i++)
                if (cost[i] < min_cost) { min_cost = cost[i]; split = i; }
            float leaf_cost = (float)count;
            if (count > max_prims || min_cost < leaf_cost) {
                // MathExt.Partition(list, start, end - 1, pred) (MathExt.cs:170-201)
                auto pred = [&](int i) { return bucket_of(i) <= split; };
                int first = t.start, last = t.end - 1;
                for (;;) {
                    bool done = false;
                    for (;;) {
                        if (first == last) { done = true; break; }
                        if (!pred(first)) break;
                        ++first;
                    }
                    if (done) break;
                    while (!pred(last)) {
                        if (first == last) { done = true; break; }
                        last--;
                    }
                    if (done) break;
                    swap_rec(first, last);
                    first++;
                }
                mid = first;
            } else {
                if (!make_leaf()) return false;
                continue;
            }
        }
        if (mid <= t.start || mid >= t.end) { *err = "BVH split produced an empty child"; return false; }
        NodeRec& r = out->nodes[idx];
        for (int k = 0; k < 3; k++) { r.mn[k] = bounds.mn[k]; r.mx[k] = bounds.mx[k]; }
        r.axis = (uint8_t)dim;
        r.count = 0;
        r.offset = -1;
        stack.push_back({mid, t.end, t.depth + 1, idx});     // right, patched later
        stack.push_back({t.start, mid, t.depth + 1, -1});    // left, next in pre-order
    }
    return true;
}

}  // namespace rd
