// rd_kernels.hip — gfx950 kernels: BVH traversal and the wavefront stages of
// PathTracer.Mis / NextEventEstimation / OnlySampleBxdf
// (Radiantium.Offline/Integrators/PathTracer.cs:43-270).
//
// Wavefront mapping of one reference bounce (PathTracer.cs:197-269):
//   trace_ext     Scene.Intersect of the extension rays (closest hit)
//   shade         emission, light selection + SampleLi, the MIS BSDF sample,
//                 the continuation BSDF sample and Russian roulette — every
//                 random draw of the bounce, in the reference's order — and
//                 appends shadow, MIS and next-extension rays to queues
//   trace_shadow  Scene.IsOccluded (any hit) of the light-sample rays
//   trace_mis     Scene.Intersect of the BSDF-sample rays + Le of the hit
//   resolve       radiance += coeff * (le_light + le_bsdf) / lightPdf
// The random stream is keyed by (seed, pixel, sample) and a per-path draw
// counter, so the kernels reproduce the CPU restatement's draws exactly.
#include <mutex>
#include <unordered_map>

#include "rd_chain.h"
#include "rd_kernels.h"
#include "rd_shading.h"
#include "rd_wave.h"

namespace rd {

static int g_num_cus = 256;
void set_device_grid(int num_cus) { g_num_cus = num_cus > 0 ? num_cus : 256; }
int persistent_blocks() { return g_num_cus * 8; }

// ---------------------------------------------------------------- traversal
// BoundingBox3F.Intersect (BoundingBox3F.cs:155-176) on a node stored as
// lo = {min.xyz, max.x}, hi = {max.y, max.z, ...}.  The reference's factor
// 1 + 2*Gamma(3) is exactly 1.0f and is omitted.
// Branch-free: the early returns become flags (once false, the later values
// no longer matter), so a wavefront's lanes never split over the test.
__device__ __forceinline__ bool box_hit(float4 lo, float4 hi, f3 o, f3 inv, bool nx, bool ny, bool nz, float ray_tmax) {
    float tmin = ((nx ? lo.w : lo.x) - o.x) * inv.x;
    float tmax = ((nx ? lo.x : lo.w) - o.x) * inv.x;
    float tymin = ((ny ? hi.x : lo.y) - o.y) * inv.y;
    float tymax = ((ny ? lo.y : hi.x) - o.y) * inv.y;
    bool ok = !(tmin > tymax) & !(tymin > tmax);
    tmin = tymin > tmin ? tymin : tmin;
    tmax = tymax < tmax ? tymax : tmax;
    float tzmin = ((nz ? hi.y : lo.z) - o.z) * inv.z;
    float tzmax = ((nz ? lo.z : hi.y) - o.z) * inv.z;
    ok = ok & !(tmin > tzmax) & !(tzmin > tmax);
    tmin = tzmin > tmin ? tzmin : tmin;
    tmax = tzmax < tmax ? tzmax : tmax;
    return ok & (tmin < ray_tmax) & (tmax > 0);
}

struct Hit {
    float t, u, v;
    int prim;   // original primitive index, -1 = miss
};

// Whole-record loads.  Left alone, the compiler sinks the parts of a node
// (or primitive) record into the branches that use them, so one node step
// becomes up to three dependent round trips (x/y bounds, then z bounds, then
// the child offset) and five load instructions.  Passing the loaded values
// through an empty asm statement makes every part live at that point: both
// 16-byte loads issue together and are waited for once.
__device__ __forceinline__ void whole(float4& a) { asm volatile("" : "+v"(a.x), "+v"(a.y), "+v"(a.z), "+v"(a.w)); }

// ---------------------------------------------------------------- traversal kernels
// One kernel template per (query class, LDS stack depth, schedule).  Every
// ray walks the reference's depth-first order (Bvh.cs:126-232) one node per
// step, which keeps hits and visit counts identical to the oracle; the
// schedules differ only in how rays are assigned to lanes.
enum : int {
    kModeRaysClosest = 0, kModeRaysAny = 1, kModeExt = 2, kModeShadow = 3, kModeMis = 4,
    kModeChainShadow = 5,   // vol_path Transmittance segment: closest hit into hits[]
    kModeChainMis = 6,      // vol_path IntersectTr segment: closest hit into hits[]
    kModeSecondary = 7,     // MIS rays (q, closest) then shadow rays (q2, any hit) in one launch
    kModeChainFullShadow = 8,   // vol_path Transmittance chains, each walked to its end by one lane
    kModeChainFullMis = 9       // vol_path IntersectTr chains, likewise
};
constexpr int kRefill = 16;

struct TraceArgs {
    const float4* rays;       // ray-array modes: 2 float4 per ray (rd_ray)
    RayQueue q;               // queue modes
    float4* hits;             // closest modes: {t, prim bits, u, v} per ray / slot
    uint8_t* occ;             // ray-array any-hit output
    PathState ps;             // shadow / MIS outputs
    int* counts;
    unsigned long long* counters;   // ray-array modes: {nodes, prims} (may be null)
    int n_fixed;              // ray-array modes: ray count
    int cnt_slot;             // queue modes: count slot
    int cursor_slot;          // work cursor, zero on entry
    int zero[3];              // count slots block 0 clears for the next stage (-1: none)
    const DevScene* scenes;   // batched queue modes: the ray's scene is scenes[ps.req[pid]]
    int* sub;                 // closest modes of instanced scenes: set triangle (DevPrim index) or -1
    const int* perm;          // extension queue: the order to trace its rays in (null = queue order)
    int* rank;                // with perm: rank[queue position] = position in that order
    int dyn;                  // straight schedule: 0 grid-stride, 1 chunks from the work cursor
    RayQueue q2;              // kModeSecondary: the shadow queue (q = the MIS queue)
    int cnt_slot2;            // kModeSecondary: its count slot
};

__device__ __forceinline__ uint64_t lanemask_lt() {
    int lane = __lane_id();
    return lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
}

// Per-lane traversal state of one ray.
struct Trav {
    f3 o, d, inv;
    float tmin, tmax;
    bool nx, ny, nz;
    int sp, cur;
    Hit best;
};

__device__ __forceinline__ void trav_begin(Trav& t, float4 ra, float4 rb) {
    t.o = mk(ra.x, ra.y, ra.z);
    t.d = mk(rb.x, rb.y, rb.z);
    t.tmin = ra.w;
    t.tmax = rb.w;
    t.inv = mk(1.0f / t.d.x, 1.0f / t.d.y, 1.0f / t.d.z);
    t.nx = t.inv.x < 0; t.ny = t.inv.y < 0; t.nz = t.inv.z < 0;
    t.sp = 0; t.cur = 0;
    t.best.prim = -1; t.best.t = t.best.u = t.best.v = 0.f;
}

// One iteration of Bvh.Intersect's loop (Bvh.cs:134-170 / 187-229): test the
// current node; a leaf tests its primitives (any-hit stops at the first, closest
// hit keeps the first strictly closer one and shrinks MaxT); an interior node
// descends to the child on the near side of its split axis and pushes the
// other.  Returns true when the ray is finished.
template <bool kAny, int kBlock>
__device__ __forceinline__ bool trav_step(const float4* __restrict__ nodes, const float4* __restrict__ prims, Trav& t,
                                          int* stk, uint32_t& nv, uint32_t& nt) {
    float4 lo = nodes[2 * t.cur], hi = nodes[2 * t.cur + 1];
    whole(lo); whole(hi);
    nv++;
    if (box_hit(lo, hi, t.o, t.inv, t.nx, t.ny, t.nz, t.tmax)) {
        int off = __float_as_int(hi.z);
        int ca = __float_as_int(hi.w);
        int cnt = ca & 0xffff;
        if (cnt > 0) {
            for (int i = 0; i < cnt; i++) {
                const float4* p = prims + 3 * (off + i);
                float4 p0 = p[0], p1 = p[1], p2 = p[2];   // a sphere's third float4 is padding
                whole(p0); whole(p1); whole(p2);
                nt++;
                float tu = 0.f, tv = 0.f, tt = 0.f;
                bool h;
                if (__float_as_int(p1.w) == kTri) {
                    h = tri_test(mk(p0.x, p0.y, p0.z), mk(p1.x, p1.y, p1.z), mk(p2.x, p2.y, p2.z), t.o, t.d, t.tmin,
                                 t.tmax, &tu, &tv, &tt);
                } else {
                    h = sphere_test(make_float4(p0.x, p0.y, p0.z, p1.x), t.o, t.d, t.tmin, t.tmax, &tt);
                }
                if (h) {
                    if (kAny) { t.best.prim = __float_as_int(p0.w); return true; }
                    if (tt < (t.best.prim >= 0 ? t.best.t : kFloatMax)) {
                        t.best.t = tt; t.best.u = tu; t.best.v = tv; t.best.prim = __float_as_int(p0.w);
                        t.tmax = tt;
                    }
                }
            }
            if (t.sp == 0) return true;
            t.cur = stk[(--t.sp) * kBlock];
        } else {
            int axis = ca >> 16;
            bool neg = axis == 0 ? t.nx : (axis == 1 ? t.ny : t.nz);
            if (neg) { stk[(t.sp++) * kBlock] = t.cur + 1; t.cur = off; }
            else { stk[(t.sp++) * kBlock] = off; t.cur = t.cur + 1; }
        }
    } else {
        if (t.sp == 0) return true;
        t.cur = stk[(--t.sp) * kBlock];
    }
    return false;
}

template <int kMode>
__device__ __forceinline__ void trav_load(const TraceArgs& a, int r, float4* ra, float4* rb) {
    if (kMode <= kModeRaysAny) { *ra = a.rays[2 * r]; *rb = a.rays[2 * r + 1]; }
    else { *ra = a.q.org[r]; *rb = a.q.dir[r]; }
}

// Writes the query result of ray / slot r.
template <int kMode>
__device__ __forceinline__ void trav_finish(const DevScene& sc, const TraceArgs& a, int r, const Trav& t) {
    const bool found = t.best.prim >= 0;
    if (kMode == kModeRaysClosest || kMode == kModeExt || kMode == kModeChainShadow || kMode == kModeChainMis) {
        a.hits[r] = found ? make_float4(t.best.t, __int_as_float(t.best.prim), t.best.u, t.best.v)
                          : make_float4(0.f, __int_as_float(-1), 0.f, 0.f);
    } else if (kMode == kModeRaysAny) {
        a.occ[r] = found ? 1 : 0;
    } else if (kMode == kModeShadow) {
        a.ps.vis[a.q.pid[r]] = found ? 0 : 1;
    } else {
        // BSDF-sampled MIS ray: Le of the hit when it is the selected light,
        // every infinite light on a miss (PathTracer.cs:377-390)
        int pid = a.q.pid[r];
        f3 li = mk(0.0f);
        if (!found) {
            li = eval_infinite(sc, t.d);
        } else {
            const DevPrim pr = sc.prim_info[t.best.prim];
            int sel = a.ps.sel_light[pid];
            if (pr.light >= 0 && pr.light == sel) {
                f3 nrm = hit_normal(sc, pr, t.o, t.d, t.best.u, t.best.v, t.best.t);
                const DevLight& L = sc.lights[sel];
                if (dot(nrm, -t.d) > 0) li = mk(L.le[0], L.le[1], L.le[2]);
            }
        }
        float4 prev = a.ps.mis_li[pid];
        a.ps.mis_li[pid] = make_float4(li.x, li.y, li.z, prev.w);
    }
}

// ---------------------------------------------------------------- two-level traversal
// Instanced scenes (Primitive.cs:144-189): a scene leaf's InstancedPrimitive
// runs its set's Bvh.Intersect on the ray transformed into model space
// (Ray3F.Transform: origin and normalised direction through WorldToModel,
// MaxT = the model distance to the image of ray.At(MaxT), MinT kept) before
// the leaf's next primitive, exactly as the reference's nested call does.  The
// set's DFS runs on the same LDS stack above the scene entries (sp_base); the
// rest of the scene leaf waits in li/le.  A set hit becomes a scene hit at
// T = |O - M2W * P_model| (P_model from the barycentrics, Triangle.cs:98-99)
// and competes with the other scene primitives on T.
struct TravI {
    Trav t;            // current ray (world, or model inside a set), stack, world best
    f3 wo, wd;         // the world ray while inside a set
    f3 winv;           // and its reciprocal direction
    float wtmax;
    int sub;           // set triangle of the world best (DevPrim index), -1
    int sp_base;       // stack depth at set entry, -1 at scene level
    int li, le;        // scene leaf: next primitive record, end
    int inst, inst_orig;
    float bt, bu, bv;  // best set hit (model t)
    int bord;          // its primitive record, -1 none
};

__device__ __forceinline__ void travi_begin(TravI& s, float4 ra, float4 rb) {
    trav_begin(s.t, ra, rb);
    s.sub = -1; s.sp_base = -1; s.li = s.le = 0; s.inst = s.inst_orig = 0;
    s.bt = s.bu = s.bv = 0.f; s.bord = -1;
    s.wo = s.t.o; s.wd = s.t.d; s.wtmax = s.t.tmax;
    s.winv = s.t.inv;
}

__device__ __forceinline__ void set_ray(Trav& t, f3 o, f3 d, float tmax) {
    t.o = o; t.d = d; t.tmax = tmax;
    t.inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    t.nx = t.inv.x < 0; t.ny = t.inv.y < 0; t.nz = t.inv.z < 0;
}

// The world ray back after a set, with the reciprocal direction kept at
// entry: the values set_ray would compute again, without its divisions.
__device__ __forceinline__ void restore_world(TravI& s) {
    Trav& t = s.t;
    t.o = s.wo; t.d = s.wd; t.tmax = s.wtmax;
    t.inv = s.winv;
    t.nx = t.inv.x < 0; t.ny = t.inv.y < 0; t.nz = t.inv.z < 0;
}

// InstancedPrimitive.Intersect: Ray3F.Transform (Ray3F.cs:27-34) and the set's root.
__device__ __forceinline__ void enter_set(const DevScene& sc, TravI& s, int inst, int orig) {
    Trav& t = s.t;
    const DevInstance& I = sc.insts[inst];
    s.wo = t.o; s.wd = t.d; s.wtmax = t.tmax; s.winv = t.inv;
    f3 mo = xform_point12(I.w2m, t.o);
    f3 md = normalize(xform_vector12(I.w2m, t.d));
    f3 mx = xform_point12(I.w2m, t.o + t.tmax * t.d);
    set_ray(t, mo, md, length(mo - mx));
    s.sp_base = t.sp;
    s.inst = inst; s.inst_orig = orig;
    s.bord = -1;
    t.cur = I.root;
}

// Back at scene level: the set's best hit (if any) as a scene hit
// (Primitive.cs:175-182 and the T comparison of Bvh.cs:196-203).
__device__ __forceinline__ void leave_set(const DevScene& sc, const float4* __restrict__ prims, TravI& s) {
    Trav& t = s.t;
    if (s.bord >= 0) {
        const float4* p = prims + 3 * s.bord;
        float4 p0 = p[0], p1 = p[1], p2 = p[2];
        float bx = 1 - (s.bu + s.bv), by = s.bu, bz = s.bv;
        f3 pm = bx * mk(p0.x, p0.y, p0.z) + by * mk(p1.x, p1.y, p1.z) + bz * mk(p2.x, p2.y, p2.z);
        f3 pw = xform_point12(sc.insts[s.inst].m2w, pm);
        float T = length(s.wo - pw);
        if (T < (t.best.prim >= 0 ? t.best.t : kFloatMax)) {
            t.best.t = T; t.best.u = s.bu; t.best.v = s.bv; t.best.prim = s.inst_orig;
            s.sub = __float_as_int(p0.w);
            s.wtmax = T;
        }
    }
    restore_world(s);
    s.sp_base = -1;
}

// The rest of a scene leaf from li: plain primitives are tested on the world
// ray, an instance enters its set (false: keep stepping).  At the leaf's end
// the scene stack is popped; true when the ray is finished.
template <bool kAny, int kBlock>
__device__ __forceinline__ bool scene_leaf(const DevScene& sc, const float4* __restrict__ prims, TravI& s, int* stk,
                                           uint32_t& nt) {
    Trav& t = s.t;
    while (s.li < s.le) {
        const float4* p = prims + 3 * s.li;
        s.li++;
        float4 p0 = p[0], p1 = p[1], p2 = p[2];   // padding for spheres and instances
        whole(p0); whole(p1); whole(p2);
        nt++;
        int kind = __float_as_int(p1.w);
        if (kind == kInst) {
            enter_set(sc, s, __float_as_int(p0.x), __float_as_int(p0.w));
            return false;
        }
        float tu = 0.f, tv = 0.f, tt = 0.f;
        bool h;
        if (kind == kTri) {
            h = tri_test(mk(p0.x, p0.y, p0.z), mk(p1.x, p1.y, p1.z), mk(p2.x, p2.y, p2.z), t.o, t.d, t.tmin, t.tmax,
                         &tu, &tv, &tt);
        } else {
            h = sphere_test(make_float4(p0.x, p0.y, p0.z, p1.x), t.o, t.d, t.tmin, t.tmax, &tt);
        }
        if (h) {
            if (kAny) { t.best.prim = __float_as_int(p0.w); return true; }
            if (tt < (t.best.prim >= 0 ? t.best.t : kFloatMax)) {
                t.best.t = tt; t.best.u = tu; t.best.v = tv; t.best.prim = __float_as_int(p0.w);
                s.sub = -1;
                t.tmax = tt;
            }
        }
    }
    if (t.sp == 0) return true;
    t.cur = stk[(--t.sp) * kBlock];
    return false;
}

// One node of the two-level DFS; true when the ray is finished.
template <bool kAny, int kBlock>
__device__ __forceinline__ bool trav_step_inst(const DevScene& sc, const float4* __restrict__ nodes,
                                               const float4* __restrict__ prims, TravI& s, int* stk, uint32_t& nv,
                                               uint32_t& nt) {
    Trav& t = s.t;
    float4 lo = nodes[2 * t.cur], hi = nodes[2 * t.cur + 1];
    whole(lo); whole(hi);
    nv++;
    bool pop;
    if (box_hit(lo, hi, t.o, t.inv, t.nx, t.ny, t.nz, t.tmax)) {
        int off = __float_as_int(hi.z);
        int ca = __float_as_int(hi.w);
        int cnt = ca & 0xffff;
        if (cnt > 0) {
            if (s.sp_base < 0) {
                s.li = off;
                s.le = off + cnt;
                return scene_leaf<kAny, kBlock>(sc, prims, s, stk, nt);
            }
            // a set leaf: its triangles on the model ray, best by model t
            for (int i = 0; i < cnt; i++) {
                const float4* p = prims + 3 * (off + i);
                float4 p0 = p[0], p1 = p[1], p2 = p[2];
                whole(p0); whole(p1); whole(p2);
                nt++;
                float tu = 0.f, tv = 0.f, tt = 0.f;
                if (tri_test(mk(p0.x, p0.y, p0.z), mk(p1.x, p1.y, p1.z), mk(p2.x, p2.y, p2.z), t.o, t.d, t.tmin,
                             t.tmax, &tu, &tv, &tt)) {
                    if (kAny) { t.best.prim = s.inst_orig; return true; }
                    if (tt < (s.bord >= 0 ? s.bt : kFloatMax)) {
                        s.bt = tt; s.bu = tu; s.bv = tv; s.bord = off + i;
                        t.tmax = tt;
                    }
                }
            }
            pop = true;
        } else {
            int axis = ca >> 16;
            bool neg = axis == 0 ? t.nx : (axis == 1 ? t.ny : t.nz);
            if (neg) { stk[(t.sp++) * kBlock] = t.cur + 1; t.cur = off; }
            else { stk[(t.sp++) * kBlock] = off; t.cur = t.cur + 1; }
            pop = false;
        }
    } else {
        pop = true;
    }
    if (!pop) return false;
    if (s.sp_base >= 0 && t.sp == s.sp_base) {
        leave_set(sc, prims, s);
        return scene_leaf<kAny, kBlock>(sc, prims, s, stk, nt);
    }
    if (t.sp == 0) return true;
    t.cur = stk[(--t.sp) * kBlock];
    return false;
}

template <int kMode>
__device__ __forceinline__ void trav_counters(const TraceArgs& a, uint32_t nrays, uint32_t nv, uint32_t nt) {
    constexpr bool kAny = (kMode == kModeRaysAny || kMode == kModeShadow);
    if (kMode <= kModeRaysAny) {
        if (a.counters) {
            add_counter(a.counters, nv);
            add_counter(a.counters + 1, nt);
        }
    } else {
        constexpr int cls = kMode == kModeExt ? 0 : ((kMode == kModeMis || kMode == kModeChainMis) ? 1 : 2);
        add_class_counters(a.counts, cls, nrays, nv, nt);
        add_counter((unsigned long long*)(a.counts + (kAny ? kCntRaysAny : kCntRaysClosest)), nrays);
    }
}

__device__ __forceinline__ int trav_count(const TraceArgs& a, bool fixed) {
    return fixed ? a.n_fixed : a.counts[a.cnt_slot];
}

// Chunks of consecutive ranks for the cursor schedules (wave-uniform), taken
// from the work cursor, so wavefronts that finish early take more rays and
// the launch ends evenly whatever the grid and the rays' costs.  Atomics on
// one address serialise (tens of ns each at the device scope), so a resident
// grid over a small or empty queue would pay one per wavefront: only the
// first ceil(n / size) wavefronts of the grid take part (together they take
// every chunk), the others exit without touching the cursor.  (Measured on
// the extension traversal of large queues: a static first chunk per
// wavefront made it 14 % slower, a plain load of the cursor before each
// atomic twice as slow.)
__device__ __forceinline__ int next_chunk(int* cursor, int size = 64) {
    int base = 0;
    if (__lane_id() == 0) base = atomicAdd(cursor, size);
    return __shfl(base, 0);
}

// Whether the calling wavefront takes part in a cursor schedule over n rays.
__device__ __forceinline__ bool chunk_wave(int n, int size = 64) {
    const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    return (int64_t)wave * size < n;
}

// Waves per SIMD that the LDS stack allows (160 KB of LDS per CU, four SIMDs;
// a wavefront's stack takes stack x 64 x 4 B), at most 8: the second
// argument of __launch_bounds__ on this target is the minimum waves per SIMD,
// so the traversal kernels are compiled for the register budget of exactly
// the occupancy the stack already sets, whatever their block size.
constexpr int lds_blocks(int stack, int block) {
    return (160 * 1024) / (stack * 64 * 4) / 4 < 8 ? (160 * 1024) / (stack * 64 * 4) / 4 : 8;
}
// The two-level kernels hold more state (the world ray, the scene leaf, the
// set's best hit): their budget is capped at 96 (straight) or 128 (sequence)
// VGPRs, above which shallow two-level trees would spill on every step.
constexpr int inst_waves(int stack, int block, int cap) {
    return lds_blocks(stack, block) < cap ? lds_blocks(stack, block) : cap;
}

// Schedule 0 / 2 ("straight"): each lane runs its ray to completion; rays
// are taken grid-stride (0) or in chunks of 64 consecutive rays per
// wavefront from the work cursor (2, a.dyn).  kBatch: each ray traverses its
// own request's scene.
template <int kMode, int kStack, int kBlock, bool kBatch>
__global__ __launch_bounds__(kBlock, lds_blocks(kStack, kBlock)) void k_trace_straight(DevScene sc, TraceArgs a) {
    constexpr bool kAny = (kMode == kModeRaysAny || kMode == kModeShadow);
    __shared__ int lds[kStack * kBlock];
    int* stk = lds + threadIdx.x;
    const int n = trav_count(a, kMode <= kModeRaysAny);
    if (blockIdx.x == 0 && threadIdx.x < 3 && a.zero[threadIdx.x] >= 0) a.counts[a.zero[threadIdx.x]] = 0;
    int* cursor = a.counts + a.cursor_slot;
    uint32_t nv = 0, nt = 0, nrays = 0;
    // one ray per lane; the chunk fetch stays outside it, where the whole wavefront is active
    auto one = [&](int r0) {
        const int r = a.perm ? a.perm[r0] : r0;
        if (a.rank) a.rank[r] = r0;
        // ranked shadow queues leave holes (pid -1) for paths without a shadow ray
        if (kMode == kModeShadow && a.q.pid[r] < 0) return;
        float4 ra, rb;
        trav_load<kMode>(a, r, &ra, &rb);
        Trav t;
        trav_begin(t, ra, rb);
        if constexpr (kBatch) {
            const DevScene& s2 = a.scenes[a.ps.req[a.q.pid[r]]];
            const float4* nodes = s2.nodes;   // per-ray scene: pointers loaded once per ray
            const float4* prims = s2.prims;
            while (!trav_step<kAny, kBlock>(nodes, prims, t, stk, nv, nt)) {}
            trav_finish<kMode>(s2, a, r, t);
        } else {
            while (!trav_step<kAny, kBlock>(sc.nodes, sc.prims, t, stk, nv, nt)) {}
            trav_finish<kMode>(sc, a, r, t);
        }
        nrays++;
    };
    if (a.dyn) {
        if (!chunk_wave(n)) return;
        for (int c = next_chunk(cursor); c < n; c = next_chunk(cursor))
            if (c + __lane_id() < n) one(c + __lane_id());
    } else {
        for (int r0 = blockIdx.x * kBlock + threadIdx.x; r0 < n; r0 += gridDim.x * kBlock) one(r0);
    }
    trav_counters<kMode>(a, nrays, nv, nt);
}

// Instanced scenes: straight schedule over the two-level DFS; closest modes
// also write the set triangle of each hit to a.sub.  Compiled for the
// occupancy the LDS stack allows, as the one-level kernels are.
template <int kMode, int kStack, int kBlock, bool kBatch>
__global__ __launch_bounds__(kBlock, inst_waves(kStack, kBlock, 5)) void k_trace_inst(DevScene sc, TraceArgs a) {
    constexpr bool kAny = (kMode == kModeRaysAny || kMode == kModeShadow);
    __shared__ int lds[kStack * kBlock];
    int* stk = lds + threadIdx.x;
    const int n = trav_count(a, kMode <= kModeRaysAny);
    if (blockIdx.x == 0 && threadIdx.x < 3 && a.zero[threadIdx.x] >= 0) a.counts[a.zero[threadIdx.x]] = 0;
    int* cursor = a.counts + a.cursor_slot;
    uint32_t nv = 0, nt = 0, nrays = 0;
    auto one = [&](int r0) {
        const int r = a.perm ? a.perm[r0] : r0;
        if (a.rank) a.rank[r] = r0;
        if (kMode == kModeShadow && a.q.pid[r] < 0) return;
        float4 ra, rb;
        trav_load<kMode>(a, r, &ra, &rb);
        TravI s;
        travi_begin(s, ra, rb);
        const DevScene& s2 = kBatch ? a.scenes[a.ps.req[a.q.pid[r]]] : sc;
        const float4* nodes = s2.nodes;
        const float4* prims = s2.prims;
        if (s2.num_nodes > 0)
            while (!trav_step_inst<kAny, kBlock>(s2, nodes, prims, s, stk, nv, nt)) {}
        if (s.sp_base >= 0) restore_world(s);   // any-hit ended inside a set
        trav_finish<kMode>(s2, a, r, s.t);
        if (!kAny && a.sub) a.sub[r] = s.t.best.prim >= 0 ? s.sub : -1;
        nrays++;
    };
    if (a.dyn) {
        if (!chunk_wave(n)) return;
        for (int c = next_chunk(cursor); c < n; c = next_chunk(cursor))
            if (c + __lane_id() < n) one(c + __lane_id());
    } else {
        for (int r0 = blockIdx.x * kBlock + threadIdx.x; r0 < n; r0 += gridDim.x * kBlock) one(r0);
    }
    trav_counters<kMode>(a, nrays, nv, nt);
}

// Schedule 3 ("sequences"): a wavefront takes kSeq x 64 consecutive rays
// from the work cursor and lane l traces rays l, l + 64, ... of them one after
// another, starting its next ray as soon as one ends, so a wavefront waits
// for its slowest lane once per kSeq rays instead of once per ray.
constexpr int kSeq = 4;

// kBatch: each ray's scene is a.scenes[ps.req[pid]]; a lane keeps its
// current ray's node and primitive arrays in registers.
template <int kMode, int kStack, int kBlock, bool kBatch>
__global__ __launch_bounds__(kBlock, lds_blocks(kStack, kBlock)) void k_trace_seq(DevScene sc, TraceArgs a) {
    constexpr bool kAny = (kMode == kModeRaysAny || kMode == kModeShadow);
    __shared__ int lds[kStack * kBlock];
    int* stk = lds + threadIdx.x;
    const int n = trav_count(a, kMode <= kModeRaysAny);
    if (blockIdx.x == 0 && threadIdx.x < 3 && a.zero[threadIdx.x] >= 0) a.counts[a.zero[threadIdx.x]] = 0;
    int* cursor = a.counts + a.cursor_slot;
    uint32_t nv = 0, nt = 0, nrays = 0;
    if (!chunk_wave(n, 64 * kSeq)) return;
    for (int c = next_chunk(cursor, 64 * kSeq); c < n; c = next_chunk(cursor, 64 * kSeq)) {
        int j = 0, r = 0, req = 0;
        const float4* nodes = sc.nodes;
        const float4* prims = sc.prims;
        Trav t;
        // the lane's next ray of the chunk that needs tracing; false when none is left
        auto start = [&]() {
            while (j < kSeq) {
                const int r0 = c + __lane_id() + 64 * j;
                j++;
                if (r0 >= n) break;
                r = a.perm ? a.perm[r0] : r0;
                if (a.rank) a.rank[r] = r0;
                if (kMode == kModeShadow && a.q.pid[r] < 0) continue;   // a hole of the ranked queue
                float4 ra, rb;
                trav_load<kMode>(a, r, &ra, &rb);
                trav_begin(t, ra, rb);
                if constexpr (kBatch) {
                    req = a.ps.req[a.q.pid[r]];
                    nodes = a.scenes[req].nodes;
                    prims = a.scenes[req].prims;
                }
                return true;
            }
            j = kSeq;
            return false;
        };
        bool active = start();
        while (__ballot(active)) {
            if (active && trav_step<kAny, kBlock>(nodes, prims, t, stk, nv, nt)) {
                if constexpr (kBatch) trav_finish<kMode>(a.scenes[req], a, r, t);
                else trav_finish<kMode>(sc, a, r, t);
                nrays++;
                active = start();
            }
        }
    }
    trav_counters<kMode>(a, nrays, nv, nt);
}

// MIS and shadow rays of one bounce in one launch: the concatenated queue
// [MIS rays | shadow rays] is taken like a straight schedule, so the MIS
// rays (few, often long) run beside the shadow rays instead of leaving the
// chip idle behind their longest one.  Each ray's query and result are those
// of the separate MIS (closest hit + Le) and shadow (any hit) launches.
template <int kStack, int kBlock, bool kBatch, bool kInst>
__global__ __launch_bounds__(kBlock, kInst ? inst_waves(kStack, kBlock, 4) : lds_blocks(kStack, kBlock))
void k_trace_secondary(DevScene sc, TraceArgs a) {
    __shared__ int lds[kStack * kBlock];
    int* stk = lds + threadIdx.x;
    const int n1 = a.counts[a.cnt_slot];
    const int n = n1 + a.counts[a.cnt_slot2];
    int* cursor = a.counts + a.cursor_slot;
    uint32_t nv_m = 0, nt_m = 0, nr_m = 0, nv_s = 0, nt_s = 0, nr_s = 0;
    auto one = [&](int r0) {
        const bool mis = r0 < n1;
        const int r = mis ? r0 : r0 - n1;
        const RayQueue& q = mis ? a.q : a.q2;
        const int pid = q.pid[r];
        if (pid < 0) return;   // a hole of the ranked shadow queue
        const DevScene& s2 = kBatch ? a.scenes[a.ps.req[pid]] : sc;
        const float4* nodes = s2.nodes;
        const float4* prims = s2.prims;
        if constexpr (kInst) {   // the two-level DFS (Primitive.cs:144-189)
            TravI ti;
            travi_begin(ti, q.org[r], q.dir[r]);
            if (mis) {
                if (s2.num_nodes > 0)
                    while (!trav_step_inst<false, kBlock>(s2, nodes, prims, ti, stk, nv_m, nt_m)) {}
                trav_finish<kModeMis>(s2, a, r, ti.t);
                nr_m++;
            } else {
                if (s2.num_nodes > 0)
                    while (!trav_step_inst<true, kBlock>(s2, nodes, prims, ti, stk, nv_s, nt_s)) {}
                a.ps.vis[pid] = ti.t.best.prim >= 0 ? 0 : 1;
                nr_s++;
            }
            return;
        }
        Trav t;
        trav_begin(t, q.org[r], q.dir[r]);
        if (mis) {
            while (!trav_step<false, kBlock>(nodes, prims, t, stk, nv_m, nt_m)) {}
            trav_finish<kModeMis>(s2, a, r, t);
            nr_m++;
        } else {
            while (!trav_step<true, kBlock>(nodes, prims, t, stk, nv_s, nt_s)) {}
            a.ps.vis[pid] = t.best.prim >= 0 ? 0 : 1;
            nr_s++;
        }
    };
    if (a.dyn) {
        // 128-ray chunks traced as two rounds of 64: the second round is the
        // neighbouring rays of the shadow order (at the first bounce the
        // same tile's next sample), whose nodes the first round left in L1
        if (!chunk_wave(n, 128)) return;
        for (int c = next_chunk(cursor, 128); c < n; c = next_chunk(cursor, 128))
            for (int k = 0; k < 2; k++)
                if (c + 64 * k + __lane_id() < n) one(c + 64 * k + __lane_id());
    } else {
        for (int r0 = blockIdx.x * kBlock + threadIdx.x; r0 < n; r0 += gridDim.x * kBlock) one(r0);
    }
    add_class_counters(a.counts, 1, nr_m, nv_m, nt_m);
    add_counter((unsigned long long*)(a.counts + kCntRaysClosest), nr_m);
    add_class_counters(a.counts, 2, nr_s, nv_s, nt_s);
    add_counter((unsigned long long*)(a.counts + kCntRaysAny), nr_s);
}

// vol_path chains walked to the end per lane: a lane takes a chain ray of
// the queue (64-ray chunks from the work cursor), traces its segment, applies
// chain_step and, while the chain crosses medium boundaries, traces the next
// segment itself.  Each segment is the same closest-hit query and the same
// chain step as the lockstep kernels' (k_trace_* in a chain mode +
// k_chain_advance), so results and query counts are theirs; no host loop
// over the steps.
template <bool kMis, int kStack, int kBlock, bool kBatch, bool kInst>
__global__ __launch_bounds__(kBlock, kInst ? inst_waves(kStack, kBlock, 4) : lds_blocks(kStack, kBlock))
void k_chain_full(DevScene sc, TraceArgs a) {
    constexpr int kChainMode = kMis ? kModeChainMis : kModeChainShadow;
    __shared__ int lds[kStack * kBlock];
    int* stk = lds + threadIdx.x;
    const int n = a.counts[a.cnt_slot];
    int* cursor = a.counts + a.cursor_slot;
    uint32_t nv = 0, nt = 0, nrays = 0;
    auto one = [&](int r) {
        const int pid = a.q.pid[r];
        const DevScene& s2 = kBatch ? a.scenes[a.ps.req[pid]] : sc;
        float4 ra = a.q.org[r], rb = a.q.dir[r];
        for (int step = 0;; step++) {
            if (step >= kMaxChainSteps) {   // the lockstep loop's bound: the render fails
                atomicOr(a.counts + kCntChainStuck, 1);
                break;
            }
            float4 h;
            int hsub = -1;
            if constexpr (kInst) {
                TravI ti;
                travi_begin(ti, ra, rb);
                if (s2.num_nodes > 0)
                    while (!trav_step_inst<false, kBlock>(s2, s2.nodes, s2.prims, ti, stk, nv, nt)) {}
                const Trav& t = ti.t;
                h = t.best.prim >= 0 ? make_float4(t.best.t, __int_as_float(t.best.prim), t.best.u, t.best.v)
                                     : make_float4(0.f, __int_as_float(-1), 0.f, 0.f);
                hsub = t.best.prim >= 0 ? ti.sub : -1;
            } else {
                Trav t;
                trav_begin(t, ra, rb);
                while (!trav_step<false, kBlock>(s2.nodes, s2.prims, t, stk, nv, nt)) {}
                h = t.best.prim >= 0 ? make_float4(t.best.t, __int_as_float(t.best.prim), t.best.u, t.best.v)
                                     : make_float4(0.f, __int_as_float(-1), 0.f, 0.f);
            }
            nrays++;
            f3 no, nd;
            float ntmin, ntmax;
            if (!chain_step(s2, kMis, pid, ra, rb, h, hsub, a.ps, &no, &nd, &ntmin, &ntmax)) break;
            ra = make_float4(no.x, no.y, no.z, ntmin);
            rb = make_float4(nd.x, nd.y, nd.z, ntmax);
        }
    };
    // a large queue: 256 chains per cursor atomic, walked as four
    // wavefront-wide rounds (a chain of the small media scenes is a few cheap
    // queries, and one atomic per 64 chains cost more than the balance it
    // bought); a queue that would not give every wavefront such a chunk keeps
    // 64-chain chunks, so more wavefronts share it
    const int waves = gridDim.x * (kBlock / 64);
    const int rounds = n >= 256 * waves ? 4 : 1;
    if (!chunk_wave(n, 64 * rounds)) return;
    for (int c = next_chunk(cursor, 64 * rounds); c < n; c = next_chunk(cursor, 64 * rounds))
        for (int k = 0; k < rounds; k++)
            if (c + 64 * k + __lane_id() < n) one(c + 64 * k + __lane_id());
    trav_counters<kChainMode>(a, nrays, nv, nt);
}

// Schedule 1 ("refill"): lanes that finish are refilled from an atomic ray
// cursor once at least kRefill lanes of the wavefront are idle.
template <int kMode, int kStack, int kBlock>
__global__ __launch_bounds__(kBlock, lds_blocks(kStack, kBlock)) void k_trace_refill(DevScene sc, TraceArgs a) {
    constexpr bool kAny = (kMode == kModeRaysAny || kMode == kModeShadow);
    __shared__ int lds[kStack * kBlock];
    int* stk = lds + threadIdx.x;
    const int n = trav_count(a, kMode <= kModeRaysAny);
    if (blockIdx.x == 0 && threadIdx.x < 3 && a.zero[threadIdx.x] >= 0) a.counts[a.zero[threadIdx.x]] = 0;
    int* cursor = a.counts + a.cursor_slot;
    uint32_t nv = 0, nt = 0, nrays = 0;
    bool active = false, drained = false;
    int r = 0;
    Trav t;
    trav_begin(t, make_float4(0.f, 0.f, 0.f, 0.f), make_float4(1.f, 1.f, 1.f, 0.f));
    while (true) {
        if (!drained) {
            uint64_t idle = __ballot(!active);
            int nidle = __popcll(idle);
            if (nidle >= kRefill) {
                int leader = __ffsll((unsigned long long)idle) - 1;
                int base = 0;
                if (__lane_id() == leader) base = atomicAdd(cursor, nidle);
                base = __shfl(base, leader);
                if (base + nidle >= n) drained = true;
                if (!active) {
                    int rr = base + __popcll(idle & lanemask_lt());
                    if (rr < n) {
                        r = a.perm ? a.perm[rr] : rr;
                        if (a.rank) a.rank[r] = rr;
                    }
                    if (rr < n && !(kMode == kModeShadow && a.q.pid[r] < 0)) {
                        float4 ra, rb;
                        trav_load<kMode>(a, r, &ra, &rb);
                        trav_begin(t, ra, rb);
                        active = true;
                        nrays++;
                    }
                }
            }
        }
        if (__ballot(active) == 0) {
            if (drained) break;
            continue;   // a chunk of holes only (ranked shadow queue): take the next one
        }
        if (!active) continue;
        // a few steps between refill checks
        bool done = false;
        for (int k = 0; k < 4 && !done; k++) done = trav_step<kAny, kBlock>(sc.nodes, sc.prims, t, stk, nv, nt);
        if (!done) continue;
        active = false;
        trav_finish<kMode>(sc, a, r, t);
    }
    trav_counters<kMode>(a, nrays, nv, nt);
}


// LDS stack variants: the DFS stack never holds more entries than the
// deepest leaf's depth, so the smallest variant that covers the tree is
// launched (64 = the reference's Bvh.StackSize).  LDS per wavefront
// (entries x 64 x 4 B) sets the occupancy: 24 or 26 entries give six
// wavefronts per SIMD, 32 give five.
// Blocks of `kernel` that the device holds at once (registers and LDS), at
// least one per CU: grids beyond it would run as a second, partial round.
static int resident_blocks(const void* kernel, int block) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    return per_cu * g_num_cus;
}

static void launch_resident(void (*kernel)(DevScene, TraceArgs), int blocks, int block, hipStream_t s,
                            const DevScene& sc, const TraceArgs& a) {
    static std::mutex mu;
    static std::unordered_map<const void*, int> cache;   // per kernel
    const void* key = reinterpret_cast<const void*>(kernel);
    int res;
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find(key);
        if (it == cache.end()) it = cache.emplace(key, resident_blocks(key, block)).first;
        res = it->second;
    }
    hipLaunchKernelGGL(kernel, dim3(std::min(blocks, res)), dim3(block), 0, s, sc, a);
}

template <int kMode, int kStack, int kBlock>
static void launch_trace_variant(const DevScene& sc, const TraceArgs& a, int blocks, int schedule, hipStream_t s) {
    if constexpr (kMode == kModeChainFullShadow || kMode == kModeChainFullMis) {
        constexpr bool kMis = kMode == kModeChainFullMis;
        if (sc.num_insts > 0) {
            if (a.scenes) launch_resident(k_chain_full<kMis, kStack, kBlock, true, true>, blocks, kBlock, s, sc, a);
            else launch_resident(k_chain_full<kMis, kStack, kBlock, false, true>, blocks, kBlock, s, sc, a);
        } else {
            if (a.scenes) launch_resident(k_chain_full<kMis, kStack, kBlock, true, false>, blocks, kBlock, s, sc, a);
            else launch_resident(k_chain_full<kMis, kStack, kBlock, false, false>, blocks, kBlock, s, sc, a);
        }
        return;
    } else if constexpr (kMode == kModeSecondary) {
        if (sc.num_insts > 0) {
            if (a.scenes) launch_resident(k_trace_secondary<kStack, kBlock, true, true>, blocks, kBlock, s, sc, a);
            else launch_resident(k_trace_secondary<kStack, kBlock, false, true>, blocks, kBlock, s, sc, a);
        } else {
            if (a.scenes) launch_resident(k_trace_secondary<kStack, kBlock, true, false>, blocks, kBlock, s, sc, a);
            else launch_resident(k_trace_secondary<kStack, kBlock, false, false>, blocks, kBlock, s, sc, a);
        }
        return;
    } else {
    // instanced scenes: the two-level kernel, sequences taken as chunks (a
    // lane's rays back to back measured slower there: its lanes then diverge
    // between scene and set steps)
    if (sc.num_insts > 0) {
        if (kMode >= kModeExt && a.scenes)
            launch_resident(k_trace_inst<kMode, kStack, kBlock, kMode >= kModeExt>, blocks, kBlock, s, sc, a);
        else
            launch_resident(k_trace_inst<kMode, kStack, kBlock, false>, blocks, kBlock, s, sc, a);
        return;
    }
    if constexpr (kMode >= kModeExt) {
        if (a.scenes) {   // batched: per-ray scene; sequences or the straight schedule
            if (schedule == 3) launch_resident(k_trace_seq<kMode, kStack, kBlock, true>, blocks, kBlock, s, sc, a);
            else launch_resident(k_trace_straight<kMode, kStack, kBlock, true>, blocks, kBlock, s, sc, a);
            return;
        }
    }
    if (schedule == 1) launch_resident(k_trace_refill<kMode, kStack, kBlock>, blocks, kBlock, s, sc, a);
    else if (schedule == 3) launch_resident(k_trace_seq<kMode, kStack, kBlock, false>, blocks, kBlock, s, sc, a);
    else launch_resident(k_trace_straight<kMode, kStack, kBlock, false>, blocks, kBlock, s, sc, a);
    }
}

// ---------------------------------------------------------------- octree traversal
// Octree.Intersect (Accelerators/Octree.cs:95-196): the root box, then a
// LIFO stack of the nodes whose boxes the ray met when their parent was
// popped (children tested and pushed in index order, popped in reverse; a
// popped node is not tested again); a leaf tests its primitives in order (a
// primitive may sit in several leaves): closest hit keeps the first strictly
// closer hit and shrinks MaxT, any hit stops at the first.  Node records:
// lo = {min, max.x}, hi = {max.y, max.z, offset, count}: a leaf has count > 0
// primitive records from `offset`, an interior node -count children from
// node `offset`.  The stack lives in private memory (kStack entries, at
// least the tree's worst case: the host computes it from the flattened tree).
template <bool kAny, int kStack>
__device__ __forceinline__ void oct_trace(const float4* __restrict__ nodes, const float4* __restrict__ prims,
                                          int num_nodes, Trav& t, uint32_t& nv, uint32_t& nt) {
    if (num_nodes == 0) return;
    nv++;
    if (!box_hit(nodes[0], nodes[1], t.o, t.inv, t.nx, t.ny, t.nz, t.tmax)) return;
    int stk[kStack];
    int sp = 0;
    stk[sp++] = 0;
    while (sp > 0) {
        const int idx = stk[--sp];
        const float4 hi = nodes[2 * idx + 1];
        const int off = __float_as_int(hi.z), cnt = __float_as_int(hi.w);
        if (cnt > 0) {
            for (int i = 0; i < cnt; i++) {
                const float4* p = prims + 3 * (off + i);
                float4 p0 = p[0], p1 = p[1], p2 = p[2];
                whole(p0); whole(p1); whole(p2);
                nt++;
                float tu = 0.f, tv = 0.f, tt = 0.f;
                bool h;
                if (__float_as_int(p1.w) == kTri) {
                    h = tri_test(mk(p0.x, p0.y, p0.z), mk(p1.x, p1.y, p1.z), mk(p2.x, p2.y, p2.z), t.o, t.d, t.tmin,
                                 t.tmax, &tu, &tv, &tt);
                } else {
                    h = sphere_test(make_float4(p0.x, p0.y, p0.z, p1.x), t.o, t.d, t.tmin, t.tmax, &tt);
                }
                if (h) {
                    if (kAny) { t.best.prim = __float_as_int(p0.w); return; }
                    if (tt < (t.best.prim >= 0 ? t.best.t : kFloatMax)) {
                        t.best.t = tt; t.best.u = tu; t.best.v = tv; t.best.prim = __float_as_int(p0.w);
                        t.tmax = tt;
                    }
                }
            }
        } else {
            for (int i = 0; i < -cnt; i++) {
                const int ch = off + i;
                nv++;
                if (box_hit(nodes[2 * ch], nodes[2 * ch + 1], t.o, t.inv, t.nx, t.ny, t.nz, t.tmax) && sp < kStack)
                    stk[sp++] = ch;
            }
        }
    }
}

// Every query mode over an octree scene, one ray (or, for the full chain
// modes, one chain) per lane, grid-stride.
template <int kMode, int kStack>
__global__ __launch_bounds__(256) void k_trace_oct(DevScene sc, TraceArgs a) {
    constexpr bool kAny = (kMode == kModeRaysAny || kMode == kModeShadow);
    constexpr bool kFull = (kMode == kModeChainFullShadow || kMode == kModeChainFullMis);
    const int n = kFull ? a.counts[a.cnt_slot] : trav_count(a, kMode <= kModeRaysAny);
    if (blockIdx.x == 0 && threadIdx.x < 3 && a.zero[threadIdx.x] >= 0) a.counts[a.zero[threadIdx.x]] = 0;
    uint32_t nv = 0, nt = 0, nrays = 0;
    for (int r0 = blockIdx.x * 256 + threadIdx.x; r0 < n; r0 += gridDim.x * 256) {
        if constexpr (kFull) {
            constexpr bool kMis = kMode == kModeChainFullMis;
            const int pid = a.q.pid[r0];
            float4 ra = a.q.org[r0], rb = a.q.dir[r0];
            for (int step = 0;; step++) {
                if (step >= kMaxChainSteps) {
                    atomicOr(a.counts + kCntChainStuck, 1);
                    break;
                }
                Trav t;
                trav_begin(t, ra, rb);
                oct_trace<false, kStack>(sc.nodes, sc.prims, sc.num_nodes, t, nv, nt);
                const float4 h = t.best.prim >= 0 ? make_float4(t.best.t, __int_as_float(t.best.prim), t.best.u, t.best.v)
                                                  : make_float4(0.f, __int_as_float(-1), 0.f, 0.f);
                nrays++;
                f3 no, nd;
                float ntmin, ntmax;
                if (!chain_step(sc, kMis, pid, ra, rb, h, -1, a.ps, &no, &nd, &ntmin, &ntmax)) break;
                ra = make_float4(no.x, no.y, no.z, ntmin);
                rb = make_float4(nd.x, nd.y, nd.z, ntmax);
            }
        } else {
            const int r = a.perm ? a.perm[r0] : r0;
            if (a.rank) a.rank[r] = r0;
            if (kMode == kModeShadow && a.q.pid[r] < 0) continue;   // a hole of a ranked shadow queue
            float4 ra, rb;
            trav_load<kMode>(a, r, &ra, &rb);
            Trav t;
            trav_begin(t, ra, rb);
            oct_trace<kAny, kStack>(sc.nodes, sc.prims, sc.num_nodes, t, nv, nt);
            trav_finish<kMode>(sc, a, r, t);
            nrays++;
        }
    }
    if constexpr (kFull) trav_counters<kMode == kModeChainFullMis ? kModeChainMis : kModeChainShadow>(a, nrays, nv, nt);
    else trav_counters<kMode>(a, nrays, nv, nt);
}

template <int kMode>
static hipError_t launch_trace_mode(const DevScene& sc, const TraceArgs& a0, int blocks_hint, int sched,
                                    hipStream_t s) {
    if (sc.octree) {
        // octree scenes: one ray per lane; batches, instancing and the fused
        // MIS + shadow launch are not built for them (the callers avoid them)
        if constexpr (kMode == kModeSecondary) return hipErrorInvalidValue;
        if (a0.scenes) return hipErrorInvalidValue;
        const int pb = persistent_blocks();
        const int b = blocks_hint > 0 ? std::min(blocks_hint, pb) : pb;
        if (sc.stack_needed <= 32) hipLaunchKernelGGL((k_trace_oct<kMode, 32>), dim3(b), dim3(256), 0, s, sc, a0);
        else if (sc.stack_needed <= 96) hipLaunchKernelGGL((k_trace_oct<kMode, 96>), dim3(b), dim3(256), 0, s, sc, a0);
        else hipLaunchKernelGGL((k_trace_oct<kMode, 256>), dim3(b), dim3(256), 0, s, sc, a0);
        return hipGetLastError();
    }
    int need = sc.stack_needed;
    int pb = persistent_blocks();
    if (kMode <= kModeRaysAny || sched < 0 || sched >= kNumSchedules) sched = 0;
    TraceArgs a = a0;
    // straight kernels: 2 = cursor chunks; 3 = sequences
    a.dyn = sched == 2 ? 1 : (sched == 3 ? 2 : 0);
    if (need <= 24) {
        int b = blocks_hint > 0 ? std::min(blocks_hint, pb) : pb;
        launch_trace_variant<kMode, 24, 256>(sc, a, b, sched, s);
    } else if (need <= 26) {
        // 26 entries x 256 lanes x 4 B: six blocks fill the 160 KB of LDS, six waves per SIMD
        int b = blocks_hint > 0 ? std::min(blocks_hint, pb) : pb;
        launch_trace_variant<kMode, 26, 256>(sc, a, b, sched, s);
    } else if (need <= 32) {
        int b = blocks_hint > 0 ? std::min(blocks_hint, pb) : pb;
        launch_trace_variant<kMode, 32, 256>(sc, a, b, sched, s);
    } else if (need <= 40) {
        // 40 entries x 128 lanes x 4 B: eight blocks, four waves per SIMD (the
        // two-level trees of a 14-deep scene over 26-deep instance sets)
        int b = blocks_hint > 0 ? std::min(2 * blocks_hint, 2 * pb) : 2 * pb;
        launch_trace_variant<kMode, 40, 128>(sc, a, b, sched, s);
    } else if (need <= 48) {
        int b = blocks_hint > 0 ? std::min(2 * blocks_hint, 2 * pb) : 2 * pb;
        launch_trace_variant<kMode, 48, 128>(sc, a, b, sched, s);
    } else {
        int b = blocks_hint > 0 ? std::min(2 * blocks_hint, 2 * pb) : 2 * pb;
        launch_trace_variant<kMode, 64, 128>(sc, a, b, sched, s);
    }
    return hipGetLastError();
}

static TraceArgs trace_args(int* counts) {
    TraceArgs a{};
    a.perm = nullptr;
    a.rank = nullptr;
    a.counts = counts;
    a.zero[0] = a.zero[1] = a.zero[2] = -1;
    a.scenes = nullptr;
    a.dyn = 0;
    return a;
}

__global__ void k_zero_slot(int* c, int slot) {
    if (threadIdx.x == 0) c[slot] = 0;
}

__global__ void k_zero_slots(int* c, int a, int b) {
    if (threadIdx.x == 0) { c[a] = 0; c[b] = 0; }
}

hipError_t launch_zero_slots(int* counts, int a, int b, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_slots, dim3(1), dim3(64), 0, s, counts, a, b);
    return hipGetLastError();
}

hipError_t launch_trace_rays(const DevScene& sc, const float4* rays, int n, float4* hits, uint8_t* occ, int* sub,
                             unsigned long long* counters, bool any, int* counts, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_slot, dim3(1), dim3(64), 0, s, counts, kCntWork + 3);
    TraceArgs a = trace_args(counts);
    a.rays = rays; a.hits = hits; a.occ = occ; a.counters = counters; a.n_fixed = n; a.sub = sub;
    a.cursor_slot = kCntWork + 3;
    int blocks = (n + 255) / 256;
    return any ? launch_trace_mode<kModeRaysAny>(sc, a, blocks, 0, s)
               : launch_trace_mode<kModeRaysClosest>(sc, a, blocks, 0, s);
}
// ---------------------------------------------------------------- wavefront
// Vector3.Transform / TransformNormal with a row-major 4x4 (row-vector convention).
__device__ __forceinline__ f3 xform_point(const float* M, f3 v) {
    return mk(v.x * M[0] + v.y * M[4] + v.z * M[8] + M[12], v.x * M[1] + v.y * M[5] + v.z * M[9] + M[13],
              v.x * M[2] + v.y * M[6] + v.z * M[10] + M[14]);
}
__device__ __forceinline__ f3 xform_vector(const float* M, f3 v) {
    return mk(v.x * M[0] + v.y * M[4] + v.z * M[8], v.x * M[1] + v.y * M[5] + v.z * M[9],
              v.x * M[2] + v.y * M[6] + v.z * M[10]);
}

// ---------------------------------------------------------------- path pool
// Paths stream through a fixed pool of slots: every bounce, the slots of the
// paths that finished are refilled with the next samples of the pass, so the
// bounce kernels see a full queue until the pass runs dry.  Pass path g lives
// in the segment whose [path_offset, path_offset + nsamples * npix) holds it
// (tile-major within a segment: seg_path below).
__device__ __forceinline__ int find_seg(const BatchSeg* segs, int nseg, int p) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (segs[mid].path_offset <= p) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Pixel q (< w*h) of a crop in t x t tiles (t = 8 by default,
// rd_tuning.tile): tile rows of t pixel rows top to bottom, tiles left to
// right, pixels row-major inside a tile (the last tile row and column may be
// narrower).  Consecutive paths then start camera rays over a square of the
// image, which share more BVH nodes per wavefront load than a row's (about
// 28 % fewer distinct node lines on the 1080p workload at t = 8).  The
// accumulation buffers are kept in the same order.
__device__ __forceinline__ void tile_pixel(int q, int w, int h, int t, int* px, int* py) {
    const int tr = q / (t * w);
    const int th = min(t, h - tr * t);
    const int r = q - tr * t * w;
    const int ntc = (w + t - 1) / t;
    const int j = min(r / (t * th), ntc - 1);
    const int o = r - j * t * th;
    const int tw = min(t, w - j * t);
    *px = j * t + o % tw;
    *py = tr * t + o / tw;
}

// Path order inside a segment: tile by tile (tile_pixel's order), and inside
// a tile, sample by sample with the tile's pixels in order.  A wavefront's 64
// consecutive paths are one 8x8 tile at one sample, as in a sample-by-sample
// sweep, but all samples of a tile now run back to back: the chip's rays in
// flight cover about a hundred tiles instead of a band of the image per
// sample, so the BVH nodes and triangles under them are fetched into L2 once
// rather than once per sample.  seg_path: segment-local path index ->
// (pixel q in tile order, sample si); seg_local is its inverse.  The order is
// a schedule only: a path's key is its pixel and sample, and every pixel's
// samples are summed in sample order, so images do not depend on t.
__device__ __forceinline__ void seg_path(int local, int w, int h, int ns, int t, int* q, int* si) {
    // a segment holds fewer than 2^31 paths, but a full tile row of a crop a
    // few pixels tall can be counted past that
    const long long row = (long long)t * w * ns;   // paths per full tile row
    const int tr = (int)(local / row);
    const int th = min(t, h - tr * t);
    const int r = (int)(local - tr * row);
    const int ntc = (w + t - 1) / t;
    const int j = (int)min((long long)r / ((long long)t * th * ns), (long long)(ntc - 1));
    const int o = (int)(r - (long long)j * t * th * ns);
    const int nt = min(t, w - j * t) * th;         // pixels in the tile
    *si = o / nt;
    *q = tr * t * w + j * t * th + (o - *si * nt);
}
__device__ __forceinline__ int seg_local(int q, int si, int w, int h, int ns, int t) {
    const int tr = q / (t * w);
    const int th = min(t, h - tr * t);
    const int r = q - tr * t * w;
    const int ntc = (w + t - 1) / t;
    const int j = min(r / (t * th), ntc - 1);
    const int o = r - j * t * th;                  // pixel within the tile
    const int nt = min(t, w - j * t) * th;
    return (q - o) * ns + si * nt + o;
}

// Pixel sample (x,y)+U^2 and PerspectiveCamera.SampleRay (Camera.cs:182-191),
// BlockBasedRenderer.RenderTask (BlockBasedRenderer.cs:136-139), into slot `slot`.
__device__ __forceinline__ void start_path(const DevScene& sc, const BatchSeg& sg, int g, int slot, int qi, RayQueue q,
                                           PathState ps) {
    int si, qq, tx, ty;
    seg_path(g - sg.path_offset, sg.crop_w, sg.npix / sg.crop_w, sg.nsamples, sg.tile, &qq, &si);
    tile_pixel(qq, sg.crop_w, sg.npix / sg.crop_w, sg.tile, &tx, &ty);
    int x = sg.x0 + tx, y = sg.y0 + ty;
    const uint64_t key = path_key(sg.seed, (uint32_t)(y * sg.width + x), (uint32_t)(sg.sample_base + si));
    Rng rng{key, 0u};
    float ox = rng.next_float(), oy = rng.next_float();
    float sx = (float)x + ox, sy = (float)y + oy;
    const DevCamera& c = sc.cam;
    f3 ndc = mk(sx * c.inv_res_x, sy * c.inv_res_y, 0.0f);
    f3 nearp = xform_point(c.clip_to_camera, ndc);
    f3 dir = normalize(nearp);
    float inv_z = 1.0f / dir.z;
    f3 o = xform_point(c.camera_to_world, mk(0.0f));
    f3 d = normalize(xform_vector(c.camera_to_world, dir));
    q.pid[qi] = slot;
    q.org[qi] = make_float4(o.x, o.y, o.z, c.near_z * inv_z);
    q.dir[qi] = make_float4(d.x, d.y, d.z, c.far_z * inv_z);
    ps.L[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
    ps.beta[slot] = make_float4(1.f, 1.f, 1.f, 0.f);
    ps.flags[slot] = kFlagSpecular;
    ps.rng_n[slot] = rng.n;
    ps.key[slot] = key;
    ps.gid[slot] = g;
    ps.req[slot] = sg.req;
    ps.pix[slot] = x * sg.height + y;
    ps.depth[slot] = 0;
    if (ps.env) ps.env[slot] = sc.global_medium;   // vol_path: rayEnv = Scene.GlobalMedium
}

// ---- ordered compaction of the pool
// Blocks of kCompactSlots slots (256 threads x 4 runs of 256 slots).  count: per block the
// slots whose path continues and the free slots; scan: block offsets of the
// free slots and of the queue entries (continuing slots + free slots that get
// a new path, i.e. whose pool-wide free rank is below the paths left);
// scatter: queue entries in slot order, new paths started in place.
constexpr int kCompactSlots = 1024;

// Exclusive scan of one int per thread over a block of kThreads threads;
// *total receives the block sum.  tmp: kThreads / 64 ints of LDS.
template <int kThreads>
__device__ __forceinline__ int block_excl_scan(int v, int* tmp, int* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) tmp[wid] = x;
    __syncthreads();
    int before = 0, tot = 0;
    for (int w = 0; w < kThreads / 64; w++) {
        int t = tmp[w];
        if (w < wid) before += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return before + x - v;
}

__global__ __launch_bounds__(256) void k_pool_count(const int* __restrict__ alive, int pool, int2* blk) {
    __shared__ int tmp[4];
    const int base = blockIdx.x * kCompactSlots;
    int na = 0, nf = 0;
    for (int i = 0; i < kCompactSlots / 256; i++) {   // coalesced: slot base + i * 256 + thread
        int slot = base + i * 256 + threadIdx.x;
        if (slot < pool) {
            if (alive[slot]) na++;
            else nf++;
        }
    }
    int ta, tf;
    block_excl_scan<256>(na, tmp, &ta);
    block_excl_scan<256>(nf, tmp, &tf);
    if (threadIdx.x == 0) blk[blockIdx.x] = make_int2(ta, tf);
}

// Block offsets in three steps, for any number of count blocks: tiles of
// 1024 blocks are summed, one block scans the tile sums, and every tile then
// scans its blocks from its tile prefix.  The paths started are the first
// `left` free slots, so a tile starts min(max(left - free before it, 0), its
// free slots) of them.
__global__ __launch_bounds__(1024) void k_scan_tiles(const int2* __restrict__ blk, int nblk, int2* tsum) {
    __shared__ int tmp[16];
    const int b = blockIdx.x * 1024 + threadIdx.x;
    const int2 c = b < nblk ? blk[b] : make_int2(0, 0);
    int ta, tf;
    block_excl_scan<1024>(c.x, tmp, &ta);
    block_excl_scan<1024>(c.y, tmp, &tf);
    if (threadIdx.x == 0) tsum[blockIdx.x] = make_int2(ta, tf);
}

__global__ __launch_bounds__(1024) void k_scan_top(const int2* __restrict__ tsum, int ntiles, int npass, int* counts,
                                                   int cnt_q, int2* tpre, int pass_state) {
    __shared__ int tmp[16];
    const int next = counts[kCntNext];
    const int left = max(npass - next, 0);
    int frun = 0, qrun = 0;
    for (int t0 = 0; t0 < ntiles; t0 += 1024) {
        const int t = t0 + threadIdx.x;
        const int2 c = t < ntiles ? tsum[t] : make_int2(0, 0);
        int ftot;
        const int fpre = frun + block_excl_scan<1024>(c.y, tmp, &ftot);
        const int q = c.x + min(max(left - fpre, 0), c.y);
        int qtot;
        const int qpre = qrun + block_excl_scan<1024>(q, tmp, &qtot);
        if (t < ntiles) tpre[t] = make_int2(fpre, qpre);
        frun += ftot;
        qrun += qtot;
    }
    if (threadIdx.x == 0) {
        counts[cnt_q] = qrun;
        if (pass_state) {
            counts[kCntFree] = next;
            counts[kCntNext] = next + min(left, frun);
        }
    }
}

__global__ __launch_bounds__(1024) void k_scan_apply(const int2* __restrict__ blk, int nblk,
                                                     const int2* __restrict__ tpre, int npass, const int* counts,
                                                     int2* off) {
    __shared__ int tmp[16];
    const int left = max(npass - counts[kCntFree], 0);
    const int b = blockIdx.x * 1024 + threadIdx.x;
    const int2 c = b < nblk ? blk[b] : make_int2(0, 0);
    const int2 tp = tpre[blockIdx.x];
    int tf;
    const int fpre = tp.x + block_excl_scan<1024>(c.y, tmp, &tf);
    const int q = c.x + min(max(left - fpre, 0), c.y);
    int tq;
    const int qpre = tp.y + block_excl_scan<1024>(q, tmp, &tq);
    if (b < nblk) off[b] = make_int2(fpre, qpre);
}

// Queue filter: once every path of the pass has started, the next queue is
// the previous one's entries whose path continues, in the same order (slot
// order, as the pool compaction would give).  Counts per 1024 entries.
__global__ __launch_bounds__(256) void k_queue_count(const int* __restrict__ pid, int n, const int* n_dev,
                                                     const int* __restrict__ alive, int2* blk) {
    __shared__ int tmp[4];
    if (n_dev) n = min(n, *n_dev);
    const int base = blockIdx.x * kCompactSlots;
    int na = 0;
    for (int i = 0; i < kCompactSlots / 256; i++) {
        const int k = base + i * 256 + threadIdx.x;
        if (k < n && alive[pid[k]]) na++;
    }
    int ta;
    block_excl_scan<256>(na, tmp, &ta);
    if (threadIdx.x == 0) blk[blockIdx.x] = make_int2(ta, 0);
}

// Hole removal: the entries of a ranked queue that hold a ray (pid >= 0), in
// the same order.  Counts per 1024 entries.
__global__ __launch_bounds__(256) void k_hole_count(const int* __restrict__ pid, int n, const int* n_dev, int2* blk) {
    __shared__ int tmp[4];
    if (n_dev) n = min(n, *n_dev);
    const int base = blockIdx.x * kCompactSlots;
    int na = 0;
    for (int i = 0; i < kCompactSlots / 256; i++) {
        const int k = base + i * 256 + threadIdx.x;
        if (k < n && pid[k] >= 0) na++;
    }
    int ta;
    block_excl_scan<256>(na, tmp, &ta);
    if (threadIdx.x == 0) blk[blockIdx.x] = make_int2(ta, 0);
}

__global__ __launch_bounds__(256) void k_hole_scatter(RayQueue qin, int n, const int* n_dev,
                                                      const int2* __restrict__ off, RayQueue q) {
    __shared__ int tmp[4];
    if (n_dev) n = min(n, *n_dev);
    const int base = blockIdx.x * kCompactSlots;
    int pos_base = off[blockIdx.x].y;
    for (int i = 0; i < kCompactSlots / 256; i++) {
        const int k = base + i * 256 + threadIdx.x;
        const int pid = k < n ? qin.pid[k] : -1;
        int ts;
        const int pos = pos_base + block_excl_scan<256>(pid >= 0 ? 1 : 0, tmp, &ts);
        if (pid >= 0) {
            q.pid[pos] = pid;
            q.org[pos] = qin.org[k];
            q.dir[pos] = qin.dir[k];
        }
        pos_base += ts;
    }
}

__global__ __launch_bounds__(256) void k_queue_scatter(RayQueue qin, int n, const int* n_dev, PathState ps,
                                                       const int2* __restrict__ off, RayQueue q) {
    __shared__ int tmp[4];
    if (n_dev) n = min(n, *n_dev);
    const int base = blockIdx.x * kCompactSlots;
    int pos_base = off[blockIdx.x].y;
    for (int i = 0; i < kCompactSlots / 256; i++) {
        const int k = base + i * 256 + threadIdx.x;
        const int slot = k < n ? qin.pid[k] : -1;
        const bool take = slot >= 0 && ps.alive[slot] == 1;
        int ts;
        const int pos = pos_base + block_excl_scan<256>(take ? 1 : 0, tmp, &ts);
        if (take) {
            q.pid[pos] = slot;
            q.org[pos] = ps.ray_o[slot];
            q.dir[pos] = ps.ray_d[slot];
        }
        pos_base += ts;
    }
}

template <bool kBatch>
__global__ __launch_bounds__(256) void k_pool_scatter(DevScene sc0, const DevScene* __restrict__ scenes,
                                                      const BatchSeg* segs, int nseg, int npass, PathState ps,
                                                      int pool, const int2* __restrict__ off, const int* counts,
                                                      RayQueue q) {
    __shared__ int tmp[4];
    const int base = blockIdx.x * kCompactSlots;
    const int first = counts[kCntFree];
    const int left = max(npass - first, 0);
    const int2 o = off[blockIdx.x];
    int frank_base = o.x, pos_base = o.y;
    // the block's slots in four runs of 256 (coalesced accesses); within and
    // across runs, free ranks and queue positions follow slot order
    for (int i = 0; i < kCompactSlots / 256; i++) {
        const int slot = base + i * 256 + threadIdx.x;
        const int st = slot < pool ? ps.alive[slot] : -1;
        int tf;
        const int frank = frank_base + block_excl_scan<256>(st == 0 ? 1 : 0, tmp, &tf);
        const bool take = st == 1 || (st == 0 && frank < left);
        int ts;
        const int pos = pos_base + block_excl_scan<256>(take ? 1 : 0, tmp, &ts);
        if (st == 1) {
            q.pid[pos] = slot;
            q.org[pos] = ps.ray_o[slot];
            q.dir[pos] = ps.ray_d[slot];
        } else if (take) {
            const int g = first + frank;
            const BatchSeg sg = segs[find_seg(segs, nseg, g)];
            if constexpr (kBatch) start_path(scenes[sg.req], sg, g, slot, pos, q, ps);
            else start_path(sc0, sg, g, slot, pos, q, ps);
            ps.alive[slot] = 1;
        }
        frank_base += tf;
        pos_base += ts;
    }
}

// blk: nblk block counts, then nblk offsets (off), then 2 x the tile count.
// pass_state: the compaction of the pass's pool or queue (it records the
// paths it started); hole removal leaves those counters alone.
static void scan_blocks(const int2* blk, int nblk, int npass, int* counts, int cnt_q, int2* off, bool pass_state,
                        hipStream_t s) {
    const int ntiles = (nblk + 1023) / 1024;
    int2* tsum = off + nblk;
    int2* tpre = tsum + ntiles;
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(1024), 0, s, blk, nblk, tsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, tsum, ntiles, npass, counts, cnt_q, tpre,
                       pass_state ? 1 : 0);
    hipLaunchKernelGGL(k_scan_apply, dim3(ntiles), dim3(1024), 0, s, blk, nblk, tpre, npass, counts, off);
}

hipError_t launch_compact_queue(RayQueue qin, int n, int npass, int2* blk, RayQueue q, int cnt_q, PathState ps,
                                int* counts, hipStream_t s, const int* n_dev) {
    if (n <= 0) {
        hipLaunchKernelGGL(k_zero_slot, dim3(1), dim3(64), 0, s, counts, cnt_q);
        return hipGetLastError();
    }
    const int nblk = (n + kCompactSlots - 1) / kCompactSlots;
    int2* off = blk + nblk;
    hipLaunchKernelGGL(k_queue_count, dim3(nblk), dim3(256), 0, s, qin.pid, n, n_dev, ps.alive, blk);
    scan_blocks(blk, nblk, npass, counts, cnt_q, off, true, s);
    hipLaunchKernelGGL(k_queue_scatter, dim3(nblk), dim3(256), 0, s, qin, n, n_dev, ps, off, q);
    return hipGetLastError();
}

hipError_t launch_compact_holes(RayQueue qin, int n, int2* blk, RayQueue q, int cnt_q, int* counts, hipStream_t s,
                                const int* n_dev) {
    if (n <= 0) {
        hipLaunchKernelGGL(k_zero_slot, dim3(1), dim3(64), 0, s, counts, cnt_q);
        return hipGetLastError();
    }
    const int nblk = (n + kCompactSlots - 1) / kCompactSlots;
    int2* off = blk + nblk;
    hipLaunchKernelGGL(k_hole_count, dim3(nblk), dim3(256), 0, s, qin.pid, n, n_dev, blk);
    scan_blocks(blk, nblk, 0, counts, cnt_q, off, false, s);
    hipLaunchKernelGGL(k_hole_scatter, dim3(nblk), dim3(256), 0, s, qin, n, n_dev, off, q);
    return hipGetLastError();
}

// The device-driven tail's bounce count: one more bounce when its queue is
// not empty (read on the device, so the launch can sit in a graph).
__global__ void k_tail_tick(int* counts, int cnt_in) {
    if (threadIdx.x == 0 && counts[cnt_in] > 0) counts[kCntTail]++;
}
hipError_t launch_tail_tick(int* counts, int cnt_in, hipStream_t s) {
    hipLaunchKernelGGL(k_tail_tick, dim3(1), dim3(64), 0, s, counts, cnt_in);
    return hipGetLastError();
}

hipError_t launch_compact(const DevScene& sc, const DevScene* scenes, const BatchSeg* segs, int nseg, int npass,
                          int pool, int2* blk, RayQueue q, int cnt_q, PathState ps, int* counts, hipStream_t s) {
    const int nblk = (pool + kCompactSlots - 1) / kCompactSlots;
    int2* off = blk + nblk;
    hipLaunchKernelGGL(k_pool_count, dim3(nblk), dim3(256), 0, s, ps.alive, pool, blk);
    scan_blocks(blk, nblk, npass, counts, cnt_q, off, true, s);
    if (scenes)
        hipLaunchKernelGGL(k_pool_scatter<true>, dim3(nblk), dim3(256), 0, s, sc, scenes, segs, nseg, npass, ps, pool,
                           off, counts, q);
    else
        hipLaunchKernelGGL(k_pool_scatter<false>, dim3(nblk), dim3(256), 0, s, sc, scenes, segs, nseg, npass, ps, pool,
                           off, counts, q);
    return hipGetLastError();
}

// Every slot free and the pass cursor at zero.
__global__ void k_init_pool(PathState ps, int pool, int* counts) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < pool; i += gridDim.x * 256) ps.alive[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) counts[kCntNext] = 0;
}

hipError_t launch_init_pool(PathState ps, int pool, int* counts, hipStream_t s) {
    int blocks = (pool + 255) / 256;
    if (blocks > persistent_blocks()) blocks = persistent_blocks();
    hipLaunchKernelGGL(k_init_pool, dim3(blocks), dim3(256), 0, s, ps, pool, counts);
    return hipGetLastError();
}

// Closest hit of the extension queue.  It also clears the queue counters and
// work cursors the following shade / shadow / MIS stages start from.
hipError_t launch_trace_ext(const DevScene& sc, const DevScene* scenes, RayQueue q, PathState ps, float4* hits,
                            int* counts, int cnt_slot, int za, int zb, int zc, const int* perm, int* rank, int sched,
                            hipStream_t s) {
    TraceArgs a = trace_args(counts);
    a.perm = perm;
    a.rank = perm ? rank : nullptr;
    a.q = q; a.ps = ps; a.hits = hits; a.cnt_slot = cnt_slot; a.cursor_slot = kCntWork + 0; a.scenes = scenes;
    a.zero[0] = za; a.zero[1] = zb; a.zero[2] = zc;
    a.sub = ps.hsub;
    return launch_trace_mode<kModeExt>(sc, a, 0, sched, s);
}

// Any hit of the shadow queue: Scene.IsOccluded (Scene.cs:58-70).
hipError_t launch_trace_shadow(const DevScene& sc, const DevScene* scenes, RayQueue q, PathState ps, int* counts,
                               int cnt_slot, int sched, hipStream_t s) {
    TraceArgs a = trace_args(counts);
    a.q = q; a.ps = ps; a.cnt_slot = cnt_slot; a.cursor_slot = kCntWork + 1; a.scenes = scenes;
    return launch_trace_mode<kModeShadow>(sc, a, 0, sched, s);
}

// The MIS rays (closest hit + Le) and the shadow rays (any hit) of a bounce:
// one fused launch (the one- or two-level traversal) with a straight
// schedule, else the two launches in turn.
hipError_t launch_trace_secondary(const DevScene& sc, const DevScene* scenes, RayQueue qshadow, int cnt_shadow,
                                  RayQueue qmis, PathState ps, int* counts, int sched_shadow, int sched_mis,
                                  hipStream_t s) {
    if (sched_shadow == 1 || sched_mis == 1 || sched_shadow == 3 || sched_mis == 3 || sc.octree) {
        if (hipError_t e = launch_trace_mis(sc, scenes, qmis, ps, counts, sched_mis, s)) return e;
        return launch_trace_shadow(sc, scenes, qshadow, ps, counts, cnt_shadow, sched_shadow, s);
    }
    TraceArgs a = trace_args(counts);
    a.q = qmis; a.cnt_slot = kCntMis;
    a.q2 = qshadow; a.cnt_slot2 = cnt_shadow;
    a.ps = ps; a.cursor_slot = kCntWork + 1; a.scenes = scenes;
    return launch_trace_mode<kModeSecondary>(sc, a, 0, sched_shadow, s);
}

// Closest hit of the BSDF-sampled MIS rays and the radiance they carry.
hipError_t launch_trace_mis(const DevScene& sc, const DevScene* scenes, RayQueue q, PathState ps, int* counts,
                            int sched, hipStream_t s) {
    TraceArgs a = trace_args(counts);
    a.q = q; a.ps = ps; a.cnt_slot = kCntMis; a.cursor_slot = kCntWork + 2; a.scenes = scenes;
    return launch_trace_mode<kModeMis>(sc, a, 0, sched, s);
}
// vol_path Transmittance (mis = false) or IntersectTr (mis = true) chains of
// the queue, each walked to its end (k_chain_full).
hipError_t launch_chain_full(const DevScene& sc, const DevScene* scenes, RayQueue q, PathState ps, bool mis,
                             int* counts, int cnt_slot, hipStream_t s) {
    TraceArgs a = trace_args(counts);
    a.q = q; a.ps = ps; a.cnt_slot = cnt_slot; a.cursor_slot = kCntWork + (mis ? 2 : 1);
    a.scenes = scenes;
    return mis ? launch_trace_mode<kModeChainFullMis>(sc, a, 0, 2, s)
               : launch_trace_mode<kModeChainFullShadow>(sc, a, 0, 2, s);
}

// One closest-hit step of the vol_path Transmittance (mis = false) or
// IntersectTr (mis = true) chains; clears the next chain queue's counter.
hipError_t launch_trace_chain(const DevScene& sc, const DevScene* scenes, RayQueue q, PathState ps, float4* hits,
                              bool mis, int* counts, int cnt_slot, int zero_slot, int sched, hipStream_t s) {
    TraceArgs a = trace_args(counts);
    a.q = q; a.ps = ps; a.hits = hits; a.cnt_slot = cnt_slot; a.cursor_slot = kCntWork + (mis ? 2 : 1);
    a.sub = mis ? ps.hsub2 : ps.hsub;
    a.scenes = scenes;
    a.zero[0] = zero_slot;
    return mis ? launch_trace_mode<kModeChainMis>(sc, a, 0, sched, s)
               : launch_trace_mode<kModeChainShadow>(sc, a, 0, sched, s);
}

// One bounce of PathTracer.Mis / NextEventEstimation / OnlySampleBxdf for
// every path in the extension queue (PathTracer.cs:43-270).
// kAdj: the adjoint replay also records, per material vertex, the albedo R,
// the texture coordinates and the roughness ratio (plastic_rho_ratio) of its
// light-sample, BSDF-sample and continuation terms for resolve's gradients.
template <bool kBatch, bool kAdj, unsigned kM>
__global__ __launch_bounds__(256, 3) void k_shade(DevScene sc0, const DevScene* __restrict__ scenes, RayQueue qin,
                                               const float4* __restrict__ hits, RayQueue qshadow, RayQueue qmis,
                                               PathState ps, int* counts, int cnt_in, const int* __restrict__ rank,
                                               bool rank_identity) {
    __shared__ int sh_app[256 / 64 + 1];   // block_append
    int n = counts[cnt_in];
    // the traversal work cursors restart at zero for the next trace stages
    if (blockIdx.x == 0 && threadIdx.x < 3) counts[kCntWork + threadIdx.x] = 0;
    uint64_t nvert = 0;
    for (int base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {
        int k = base + threadIdx.x;
        bool want_ext = false, want_shadow = false, want_mis = false, want_sss = false;
        f3 eo = mk(0.f), ed = mk(0.f), so = mk(0.f), sd = mk(0.f), mo = mk(0.f), md = mk(0.f);
        float etmin = 0.f, etmax = 0.f, stmax = 0.f;
        int pid = 0;
        if (k < n) {
            pid = qin.pid[k];
            auto run = [&](const DevScene& sc) {
            const DevIntegrator& in = sc.integ;
            float4 a = qin.org[k], b = qin.dir[k];
            f3 ro = mk(a.x, a.y, a.z), rd = mk(b.x, b.y, b.z);
            float rtmin = a.w;
            float4 h = hits[k];
            int prim = __float_as_int(h.y);
            bool hit = prim >= 0;
            Rng rng{ps.key[pid], ps.rng_n[pid]};
            f3 beta = ld3(ps.beta, pid);
            f3 L = ld3(ps.L, pid);
            int flags = ps.flags[pid];
            bool spec = (flags & kFlagSpecular) != 0;
            int depth = ps.depth[pid];
            int nflags = 0;
            f3 wo = -rd;
            DevPrim pr;
            Inct it;
            if (hit) {
                pr = sc.prim_info[prim];
                it = make_inct_any(sc, pr, pr.kind == kInst ? ps.hsub[k] : -1, ro, rd, h.z, h.w, h.x);
            }
            auto inct_le = [&](f3 w) -> f3 {   // Intersection.Le -> DiffuseAreaLight.L
                if (pr.light < 0) return mk(0.0f);
                const DevLight& Lt = sc.lights[pr.light];
                return dot(it.sh.Z, w) > 0 ? mk(Lt.le[0], Lt.le[1], Lt.le[2]) : mk(0.0f);
            };
            bool go = true;
            if (in.method == 0) {
                // OnlySampleBxdf (PathTracer.cs:43-93)
                if (!hit) { L = L + beta * eval_infinite(sc, rd); go = false; }
                else if (in.max_depth != -1 && depth >= in.max_depth) go = false;
            } else {
                if (depth == 0 || spec) {
                    if (hit) L = L + beta * inct_le(wo);
                    else L = L + beta * eval_infinite(sc, rd);
                }
                if (!hit) go = false;
                else if (in.max_depth != -1 && depth >= in.max_depth) go = false;
            }
            if (go && pr.material < 0) {
                // medium boundary without a surface: continue straight on
                want_ext = true;
                eo = it.P; ed = rd; etmin = rtmin; etmax = kFloatMax;
                nflags = flags & kFlagSpecular;
                go = false;
            } else if (go) {
                nvert++;
                MatCtx m = load_material<kM>(sc, pr.material, it.u, it.v);
                f3 wo_l = to_local(it.sh, wo);
                // EstimateDirect takes wo = inct.Wr (PathTracer.cs:314), -ray.D
                // except on an instance, where it is mapped through the transforms
                f3 wo_d = to_local(it.sh, it.wr);
                float rho = 0.0f, r_nee = 0.0f, r_mis = 0.0f, r_cont = 0.0f;
                if constexpr (kAdj) {
                    if (m.type == kMatRoughPlastic) rho = tex_sample(sc, sc.mats[pr.material].tex[1], it.u, it.v).x;
                }
                if (in.method == 0) {
                    if (pr.light >= 0) L = L + beta * inct_le(wo);
                } else if (sc.num_lights > 0) {
                    int l = rng.next_int(sc.num_lights);
                    float selpdf = 1.0f / (float)sc.num_lights;
                    const DevLight Lt = sc.lights[l];
                    LSample ls = light_sample(sc, Lt, it.P, rng);
                    f3 nee = mk(0.0f);
                    if (in.method == 2) {
                        // EstimateDirect, light-sampling half (PathTracer.cs:317-350)
                        if (ls.pdf > 0.0f && !is_black(ls.li)) {
                            f3 wi_l = to_local(it.sh, ls.wi);
                            f3 fr = mat_fr<kM>(m, wo_d, wi_l);
                            float spdf = mat_pdf<kM>(m, wo_d, wi_l);
                            if (!is_black(fr)) {
                                f3 vec = ls.p - it.P;
                                float dis = length(vec);
                                if (!(dis < 0.001f)) {
                                    want_shadow = true;
                                    so = it.P; sd = normalize(vec); stmax = dis - 0.0001f;
                                    fr = fr * abs_cos_theta(to_local(it.sh, ls.wi));
                                    float w = power_heuristic(ls.pdf, spdf);
                                    nee = fr * ls.li * w / ls.pdf;
                                    if constexpr (kAdj) r_nee = plastic_rho_ratio(m, rho, wo_d, wi_l);
                                }
                            }
                        }
                        // EstimateDirect, BSDF-sampling half (PathTracer.cs:351-397)
                        BSample bs = mat_sample<kM>(m, wo_d, rng);
                        f3 fr2 = bs.fr * abs_cos_theta(bs.wi);
                        float sp2 = bs.pdf;
                        if (!is_black(fr2) && sp2 > 0) {
                            float w2 = 1;
                            bool ok = true;
                            f3 wi_w = to_world(it.sh, bs.wi);
                            if (!(bs.type & kSpecular)) {
                                float lp = light_pdf(sc, Lt, it.P, wi_w);
                                if (lp == 0.0f) ok = false;
                                else w2 = power_heuristic(sp2, lp);
                            }
                            if (ok) {
                                want_mis = true;
                                mo = it.P; md = wi_w;
                                st3(ps.mis_f, pid, fr2, w2);
                                ps.mis_li[pid] = make_float4(0.f, 0.f, 0.f, sp2);
                                ps.sel_light[pid] = l;
                                if constexpr (kAdj) r_mis = plastic_rho_ratio(m, rho, wo_d, bs.wi);
                            }
                        }
                    } else {
                        // NextEventEstimation.EstimateLight (PathTracer.cs:175-189)
                        if (ls.pdf > 0.0f) {
                            f3 vec = ls.p - it.P;
                            float dis = length(vec);
                            if (!(dis < 0.001f)) {
                                want_shadow = true;
                                so = it.P; sd = normalize(vec); stmax = dis - 0.0001f;
                                f3 wi_l = to_local(it.sh, ls.wi);
                                f3 fr = mat_fr<kM>(m, wo_l, wi_l);
                                nee = fr * ls.li * abs_cos_theta(wi_l) / ls.pdf;
                                if constexpr (kAdj) r_nee = plastic_rho_ratio(m, rho, wo_l, wi_l);
                            }
                        }
                    }
                    nflags |= kFlagNeePending;
                    st3(ps.nee_beta, pid, beta, selpdf);
                    st3(ps.nee_le, pid, nee);
                } else if (in.method == 2) {
                    // SampleLight returned pdf 0: the vertex still adds coeff * 0
                    nflags |= kFlagNeePending;
                    st3(ps.nee_beta, pid, beta, 1.0f);
                    st3(ps.nee_le, pid, mk(0.0f));
                }
                if (want_shadow) nflags |= kFlagShadowIssued;
                if (want_mis) nflags |= kFlagMisIssued;
                // continuation (PathTracer.cs:227-268)
                BSample bs = mat_sample<kM>(m, wo_l, rng);
                bool sspec = (bs.type & kSpecular) != 0;
                if (bs.pdf > 0.0f) {
                    if constexpr (kAdj) r_cont = plastic_rho_ratio(m, rho, wo_l, bs.wi);
                    beta = beta * (bs.fr * abs_cos_theta(bs.wi) / bs.pdf);
                    ed = to_world(it.sh, bs.wi);
                    eo = it.P; etmin = 0.0001f; etmax = kFloatMax;
                    bool alive = true;
                    if constexpr (RD_MAT(kMatDisney) || RD_MAT(kMatSubsurface)) {
                        // the sample entered a BSSRDF (PathTracer.cs:244): the
                        // BSSRDF stage samples pi, lights it and continues the
                        // path (its Russian roulette included)
                        if (in.method == 2 && ps.sss_q && (bs.type & kSubsurface) && (bs.type & kTransmission)) {
                            want_sss = true;
                            alive = false;
                            ps.sss0[pid] = make_float4(it.P.x, it.P.y, it.P.z, __int_as_float(pr.material));
                            ps.sss1[pid] = make_float4(it.sh.Z.x, it.sh.Z.y, it.sh.Z.z, it.u);
                            ps.sss2[pid] = make_float4(it.wr.x, it.wr.y, it.wr.z, it.v);
                        }
                    }
                    if (!want_sss && depth > in.min_depth) {
                        float qq = nmin(max_elem(beta), in.rr_threshold);
                        if (rng.next_float() > qq) alive = false;
                        else beta = beta / qq;
                    }
                    if (alive) { want_ext = true; depth++; }
                }
                if (sspec) nflags |= kFlagSpecular;
                if constexpr (kAdj) {
                    const bool has_params = m.type == kMatDiffuse || m.type == kMatRoughPlastic;
                    nflags |= kFlagVertex;
                    ps.adj0[pid] = make_float4(m.r.x, m.r.y, m.r.z, __int_as_float(has_params ? pr.material : -1));
                    ps.adj1[pid] = make_float4(it.u, it.v, r_nee, r_mis);
                    ps.adj2[pid] = make_float4(r_cont, 0.f, 0.f, 0.f);
                }
            }
            st3(ps.L, pid, L);
            st3(ps.beta, pid, beta);
            if (!want_ext && !want_sss) nflags |= kFlagDone;
            ps.flags[pid] = nflags;
            ps.rng_n[pid] = rng.n;
            ps.depth[pid] = depth;
            };
            if constexpr (kBatch) run(scenes[ps.req[pid]]);
            else run(sc0);
            // the continuation ray waits in the slot for the next compaction
            ps.alive[pid] = want_ext ? 1 : 0;
            if (want_ext) {
                ps.ray_o[pid] = make_float4(eo.x, eo.y, eo.z, etmin);
                ps.ray_d[pid] = make_float4(ed.x, ed.y, ed.z, etmax);
            }
        }
        if (rank || rank_identity) {
            // ranked: each path's shadow ray sits at its rank in the coherence
            // order of the extension queue (its own position when that queue
            // is traced in queue order), holes marked pid -1
            if (k < n) {
                const int r = rank ? rank[k] : k;
                qshadow.pid[r] = want_shadow ? pid : -1;
                if (want_shadow) {
                    qshadow.org[r] = make_float4(so.x, so.y, so.z, 0.0001f);
                    qshadow.dir[r] = make_float4(sd.x, sd.y, sd.z, stmax);
                }
            }
        } else {
            int slot = block_append<256>(&counts[kCntShadow], want_shadow, sh_app);
            if (want_shadow) {
                qshadow.pid[slot] = pid;
                qshadow.org[slot] = make_float4(so.x, so.y, so.z, 0.0001f);
                qshadow.dir[slot] = make_float4(sd.x, sd.y, sd.z, stmax);
            }
        }
        // MIS rays (only those that can reach the selected light) are few:
        // appended, so their queue holds no holes
        int slot = block_append<256>(&counts[kCntMis], want_mis, sh_app);
        if (want_mis) {
            qmis.pid[slot] = pid;
            qmis.org[slot] = make_float4(mo.x, mo.y, mo.z, 0.0001f);
            qmis.dir[slot] = make_float4(md.x, md.y, md.z, kFloatMax);
        }
        if constexpr (RD_MAT(kMatDisney) || RD_MAT(kMatSubsurface)) {
            if (ps.sss_q) {
                const int ss = block_append<256>(&counts[kCntSss], want_sss, sh_app);
                if (want_sss) ps.sss_q[ss] = pid;
            }
        }
    }
    if ((rank || rank_identity) && blockIdx.x == 0 && threadIdx.x == 0) counts[kCntShadow] = n;
    add_counter((unsigned long long*)(counts + kCntVertices), nvert);
}

hipError_t launch_shade(const DevScene& sc, const DevScene* scenes, RayQueue qin, const float4* hits,
                        RayQueue qshadow, RayQueue qmis, PathState ps, int* counts, int cnt_in, bool adjoint,
                        const int* rank, bool rank_identity, hipStream_t s) {
    const dim3 g(persistent_blocks()), b(256);
    if (adjoint) {
        hipLaunchKernelGGL((k_shade<false, true, kMatsAll>), g, b, 0, s, sc, scenes, qin, hits, qshadow, qmis, ps,
                           counts, cnt_in, rank, rank_identity);
        return hipGetLastError();
    }
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, g, b, 0, s, sc, scenes, qin, hits, qshadow, qmis, ps, counts, cnt_in, rank,
                           rank_identity);
    };
    const unsigned set = mat_set(sc.mat_mask);
    if (scenes) {
        if (set == kMatsDelta) go(k_shade<true, false, kMatsDelta>);
        else if (set == kMatsMicrofacet) go(k_shade<true, false, kMatsMicrofacet>);
        else go(k_shade<true, false, kMatsAll>);
    } else {
        if (set == kMatsDelta) go(k_shade<false, false, kMatsDelta>);
        else if (set == kMatsMicrofacet) go(k_shade<false, false, kMatsMicrofacet>);
        else go(k_shade<false, false, kMatsAll>);
    }
    return hipGetLastError();
}

// radiance += coeff * SampleLightToEstimateDirect (PathTracer.cs:226, :273-307)
// for Mis; radiance += coeff * EstimateLight / lightPdf (:81) for Nee.  A
// path that ended this bounce then retires its radiance into the pass's
// per-sample result (its slot is free for the next compaction).
//
// kAdj (path-replay backpropagation): the path's full radiance Ltot comes
// from the primal replay; after this vertex's terms, Lr = Ltot - L is what the
// rest of the path adds, all of it through the continuation's BSDF factor.
// With delta = dLoss/dI(pixel) / spp and the vertex's terms C_nee, C_mis:
//   dLoss/dR_c   = delta_c (C_nee_c + C_mis_c + Lr_c) / R_c
//   dLoss/drho   = sum_c delta_c (C_nee_c r_nee + C_mis_c r_mis + Lr_c r_cont)
// scattered onto the texels of the albedo / roughness lookups.
template <bool kAdj>
__global__ __launch_bounds__(256) void k_resolve(DevScene sc, int method, RayQueue qin, PathState ps, int* counts,
                                                 int cnt_in, AdjArgs aa) {
    int n = counts[cnt_in];
    for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
        int pid = qin.pid[k];
        int flags = ps.flags[pid];
        bool done = (flags & kFlagDone) != 0;
        f3 L = ld3(ps.L, pid);
        f3 c_nee = mk(0.0f), c_mis = mk(0.0f);
        if (flags & kFlagNeePending) {
            float4 nb = ps.nee_beta[pid];
            f3 beta = mk(nb.x, nb.y, nb.z);
            float selpdf = nb.w;
            bool vis = (flags & kFlagShadowIssued) && ps.vis[pid];
            if (method == 2) {
                f3 le = mk(0.0f), le_mis = mk(0.0f);
                if (vis) le = le + ld3(ps.nee_le, pid);
                if (flags & kFlagMisIssued) {
                    float4 li4 = ps.mis_li[pid];
                    f3 li = mk(li4.x, li4.y, li4.z);
                    if (!is_black(li)) {
                        float4 f4 = ps.mis_f[pid];
                        le_mis = mk(f4.x, f4.y, f4.z) * li * f4.w / li4.w;
                        le = le + le_mis;
                    }
                }
                f3 result = mk(0.0f) + le / selpdf;
                L = L + beta * result;
                if constexpr (kAdj) {
                    if (vis) c_nee = beta * (ld3(ps.nee_le, pid) / selpdf);
                    c_mis = beta * (le_mis / selpdf);
                }
            } else {
                f3 e = vis ? ld3(ps.nee_le, pid) : mk(0.0f);
                L = L + beta * e / selpdf;
                if constexpr (kAdj) c_nee = beta * e / selpdf;
            }
            if constexpr (!kAdj) st3(ps.L, pid, L);
        }
        if constexpr (kAdj) {
            if (flags & kFlagNeePending) st3(ps.L, pid, L);
            const float4 a0 = ps.adj0[pid];
            const int mat = __float_as_int(a0.w);
            const f3 ltot = ld3(ps.res, ps.gid[pid]);
            if ((flags & kFlagVertex) && mat >= 0 && valid_color(ltot)) {
                const f3 lr = ltot - L;
                const float4 a1 = ps.adj1[pid];
                const float r_cont = ps.adj2[pid].x;
                const int pix = ps.pix[pid];
                const f3 delta = mk(aa.dl[3 * pix], aa.dl[3 * pix + 1], aa.dl[3 * pix + 2]) * aa.inv_spp;
                const f3 t = c_nee + c_mis + lr;
                const f3 gr = mk(a0.x != 0 ? delta.x * t.x / a0.x : 0.0f, a0.y != 0 ? delta.y * t.y / a0.y : 0.0f,
                                 a0.z != 0 ? delta.z * t.z / a0.z : 0.0f);
                const DevMaterial& dm = sc.mats[mat];
                double* g0 = aa.grads[dm.tex[0]];
                if (g0) {
                    TexFoot f = tex_footprint(sc, dm.tex[0], a1.x, a1.y);
                    for (int i = 0; i < f.n; i++) {
                        atomicAdd(g0 + 3 * f.idx[i], (double)(f.w[i] * gr.x));
                        atomicAdd(g0 + 3 * f.idx[i] + 1, (double)(f.w[i] * gr.y));
                        atomicAdd(g0 + 3 * f.idx[i] + 2, (double)(f.w[i] * gr.z));
                    }
                }
                if (dm.type == kMatRoughPlastic) {
                    double* g1 = aa.grads[dm.tex[1]];
                    const float grho = dot(delta, c_nee * a1.z + c_mis * a1.w + lr * r_cont);
                    if (g1 && grho != 0.0f) {
                        TexFoot f = tex_footprint(sc, dm.tex[1], a1.x, a1.y);
                        for (int i = 0; i < f.n; i++) atomicAdd(g1 + 3 * f.idx[i], (double)(f.w[i] * grho));
                    }
                }
            }
        } else {
            if (done) ps.res[ps.gid[pid]] = make_float4(L.x, L.y, L.z, 0.f);
        }
    }
}

hipError_t launch_resolve(const DevScene& sc, int method, RayQueue qin, PathState ps, int* counts, int cnt_in,
                          const AdjArgs* adj, hipStream_t s) {
    if (adj)
        hipLaunchKernelGGL(k_resolve<true>, dim3(persistent_blocks()), dim3(256), 0, s, sc, method, qin, ps, counts,
                           cnt_in, *adj);
    else
        hipLaunchKernelGGL(k_resolve<false>, dim3(persistent_blocks()), dim3(256), 0, s, sc, method, qin, ps, counts,
                           cnt_in, AdjArgs{});
    return hipGetLastError();
}

// Pass accumulation: for every segment of the pass, each of its pixels adds
// the segment's samples in sample order to its request's sums
// (BlockBasedRenderer.cs:140-146: invalid samples count as 0).  Sample
// radiance is float32; the per-pixel sum is float64 (the reference sums in
// float32: here only the final per-pixel mean is kept in float64, so a
// pixel's mean does not round at the sum's ulp), the sum of squares float32.
__global__ __launch_bounds__(256) void k_accumulate_batch(const BatchSeg* segs, int nseg, PathState ps, double4* sum,
                                                          float4* sum2, int* counts) {
    uint64_t invalid = 0;
    for (int g = blockIdx.y; g < nseg; g += gridDim.y) {
        const BatchSeg sg = segs[g];
        const int h = sg.npix / sg.crop_w;
        for (int q = blockIdx.x * 256 + threadIdx.x; q < sg.npix; q += gridDim.x * 256) {
            int o = sg.sum_offset + q;
            double4 s4 = sum[o];
            float4 q4 = sum2[o];
            double sx = s4.x, sy = s4.y, sz = s4.z;
            f3 s2 = mk(q4.x, q4.y, q4.z);
            const int p0 = sg.path_offset + seg_local(q, 0, sg.crop_w, h, sg.nsamples, sg.tile);
            const int step = seg_local(q, 1, sg.crop_w, h, sg.nsamples, sg.tile) - seg_local(q, 0, sg.crop_w, h, sg.nsamples, sg.tile);
            for (int j = 0; j < sg.nsamples; j++) {
                f3 c = ld3(ps.res, p0 + j * step);
                if (!valid_color(c)) { c = mk(0.0f); invalid++; }
                sx += (double)c.x;
                sy += (double)c.y;
                sz += (double)c.z;
                s2 = s2 + c * c;
            }
            sum[o] = make_double4(sx, sy, sz, 0.0);
            sum2[o] = make_float4(s2.x, s2.y, s2.z, 0.f);
        }
    }
    add_counter((unsigned long long*)(counts + kCntInvalid), invalid);
}

hipError_t launch_accumulate_batch(const BatchSeg* segs, int nseg, int max_npix, PathState ps, double4* sum,
                                   float4* sum2, int* counts, hipStream_t s) {
    int bx = (max_npix + 255) / 256;
    if (bx > 1024) bx = 1024;
    int by = nseg < 64 ? nseg : 64;
    hipLaunchKernelGGL(k_accumulate_batch, dim3(bx, by), dim3(256), 0, s, segs, nseg, ps, sum, sum2, counts);
    return hipGetLastError();
}

// sum / spp into the column-major ColorBuffer layout (x*H + y)*3
// (ColorBuffer.cs:35, BlockBasedRenderer.cs:167-178): the mean in float64,
// delivered as float32 (out) or float64 (out64).
template <class T>
__global__ __launch_bounds__(256) void k_finalize(WaveParams wp, int height, int spp, const double4* sum,
                                                  const float4* sum2, T* out, float* out2) {
    for (int q = blockIdx.x * 256 + threadIdx.x; q < wp.npix; q += gridDim.x * 256) {
        int tx, ty;
        tile_pixel(q, wp.crop_w, wp.crop_h, wp.tile, &tx, &ty);
        int x = wp.x0 + tx, y = wp.y0 + ty;
        size_t o = ((size_t)x * height + y) * 3;
        double4 s = sum[q];
        double ds = (double)spp;
        out[o] = (T)(s.x / ds); out[o + 1] = (T)(s.y / ds); out[o + 2] = (T)(s.z / ds);
        if (out2) {
            float4 s2 = sum2[q];
            out2[o] = s2.x; out2[o + 1] = s2.y; out2[o + 2] = s2.z;
        }
    }
}

hipError_t launch_finalize(const WaveParams& wp, int height, int spp, const double4* sum, const float4* sum2,
                           float* out_rgb, float* out_sum2, double* out64, hipStream_t s) {
    int blocks = (wp.npix + 255) / 256;
    if (blocks > persistent_blocks()) blocks = persistent_blocks();
    if (out64)
        hipLaunchKernelGGL(k_finalize<double>, dim3(blocks), dim3(256), 0, s, wp, height, spp, sum, sum2, out64,
                           out_sum2);
    else
        hipLaunchKernelGGL(k_finalize<float>, dim3(blocks), dim3(256), 0, s, wp, height, spp, sum, sum2, out_rgb,
                           out_sum2);
    return hipGetLastError();
}

__global__ void k_add_f64_to_f32(const double* acc, float* out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        out[i] += (float)acc[i];
}
hipError_t launch_add_f64_to_f32(const double* acc, float* out, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > persistent_blocks()) blocks = persistent_blocks();
    hipLaunchKernelGGL(k_add_f64_to_f32, dim3((unsigned)blocks), dim3(256), 0, s, acc, out, n);
    return hipGetLastError();
}

__global__ void k_zero_d4(double4* b, int n) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) b[i] = make_double4(0.0, 0.0, 0.0, 0.0);
}
hipError_t launch_zero(double4* buf, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int blocks = (n + 255) / 256;
    if (blocks > persistent_blocks()) blocks = persistent_blocks();
    hipLaunchKernelGGL(k_zero_d4, dim3(blocks), dim3(256), 0, s, buf, n);
    return hipGetLastError();
}
__global__ void k_zero4(float4* b, int n) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) b[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}
hipError_t launch_zero(float4* buf, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int blocks = (n + 255) / 256;
    if (blocks > persistent_blocks()) blocks = persistent_blocks();
    hipLaunchKernelGGL(k_zero4, dim3(blocks), dim3(256), 0, s, buf, n);
    return hipGetLastError();
}
// ---------------------------------------------------------------- BSSRDF
// PathTracer.Mis's subsurface branch (PathTracer.cs:244-259) for the paths
// k_shade queued in ps.sss_q: k_sss_probe samples pi (SeparableBssrdf.
// SamplePi / PdfPi, Bssrdf.cs:139-200, with the normalised-diffusion profile
// of :64-124), k_sss_shade lights pi through the BssrdfAdapter
// (Bssrdf.cs:203-243) and continues the path with the adapter's sample.

// The BSSRDF's profile and eta at po: a Disney material's
// (BaseColor * (1 - Metallic), ScattingDistance, Eta) (DisneyBsdf.GetBssrdf,
// DisneyBsdf.cs:514-518), a subsurface material's (A, ScattingDistance,
// EtaB / EtaA) (Subsurface.cs:57-64); D = d / (100 (A - 0.33)^4 + 3.5).
struct SssProfile {
    f3 A, D;
    float eta;
};
__device__ __forceinline__ SssProfile sss_profile(const DevScene& sc, int mat, float u, float v) {
    const DevMaterial& dm = sc.mats[mat];
    SssProfile p;
    f3 sd;
    if (dm.type == kMatDisney) {
        const f3 col = tex_sample(sc, dm.tex[0], u, v);
        const float met = tex_sample(sc, dm.tex[1], u, v).x;
        p.A = col * (1 - met);
        sd = tex_sample(sc, dm.tex[11], u, v);
        p.eta = tex_sample(sc, dm.tex[3], u, v).x;
    } else {
        p.A = tex_sample(sc, dm.tex[2], u, v);
        sd = tex_sample(sc, dm.tex[3], u, v);
        p.eta = dm.eta_b / dm.eta_a;
    }
    const f3 t = p.A - mk(0.33f);
    const f3 s = t * t * t * t * 100.0f + mk(3.5f);
    p.D = mk(sd.x / s.x, sd.y / s.y, sd.z / s.z);
    return p;
}
__device__ __forceinline__ float sss_ch(f3 v, int c) { return c == 0 ? v.x : (c == 1 ? v.y : v.z); }
// NormalizedDiffusionRadialProfile.PdfSr
__device__ __forceinline__ float sss_pdf_sr(const SssProfile& p, int c, float distance) {
    const float chd = sss_ch(p.D, c);
    distance /= chd;
    return 0.25f * (expf(-distance) + expf(-distance / 3)) / chd;
}
// NormalizedDiffusionRadialProfile.Sr
__device__ __forceinline__ f3 sss_sr(const SssProfile& p, float distance) {
    auto one = [&](float a, float d) { return a * (expf(-distance / d) + expf(-distance / (3.0f * d))) / (8 * kPi * d); };
    return mk(one(p.A.x, p.D.x), one(p.A.y, p.D.y), one(p.A.z, p.D.z));
}
// NormalizedDiffusionRadialProfile.SampleSr: the radius from the 2048-entry table
__device__ __forceinline__ float sss_sample_sr(const SssProfile& p, int c, float u, const float* __restrict__ lut) {
    const float g = u * 2047.0f;
    const int low = min((int)g, 2046);
    const float local = g - low;
    return lerpf(local, lut[low], lut[low + 1]) * sss_ch(p.D, c);
}

// BssrdfUtility.FresnelMoment and the adapter lobe
__device__ __forceinline__ float fresnel_moment(float eta) {
    const float eta2 = eta * eta, eta3 = eta2 * eta, eta4 = eta2 * eta2, eta5 = eta2 * eta3;
    if (eta < 1)
        return 0.45966f - 1.73965f * eta + 3.37668f * eta2 - 3.904945f * eta3 + 2.49277f * eta4 - 0.68441f * eta5;
    return -4.61686f + 11.1136f * eta - 10.4646f * eta2 + 5.11455f * eta3 - 1.27198f * eta4 + 0.12746f * eta5;
}
__device__ __forceinline__ f3 adapter_fr(float eta, f3 wi) {
    const float ci = 1 - 2 * fresnel_moment(1 / eta);
    const float fr = fr_dielectric(cos_theta(wi), 1, eta);
    const f3 f = mk((1 - fr) / (ci * kPi));
    return f * (eta * eta);
}
__device__ __forceinline__ float adapter_pdf(f3 wo, f3 wi) {
    return same_hemisphere(wo, wi) ? abs_cos_theta(wi) / kPi : 0.0f;
}

// One closest-hit query of the probe chain over the scene's accelerator:
// kKind 0 the BVH, 1 the two-level tree, 2 the octree.  Returns the hit
// {t, prim, u, v} (prim -1: a miss) and, for a set triangle, *sub.
template <int kKind>
__device__ __forceinline__ float4 sss_closest(const DevScene& sc, float4 ra, float4 rb, int* stk, int* sub,
                                              uint32_t& nv, uint32_t& nt) {
    *sub = -1;
    const float4 miss = make_float4(0.f, __int_as_float(-1), 0.f, 0.f);
    if (sc.num_nodes == 0) return miss;
    if constexpr (kKind == 1) {
        TravI ti;
        travi_begin(ti, ra, rb);
        while (!trav_step_inst<false, 128>(sc, sc.nodes, sc.prims, ti, stk, nv, nt)) {}
        const Trav& t = ti.t;
        if (t.best.prim < 0) return miss;
        *sub = ti.sub;
        return make_float4(t.best.t, __int_as_float(t.best.prim), t.best.u, t.best.v);
    } else {
        Trav t;
        trav_begin(t, ra, rb);
        if constexpr (kKind == 2) oct_trace<false, 256>(sc.nodes, sc.prims, sc.num_nodes, t, nv, nt);
        else while (!trav_step<false, 128>(sc.nodes, sc.prims, t, stk, nv, nt)) {}
        if (t.best.prim < 0) return miss;
        return make_float4(t.best.t, __int_as_float(t.best.prim), t.best.u, t.best.v);
    }
}

// SeparableBssrdf.SamplePi (Bssrdf.cs:139-183) for each path of ps.sss_q:
// channel, projection axis, radius and angle from the path's stream, the
// probe chain along the axis collecting the hits on po's material (walked
// once to count them, then again to the one Next(count) picks: the same
// deterministic queries), PdfPi (:185-200), S = Sr(|po - pi|) (1 - F).  A
// path with a point takes beta *= S / pdf and goes to ps.sss_q2; the others
// end here (their radiance is final).  Also clears the shadow / MIS queue
// counters k_sss_shade appends to.
template <int kKind>
__global__ __launch_bounds__(128) void k_sss_probe(DevScene sc, PathState ps, int* counts,
                                                   const float* __restrict__ lut) {
    __shared__ int lds[64 * 128];
    int* stk = lds + threadIdx.x;
    const int n = counts[kCntSss];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counts[kCntShadow] = 0;
        counts[kCntMis] = 0;
    }
    uint32_t nv = 0, nt = 0, nq = 0;
    for (int base = blockIdx.x * 128; base < n; base += gridDim.x * 128) {
        const int k = base + threadIdx.x;
        bool ok = false;
        int pid = -1;
        if (k < n) {
            pid = ps.sss_q[k];
            const float4 a0 = ps.sss0[pid], a1 = ps.sss1[pid], a2 = ps.sss2[pid];
            const f3 P = mk(a0.x, a0.y, a0.z);
            const int mat = __float_as_int(a0.w);
            const Frame fr = make_frame(mk(a1.x, a1.y, a1.z));
            const f3 wr = mk(a2.x, a2.y, a2.z);
            const SssProfile prof = sss_profile(sc, mat, a1.w, a2.w);
            Rng rng{ps.key[pid], ps.rng_n[pid]};
            const int ch = rng.next_int(3);
            const float uax = rng.next_float();
            Frame pj = fr;
            if (uax < 0.5f) {
            } else if (uax < 0.75f) {
                pj.X = fr.X; pj.Y = fr.Z; pj.Z = fr.Y;   // Coordinate(X, Z, Y)
            } else {
                pj.X = fr.Z; pj.Y = fr.Y; pj.Z = fr.X;   // Coordinate(Z, Y, X)
            }
            const float r = sss_sample_sr(prof, ch, rng.next_float(), lut);
            const float r_max = sss_sample_sr(prof, ch, 0.996f, lut);
            if (!(r <= 0 || r > r_max)) {
                const float phi = 2 * kPi * rng.next_float();
                const float hl = sqrtf(nmax(0, sqr(r_max) - sqr(r)));
                const float sp = sinf(phi), cp = cosf(phi);
                const f3 o0 = P + to_world(pj, mk(r * cp, r * sp, hl));
                const f3 d = -pj.Z;
                const float len = 2 * hl;
                // the chain: a query, the hit counted when on po's material,
                // the origin moved 1e-4 past it and MaxT shortened to match.
                // The counting walk keeps its first two hits, so the select
                // walk (the same queries again, up to the chosen hit) runs
                // only for a later one.
                float4 c_h[2];
                int c_sub[2];
                f3 c_org[2];
                auto walk = [&](int stop, float4* hit, int* hsub, f3* horg, bool count_queries) -> int {
                    f3 o = o0;
                    float tmax = len;
                    int found = 0;
                    for (int step = 0; step < kMaxChainSteps; step++) {
                        if (0.0001f >= tmax) return found;
                        int sub;
                        const float4 h = sss_closest<kKind>(sc, make_float4(o.x, o.y, o.z, 0.0001f),
                                                            make_float4(d.x, d.y, d.z, tmax), stk, &sub, nv, nt);
                        if (count_queries) nq++;
                        const int prim = __float_as_int(h.y);
                        if (prim < 0) return found;
                        if (sc.prim_info[prim].material == mat) {
                            if (found == stop) { *hit = h; *hsub = sub; *horg = o; return found + 1; }
                            if (stop < 0 && found == 0) { c_h[0] = h; c_sub[0] = sub; c_org[0] = o; }
                            if (stop < 0 && found == 1) { c_h[1] = h; c_sub[1] = sub; c_org[1] = o; }
                            found++;
                        }
                        const float adv = h.x + 0.0001f;
                        o = o + d * adv;
                        tmax -= adv;
                    }
                    atomicOr(counts + kCntChainStuck, 1);
                    return found;
                };
                float4 h = make_float4(0.f, 0.f, 0.f, 0.f);
                int hsub = -1;
                f3 horg = o0;
                const int count = walk(-1, &h, &hsub, &horg, true);
                if (count > 0) {
                    const int sel = rng.next_int(count);
                    if (sel == 0) { h = c_h[0]; hsub = c_sub[0]; horg = c_org[0]; }
                    else if (sel == 1) { h = c_h[1]; hsub = c_sub[1]; horg = c_org[1]; }
                    else walk(sel, &h, &hsub, &horg, false);
                    const DevPrim pr = sc.prim_info[__float_as_int(h.y)];
                    const Inct pi = make_inct_any(sc, pr, pr.kind == kInst ? hsub : -1, horg, d, h.z, h.w, h.x);
                    // PdfPi
                    const f3 ld = to_local(fr, P - pi.P);
                    const f3 ln = to_local(fr, pi.sh.Z);
                    const float rproj[3] = {sqrtf(ld.y * ld.y + ld.z * ld.z), sqrtf(ld.z * ld.z + ld.x * ld.x),
                                            sqrtf(ld.x * ld.x + ld.y * ld.y)};
                    const float axis_pdf[3] = {0.25f, 0.25f, 0.5f};
                    float pdf = 0;
                    for (int axis = 0; axis < 3; axis++) {
                        const float lna = fabsf(axis == 0 ? ln.x : (axis == 1 ? ln.y : ln.z));
                        for (int c = 0; c < 3; c++) pdf += lna * sss_pdf_sr(prof, c, rproj[axis]) * axis_pdf[axis];
                    }
                    pdf = pdf * (1.0f / 3.0f);
                    pdf = pdf / (2 * kPi) / (float)count;
                    const float fro = 1 - fr_dielectric(cos_theta(to_local(fr, wr)), 1, prof.eta);
                    const f3 S = sss_sr(prof, length(P - pi.P)) * fro;
                    if (!(pdf == 0 || is_black(S))) {
                        ok = true;
                        st3(ps.beta, pid, ld3(ps.beta, pid) * (S / pdf));
                        ps.sss0[pid] = make_float4(pi.P.x, pi.P.y, pi.P.z, __int_as_float(mat));
                        ps.sss1[pid] = make_float4(pi.sh.Z.x, pi.sh.Z.y, pi.sh.Z.z, pi.u);
                        ps.sss2[pid] = make_float4(pi.wr.x, pi.wr.y, pi.wr.z, pi.v);
                    }
                }
            }
            ps.rng_n[pid] = rng.n;
            if (!ok) {
                // the path ends at po: its radiance (po's light included) is final
                ps.flags[pid] = kFlagDone;
                ps.alive[pid] = 0;
                ps.res[ps.gid[pid]] = ps.L[pid];
            }
        }
        const int slot = wave_append(&counts[kCntSssB], ok);
        if (ok) ps.sss_q2[slot] = pid;
    }
    add_class_counters(counts, 0, nq, nv, nt);
    add_counter((unsigned long long*)(counts + kCntRaysClosest), nq);
}

// Direct light at pi through the BssrdfAdapter and the adapter's continuation
// (PathTracer.cs:252-258, EstimateDirect :309-389 with bssrdfAdapter): the
// same light pick, light sample, shadow / MIS rays and Russian roulette as
// k_shade's Mis vertex, with the adapter's Fr / Pdf / Sample at pi (its eta
// sampled at pi's UV for a Disney material).  Resets the traversal work
// cursors for the shadow / MIS launch that follows.
__global__ __launch_bounds__(256) void k_sss_shade(DevScene sc, RayQueue qshadow, RayQueue qmis, PathState ps,
                                                   int* counts) {
    const int n = counts[kCntSssB];
    if (blockIdx.x == 0 && threadIdx.x < 3) counts[kCntWork + threadIdx.x] = 0;
    const DevIntegrator& in = sc.integ;
    for (int base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {
        const int k = base + threadIdx.x;
        bool want_shadow = false, want_mis = false;
        f3 so = mk(0.f), sd = mk(0.f), mo = mk(0.f), md = mk(0.f);
        float stmax = 0.f;
        int pid = 0;
        if (k < n) {
            pid = ps.sss_q2[k];
            const float4 a0 = ps.sss0[pid], a1 = ps.sss1[pid], a2 = ps.sss2[pid];
            const f3 P = mk(a0.x, a0.y, a0.z);
            const DevMaterial& dm = sc.mats[__float_as_int(a0.w)];
            const float eta = dm.type == kMatDisney ? tex_sample(sc, dm.tex[3], a1.w, a2.w).x : dm.eta_b / dm.eta_a;
            const Frame fr = make_frame(mk(a1.x, a1.y, a1.z));
            const f3 wo_d = to_local(fr, mk(a2.x, a2.y, a2.z));
            Rng rng{ps.key[pid], ps.rng_n[pid]};
            f3 beta = ld3(ps.beta, pid);
            int depth = ps.depth[pid];
            int nflags = 0;
            if (sc.num_lights > 0) {
                const int l = rng.next_int(sc.num_lights);
                const float selpdf = 1.0f / (float)sc.num_lights;
                const DevLight Lt = sc.lights[l];
                const LSample ls = light_sample(sc, Lt, P, rng);
                f3 nee = mk(0.0f);
                if (ls.pdf > 0.0f && !is_black(ls.li)) {
                    const f3 wi_l = to_local(fr, ls.wi);
                    f3 f = adapter_fr(eta, wi_l);
                    const float spdf = adapter_pdf(wo_d, wi_l);
                    if (!is_black(f)) {
                        const f3 vec = ls.p - P;
                        const float dis = length(vec);
                        if (!(dis < 0.001f)) {
                            want_shadow = true;
                            so = P; sd = normalize(vec); stmax = dis - 0.0001f;
                            f = f * abs_cos_theta(to_local(fr, ls.wi));
                            const float w = power_heuristic(ls.pdf, spdf);
                            nee = f * ls.li * w / ls.pdf;
                        }
                    }
                }
                // the adapter's BSDF-sampling half
                f3 wi2;
                if (cosine_dir(wo_d, rng, &wi2)) {
                    const f3 fr2 = adapter_fr(eta, wi2) * abs_cos_theta(wi2);
                    const float sp2 = adapter_pdf(wo_d, wi2);
                    if (!is_black(fr2) && sp2 > 0) {
                        const f3 wi_w = to_world(fr, wi2);
                        const float lp = light_pdf(sc, Lt, P, wi_w);
                        if (lp != 0.0f) {
                            want_mis = true;
                            mo = P; md = wi_w;
                            st3(ps.mis_f, pid, fr2, power_heuristic(sp2, lp));
                            ps.mis_li[pid] = make_float4(0.f, 0.f, 0.f, sp2);
                            ps.sel_light[pid] = l;
                        }
                    }
                }
                nflags |= kFlagNeePending;
                st3(ps.nee_beta, pid, beta, selpdf);
                st3(ps.nee_le, pid, nee);
            } else {
                nflags |= kFlagNeePending;
                st3(ps.nee_beta, pid, beta, 1.0f);
                st3(ps.nee_le, pid, mk(0.0f));
            }
            if (want_shadow) nflags |= kFlagShadowIssued;
            if (want_mis) nflags |= kFlagMisIssued;
            // the adapter's sample continues the path from pi
            bool alive = false;
            f3 wi;
            if (cosine_dir(wo_d, rng, &wi)) {
                const f3 f = adapter_fr(eta, wi);
                const float pdf = adapter_pdf(wo_d, wi);
                if (!(pdf == 0 || is_black(f))) {
                    beta = beta * (f * abs_cos_theta(wi) / pdf);
                    alive = true;
                    if (depth > in.min_depth) {
                        const float qq = nmin(max_elem(beta), in.rr_threshold);
                        if (rng.next_float() > qq) alive = false;
                        else beta = beta / qq;
                    }
                    if (alive) {
                        depth++;
                        const f3 ed = to_world(fr, wi);
                        ps.ray_o[pid] = make_float4(P.x, P.y, P.z, 0.0001f);
                        ps.ray_d[pid] = make_float4(ed.x, ed.y, ed.z, kFloatMax);
                    }
                }
            }
            if (!alive) nflags |= kFlagDone;
            st3(ps.beta, pid, beta);
            ps.flags[pid] = nflags;
            ps.rng_n[pid] = rng.n;
            ps.depth[pid] = depth;
            ps.alive[pid] = alive ? 1 : 0;
        }
        int slot = wave_append(&counts[kCntShadow], want_shadow);
        if (want_shadow) {
            qshadow.pid[slot] = pid;
            qshadow.org[slot] = make_float4(so.x, so.y, so.z, 0.0001f);
            qshadow.dir[slot] = make_float4(sd.x, sd.y, sd.z, stmax);
        }
        slot = wave_append(&counts[kCntMis], want_mis);
        if (want_mis) {
            qmis.pid[slot] = pid;
            qmis.org[slot] = make_float4(mo.x, mo.y, mo.z, 0.0001f);
            qmis.dir[slot] = make_float4(md.x, md.y, md.z, kFloatMax);
        }
    }
}

hipError_t launch_sss_probe(const DevScene& sc, PathState ps, int* counts, const float* lut, hipStream_t s) {
    const dim3 g(2 * persistent_blocks()), b(128);
    if (sc.octree) hipLaunchKernelGGL(k_sss_probe<2>, g, b, 0, s, sc, ps, counts, lut);
    else if (sc.num_insts > 0) hipLaunchKernelGGL(k_sss_probe<1>, g, b, 0, s, sc, ps, counts, lut);
    else hipLaunchKernelGGL(k_sss_probe<0>, g, b, 0, s, sc, ps, counts, lut);
    return hipGetLastError();
}

hipError_t launch_sss_shade(const DevScene& sc, RayQueue qshadow, RayQueue qmis, PathState ps, int* counts,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_sss_shade, dim3(persistent_blocks()), dim3(256), 0, s, sc, qshadow, qmis, ps, counts);
    return hipGetLastError();
}

__global__ void k_zero_counts(int* c, int n) {
    int i = threadIdx.x;
    if (i < n) c[i] = 0;
}
hipError_t launch_zero_counts(int* counts, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(64), 0, s, counts, n);
    return hipGetLastError();
}

}  // namespace rd
