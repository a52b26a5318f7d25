// rd_api.cpp — the C ABI (include/radiantium_hip.h): context, scene upload,
// traversal hooks and the wavefront render loop that replaces
// BlockBasedRenderer (Radiantium.Offline/Renderers/BlockBasedRenderer.cs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/radiantium_hip.h"
#include "rd_bvh_build.h"
#include "rd_camera.h"
#include "rd_octree.h"
#include "rd_envmap.h"
#include "rd_kernels.h"
#include "rd_scratch.h"

using namespace rd;

static thread_local std::string g_err;
static int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
namespace rd {
void set_last_error(const std::string& msg) { g_err = msg; }
}
#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return fail(RD_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

namespace {

struct StageTiming {
    double total_ms = 0;
    int64_t count = 0;
};

// Device buffers of the wavefront, reused across renders and grown on demand.
// Most segments one pass holds (the host-built segment table).
constexpr int kMaxSegs = 1024;

struct WaveBuffers {
    size_t cap = 0;        // path pool slots
    size_t res_cap = 0;    // per-sample results of one pass
    size_t pix_cap = 0;    // accumulation pixels (all requests of a render)
    size_t scene_cap = 0;  // batched: device scene table entries
    bool vol = false;
    std::vector<void*> pool_allocs, res_allocs, pix_allocs, scene_allocs, seg_allocs;
    RayQueue qa{}, qb{}, qshadow{}, qmis{};
    RayQueue qshadow2{}, qmis2{};   // vol_path chain queues (second buffers)
    float4* hits = nullptr;
    float4* hits2 = nullptr;        // vol_path: IntersectTr chain hits
    PathState ps{};
    double4* sum = nullptr;   // float64 per-pixel sums
    float4* sum2 = nullptr;
    DevScene* scenes = nullptr;
    BatchSeg* segs = nullptr;       // device segment table
    BatchSeg* host_segs = nullptr;  // pinned staging for it
    int2* blk = nullptr;            // pool compaction scratch
    int2* blk2 = nullptr;           // shadow-queue compaction scratch (second stream)
    size_t adj_cap = 0;             // adjoint per-slot arrays
    std::vector<void*> adj_allocs;
    void** grad_ptrs = nullptr;     // adjoint: device table of gradient accumulators
    size_t grad_cap = 0;
    // extension-queue coherence sort
    size_t sort_cap = 0;
    std::vector<void*> sort_allocs;
    uint32_t* sort_keys = nullptr;
    uint32_t* sort_keys_alt = nullptr;
    int* sort_vals = nullptr;
    int* sort_perm = nullptr;
    int* sort_rank = nullptr;       // queue position -> place in the traced order
    void* sort_temp = nullptr;
    size_t sort_temp_bytes = 0;
    // BSSRDF queues and points (path Mis over subsurface materials)
    size_t sss_cap = 0;
    std::vector<void*> sss_allocs;
    int* sss_q = nullptr;
    int* sss_q2 = nullptr;
    float4* sss0 = nullptr;
    float4* sss1 = nullptr;
    float4* sss2 = nullptr;
    static void free_list(std::vector<void*>& v) {
        for (void* p : v) (void)hipFree(p);
        v.clear();
    }
    void release() {
        free_list(pool_allocs);
        free_list(res_allocs);
        free_list(pix_allocs);
        free_list(scene_allocs);
        free_list(seg_allocs);
        free_list(adj_allocs);
        free_list(sort_allocs);
        free_list(sss_allocs);
        sss_cap = 0;
        sss_q = sss_q2 = nullptr;
        sss0 = sss1 = sss2 = nullptr;
        sort_cap = 0;
        sort_keys = sort_keys_alt = nullptr;
        sort_vals = sort_perm = sort_rank = nullptr;
        sort_temp = nullptr;
        adj_cap = grad_cap = 0;
        grad_ptrs = nullptr;
        if (host_segs) (void)hipHostFree(host_segs);
        host_segs = nullptr;
        segs = nullptr;
        scenes = nullptr;
        cap = res_cap = pix_cap = scene_cap = 0;
        vol = false;
    }
};

}  // namespace

struct rd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream: a bounce's shadow / MIS traversal and resolve run on it
    // beside the next bounce's compaction, sort and extension traversal
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_shaded = nullptr, ev_resolved = nullptr;
    std::mutex mu;
    int* counts = nullptr;        // device counters (kCntSlots ints)
    float* sss_lut = nullptr;     // the BSSRDF's 2048 radii, uploaded on first use
    int* host_counts = nullptr;   // pinned mirror
    WaveBuffers wb;
    // rd_render_device_noalloc: this device's fixed work area (device
    // address from hipGetSymbolAddress), the buffers carved from it for the
    // current call, and the segment table's pinned staging
    void* scratch = nullptr;
    WaveBuffers wb_na;
    BatchSeg* na_host_segs = nullptr;
    hipEvent_t ev_t0 = nullptr, ev_t1 = nullptr;   // a render's device time
    rd_render_stats stats{};
    std::map<std::string, StageTiming> timing;
    bool timing_enabled = true;
    // order extension rays by direction octant and origin before tracing
    // them (bounces after the first; RD_RAY_SORT=0 turns it off)
    bool ray_sort = true;
    int ray_sort_bits = 15;       // key bits (RD_RAY_SORT_BITS): octant + Morton prefix
    int ray_sort_min = 1 << 16;   // smaller queues are traced in queue order
    // shadow queue in the sorted order too (RD_RAY_SORT_SHADOW=1), or in the
    // extension queue's own order (0: on the workload 2 % faster MIS + shadow
    // launches than the sorted order, whose key is the extension ray's)
    bool ray_sort_shadow = false;
    // traversal schedule per query class [extension, shadow, MIS]
    // (RD_TRACE_SCHEDULE="e,s,m"): sequences for the extension rays (in the
    // tile-major path order a lane's four rays are four samples of one pixel
    // at the first bounce, so the later three find their nodes in L1), chunks
    // for the shadow and MIS rays (which the fused launch then traces together)
    int trace_schedule[3] = {3, 2, 2};
    bool overlap = true;   // RD_OVERLAP=0 turns the two-stream bounce overlap off
    // the tail of a pass device-driven (RD_DEVICE_LOOP: 0 off, 1 launches,
    // 2 captured as a hipGraph), from queues of at most tail_rays rays
    // (RD_TAIL_RAYS); vol_path chains in lockstep launches (RD_CHAIN_LOCKSTEP)
    int device_loop = 2;
    int tail_rays = 1 << 18;
    bool chain_lockstep = false;
    // the last tail graph instantiated, and the bytes of every value its
    // launches were recorded with: a pass whose tail would record the same
    // launches replays it without capturing again
    hipGraphExec_t tail_exec = nullptr;
    std::string tail_key;
    // slot pool cap (log2; 0: the memory rule), path tile edge, least LDS
    // stack variant (RD_POOL_LOG2, RD_TILE, RD_STACK_MIN)
    int pool_log2 = 0;
    int tile = 8;
    int stack_min = 0;
    // scene trees: 0 host, 1 device from kDeviceBuildMin primitives, 2 device (RD_BVH_BUILD)
    int bvh_build = 1;
    // scenes created on this context and not yet destroyed; rd_destroy with
    // live scenes releases the wavefront buffers and leaves the rest to the
    // last rd_scene_destroy, so destruction order does not matter
    int live_scenes = 0;
    bool closed = false;
};

struct rd_scene {
    rd_ctx* ctx = nullptr;
    DevScene dev{};
    OctResult oct;                      // RD_ACCEL_OCTREE scenes: the tree as built
    bool sss = false;                   // a material may sample a BSSRDF (subsurface, Disney with scatting)
    LtCamera lt_cam{};                  // SampleWi / Importance state (light paths)
    DevIntegrator integ{};
    std::vector<void*> allocs;
    std::vector<NodeRec> nodes;
    std::vector<int32_t> order;
    rd_scene_info info{};
    int width = 0, height = 0;
    int num_textures = 0;
    std::vector<DevTexture> texs;       // host copy of the texture table
    std::vector<size_t> tex_elems;      // texels per texture (1 for constants)
    std::vector<int> env_textures;      // textures an environment light samples
    std::vector<int> set_tri_base;      // DevPrim index of set i's first triangle, minus num_top_prims
    int num_top_prims = 0;
    ~rd_scene() {
        for (void* p : allocs) (void)hipFree(p);
        if (ctx) ctx->live_scenes--;
    }
};

namespace {

// RAII "one host thread at a time" guard (the reference's renderer is driven
// from one thread; concurrent use of a context is rejected, not serialised).
struct CtxLock {
    std::unique_lock<std::mutex> lk;
    explicit CtxLock(rd_ctx* c) : lk(c->mu, std::try_to_lock) {}
    bool ok() const { return lk.owns_lock(); }
};

template <class T>
int upload(rd_scene* s, const T* host, size_t count, const T** dst) {
    if (count == 0) { *dst = nullptr; return RD_OK; }
    void* d = nullptr;
    hipError_t e = hipMalloc(&d, count * sizeof(T));
    if (e != hipSuccess) return fail(RD_ERR_OOM, "hipMalloc failed for scene buffer");
    s->allocs.push_back(d);
    HIP_TRY(hipMemcpy(d, host, count * sizeof(T), hipMemcpyHostToDevice));
    *dst = (const T*)d;
    return RD_OK;
}

// HIP-event timing of one launch on the context stream.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    std::string stage;
};

struct Timer {
    rd_ctx* ctx;
    std::vector<EventPair> pairs;
    std::vector<EventPair> pool;
    explicit Timer(rd_ctx* c) : ctx(c) {}
    ~Timer() {
        for (auto& p : pairs) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
        for (auto& p : pool) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    }
    EventPair get() {
        if (!pool.empty()) { EventPair p = pool.back(); pool.pop_back(); return p; }
        EventPair p;
        (void)hipEventCreate(&p.a);
        (void)hipEventCreate(&p.b);
        return p;
    }
    template <class F>
    hipError_t run(const char* stage, F&& launch, hipStream_t s = nullptr) {
        if (!ctx->timing_enabled) return launch();
        if (!s) s = ctx->stream;
        EventPair p = get();
        p.stage = stage;
        (void)hipEventRecord(p.a, s);
        hipError_t e = launch();
        (void)hipEventRecord(p.b, s);
        pairs.push_back(p);
        return e;
    }
    // Called after a stream synchronize: folds elapsed times into ctx->timing.
    void collect() {
        for (auto& p : pairs) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                StageTiming& t = ctx->timing[p.stage];
                t.total_ms += ms;
                t.count++;
            }
            pool.push_back(p);
        }
        pairs.clear();
    }
};

// Points the path pool of `w` (cap slots: path state, ray queues, hits,
// compaction scratch, + the vol_path arrays) at memory from `alloc`
// (bytes -> pointer, null when it has none left).  False if any is missing.
template <class Alloc>
bool assign_pool(rd_ctx* c, WaveBuffers& w, size_t cap, bool vol, Alloc&& alloc) {
    auto mkq = [&](RayQueue& q, int slot) -> bool {
        q.count = c->counts + slot;
        q.pid = (int*)alloc(cap * sizeof(int));
        q.org = (float4*)alloc(cap * sizeof(float4));
        q.dir = (float4*)alloc(cap * sizeof(float4));
        return q.pid && q.org && q.dir;
    };
    // qshadow2: vol_path's second Transmittance chain queue; the path
    // integrator's ranked shadow queue without its holes
    bool ok = mkq(w.qa, kCntExtA) && mkq(w.qb, kCntExtB) && mkq(w.qshadow, kCntShadow) && mkq(w.qmis, kCntMis) &&
              mkq(w.qshadow2, kCntShadowB);
    w.hits = (float4*)alloc(cap * sizeof(float4));
    PathState& ps = w.ps;
    ps.L = (float4*)alloc(cap * sizeof(float4));
    ps.beta = (float4*)alloc(cap * sizeof(float4));
    ps.nee_beta = (float4*)alloc(cap * sizeof(float4));
    ps.nee_le = (float4*)alloc(cap * sizeof(float4));
    ps.mis_f = (float4*)alloc(cap * sizeof(float4));
    ps.mis_li = (float4*)alloc(cap * sizeof(float4));
    ps.flags = (int*)alloc(cap * sizeof(int));
    ps.sel_light = (int*)alloc(cap * sizeof(int));
    ps.key = (uint64_t*)alloc(cap * sizeof(uint64_t));
    ps.req = (int*)alloc(cap * sizeof(int));
    ps.gid = (int*)alloc(cap * sizeof(int));
    ps.pix = (int*)alloc(cap * sizeof(int));
    ps.alive = (int*)alloc(cap * sizeof(int));
    ps.ray_o = (float4*)alloc(cap * sizeof(float4));
    ps.ray_d = (float4*)alloc(cap * sizeof(float4));
    w.blk = (int2*)alloc(compact_scratch_int2(cap) * sizeof(int2));
    w.blk2 = (int2*)alloc(compact_scratch_int2(cap) * sizeof(int2));
    ps.rng_n = (uint32_t*)alloc(cap * sizeof(uint32_t));
    ps.depth = (int*)alloc(cap * sizeof(int));
    ps.vis = (uint8_t*)alloc(cap);
    ps.hsub = (int*)alloc(cap * sizeof(int));
    ok = ok && w.hits && ps.L && ps.beta && ps.nee_beta && ps.nee_le && ps.mis_f && ps.mis_li && ps.flags &&
         ps.sel_light && ps.key && ps.req && ps.gid && ps.pix && ps.alive && ps.ray_o && ps.ray_d && w.blk && w.blk2 &&
         ps.rng_n &&
         ps.depth && ps.vis && ps.hsub;
    ps.env = nullptr; ps.nee_li = ps.sh_to = ps.sh_tr = ps.mis_tr = nullptr;
    ps.hsub2 = nullptr;
    if (ok && vol) {
        ok = mkq(w.qmis2, kCntMisB);
        w.hits2 = (float4*)alloc(cap * sizeof(float4));
        ps.env = (int*)alloc(cap * sizeof(int));
        ps.nee_li = (float4*)alloc(cap * sizeof(float4));
        ps.sh_to = (float4*)alloc(cap * sizeof(float4));
        ps.sh_tr = (float4*)alloc(cap * sizeof(float4));
        ps.mis_tr = (float4*)alloc(cap * sizeof(float4));
        ps.hsub2 = (int*)alloc(cap * sizeof(int));
        ok = ok && w.hits2 && ps.env && ps.nee_li && ps.sh_to && ps.sh_tr && ps.mis_tr && ps.hsub2;
    }
    ps.adj0 = ps.adj1 = ps.adj2 = nullptr;
    ps.sss_q = ps.sss_q2 = nullptr;
    ps.sss0 = ps.sss1 = ps.sss2 = nullptr;
    return ok;
}

// Grows the render buffers to hold `pool` path slots (+ the vol_path arrays),
// `results` per-sample results and `pixels` accumulation pixels.
int ensure_wave_buffers(rd_ctx* c, size_t pool, size_t results, size_t pixels, bool vol) {
    WaveBuffers& w = c->wb;
    auto alloc_in = [&](std::vector<void*>& list, size_t bytes) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        list.push_back(p);
        return p;
    };
    if (!w.segs) {
        w.segs = (BatchSeg*)alloc_in(w.seg_allocs, sizeof(BatchSeg) * kMaxSegs);
        if (hipHostMalloc((void**)&w.host_segs, sizeof(BatchSeg) * kMaxSegs, hipHostMallocDefault) != hipSuccess)
            w.host_segs = nullptr;
        if (!w.segs || !w.host_segs) { w.release(); return fail(RD_ERR_OOM, "segment table allocation failed"); }
    }
    if (pool > w.cap || (vol && !w.vol)) {
        WaveBuffers::free_list(w.pool_allocs);
        w.cap = 0;
        w.vol = false;
        size_t cap = std::max(pool, (size_t)1024);
        bool ok = assign_pool(c, w, cap, vol, [&](size_t bytes) { return alloc_in(w.pool_allocs, bytes); });
        if (!ok) { w.release(); return fail(RD_ERR_OOM, "hipMalloc failed for the path pool"); }
        w.cap = cap;
        w.vol = vol;
        w.adj_cap = 0;
    }
    if (results > w.res_cap) {
        WaveBuffers::free_list(w.res_allocs);
        w.res_cap = 0;
        w.ps.res = (float4*)alloc_in(w.res_allocs, results * sizeof(float4));
        if (!w.ps.res) { w.release(); return fail(RD_ERR_OOM, "hipMalloc failed for per-sample results"); }
        w.res_cap = results;
    }
    if (pixels > w.pix_cap) {
        WaveBuffers::free_list(w.pix_allocs);
        w.pix_cap = 0;
        size_t pcap = std::max(pixels, (size_t)1024);
        w.sum = (double4*)alloc_in(w.pix_allocs, pcap * sizeof(double4));
        w.sum2 = (float4*)alloc_in(w.pix_allocs, pcap * sizeof(float4));
        if (!w.sum || !w.sum2) { w.release(); return fail(RD_ERR_OOM, "hipMalloc failed for accumulation"); }
        w.pix_cap = pcap;
    }
    return RD_OK;
}

// BssrdfUtility.ScattingLUT (Bssrdf.cs): r_i with 1 - e^(-r)/4 - 3 e^(-r/3)/4
// = i / 2047 for i < 2047, bisected in double to the float nearest the root;
// the last entry (no root) is the reference's 111.68667351, reached by no
// accepted sample.
static std::vector<float> scatter_radii() {
    std::vector<float> t(2048);
    for (int i = 0; i < 2047; i++) {
        const double target = (double)i / 2047.0;
        double lo = 0, hi = 64;
        for (int it = 0; it < 200; it++) {
            const double mid = 0.5 * (lo + hi);
            if (1 - 0.25 * std::exp(-mid) - 0.75 * std::exp(-mid / 3) < target) lo = mid;
            else hi = mid;
        }
        t[i] = (float)(0.5 * (lo + hi));
    }
    t[2047] = 111.68667351f;
    return t;
}

// The BSSRDF queues and points for `pool` slots (kept across renders) and the
// radius table (uploaded once per context); points w.ps at them.
int ensure_sss_buffers(rd_ctx* c, WaveBuffers& w, size_t pool) {
    if (!c->sss_lut) {
        const std::vector<float> lut = scatter_radii();
        if (hipMalloc((void**)&c->sss_lut, lut.size() * sizeof(float)) != hipSuccess) {
            c->sss_lut = nullptr;
            return fail(RD_ERR_OOM, "hipMalloc failed for the BSSRDF table");
        }
        if (hipMemcpy(c->sss_lut, lut.data(), lut.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            return fail(RD_ERR_DEVICE, "BSSRDF table upload failed");
    }
    if (pool > w.sss_cap) {
        WaveBuffers::free_list(w.sss_allocs);
        w.sss_cap = 0;
        auto get = [&](size_t bytes) -> void* {
            void* p = nullptr;
            if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
            w.sss_allocs.push_back(p);
            return p;
        };
        w.sss_q = (int*)get(pool * sizeof(int));
        w.sss_q2 = (int*)get(pool * sizeof(int));
        w.sss0 = (float4*)get(pool * sizeof(float4));
        w.sss1 = (float4*)get(pool * sizeof(float4));
        w.sss2 = (float4*)get(pool * sizeof(float4));
        if (!w.sss_q || !w.sss_q2 || !w.sss0 || !w.sss1 || !w.sss2) {
            WaveBuffers::free_list(w.sss_allocs);
            w.sss_q = w.sss_q2 = nullptr;
            w.sss0 = w.sss1 = w.sss2 = nullptr;
            return fail(RD_ERR_OOM, "hipMalloc failed for the BSSRDF queues");
        }
        w.sss_cap = pool;
    }
    w.ps.sss_q = w.sss_q;
    w.ps.sss_q2 = w.sss_q2;
    w.ps.sss0 = w.sss0;
    w.ps.sss1 = w.sss1;
    w.ps.sss2 = w.sss2;
    return RD_OK;
}

// rd_render_device_noalloc: the render's buffers carved from the context's
// fixed work area, in 256-byte aligned pieces: segment table, per-pixel sums,
// path pool, per-sample results.  With base == null nothing is assigned for
// real and the carve only measures (Bump::used).
struct Bump {
    char* base;
    size_t size;
    size_t used = 0;
    void* take(size_t bytes) {
        const size_t at = (used + 255) & ~(size_t)255;
        if (at > size || bytes > size - at) { used = size + 1; return nullptr; }
        used = at + bytes;
        return base ? (void*)(base + at) : (void*)(uintptr_t)(at + 256);   // measuring: any non-null
    }
};

bool carve_scratch(rd_ctx* c, WaveBuffers& w, Bump& b, size_t pool, size_t results, size_t pixels, bool vol) {
    w.segs = (BatchSeg*)b.take(sizeof(BatchSeg) * kMaxSegs);
    w.sum = (double4*)b.take(pixels * sizeof(double4));
    w.sum2 = (float4*)b.take(pixels * sizeof(float4));
    bool ok = w.segs && w.sum && w.sum2 && assign_pool(c, w, pool, vol, [&](size_t bytes) { return b.take(bytes); });
    w.ps.res = (float4*)b.take(results * sizeof(float4));
    ok = ok && w.ps.res;
    w.host_segs = c->na_host_segs;
    w.cap = pool;
    w.res_cap = results;
    w.pix_cap = pixels;
    w.vol = vol;
    w.sort_cap = 0;
    w.sort_keys = w.sort_keys_alt = nullptr;
    w.sort_vals = w.sort_perm = w.sort_rank = nullptr;
    w.sort_temp = nullptr;
    w.scenes = nullptr;
    return ok;
}

size_t scratch_need(rd_ctx* c, size_t pool, size_t results, size_t pixels, bool vol) {
    WaveBuffers tmp;
    Bump b{nullptr, (size_t)-1 / 2};
    carve_scratch(c, tmp, b, pool, results, pixels, vol);
    tmp.host_segs = nullptr;   // not owned
    return b.used;
}

// Smallest pool the work area must hold beside the sums and one sample's results.
constexpr size_t kNoAllocMinPool = 1024;
constexpr size_t kNoAllocMaxPool = (size_t)1 << 17;

// Adjoint replay: per-slot derivative records and a table of `ntex` gradient pointers.
// Buffers of the extension-queue sort, sized for the pool.
int ensure_sort_buffers(rd_ctx* c, size_t pool) {
    WaveBuffers& w = c->wb;
    if (w.sort_cap >= pool) return RD_OK;
    WaveBuffers::free_list(w.sort_allocs);
    w.sort_cap = 0;
    auto alloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        w.sort_allocs.push_back(p);
        return p;
    };
    w.sort_keys = (uint32_t*)alloc(pool * 4);
    w.sort_keys_alt = (uint32_t*)alloc(pool * 4);
    w.sort_vals = (int*)alloc(pool * 4);
    w.sort_perm = (int*)alloc(pool * 4);
    w.sort_rank = (int*)alloc(pool * 4);
    w.sort_temp_bytes = ray_sort_temp_bytes((int)pool);
    if (w.sort_temp_bytes == 0) {
        WaveBuffers::free_list(w.sort_allocs);
        return fail(RD_ERR_DEVICE, "ray sort workspace size query failed");
    }
    w.sort_temp = alloc(w.sort_temp_bytes);
    if (!w.sort_keys || !w.sort_keys_alt || !w.sort_vals || !w.sort_perm || !w.sort_rank || !w.sort_temp) {
        WaveBuffers::free_list(w.sort_allocs);
        return fail(RD_ERR_OOM, "hipMalloc failed for the ray sort");
    }
    w.sort_cap = pool;
    return RD_OK;
}

int ensure_adjoint_buffers(rd_ctx* c, size_t ntex) {
    WaveBuffers& w = c->wb;
    if (w.adj_cap < w.cap || w.grad_cap < ntex) {
        WaveBuffers::free_list(w.adj_allocs);
        w.adj_cap = w.grad_cap = 0;
        auto alloc = [&](size_t bytes) -> void* {
            void* p = nullptr;
            if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
            w.adj_allocs.push_back(p);
            return p;
        };
        w.ps.adj0 = (float4*)alloc(w.cap * sizeof(float4));
        w.ps.adj1 = (float4*)alloc(w.cap * sizeof(float4));
        w.ps.adj2 = (float4*)alloc(w.cap * sizeof(float4));
        w.grad_ptrs = (void**)alloc(std::max<size_t>(ntex, 1) * sizeof(void*));
        if (!w.ps.adj0 || !w.ps.adj1 || !w.ps.adj2 || !w.grad_ptrs) {
            WaveBuffers::free_list(w.adj_allocs);
            return fail(RD_ERR_OOM, "hipMalloc failed for adjoint buffers");
        }
        w.adj_cap = w.cap;
        w.grad_cap = std::max<size_t>(ntex, 1);
    }
    return RD_OK;
}

int ensure_batch_buffers(rd_ctx* c, size_t nscenes) {
    WaveBuffers& w = c->wb;
    if (nscenes <= w.scene_cap) return RD_OK;
    WaveBuffers::free_list(w.scene_allocs);
    w.scene_cap = 0;
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(DevScene) * nscenes) != hipSuccess) return fail(RD_ERR_OOM, "hipMalloc scene table");
    w.scene_allocs.push_back(p);
    w.scenes = (DevScene*)p;
    w.scene_cap = nscenes;
    return RD_OK;
}

// Device buffers of one render that are freed when it returns.
struct ScopedAllocs {
    std::vector<void*> ptrs;
    ~ScopedAllocs() { for (void* p : ptrs) (void)hipFree(p); }
    template <class T>
    T* get(size_t count) {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T*)p;
    }
};

int check_stack(const rd_scene* s) {
    if (s->dev.octree) return s->dev.stack_needed > 256 ? fail(RD_ERR_STACK, "octree stack over 256 entries") : RD_OK;
    if (s->dev.stack_needed > 64)
        return fail(RD_ERR_STACK, "BVH deeper than the 64-entry traversal stack (Bvh.StackSize)");
    return RD_OK;
}

}  // namespace

extern "C" {

int rd_abi_version(void) { return RD_ABI_VERSION; }
const char* rd_last_error(void) { return g_err.c_str(); }

static void free_ctx(rd_ctx* c);

rd_ctx* rd_create(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) { g_err = "no HIP device"; return nullptr; }
    if (device < 0 || device >= n) { g_err = "device index out of range"; return nullptr; }
    if (hipSetDevice(device) != hipSuccess) { g_err = "hipSetDevice failed"; return nullptr; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { g_err = "hipGetDeviceProperties failed"; return nullptr; }
    rd_ctx* c = new rd_ctx();
    c->device = device;
    set_device_grid(prop.multiProcessorCount);
    if (const char* env = getenv("RD_TRACE_SCHEDULE")) {   // "ext,shadow,mis", e.g. "2,2,1"
        int v[3] = {0, 0, 0};
        if (sscanf(env, "%d,%d,%d", &v[0], &v[1], &v[2]) == 3)
            for (int k = 0; k < 3; k++)
                if (v[k] >= 0 && v[k] < kNumSchedules) c->trace_schedule[k] = v[k];
    }
    if (const char* env = getenv("RD_RAY_SORT")) c->ray_sort = atoi(env) != 0;
    if (const char* env = getenv("RD_RAY_SORT_SHADOW")) c->ray_sort_shadow = atoi(env) != 0;
    if (const char* env = getenv("RD_RAY_SORT_BITS")) c->ray_sort_bits = std::max(3, std::min(30, atoi(env)));
    if (const char* env = getenv("RD_OVERLAP")) c->overlap = atoi(env) != 0;
    if (const char* env = getenv("RD_DEVICE_LOOP")) c->device_loop = std::max(0, std::min(2, atoi(env)));
    if (const char* env = getenv("RD_TAIL_RAYS")) c->tail_rays = std::max(0, atoi(env));
    if (const char* env = getenv("RD_CHAIN_LOCKSTEP")) c->chain_lockstep = atoi(env) != 0;
    if (const char* env = getenv("RD_POOL_LOG2")) { int v = atoi(env); c->pool_log2 = v == 0 ? 0 : std::max(10, std::min(27, v)); }
    if (const char* env = getenv("RD_TILE")) c->tile = std::max(1, std::min(64, atoi(env)));
    if (const char* env = getenv("RD_STACK_MIN")) c->stack_min = std::max(0, std::min(64, atoi(env)));
    if (const char* env = getenv("RD_BVH_BUILD")) c->bvh_build = std::max(0, std::min(2, atoi(env)));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        g_err = "hipStreamCreate failed"; delete c; return nullptr;
    }
    if (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_shaded, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_resolved, hipEventDisableTiming) != hipSuccess) {
        g_err = "hipStreamCreate failed";
        if (c->stream2) (void)hipStreamDestroy(c->stream2);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return nullptr;
    }
    if (hipMalloc(&c->counts, kCntSlots * sizeof(int)) != hipSuccess ||
        hipHostMalloc(&c->host_counts, kCntSlots * sizeof(int), hipHostMallocDefault) != hipSuccess) {
        g_err = "counter allocation failed";
        if (c->counts) (void)hipFree(c->counts);
        (void)hipEventDestroy(c->ev_shaded);
        (void)hipEventDestroy(c->ev_resolved);
        (void)hipStreamDestroy(c->stream2);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return nullptr;
    }
    (void)hipMemset(c->counts, 0, kCntSlots * sizeof(int));
    // set-up for rd_render_device_noalloc, so that it allocates nothing: this
    // device's work-area address, the segment table's pinned staging, the
    // render-time events
    hipError_t se = rd::scratch_address(device, &c->scratch);
    if (se != hipSuccess ||
        hipHostMalloc((void**)&c->na_host_segs, sizeof(BatchSeg) * kMaxSegs, hipHostMallocDefault) != hipSuccess ||
        hipEventCreate(&c->ev_t0) != hipSuccess || hipEventCreate(&c->ev_t1) != hipSuccess) {
        g_err = se != hipSuccess ? std::string("work-area lookup failed: ") + hipGetErrorString(se)
                                 : std::string("context set-up failed");
        free_ctx(c);
        return nullptr;
    }
    return c;
}

static void free_ctx(rd_ctx* c) {
    if (c->counts) (void)hipFree(c->counts);
    if (c->host_counts) (void)hipHostFree(c->host_counts);
    if (c->na_host_segs) (void)hipHostFree(c->na_host_segs);
    if (c->ev_t0) (void)hipEventDestroy(c->ev_t0);
    if (c->ev_t1) (void)hipEventDestroy(c->ev_t1);
    if (c->tail_exec) (void)hipGraphExecDestroy(c->tail_exec);
    if (c->sss_lut) (void)hipFree(c->sss_lut);
    (void)hipEventDestroy(c->ev_shaded);
    (void)hipEventDestroy(c->ev_resolved);
    (void)hipStreamDestroy(c->stream2);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

void rd_destroy(rd_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->wb.release();
    c->closed = true;
    if (c->live_scenes == 0) free_ctx(c);
}

// Trees of at least this many primitives are built on the device when
// rd_tuning.bvh_build is 1: below it the host builds faster than the
// device's per-level launches and synchronizations.
constexpr int kDeviceBuildMin = 1 << 15;

rd_scene* rd_scene_create(rd_ctx* c, const rd_scene_desc* d) {
    if (!c || !d) { g_err = "null argument"; return nullptr; }
    CtxLock lock(c);
    if (!lock.ok()) { g_err = "context in use by another thread"; return nullptr; }
    (void)hipSetDevice(c->device);
    auto bad = [](const char* m) -> rd_scene* { g_err = m; return nullptr; };
    if (d->num_vertices < 0 || d->num_triangles < 0 || d->num_spheres < 0 || d->num_prims < 0) return bad("negative count");
    if (d->num_triangles > 0 && (!d->positions || !d->indices)) return bad("triangles without positions/indices");
    if (d->num_spheres > 0 && !d->spheres) return bad("spheres without data");
    if (d->num_prims > 0 && !d->prims) return bad("prims missing");
    if (d->camera.width <= 0 || d->camera.height <= 0) return bad("camera screen must be positive");
    if (d->integrator.type < RD_INTEGRATOR_PATH || d->integrator.type > RD_INTEGRATOR_BDPT)
        return bad("unknown integrator type");
    if (d->global_medium < -1 || d->global_medium >= d->num_media) return bad("global medium out of range");
    if (d->num_instance_sets < 0 || d->num_instances < 0) return bad("negative count");
    if (d->num_instance_sets > 0 && !d->instance_sets) return bad("instance sets missing");
    if (d->num_instances > 0 && !d->instances) return bad("instances missing");
    for (int i = 0; i < d->num_instance_sets; i++) {
        const rd_instance_set_desc& sd = d->instance_sets[i];
        if (sd.first_triangle < 0 || sd.num_triangles <= 0 || sd.num_triangles > d->num_triangles - sd.first_triangle)
            return bad("instance set triangle range out of range");
    }
    for (int i = 0; i < d->num_instances; i++) {
        const rd_instance_desc& id = d->instances[i];
        if (id.set < 0 || id.set >= d->num_instance_sets) return bad("instance set index out of range");
        float inv[16];
        if (!invert4(id.to_world, inv)) return bad("instance: invalid transform matrix");   // InstancedTransform ctor
    }
    for (int i = 0; i < d->num_prims; i++) {
        const rd_prim_desc& pd = d->prims[i];
        if (pd.medium_inside < -1 || pd.medium_inside >= d->num_media || pd.medium_outside < -1 ||
            pd.medium_outside >= d->num_media)
            return bad("primitive medium out of range");
        if (pd.kind == RD_PRIM_INSTANCE) {
            if (pd.index < 0 || pd.index >= d->num_instances) return bad("primitive instance index out of range");
            if (pd.light != -1) return bad("an instanced primitive cannot be a light");
        }
    }
    if (d->integrator.strategy != RD_STRATEGY_UNIFORM) return bad("light strategy All is not built for the GPU path yet");
    for (int t = 0; t < d->num_triangles; t++) {
        uint8_t fl = d->tri_flags ? d->tri_flags[t] : 0;
        if ((fl & 1) && !d->normals) return bad("triangle flags normals but none given");
        if ((fl & 2) && !d->uvs) return bad("triangle flags uvs but none given");
        for (int k = 0; k < 3; k++) {
            int v = d->indices[3 * t + k];
            if (v < 0 || v >= d->num_vertices) return bad("triangle index out of range");
        }
    }
    bool has_sss = false;
    for (int i = 0; i < d->num_materials; i++) {
        int ty = d->materials[i].type;
        if (ty < RD_MAT_DIFFUSE || ty > RD_MAT_SUBSURFACE) return bad("unknown material type");
        if (d->num_textures == 0) return bad("materials need textures (constants count)");
        for (int k = 0; k < 16; k++) {
            int tx = d->materials[i].tex[k];
            if (tx < 0 || tx >= d->num_textures) return bad("material texture index out of range");
        }
        if ((ty == RD_MAT_DISNEY && !d->materials[i].thin) || ty == RD_MAT_SUBSURFACE) {
            // a scattering distance that is not black somewhere samples the
            // BSSRDF (an image texture may be replaced later: it counts)
            const rd_texture_desc& sd = d->textures[d->materials[i].tex[ty == RD_MAT_DISNEY ? 11 : 3]];
            if (sd.width > 0 || sd.value[0] != 0 || sd.value[1] != 0 || sd.value[2] != 0) has_sss = true;
        }
    }
    for (int i = 0; i < d->num_lights; i++) {
        const rd_light_desc& l = d->lights[i];
        if (l.type == RD_LIGHT_INFINITE) {
            if (l.tex < 0 || l.tex >= d->num_textures) return bad("infinite light texture out of range");
            float inv[16];
            if (!invert4(l.to_world, inv)) return bad("infinite light: invalid matrix");
            continue;
        }
        if (l.type != RD_LIGHT_AREA) return bad("unknown light type");
        if (l.prim < 0 || l.prim >= d->num_prims) return bad("light primitive out of range");
        if (d->prims[l.prim].kind == RD_PRIM_INSTANCE) return bad("an instanced primitive cannot be a light");
    }

    if (c->closed) { g_err = "context destroyed"; return nullptr; }
    rd_scene* s = new rd_scene();
    s->ctx = c;
    s->sss = has_sss;
    c->live_scenes++;
    s->width = d->camera.width;
    s->height = d->camera.height;
    auto t0 = std::chrono::steady_clock::now();
    const int np = d->num_prims;
    auto P = [&](int v, int k) { return d->positions[3 * (size_t)v + k]; };
    // Triangle ctor: bounds of the three vertices (Triangle.cs:83-94)
    auto tri_bounds = [&](int tri, float* lo3, float* hi3) {
        const int* ix = d->indices + 3 * (size_t)tri;
        for (int k = 0; k < 3; k++) {
            float a = P(ix[0], k), b = P(ix[1], k), cc = P(ix[2], k);
            float lo = a, hi = a;
            lo = lo < b ? lo : b; hi = hi > b ? hi : b;
            lo = lo < cc ? lo : cc; hi = hi > cc ? hi : cc;
            lo3[k] = lo;
            hi3[k] = hi;
        }
    };
    // the host builder or the same tree built on the device (rd_tuning.bvh_build)
    auto tree = [&](int n, const float* lo, const float* hi, int mp, BvhResult* r, std::string* e) {
        if (c->bvh_build == 2 || (c->bvh_build == 1 && n >= kDeviceBuildMin))
            return gpu_build_bvh(n, lo, hi, mp, c->stream, r, e);
        return build_bvh(n, lo, hi, mp, r, e);
    };
    // one BVH per instance set over its ShapeWrapperPrimitives (RendererBuilder.cs:895-921)
    const int nsets = d->num_instance_sets;
    std::vector<BvhResult> blas(nsets);
    std::vector<int> set_tri_base(nsets + 1, 0);
    for (int si = 0; si < nsets; si++) {
        const rd_instance_set_desc& sd = d->instance_sets[si];
        std::vector<float> lo(3 * (size_t)sd.num_triangles), hi(3 * (size_t)sd.num_triangles);
        for (int j = 0; j < sd.num_triangles; j++) tri_bounds(sd.first_triangle + j, &lo[3 * j], &hi[3 * j]);
        std::string berr;
        if (!tree(sd.num_triangles, lo.data(), hi.data(), sd.bvh_max_prim > 0 ? sd.bvh_max_prim : 1, &blas[si], &berr)) {
            delete s;
            g_err = berr;
            return nullptr;
        }
        set_tri_base[si + 1] = set_tri_base[si] + sd.num_triangles;
    }
    std::vector<float> bmin(3 * (size_t)np), bmax(3 * (size_t)np);
    std::vector<DevPrim> info((size_t)np + set_tri_base[nsets]);
    for (int p = 0; p < np; p++) {
        const rd_prim_desc& pd = d->prims[p];
        DevPrim& di = info[p];
        di.kind = pd.kind; di.index = pd.index; di.material = pd.material; di.light = pd.light;
        di.med_in = pd.medium_inside; di.med_out = pd.medium_outside; di.flags = 0;
        if (pd.material >= d->num_materials || pd.light >= d->num_lights) { delete s; return bad("primitive material/light out of range"); }
        if (pd.kind == RD_PRIM_TRIANGLE) {
            if (pd.index < 0 || pd.index >= d->num_triangles) { delete s; return bad("primitive triangle index out of range"); }
            const int* ix = d->indices + 3 * (size_t)pd.index;
            f3 A = mk(P(ix[0], 0), P(ix[0], 1), P(ix[0], 2));
            f3 B = mk(P(ix[1], 0), P(ix[1], 1), P(ix[1], 2));
            f3 C = mk(P(ix[2], 0), P(ix[2], 1), P(ix[2], 2));
            // area = |cross|/2 (Triangle.cs:83-94)
            tri_bounds(pd.index, &bmin[3 * p], &bmax[3 * p]);
            di.area = 0.5f * length(cross(B - A, C - A));
            di.flags = d->tri_flags ? d->tri_flags[pd.index] : 0;
        } else if (pd.kind == RD_PRIM_INSTANCE) {
            // InstancedPrimitive.WorldBound = BoundingBox3F.Transform(set root, ModelToWorld)
            // (Primitive.cs:154-161, BoundingBox3F.cs:232-243)
            const rd_instance_desc& id = d->instances[pd.index];
            const NodeRec& root = blas[id.set].nodes[0];
            const float* M = id.to_world;
            float lo[3], hi[3];
            // corners in the reference's ExpendBy order: bit 0 = max x, bit 1 = max y, bit 2 = max z
            static const int kCorner[8] = {0, 1, 2, 4, 6, 3, 5, 7};
            for (int c = 0; c < 8; c++) {
                int b = kCorner[c];
                float x = (b & 1) ? root.mx[0] : root.mn[0];
                float y = (b & 2) ? root.mx[1] : root.mn[1];
                float z = (b & 4) ? root.mx[2] : root.mn[2];
                float w[3];
                for (int k = 0; k < 3; k++) w[k] = x * M[k] + y * M[4 + k] + z * M[8 + k] + M[12 + k];
                for (int k = 0; k < 3; k++) {
                    if (c == 0) { lo[k] = w[k]; hi[k] = w[k]; }
                    else { lo[k] = lo[k] < w[k] ? lo[k] : w[k]; hi[k] = hi[k] > w[k] ? hi[k] : w[k]; }
                }
            }
            for (int k = 0; k < 3; k++) { bmin[3 * p + k] = lo[k]; bmax[3 * p + k] = hi[k]; }
            di.area = 0.0f;
        } else if (pd.kind == RD_PRIM_SPHERE) {
            if (pd.index < 0 || pd.index >= d->num_spheres) { delete s; return bad("primitive sphere index out of range"); }
            const float* S = d->spheres + 4 * (size_t)pd.index;
            for (int k = 0; k < 3; k++) { bmin[3 * p + k] = S[k] - S[3]; bmax[3 * p + k] = S[k] + S[3]; }
            di.area = 4 * S[3] * S[3] * kPi;   // Sphere.SurfaceArea (Sphere.cs:18)
        } else {
            delete s;
            return bad("unknown primitive kind");
        }
    }
    BvhResult bvh;
    std::string err;
    const bool use_octree = d->accel == RD_ACCEL_OCTREE;
    if (d->accel != RD_ACCEL_BVH && !use_octree) { delete s; return bad("unknown accelerator"); }
    if (use_octree && nsets > 0) { delete s; return bad("the octree accelerator holds no instanced primitives"); }
    if (use_octree) {
        if (!build_octree(np, bmin.data(), bmax.data(), d->octree_max_depth, d->octree_max_count,
                          d->octree_out_bound, &s->oct, &err)) {
            delete s;
            g_err = err;
            return nullptr;
        }
        // the traversal's LIFO at worst (every box hit) must fit Octree.StackSize
        if (s->oct.stack_needed > 256) {
            delete s;
            g_err = "octree whose traversal can hold over 256 entries (Octree.StackSize)";
            return nullptr;
        }
        // the rest of the creation reads the tree through bvh: its root box
        // (the scene bound), leaf order and depth
        for (const OctNode& o : s->oct.nodes) {
            NodeRec r{};
            for (int k = 0; k < 3; k++) { r.mn[k] = o.mn[k]; r.mx[k] = o.mx[k]; }
            bvh.nodes.push_back(r);
        }
        bvh.order = s->oct.leaf_prims;
        bvh.leaves = (int)s->oct.leaf_first.size();
        bvh.max_depth = s->oct.max_depth;
    } else if (!tree(np, bmin.data(), bmax.data(), d->bvh_max_prim > 0 ? d->bvh_max_prim : 1, &bvh, &err)) {
        delete s;
        g_err = err;
        return nullptr;
    }
    auto t1 = std::chrono::steady_clock::now();
    s->nodes = bvh.nodes;
    s->order = bvh.order;
    const int n_ordered = (int)bvh.order.size();   // the octree's leaves may repeat primitives

    // node and ordered-primitive records: the scene tree, then every set's
    // tree with its child / primitive offsets made absolute
    auto ibits = [](int v) { float f; std::memcpy(&f, &v, 4); return f; };
    std::vector<float4> nodes;
    auto add_nodes = [&](const BvhResult& b, int node_base, int prim_base) {
        for (const NodeRec& r : b.nodes) {
            int packed = (int)r.count | ((int)r.axis << 16);
            int off = r.count > 0 ? r.offset + prim_base : r.offset + node_base;
            nodes.push_back(make_float4(r.mn[0], r.mn[1], r.mn[2], r.mx[0]));
            nodes.push_back(make_float4(r.mx[1], r.mx[2], ibits(off), ibits(packed)));
        }
    };
    if (use_octree) {
        // lo = {min, max.x}, hi = {max.y, max.z, offset, count}: a leaf's
        // primitive records, an interior node's -NodeCount children
        for (const OctNode& o : s->oct.nodes) {
            const bool leaf = o.next_head < 0;
            const int off = leaf ? s->oct.leaf_first[o.count_or_leaf] : o.next_head;
            const int cnt = leaf ? s->oct.leaf_count[o.count_or_leaf] : -o.count_or_leaf;
            nodes.push_back(make_float4(o.mn[0], o.mn[1], o.mn[2], o.mx[0]));
            nodes.push_back(make_float4(o.mx[1], o.mx[2], ibits(off), ibits(cnt)));
        }
    } else {
        add_nodes(bvh, 0, 0);
    }
    std::vector<float4> prims(3 * (size_t)n_ordered);
    auto put_tri = [&](size_t i, int tri, int orig) {
        const int* ix = d->indices + 3 * (size_t)tri;
        prims[3 * i] = make_float4(P(ix[0], 0), P(ix[0], 1), P(ix[0], 2), ibits(orig));
        prims[3 * i + 1] = make_float4(P(ix[1], 0), P(ix[1], 1), P(ix[1], 2), ibits(kTri));
        prims[3 * i + 2] = make_float4(P(ix[2], 0), P(ix[2], 1), P(ix[2], 2), 0.0f);
    };
    for (int i = 0; i < n_ordered; i++) {
        int orig = bvh.order[i];
        const rd_prim_desc& pd = d->prims[orig];
        if (pd.kind == RD_PRIM_TRIANGLE) {
            put_tri(i, pd.index, orig);
        } else if (pd.kind == RD_PRIM_INSTANCE) {
            prims[3 * i] = make_float4(ibits(pd.index), 0.f, 0.f, ibits(orig));
            prims[3 * i + 1] = make_float4(0.f, 0.f, 0.f, ibits(kInst));
            prims[3 * i + 2] = make_float4(0.f, 0.f, 0.f, 0.f);
        } else {
            const float* S = d->spheres + 4 * (size_t)pd.index;
            prims[3 * i] = make_float4(S[0], S[1], S[2], ibits(orig));
            prims[3 * i + 1] = make_float4(S[3], 0.f, 0.f, ibits(kSphere));
            prims[3 * i + 2] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
    std::vector<int> set_root(nsets);
    int blas_depth = 0;
    for (int si = 0; si < nsets; si++) {
        const rd_instance_set_desc& sd = d->instance_sets[si];
        const BvhResult& b = blas[si];
        set_root[si] = (int)(nodes.size() / 2);
        int prim_base = (int)(prims.size() / 3);
        add_nodes(b, set_root[si], prim_base);
        prims.resize(prims.size() + 3 * (size_t)sd.num_triangles);
        for (int k = 0; k < sd.num_triangles; k++) {
            int j = b.order[k];   // index within the set
            put_tri((size_t)prim_base + k, sd.first_triangle + j, np + set_tri_base[si] + j);
        }
        for (int j = 0; j < sd.num_triangles; j++) {
            int tri = sd.first_triangle + j;
            DevPrim& di = info[(size_t)np + set_tri_base[si] + j];
            const int* ix = d->indices + 3 * (size_t)tri;
            f3 A = mk(P(ix[0], 0), P(ix[0], 1), P(ix[0], 2));
            f3 B = mk(P(ix[1], 0), P(ix[1], 1), P(ix[1], 2));
            f3 C = mk(P(ix[2], 0), P(ix[2], 1), P(ix[2], 2));
            di.kind = kTri; di.index = tri; di.material = -1; di.light = -1; di.med_in = -1; di.med_out = -1;
            di.area = 0.5f * length(cross(B - A, C - A));
            di.flags = d->tri_flags ? d->tri_flags[tri] : 0;
        }
        blas_depth = std::max(blas_depth, b.max_depth);
    }
    std::vector<DevInstance> insts(d->num_instances);
    for (int i = 0; i < d->num_instances; i++) {
        const rd_instance_desc& id = d->instances[i];
        DevInstance& o = insts[i];
        std::memset(&o, 0, sizeof(o));
        float inv[16];
        invert4(id.to_world, inv);   // InstancedTransform: Matrix4x4.Invert (Primitive.cs:99-106)
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 3; c++) { o.m2w[3 * r + c] = id.to_world[4 * r + c]; o.w2m[3 * r + c] = inv[4 * r + c]; }
        o.root = set_root[id.set];
    }
    s->set_tri_base = set_tri_base;
    s->num_top_prims = np;
    std::vector<float4> spheres(d->num_spheres);
    for (int i = 0; i < d->num_spheres; i++)
        spheres[i] = make_float4(d->spheres[4 * i], d->spheres[4 * i + 1], d->spheres[4 * i + 2], d->spheres[4 * i + 3]);
    std::vector<DevMaterial> mats(d->num_materials);
    for (int i = 0; i < d->num_materials; i++) {
        const rd_material_desc& m = d->materials[i];
        DevMaterial& o = mats[i];
        std::memset(&o, 0, sizeof(o));
        o.type = m.type; o.two_sided = m.two_sided; o.dist = m.dist; o.thin = m.thin;
        for (int k = 0; k < 16; k++) o.tex[k] = m.tex[k];
        o.eta_a = m.eta_a; o.eta_b = m.eta_b; o.kd = m.kd; o.ks = m.ks;
        for (int k = 0; k < 3; k++) { o.eta[k] = m.eta[k]; o.k[k] = m.k[k]; }
    }
    std::vector<DevTexture> texs(d->num_textures);
    std::vector<float> texels;
    for (int i = 0; i < d->num_textures; i++) {
        const rd_texture_desc& t = d->textures[i];
        DevTexture& o = texs[i];
        std::memset(&o, 0, sizeof(o));
        o.w = t.width; o.h = t.height; o.wrap = t.wrap; o.filter = t.filter;
        for (int k = 0; k < 3; k++) o.v[k] = t.value[k];
        o.offset = (int64_t)texels.size();
        if (t.width > 0) {
            if (t.height <= 0 || !t.data) { delete s; return bad("image texture without data"); }
            texels.insert(texels.end(), t.data, t.data + (size_t)t.width * t.height * 3);
        }
    }
    s->num_textures = d->num_textures;
    s->texs = texs;
    for (int i = 0; i < d->num_textures; i++)
        s->tex_elems.push_back(texs[i].w > 0 ? (size_t)texs[i].w * texs[i].h : 1);
    for (int i = 0; i < d->num_lights; i++)
        if (d->lights[i].type == RD_LIGHT_INFINITE) s->env_textures.push_back(d->lights[i].tex);
    std::vector<DevLight> lights(d->num_lights);
    std::vector<DevEnv> envs;
    std::vector<EnvTables> env_tables;
    for (int i = 0; i < d->num_lights; i++) {
        const rd_light_desc& l = d->lights[i];
        DevLight& o = lights[i];
        std::memset(&o, 0, sizeof(o));
        o.type = l.type;
        if (l.type == RD_LIGHT_INFINITE) {
            // InfiniteAreaLight ctor + InfiniteLight.Preprocess (InfiniteAreaLight.cs:21-54,
            // Light.cs:181-187): tables on the host, bounding sphere of the BVH root
            o.prim = -1;
            o.tex = (int)envs.size();
            DevEnv e;
            std::memset(&e, 0, sizeof(e));
            e.tex = l.tex;
            float inv[16];
            invert4(l.to_world, inv);
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) { e.m2w[3 * r + c] = l.to_world[4 * r + c]; e.w2m[3 * r + c] = inv[4 * r + c]; }
            NodeRec root{};
            if (!bvh.nodes.empty()) root = bvh.nodes[0];
            f3 mn = mk(root.mn[0], root.mn[1], root.mn[2]), mx = mk(root.mx[0], root.mx[1], root.mx[2]);
            f3 c = (mx + mn) * 0.5f;
            e.center[0] = c.x; e.center[1] = c.y; e.center[2] = c.z;
            e.radius = length(mx - mn) / 2.0f;
            env_tables.emplace_back();
            build_env_tables(d->textures[l.tex], &env_tables.back());
            e.w = env_tables.back().w;
            e.h = env_tables.back().h;
            e.marg_int = env_tables.back().marg_int;
            envs.push_back(e);
            continue;
        }
        o.prim = l.prim; o.tex = l.tex;
        for (int k = 0; k < 3; k++) o.le[k] = l.le[k];
        o.area = info[l.prim].area;
    }
    std::vector<DevMedium> media(d->num_media);
    for (int i = 0; i < d->num_media; i++) {
        const rd_medium_desc& m = d->media[i];
        for (int k = 0; k < 3; k++) {
            media[i].sa[k] = m.sigma_a[k] * m.scale;
            media[i].ss[k] = m.sigma_s[k] * m.scale;
            media[i].st[k] = media[i].sa[k] + media[i].ss[k];
        }
        media[i].g = m.g;
    }

    DevScene& ds = s->dev;
    if (!setup_camera(d->camera, &ds.cam, &s->lt_cam)) { delete s; return bad("camera matrix not invertible"); }
    ds.num_nodes = (int)(nodes.size() / 2);
    ds.num_prims = np;
    // a set's traversal runs on top of the scene tree's stack entries
    ds.stack_needed = use_octree ? s->oct.stack_needed : bvh.max_depth + (nsets > 0 ? blas_depth : 0);
    ds.octree = use_octree ? 1 : 0;
    ds.num_insts = d->num_instances;
    ds.num_lights = d->num_lights;
    ds.global_medium = d->global_medium;
    ds.mat_mask = 0;
    for (const DevMaterial& m : mats) ds.mat_mask |= mat_bit(m.type);
    int rc = RD_OK;
    if (!rc) rc = upload(s, nodes.data(), nodes.size(), &ds.nodes);
    if (!rc) rc = upload(s, prims.data(), prims.size(), &ds.prims);
    if (!rc) rc = upload(s, info.data(), info.size(), &ds.prim_info);
    if (!rc) rc = upload(s, d->indices, 3 * (size_t)d->num_triangles, &ds.tri_idx);
    if (!rc) rc = upload(s, d->positions, 3 * (size_t)d->num_vertices, &ds.pos);
    if (!rc && d->normals) rc = upload(s, d->normals, 3 * (size_t)d->num_vertices, &ds.nrm);
    if (!rc && d->uvs) rc = upload(s, d->uvs, 2 * (size_t)d->num_vertices, &ds.uv);
    if (!rc) rc = upload(s, spheres.data(), spheres.size(), &ds.spheres);
    if (!rc && d->sphere_w2m && d->num_spheres > 0) rc = upload(s, d->sphere_w2m, 9 * (size_t)d->num_spheres, &ds.sphere_w2m);
    if (!rc) rc = upload(s, mats.data(), mats.size(), &ds.mats);
    if (!rc) rc = upload(s, texs.data(), texs.size(), &ds.texs);
    if (!rc) rc = upload(s, texels.data(), texels.size(), &ds.texels);
    if (!rc) rc = upload(s, lights.data(), lights.size(), &ds.lights);
    if (!rc) rc = upload(s, media.data(), media.size(), &ds.media);
    if (!rc) rc = upload(s, insts.data(), insts.size(), &ds.insts);
    for (size_t i = 0; i < envs.size() && !rc; i++) {
        const EnvTables& t = env_tables[i];
        rc = upload(s, t.cond_func.data(), t.cond_func.size(), &envs[i].cond_func);
        if (!rc) rc = upload(s, t.cond_cdf.data(), t.cond_cdf.size(), &envs[i].cond_cdf);
        if (!rc) rc = upload(s, t.marg_func.data(), t.marg_func.size(), &envs[i].marg_func);
        if (!rc) rc = upload(s, t.marg_cdf.data(), t.marg_cdf.size(), &envs[i].marg_cdf);
    }
    ds.num_envs = (int)envs.size();
    if (!rc) rc = upload(s, envs.data(), envs.size(), &ds.envs);
    if (rc) { delete s; return nullptr; }

    const rd_integrator_desc& I = d->integrator;
    ds.integ.type = I.type;
    ds.integ.max_depth = I.max_depth;
    ds.integ.min_depth = I.min_depth;
    ds.integ.method = I.method;
    ds.integ.strategy = I.strategy;
    ds.integ.rr_threshold = I.rr_threshold;
    s->integ.type = I.type;
    s->integ.max_depth = I.max_depth;
    s->integ.min_depth = I.min_depth;
    s->integ.method = I.method;
    s->integ.strategy = I.strategy;
    s->integ.rr_threshold = I.rr_threshold;

    s->info.num_nodes = (int)bvh.nodes.size();   // the scene tree (rd_scene_export_bvh)
    s->info.num_leaves = bvh.leaves;
    s->info.max_depth = bvh.max_depth;
    s->info.num_prims = np;
    s->info.build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    uint64_t bytes = nodes.size() * 16 + prims.size() * 16 + info.size() * sizeof(DevPrim) +
                     12ull * d->num_triangles + 12ull * d->num_vertices + texels.size() * 4;
    s->info.device_bytes = bytes;
    return s;
}

void rd_scene_destroy(rd_scene* s) {
    if (!s) return;
    rd_ctx* c = s->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    delete s;
    if (c->live_scenes == 0 && c->closed) free_ctx(c);
}

int rd_scene_get_info(const rd_scene* s, rd_scene_info* info) {
    if (!s || !info) return fail(RD_ERR_INVALID, "null argument");
    *info = s->info;
    return RD_OK;
}

int rd_scene_export_bvh(const rd_scene* s, void* nodes32, int32_t max_nodes, int32_t* prim_order, int32_t max_prims) {
    if (!s) return fail(RD_ERR_INVALID, "null scene");
    if (s->dev.octree) return fail(RD_ERR_INVALID, "the scene's accelerator is the octree (rd_scene_export_octree)");
    if (nodes32) {
        if (max_nodes < (int32_t)s->nodes.size()) return fail(RD_ERR_INVALID, "node buffer too small");
        std::memcpy(nodes32, s->nodes.data(), s->nodes.size() * sizeof(NodeRec));
    }
    if (prim_order) {
        if (max_prims < (int32_t)s->order.size()) return fail(RD_ERR_INVALID, "order buffer too small");
        std::memcpy(prim_order, s->order.data(), s->order.size() * sizeof(int32_t));
    }
    return RD_OK;
}

// Copies an octree into the caller's buffers (sizes first; a null or short
// buffer only sizes).
static int copy_octree(const OctResult& o, void* nodes32, int32_t max_nodes, int32_t* leaf_first, int32_t* leaf_count,
                       int32_t max_leaves, int32_t* leaf_prims, int32_t max_leaf_prims, int32_t* sizes3) {
    if (sizes3) {
        sizes3[0] = (int32_t)o.nodes.size();
        sizes3[1] = (int32_t)o.leaf_first.size();
        sizes3[2] = (int32_t)o.leaf_prims.size();
    }
    if (!nodes32 || !leaf_first || !leaf_count || !leaf_prims) return RD_OK;
    if (max_nodes < (int32_t)o.nodes.size() || max_leaves < (int32_t)o.leaf_first.size() ||
        max_leaf_prims < (int32_t)o.leaf_prims.size())
        return fail(RD_ERR_INVALID, "octree buffers too small");
    std::memcpy(nodes32, o.nodes.data(), o.nodes.size() * sizeof(OctNode));
    std::memcpy(leaf_first, o.leaf_first.data(), o.leaf_first.size() * sizeof(int32_t));
    std::memcpy(leaf_count, o.leaf_count.data(), o.leaf_count.size() * sizeof(int32_t));
    std::memcpy(leaf_prims, o.leaf_prims.data(), o.leaf_prims.size() * sizeof(int32_t));
    return RD_OK;
}

int rd_octree_build_host(const float* bmin, const float* bmax, int32_t n, int32_t max_depth, int32_t max_count,
                         float out_bound, void* nodes32, int32_t max_nodes, int32_t* leaf_first, int32_t* leaf_count,
                         int32_t max_leaves, int32_t* leaf_prims, int32_t max_leaf_prims, int32_t* sizes3) {
    if (n < 0 || (n > 0 && (!bmin || !bmax))) return fail(RD_ERR_INVALID, "bad bounds");
    OctResult o;
    std::string err;
    if (!build_octree(n, bmin, bmax, max_depth, max_count, out_bound, &o, &err)) return fail(RD_ERR_INVALID, err);
    return copy_octree(o, nodes32, max_nodes, leaf_first, leaf_count, max_leaves, leaf_prims, max_leaf_prims, sizes3);
}

int rd_scene_export_octree(const rd_scene* s, void* nodes32, int32_t max_nodes, int32_t* leaf_first,
                           int32_t* leaf_count, int32_t max_leaves, int32_t* leaf_prims, int32_t max_leaf_prims,
                           int32_t* sizes3) {
    if (!s) return fail(RD_ERR_INVALID, "null scene");
    if (!s->dev.octree) return fail(RD_ERR_INVALID, "the scene's accelerator is the BVH");
    return copy_octree(s->oct, nodes32, max_nodes, leaf_first, leaf_count, max_leaves, leaf_prims, max_leaf_prims,
                       sizes3);
}

int rd_bvh_build_host(const float* bmin, const float* bmax, int32_t n, int32_t max_prim, void* nodes32,
                      int32_t max_nodes, int32_t* prim_order, int32_t* num_nodes, int32_t* max_depth) {
    if (n < 0 || (n > 0 && (!bmin || !bmax))) return fail(RD_ERR_INVALID, "bad argument");
    BvhResult r;
    std::string err;
    if (!build_bvh(n, bmin, bmax, max_prim, &r, &err)) return fail(RD_ERR_INVALID, err);
    if (num_nodes) *num_nodes = (int32_t)r.nodes.size();
    if (max_depth) *max_depth = r.max_depth;
    if (nodes32) {
        if (max_nodes < (int32_t)r.nodes.size()) return fail(RD_ERR_INVALID, "node buffer too small");
        std::memcpy(nodes32, r.nodes.data(), r.nodes.size() * sizeof(NodeRec));
    }
    if (prim_order) std::memcpy(prim_order, r.order.data(), r.order.size() * sizeof(int32_t));
    return RD_OK;
}

int rd_bvh_build_device(rd_ctx* c, const float* d_bmin, const float* d_bmax, int32_t n, int32_t max_prim,
                        void* d_nodes32, int32_t max_nodes, int32_t* d_prim_order, int32_t* num_nodes,
                        int32_t* max_depth) {
    if (!c || (n > 0 && (!d_bmin || !d_bmax || !d_nodes32 || !d_prim_order)) || !num_nodes || !max_depth)
        return fail(RD_ERR_INVALID, "null argument");
    if (n < 0) return fail(RD_ERR_INVALID, "negative primitive count");
    if (n > 0 && (int64_t)max_nodes < 2 * (int64_t)n - 1) return fail(RD_ERR_INVALID, "max_nodes below 2n - 1");
    CtxLock lock(c);
    if (!lock.ok()) return fail(RD_ERR_BUSY, "context in use by another thread");
    (void)hipSetDevice(c->device);
    std::string err;
    int depth = 0;
    const int nn = gpu_build_bvh_device(n, d_bmin, d_bmax, max_prim, (NodeRec*)d_nodes32, d_prim_order, c->stream,
                                        &depth, &err);
    if (nn < 0) return fail(RD_ERR_INVALID, err);
    *num_nodes = nn;
    *max_depth = depth;
    return RD_OK;
}

int rd_trace_closest_device(rd_ctx* c, rd_scene* s, const rd_ray* d_rays, int32_t n, rd_hit* d_hits,
                            uint64_t* d_counters, void* stream) {
    if (!c || !s || n < 0 || (n > 0 && (!d_rays || !d_hits))) return fail(RD_ERR_INVALID, "bad argument");
    if (int rc = check_stack(s)) return rc;
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    Timer tm(c);
    HIP_TRY(tm.run("trace_rays_closest", [&] {
        return launch_trace_rays(s->dev, (const float4*)d_rays, n, (float4*)d_hits, nullptr, nullptr,
                                 (unsigned long long*)d_counters, false, c->counts, st);
    }));
    if (!stream) {
        HIP_TRY(hipStreamSynchronize(st));
        tm.collect();
    }
    return RD_OK;
}

int rd_trace_any_device(rd_ctx* c, rd_scene* s, const rd_ray* d_rays, int32_t n, uint8_t* d_occ,
                        uint64_t* d_counters, void* stream) {
    if (!c || !s || n < 0 || (n > 0 && (!d_rays || !d_occ))) return fail(RD_ERR_INVALID, "bad argument");
    if (int rc = check_stack(s)) return rc;
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    Timer tm(c);
    HIP_TRY(tm.run("trace_rays_any", [&] {
        return launch_trace_rays(s->dev, (const float4*)d_rays, n, nullptr, d_occ, nullptr,
                                 (unsigned long long*)d_counters, true, c->counts, st);
    }));
    if (!stream) {
        HIP_TRY(hipStreamSynchronize(st));
        tm.collect();
    }
    return RD_OK;
}

int rd_trace_closest(rd_ctx* c, rd_scene* s, const rd_ray* rays, int32_t n, rd_hit* hits) {
    if (!c || !s || n < 0 || (n > 0 && (!rays || !hits))) return fail(RD_ERR_INVALID, "bad argument");
    if (n == 0) return RD_OK;
    CtxLock lock(c);
    if (!lock.ok()) return fail(RD_ERR_BUSY, "context in use by another thread");
    (void)hipSetDevice(c->device);
    void *dr = nullptr, *dh = nullptr;
    if (hipMalloc(&dr, sizeof(rd_ray) * (size_t)n) != hipSuccess) return fail(RD_ERR_OOM, "hipMalloc rays");
    if (hipMalloc(&dh, sizeof(rd_hit) * (size_t)n) != hipSuccess) { (void)hipFree(dr); return fail(RD_ERR_OOM, "hipMalloc hits"); }
    int rc = RD_OK;
    if (hipMemcpyAsync(dr, rays, sizeof(rd_ray) * (size_t)n, hipMemcpyHostToDevice, c->stream) != hipSuccess) rc = fail(RD_ERR_DEVICE, "copy rays");
    if (!rc) rc = rd_trace_closest_device(c, s, (const rd_ray*)dr, n, (rd_hit*)dh, nullptr, nullptr);
    if (!rc && hipMemcpy(hits, dh, sizeof(rd_hit) * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess) rc = fail(RD_ERR_DEVICE, "copy hits");
    (void)hipFree(dr);
    (void)hipFree(dh);
    return rc;
}

int rd_trace_closest_sub(rd_ctx* c, rd_scene* s, const rd_ray* rays, int32_t n, rd_hit* hits, int32_t* sub) {
    if (!c || !s || n < 0 || (n > 0 && (!rays || !hits || !sub))) return fail(RD_ERR_INVALID, "bad argument");
    if (n == 0) return RD_OK;
    if (int rc = check_stack(s)) return rc;
    CtxLock lock(c);
    if (!lock.ok()) return fail(RD_ERR_BUSY, "context in use by another thread");
    (void)hipSetDevice(c->device);
    void *dr = nullptr, *dh = nullptr, *ds = nullptr;
    if (hipMalloc(&dr, sizeof(rd_ray) * (size_t)n) != hipSuccess) return fail(RD_ERR_OOM, "hipMalloc rays");
    int rc = RD_OK;
    if (hipMalloc(&dh, sizeof(rd_hit) * (size_t)n) != hipSuccess) rc = fail(RD_ERR_OOM, "hipMalloc hits");
    if (!rc && hipMalloc(&ds, sizeof(int32_t) * (size_t)n) != hipSuccess) rc = fail(RD_ERR_OOM, "hipMalloc sub");
    hipStream_t st = c->stream;
    if (!rc && hipMemcpyAsync(dr, rays, sizeof(rd_ray) * (size_t)n, hipMemcpyHostToDevice, st) != hipSuccess)
        rc = fail(RD_ERR_DEVICE, "copy rays");
    if (!rc && hipMemsetAsync(ds, 0xff, sizeof(int32_t) * (size_t)n, st) != hipSuccess) rc = fail(RD_ERR_DEVICE, "clear sub");
    if (!rc && launch_trace_rays(s->dev, (const float4*)dr, n, (float4*)dh, nullptr, (int*)ds, nullptr, false, c->counts,
                                 st) != hipSuccess)
        rc = fail(RD_ERR_DEVICE, "trace launch");
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = fail(RD_ERR_DEVICE, "trace");
    if (!rc && hipMemcpy(hits, dh, sizeof(rd_hit) * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess) rc = fail(RD_ERR_DEVICE, "copy hits");
    if (!rc && hipMemcpy(sub, ds, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess) rc = fail(RD_ERR_DEVICE, "copy sub");
    // DevPrim index of a set triangle -> its index within the set
    for (int32_t i = 0; i < n && !rc; i++) {
        int g = sub[i];
        if (g < 0) { sub[i] = -1; continue; }
        int k = g - s->num_top_prims;
        auto it = std::upper_bound(s->set_tri_base.begin(), s->set_tri_base.end(), k);
        sub[i] = k - *(it - 1);
    }
    (void)hipFree(dr);
    (void)hipFree(dh);
    (void)hipFree(ds);
    return rc;
}

int rd_trace_any(rd_ctx* c, rd_scene* s, const rd_ray* rays, int32_t n, uint8_t* occ) {
    if (!c || !s || n < 0 || (n > 0 && (!rays || !occ))) return fail(RD_ERR_INVALID, "bad argument");
    if (n == 0) return RD_OK;
    CtxLock lock(c);
    if (!lock.ok()) return fail(RD_ERR_BUSY, "context in use by another thread");
    (void)hipSetDevice(c->device);
    void *dr = nullptr, *dh = nullptr;
    if (hipMalloc(&dr, sizeof(rd_ray) * (size_t)n) != hipSuccess) return fail(RD_ERR_OOM, "hipMalloc rays");
    if (hipMalloc(&dh, (size_t)n) != hipSuccess) { (void)hipFree(dr); return fail(RD_ERR_OOM, "hipMalloc occluded"); }
    int rc = RD_OK;
    if (hipMemcpyAsync(dr, rays, sizeof(rd_ray) * (size_t)n, hipMemcpyHostToDevice, c->stream) != hipSuccess) rc = fail(RD_ERR_DEVICE, "copy rays");
    if (!rc) rc = rd_trace_any_device(c, s, (const rd_ray*)dr, n, (uint8_t*)dh, nullptr, nullptr);
    if (!rc && hipMemcpy(occ, dh, (size_t)n, hipMemcpyDeviceToHost) != hipSuccess) rc = fail(RD_ERR_DEVICE, "copy occluded");
    (void)hipFree(dr);
    (void)hipFree(dh);
    return rc;
}

namespace {

// What one pass's bounce loop launches: a single scene (kernel argument) or a
// batch (device array of scenes, the path's request selects one); `sc` is
// also the traversal variant's reference for the LDS stack depth.
struct PassLaunch {
    DevScene sc;
    const DevScene* scenes;
    const AdjArgs* adj;   // adjoint replay (single scene, path integrator), else null
    int method;
    bool vol;
    const BatchSeg* segs;
    int nseg;
    int npass;     // paths in the pass
    int pool;      // path slots in flight
    WaveBuffers* wb = nullptr;   // the pass's buffers (null: the context's own)
    bool sort_ok = true;         // the coherence sort may run (its buffers are allocated on demand)
    bool graph_ok = true;        // the device-driven tail may be captured as a hipGraph
    bool sss = false;            // path Mis over a BSSRDF: each bounce runs the BSSRDF stage
    const float* sss_lut = nullptr;
};

// The BSSRDF stage of a bounce (PathTracer.cs:244-259), after the bounce's
// resolve and before the compaction that reads `alive`: the probe for every
// path k_shade queued, pi's direct light and continuation, their shadow / MIS
// rays and resolve; then the two queues are emptied for the next bounce.
hipError_t sss_phase(rd_ctx* c, const PassLaunch& L, WaveBuffers& wb, hipStream_t st) {
    hipError_t e;
    if ((e = launch_sss_probe(L.sc, wb.ps, c->counts, L.sss_lut, st))) return e;
    if ((e = launch_sss_shade(L.sc, wb.qshadow, wb.qmis, wb.ps, c->counts, st))) return e;
    if ((e = launch_trace_secondary(L.sc, L.scenes, wb.qshadow, kCntShadow, wb.qmis, wb.ps, c->counts,
                                    c->trace_schedule[1], c->trace_schedule[2], st)))
        return e;
    RayQueue q2{nullptr, wb.ps.sss_q2, nullptr, nullptr};
    if ((e = launch_resolve(L.sc, L.method, q2, wb.ps, c->counts, kCntSssB, nullptr, st))) return e;
    return launch_zero_slots(c->counts, kCntSss, kCntSssB, st);
}

// Scenes with fewer BVH nodes than this (a 128 KB tree) trace unsorted.
constexpr int kSortMinNodes = 1 << 12;

// One bounce of the device-driven tail: the launches of run_pass's bounce
// once every path has started (the same stages and streams, the shadow / MIS
// traversal and resolve beside the next bounce's compaction and extension
// traversal on stream2), without the sort, and with every kernel reading its
// queue size on the device; the grids of the compactions are sized for
// `bound` rays (once every path of the pass has started a queue only
// shrinks).  *pending: this bounce's resolve is still queued on stream2.
hipError_t tail_bounce(rd_ctx* c, const PassLaunch& L, WaveBuffers& wb, RayQueue qin, RayQueue qout, int cin, int cout,
                       int bound, bool* pending) {
    hipStream_t st = c->stream;
    hipError_t e;
#define TAIL_TRY(x) do { if ((e = (x)) != hipSuccess) return e; } while (0)
    TAIL_TRY(launch_tail_tick(c->counts, cin, st));
    const int za = *pending ? -1 : cout, zb = *pending ? -1 : kCntShadow, zc = *pending ? -1 : kCntMis;
    TAIL_TRY(launch_trace_ext(L.sc, L.scenes, qin, wb.ps, wb.hits, c->counts, cin, za, zb, zc, nullptr, nullptr,
                              c->trace_schedule[0], st));
    if (L.vol) {
        TAIL_TRY(launch_shade_vol(L.sc, L.scenes, qin, wb.hits, wb.qshadow, wb.qmis, wb.ps, c->counts, cin, st));
        TAIL_TRY(launch_chain_full(L.sc, L.scenes, wb.qshadow, wb.ps, false, c->counts, kCntShadow, st));
        TAIL_TRY(launch_chain_full(L.sc, L.scenes, wb.qmis, wb.ps, true, c->counts, kCntMis, st));
        TAIL_TRY(launch_resolve_vol(qin, wb.ps, c->counts, cin, st));
    } else {
        if (*pending) {   // shade adds to L, which the last bounce's resolve also writes
            TAIL_TRY(hipStreamWaitEvent(st, c->ev_resolved, 0));
            *pending = false;
        }
        const bool ranked = L.method != RD_METHOD_BXDF;
        TAIL_TRY(launch_shade(L.sc, L.scenes, qin, wb.hits, wb.qshadow, wb.qmis, wb.ps, c->counts, cin, false, nullptr,
                              ranked, st));
        const bool overlap = c->overlap && !L.sss;
        hipStream_t s2 = overlap ? c->stream2 : st;
        if (overlap) {
            TAIL_TRY(hipEventRecord(c->ev_shaded, st));
            TAIL_TRY(hipStreamWaitEvent(s2, c->ev_shaded, 0));
        }
        RayQueue qsh = wb.qshadow;
        int csh = kCntShadow;
        if (ranked) {
            TAIL_TRY(launch_compact_holes(wb.qshadow, bound, overlap ? wb.blk2 : wb.blk, wb.qshadow2, kCntShadowB,
                                          c->counts, s2, c->counts + cin));
            qsh = wb.qshadow2;
            csh = kCntShadowB;
        }
        if (L.method == RD_METHOD_MIS)
            TAIL_TRY(launch_trace_secondary(L.sc, L.scenes, qsh, csh, wb.qmis, wb.ps, c->counts, c->trace_schedule[1],
                                            c->trace_schedule[2], s2));
        else if (L.method == RD_METHOD_NEE)
            TAIL_TRY(launch_trace_shadow(L.sc, L.scenes, qsh, wb.ps, c->counts, csh, c->trace_schedule[1], s2));
        TAIL_TRY(launch_resolve(L.sc, L.method, qin, wb.ps, c->counts, cin, nullptr, s2));
        if (L.sss) TAIL_TRY(sss_phase(c, L, wb, st));
        if (overlap) {
            TAIL_TRY(launch_zero_slots(c->counts, kCntShadow, kCntMis, s2));
            TAIL_TRY(hipEventRecord(c->ev_resolved, s2));
            *pending = true;
        }
    }
    TAIL_TRY(launch_compact_queue(qin, bound, L.npass, wb.blk, qout, cout, wb.ps, c->counts, st, c->counts + cin));
#undef TAIL_TRY
    return hipSuccess;
}

// The rest of a pass once every path has started and the queue holds at most
// c->tail_rays rays, in groups of kTailBounces bounces (an even number, so
// the queue roles qin -> qout -> qin repeat from group to group) queued
// without a host synchronization; the host reads the queue size once per
// group.  The first group runs as plain launches; the next ones replay it as
// a hipGraph (device_loop 2: captured once and kept on the context for
// passes that would record the same launches) or run as plain launches
// again (1).  Bounces past the end of the pass run on empty queues and do
// nothing.
constexpr int kTailBounces = 8;
constexpr int kMaxTailGroups = 1 << 16;

extern "C++" {
template <class T>
static void key_add(std::string& k, const T& v) {
    static_assert(std::is_trivially_copyable<T>::value, "key fields are plain values");
    k.append(reinterpret_cast<const char*>(&v), sizeof(T));
}
}

int run_tail(rd_ctx* c, Timer& tm, const PassLaunch& L, WaveBuffers& wb, RayQueue qin, RayQueue qout, int cin,
             int cout, int bound, int first_bounce, int* max_bounce, volatile int32_t* cancel) {
    hipStream_t st = c->stream;
    HIP_TRY(launch_zero_slots(c->counts, kCntTail, kCntTail, st));
    auto group = [&]() -> hipError_t {
        bool pending = false;
        for (int i = 0; i < kTailBounces; i += 2) {
            hipError_t e = tail_bounce(c, L, wb, qin, qout, cin, cout, bound, &pending);
            if (e == hipSuccess) e = tail_bounce(c, L, wb, qout, qin, cout, cin, bound, &pending);
            if (e != hipSuccess) return e;
        }
        return pending ? hipStreamWaitEvent(st, c->ev_resolved, 0) : hipSuccess;
    };
    auto read_counts = [&]() -> hipError_t {
        hipError_t e = hipMemcpyAsync(c->host_counts + cin, c->counts + cin, sizeof(int), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(c->host_counts + kCntTail, c->counts + kCntTail, sizeof(int), hipMemcpyDeviceToHost, st);
        return e == hipSuccess ? hipStreamSynchronize(st) : e;
    };
    int rc = RD_OK;
    hipError_t e = tm.run("tail", group);
    if (e == hipSuccess) e = read_counts();
    if (e != hipSuccess) rc = fail(RD_ERR_DEVICE, std::string("device-driven tail: ") + hipGetErrorString(e));
    if (!rc && c->host_counts[cin] > 0 && c->device_loop == 2 && L.graph_ok) {
        // every value the group's launches are recorded with
        std::string key;
        key_add(key, L.sc); key_add(key, L.scenes); key_add(key, L.method); key_add(key, L.vol);
        key_add(key, L.npass); key_add(key, bound); key_add(key, wb.ps); key_add(key, wb.hits);
        key_add(key, wb.qshadow); key_add(key, wb.qshadow2); key_add(key, wb.qmis); key_add(key, wb.blk);
        key_add(key, wb.blk2); key_add(key, qin); key_add(key, qout); key_add(key, cin); key_add(key, cout);
        key_add(key, c->trace_schedule); key_add(key, c->overlap); key_add(key, c->counts);
        key_add(key, L.sss); key_add(key, L.sss_lut);
        if (!c->tail_exec || key != c->tail_key) {
            if (c->tail_exec) (void)hipGraphExecDestroy(c->tail_exec);
            c->tail_exec = nullptr;
            c->tail_key.clear();
            hipGraph_t graph = nullptr;
            e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                hipError_t e1 = group();
                hipError_t e2 = hipStreamEndCapture(st, &graph);
                e = e1 != hipSuccess ? e1 : e2;
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&c->tail_exec, graph, nullptr, nullptr, 0);
            if (graph) (void)hipGraphDestroy(graph);
            if (e == hipSuccess) c->tail_key = key;
            else {
                c->tail_exec = nullptr;
                rc = fail(RD_ERR_DEVICE, std::string("tail graph capture: ") + hipGetErrorString(e));
            }
        }
    }
    for (int g = 1; !rc && c->host_counts[cin] > 0; g++) {
        if (cancel && *cancel) { rc = fail(RD_ERR_CANCELLED, "render cancelled"); break; }
        if (g >= kMaxTailGroups) { rc = fail(RD_ERR_DEVICE, "pass did not end"); break; }
        const bool graph = c->device_loop == 2 && L.graph_ok;
        e = tm.run("tail", [&] { return graph ? hipGraphLaunch(c->tail_exec, st) : group(); });
        if (e == hipSuccess) e = read_counts();
        if (e != hipSuccess) rc = fail(RD_ERR_DEVICE, std::string("device-driven tail: ") + hipGetErrorString(e));
    }
    *max_bounce = std::max(*max_bounce, first_bounce + c->host_counts[kCntTail] - 1);
    return rc;
}

// The bounce loop of one pass: every bounce traces and shades the queue,
// retires finished paths, then compacts the pool into the next queue (slot
// order) while starting the pass's next paths in the free slots (once all
// have started, by filtering the last queue, which gives the same queue);
// one sync per bounce reads the next queue size (for vol_path also the chain
// queue sizes).  Ends when the queue is empty (the pass has run dry).
int run_pass(rd_ctx* c, Timer& tm, const PassLaunch& L, int* max_bounce, volatile int32_t* cancel) {
    WaveBuffers& wb = L.wb ? *L.wb : c->wb;
    hipStream_t st = c->stream;
    RayQueue qin = wb.qa, qout = wb.qb;
    int cin = kCntExtA, cout = kCntExtB;
    HIP_TRY(launch_init_pool(wb.ps, L.pool, c->counts, st));
    HIP_TRY(tm.run("compact", [&] {
        return launch_compact(L.sc, L.scenes, L.segs, L.nseg, L.npass, L.pool, wb.blk, qin, cin, wb.ps, c->counts, st);
    }));
    // the first queue's size: the first bounce's queue filter and shadow-hole
    // compaction are sized on the host
    HIP_TRY(hipMemcpyAsync(c->host_counts + cin, c->counts + cin, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    int rc = RD_OK;
    // known on the host after each bounce's compaction; the first one starts
    // a path in every slot (all are free), so it starts them all when the
    // pass fits the pool
    bool all_started = L.npass <= L.pool;
    bool resolve_pending = false;   // the last bounce's resolve is still queued on stream2
    for (int bounce = 0; !rc; bounce++) {
        // the queue size is known on the host from the second bounce on
        const int* perm = nullptr;
        // the coherence sort pays where traversal is memory-bound; a tree this
        // small stays in the caches whatever the ray order (the 16-request
        // vol_path batch's largest request: 610 ms sorted, 571 ms unsorted)
        if (L.sort_ok && c->ray_sort && bounce > 0 && c->host_counts[cin] >= c->ray_sort_min &&
            L.sc.num_nodes >= kSortMinNodes) {
            if (int e = ensure_sort_buffers(c, (size_t)L.pool)) return e;
            HIP_TRY(tm.run("sort", [&] {
                return launch_ray_sort(L.sc.nodes, qin, c->host_counts[cin], c->ray_sort_bits, wb.sort_keys,
                                       wb.sort_keys_alt,
                                       wb.sort_vals, wb.sort_perm, wb.sort_temp, wb.sort_temp_bytes, st);
            }));
            perm = wb.sort_perm;
        }
        // the counters the next stages start from; after an overlapped bounce
        // stream2 clears the shadow / MIS ones behind its resolve, which
        // still reads this queue's old count slot
        const int za = resolve_pending ? -1 : cout, zb = resolve_pending ? -1 : kCntShadow;
        const int zc = resolve_pending ? -1 : kCntMis;
        HIP_TRY(tm.run("trace_closest", [&] {
            return launch_trace_ext(L.sc, L.scenes, qin, wb.ps, wb.hits, c->counts, cin, za, zb, zc, perm,
                                    wb.sort_rank, c->trace_schedule[0], st);
        }));
        // the shadow rays follow the order the extension rays were traced in
        // (the sorted order's ranks; ray_sort_shadow 0: the queue order
        // whatever the sort), or the queue order (tile-major path order at
        // the first bounce) where there was no sort
        const int* rank = (perm && c->ray_sort_shadow) ? wb.sort_rank : nullptr;
        const bool rank_identity = !rank && L.method != RD_METHOD_BXDF;
        if (L.vol) {
            // VolumetricPathTracer.Li: shade, then the Transmittance / IntersectTr
            // chains step by step until both chain queues drain
            HIP_TRY(tm.run("shade", [&] {
                return launch_shade_vol(L.sc, L.scenes, qin, wb.hits, wb.qshadow, wb.qmis, wb.ps, c->counts, cin, st);
            }));
            RayQueue sa = wb.qshadow, sb = wb.qshadow2, ma = wb.qmis, mb = wb.qmis2;
            int csa = kCntShadow, csb = kCntShadowB, cma = kCntMis, cmb = kCntMisB;
            if (!c->chain_lockstep) {
                // every chain walked to its end by one lane: no host loop
                HIP_TRY(tm.run("trace_chain", [&] {
                    return launch_chain_full(L.sc, L.scenes, sa, wb.ps, false, c->counts, csa, st);
                }));
                HIP_TRY(tm.run("trace_chain", [&] {
                    return launch_chain_full(L.sc, L.scenes, ma, wb.ps, true, c->counts, cma, st);
                }));
            }
            for (int step = 0; c->chain_lockstep; step++) {
                if (step >= kMaxChainSteps) return fail(RD_ERR_DEVICE, "medium chain did not terminate");
                HIP_TRY(tm.run("trace_chain", [&] {
                    return launch_trace_chain(L.sc, L.scenes, sa, wb.ps, wb.hits, false, c->counts, csa, csb,
                                              c->trace_schedule[1], st);
                }));
                HIP_TRY(tm.run("trace_chain", [&] {
                    return launch_trace_chain(L.sc, L.scenes, ma, wb.ps, wb.hits2, true, c->counts, cma, cmb,
                                              c->trace_schedule[2], st);
                }));
                HIP_TRY(tm.run("chain_advance", [&] {
                    return launch_chain_advance(L.sc, L.scenes, sa, wb.hits, sb, false, wb.ps, c->counts, csa, csb, st);
                }));
                HIP_TRY(tm.run("chain_advance", [&] {
                    return launch_chain_advance(L.sc, L.scenes, ma, wb.hits2, mb, true, wb.ps, c->counts, cma, cmb, st);
                }));
                HIP_TRY(hipMemcpyAsync(c->host_counts + kCntShadowB, c->counts + kCntShadowB, 2 * sizeof(int),
                                       hipMemcpyDeviceToHost, st));
                HIP_TRY(hipMemcpyAsync(c->host_counts + kCntShadow, c->counts + kCntShadow, 2 * sizeof(int),
                                       hipMemcpyDeviceToHost, st));
                HIP_TRY(hipStreamSynchronize(st));
                if (c->host_counts[csb] == 0 && c->host_counts[cmb] == 0) break;
                std::swap(sa, sb); std::swap(csa, csb);
                std::swap(ma, mb); std::swap(cma, cmb);
            }
            HIP_TRY(tm.run("resolve", [&] { return launch_resolve_vol(qin, wb.ps, c->counts, cin, st); }));
        } else {
            // shade adds to L, which the last bounce's resolve also writes
            if (resolve_pending) {
                HIP_TRY(hipStreamWaitEvent(st, c->ev_resolved, 0));
                resolve_pending = false;
            }
            HIP_TRY(tm.run("shade", [&] {
                return launch_shade(L.sc, L.scenes, qin, wb.hits, wb.qshadow, wb.qmis, wb.ps, c->counts, cin,
                                    L.adj != nullptr, rank, rank_identity, st);
            }));
            // Once every path of the pass has started, the next compaction
            // only filters this queue by `alive` (set by shade), so the
            // shadow / MIS traversal and resolve of this bounce (which touch
            // neither `alive` nor the queues the next bounce builds) run on
            // stream2 beside the next bounce's compaction, sort and
            // extension traversal, filling each other's drain tails.
            // (not at the first bounce: its MIS + shadow launch is the
            // frame's longest after the first extension trace, and beside the
            // second bounce's extension trace both lose more to each other
            // than the overlap fills)
            const bool overlap = c->overlap && all_started && bounce > 0 && !L.adj && !L.sss;
            hipStream_t s2 = overlap ? c->stream2 : st;
            if (overlap) {
                HIP_TRY(hipEventRecord(c->ev_shaded, st));
                HIP_TRY(hipStreamWaitEvent(s2, c->ev_shaded, 0));
            }
            // a ranked shadow queue holds a hole for every path without a
            // shadow ray: trace it without them, in the same order
            RayQueue qsh = wb.qshadow;
            int csh = kCntShadow;
            if ((rank || rank_identity) && L.method != RD_METHOD_BXDF) {
                HIP_TRY(tm.run("compact_shadow", [&] {
                    return launch_compact_holes(wb.qshadow, c->host_counts[cin], overlap ? wb.blk2 : wb.blk,
                                                wb.qshadow2, kCntShadowB, c->counts, s2);
                }, s2));
                qsh = wb.qshadow2;
                csh = kCntShadowB;
            }
            if (L.method == RD_METHOD_MIS) {
                HIP_TRY(tm.run("trace_secondary", [&] {
                    return launch_trace_secondary(L.sc, L.scenes, qsh, csh, wb.qmis, wb.ps, c->counts,
                                                  c->trace_schedule[1], c->trace_schedule[2], s2);
                }, s2));
            } else if (L.method == RD_METHOD_NEE) {
                HIP_TRY(tm.run("trace_any", [&] {
                    return launch_trace_shadow(L.sc, L.scenes, qsh, wb.ps, c->counts, csh, c->trace_schedule[1], s2);
                }, s2));
            }
            HIP_TRY(tm.run(L.adj ? "resolve_adjoint" : "resolve", [&] {
                return launch_resolve(L.sc, L.method, qin, wb.ps, c->counts, cin, L.adj, s2);
            }, s2));
            if (L.sss) HIP_TRY(tm.run("sss", [&] { return sss_phase(c, L, wb, st); }));
            if (overlap) {
                HIP_TRY(launch_zero_slots(c->counts, kCntShadow, kCntMis, s2));
                HIP_TRY(hipEventRecord(c->ev_resolved, s2));
                resolve_pending = true;
            }
        }
        HIP_TRY(tm.run("compact", [&] {
            // once every path of the pass has started, only the entries of
            // this queue can continue: filter it instead of scanning the pool
            if (all_started)
                return launch_compact_queue(qin, c->host_counts[cin], L.npass, wb.blk, qout, cout, wb.ps, c->counts,
                                            st);
            return launch_compact(L.sc, L.scenes, L.segs, L.nseg, L.npass, L.pool, wb.blk, qout, cout, wb.ps,
                                  c->counts, st);
        }));
        HIP_TRY(hipMemcpyAsync(c->host_counts + cout, c->counts + cout, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(c->host_counts + kCntNext, c->counts + kCntNext, sizeof(int), hipMemcpyDeviceToHost,
                               st));
        HIP_TRY(hipStreamSynchronize(st));
        all_started = c->host_counts[kCntNext] >= L.npass;
        if (bounce > *max_bounce) *max_bounce = bounce;
        if (c->host_counts[cout] == 0) break;
        if (cancel && *cancel) rc = fail(RD_ERR_CANCELLED, "render cancelled");
        std::swap(qin, qout);
        std::swap(cin, cout);
        if (!rc && c->device_loop > 0 && !L.adj && all_started && c->host_counts[cin] <= c->tail_rays) {
            if (resolve_pending) {
                HIP_TRY(hipStreamWaitEvent(st, c->ev_resolved, 0));
                resolve_pending = false;
            }
            rc = run_tail(c, tm, L, wb, qin, qout, cin, cout, c->host_counts[cin], bounce + 1, max_bounce, cancel);
            break;
        }
    }
    if (resolve_pending) HIP_TRY(hipStreamWaitEvent(st, c->ev_resolved, 0));
    return rc;
}

// Device counters -> rd_render_stats after the render's final synchronize.
void collect_stats(rd_ctx* c, uint64_t paths, float ms, int passes, int max_bounce) {
    auto u64 = [&](int slot) { uint64_t v; std::memcpy(&v, c->host_counts + slot, 8); return v; };
    c->stats.rays_closest = u64(kCntRaysClosest);
    c->stats.rays_any = u64(kCntRaysAny);
    c->stats.vertices = u64(kCntVertices);
    c->stats.invalid = u64(kCntInvalid);
    c->stats.paths = paths;
    c->stats.render_ms = ms;
    c->stats.waves = passes;
    c->stats.bounces = max_bounce;
    for (int k = 0; k < 3; k++) {
        c->stats.class_rays[k] = u64(kCntClassRays + 2 * k);
        c->stats.class_nodes[k] = u64(kCntClassNodes + 2 * k);
        c->stats.class_prims[k] = u64(kCntClassPrims + 2 * k);
    }
}

struct Crop {
    int x0, y0, x1, y1;
};
int resolve_crop(const rd_scene* s, const rd_render_params* p, Crop* out) {
    int x0 = p->x0, y0 = p->y0, x1 = p->x1, y1 = p->y1;
    if (x0 == 0 && y0 == 0 && x1 == 0 && y1 == 0) { x1 = s->width; y1 = s->height; }
    if (x0 < 0 || y0 < 0 || x1 > s->width || y1 > s->height || x0 >= x1 || y0 >= y1)
        return fail(RD_ERR_INVALID, "bad crop");
    *out = Crop{x0, y0, x1, y1};
    return RD_OK;
}

// Paths of one pass (their per-sample results, 16 B each, stay in HBM until
// the pass's accumulation): as many as an eighth of the free device memory
// holds, a power of two in [kMinPass, kMaxPass] (2^29 paths = 8 GiB of
// results).  A pass ends with the drain of its last bounces, so larger passes
// have fewer of those.
constexpr size_t kMinPass = (size_t)1 << 24;
constexpr size_t kMaxPass = (size_t)1 << 29;

size_t pass_limit() {
    size_t free_b = 0, total_b = 0;
    size_t pass = kMinPass;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
        while (pass * 2 <= free_b / 8 / 16 && pass * 2 <= kMaxPass) pass *= 2;
    return pass;
}

// Slot pool when the caller leaves wave_paths at 0: as many paths in flight
// as a quarter of the device memory (and half of what is free) holds at
// kPoolSlotBytes per slot (the path-state, queue and sort arrays of a slot,
// vol_path's included), a power of two in [kMinPool, kMaxPool], never more
// than the render's paths.  Larger pools mean fewer, longer bounce launches,
// and every launch ends with a tail in which the chip drains while its
// longest rays finish: on the 1.31M-triangle 1080p workload 2^22 slots give
// 1323 Mrays/s, 2^27 (every path of the frame at once, 50 GB) 1824.
constexpr size_t kMinPool = (size_t)1 << 22;
constexpr size_t kMaxPool = (size_t)1 << 27;
constexpr size_t kPoolSlotBytes = 512;

size_t default_pool(size_t paths, int cap_log2) {
    size_t free_b = 0, total_b = 0;
    size_t pool = kMinPool;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
        const size_t slots = std::min(total_b / 4, free_b / 2) / kPoolSlotBytes;
        while (pool * 2 <= slots && pool * 2 <= kMaxPool) pool *= 2;
    }
    if (cap_log2 > 0) pool = std::min(pool, (size_t)1 << cap_log2);   // rd_tuning.pool_log2
    return std::min(pool, std::max<size_t>(paths, 1));
}

// Renders requests with device outputs.  single = true: one request rendered
// with the scene as a kernel argument; otherwise requests are grouped by
// integrator kind and their paths share passes, each path reading its
// request's scene.  Passes pack whole samples of requests in request order
// (a pass may hold several requests, a request may span several passes);
// within a pass paths stream through the slot pool.
int render_requests(rd_ctx* c, int32_t n, const rd_request* reqs, bool single, int32_t pool_paths,
                    volatile int32_t* cancel, volatile int32_t* progress, bool noalloc = false,
                    double* out64 = nullptr) {
    std::vector<Crop> crops(n);
    for (int i = 0; i < n; i++) {
        const rd_request& r = reqs[i];
        if (!r.scene || (!r.out_rgb && !(single && out64))) return fail(RD_ERR_INVALID, "request without scene or output");
        if (r.scene->ctx != c) return fail(RD_ERR_INVALID, "request scene belongs to another context");
        if (r.params.spp <= 0) return fail(RD_ERR_INVALID, "spp must be positive");
        if (!single && r.scene->dev.octree)
            return fail(RD_ERR_UNSUPPORTED, "batched rendering covers BVH scenes (an octree scene renders alone)");
        const int it = r.scene->dev.integ.type;
        if (r.scene->sss && it == RD_INTEGRATOR_PATH && r.scene->dev.integ.method == RD_METHOD_MIS) {
            if (!single)
                return fail(RD_ERR_UNSUPPORTED, "batched rendering covers scenes without a BSSRDF (such a scene renders alone)");
            if (noalloc) return fail(RD_ERR_UNSUPPORTED, "rd_render_device_noalloc covers scenes without a BSSRDF");
        }
        if (it != RD_INTEGRATOR_PATH && it != RD_INTEGRATOR_VOL_PATH)
            return fail(RD_ERR_UNSUPPORTED, single ? "this entry point renders path and vol_path scenes"
                                                   : "batched rendering covers path and vol_path scenes");
        if (int rc = check_stack(r.scene)) return rc;
        if (int rc = resolve_crop(r.scene, &r.params, &crops[i])) return rc;
    }
    std::vector<std::vector<int>> groups;
    std::vector<std::pair<int, int>> kinds;
    for (int i = 0; i < n; i++) {
        const DevIntegrator& I = reqs[i].scene->dev.integ;
        std::pair<int, int> k{I.type, I.type == RD_INTEGRATOR_VOL_PATH ? 0 : I.method};
        size_t g = 0;
        while (g < kinds.size() && kinds[g] != k) g++;
        if (g == kinds.size()) { kinds.push_back(k); groups.emplace_back(); }
        groups[g].push_back(i);
    }
    size_t max_npix = 0, total_pix = 0, total_paths = 0;
    for (int i = 0; i < n; i++) {
        size_t np = (size_t)(crops[i].x1 - crops[i].x0) * (crops[i].y1 - crops[i].y0);
        max_npix = std::max(max_npix, np);
        total_pix += np;
        total_paths += np * (size_t)reqs[i].params.spp;
    }
    bool any_vol = false;
    for (auto& k : kinds) any_vol = any_vol || k.first == RD_INTEGRATOR_VOL_PATH;
    size_t pool = 0, pass_cap = 0;
    if (!noalloc) {
        pool = pool_paths > 0 ? std::min((size_t)pool_paths, std::max<size_t>(total_paths, 1))
                              : default_pool(total_paths, c->pool_log2);
        pass_cap = std::max(std::min(pass_limit(), total_paths), max_npix);
    } else {
        // everything from the work area: the sums and one sample's results
        // must fit beside the smallest pool; then the largest pool up to
        // kNoAllocMaxPool (or the caller's cap), then results for the rest
        if (!c->scratch) return fail(RD_ERR_DEVICE, "no work area on this device");
        if (total_pix > kScratchBytes || scratch_need(c, kNoAllocMinPool, max_npix, total_pix, any_vol) > kScratchBytes)
            return fail(RD_ERR_INVALID, "render of " + std::to_string(total_pix) + " pixels does not fit the " +
                                            std::to_string(kScratchBytes >> 20) +
                                            " MiB work area of rd_render_device_noalloc (64 B per pixel)");
        size_t cap = kNoAllocMaxPool;
        if (pool_paths > 0) cap = std::min(cap, (size_t)pool_paths);
        cap = std::max(std::min(cap, total_paths), kNoAllocMinPool);
        while (cap > kNoAllocMinPool && scratch_need(c, cap, max_npix, total_pix, any_vol) > kScratchBytes) cap /= 2;
        pool = std::max(cap, kNoAllocMinPool);
        const size_t fixed = scratch_need(c, pool, 0, total_pix, any_vol);
        pass_cap = std::max(std::min((kScratchBytes - std::min(fixed, kScratchBytes)) / sizeof(float4), total_paths),
                            max_npix);
        while (pass_cap > max_npix && scratch_need(c, pool, pass_cap, total_pix, any_vol) > kScratchBytes) pass_cap--;
    }
    if (pass_cap > (size_t)INT32_MAX / 2 || total_pix > (size_t)INT32_MAX / 2) return fail(RD_ERR_INVALID, "render too large");
    if (noalloc) {
        Bump b{(char*)c->scratch, kScratchBytes};
        if (!carve_scratch(c, c->wb_na, b, pool, pass_cap, total_pix, any_vol))
            return fail(RD_ERR_INVALID, "render does not fit the work area of rd_render_device_noalloc");
    } else if (int rc = ensure_wave_buffers(c, pool, pass_cap, total_pix, any_vol)) {
        return rc;
    }
    WaveBuffers& wb = noalloc ? c->wb_na : c->wb;
    hipStream_t st = c->stream;
    // path Mis over a BSSRDF (one scene: batches and the work area refuse it above)
    const bool need_sss = single && reqs[0].scene->sss && reqs[0].scene->dev.integ.type == RD_INTEGRATOR_PATH &&
                          reqs[0].scene->dev.integ.method == RD_METHOD_MIS;
    if (need_sss) {
        if (int rc = ensure_sss_buffers(c, wb, pool)) return rc;
    } else {
        wb.ps.sss_q = wb.ps.sss_q2 = nullptr;
        wb.ps.sss0 = wb.ps.sss1 = wb.ps.sss2 = nullptr;
    }
    std::vector<DevScene> dscenes(n);
    for (int i = 0; i < n; i++) dscenes[i] = reqs[i].scene->dev;
    const DevScene* d_scenes = nullptr;
    if (!single) {
        if (noalloc) return fail(RD_ERR_INVALID, "batches allocate their scene table");
        if (int rc = ensure_batch_buffers(c, (size_t)n)) return rc;
        HIP_TRY(hipMemcpyAsync(wb.scenes, dscenes.data(), sizeof(DevScene) * n, hipMemcpyHostToDevice, st));
        d_scenes = wb.scenes;
    }
    std::vector<size_t> sum_off(n);
    {
        size_t o = 0;
        for (int i = 0; i < n; i++) { sum_off[i] = o; o += (size_t)(crops[i].x1 - crops[i].x0) * (crops[i].y1 - crops[i].y0); }
    }
    // stage timing creates events on demand: off without allocation
    struct TimingOff {
        rd_ctx* c;
        bool was;
        ~TimingOff() { c->timing_enabled = was; }
    } timing_guard{c, c->timing_enabled};
    if (noalloc) c->timing_enabled = false;
    Timer tm(c);
    std::memset(&c->stats, 0, sizeof(c->stats));
    HIP_TRY(launch_zero_counts(c->counts, kCntSlots, st));
    HIP_TRY(launch_zero(wb.sum, (int)total_pix, st));
    HIP_TRY(launch_zero(wb.sum2, (int)total_pix, st));
    hipEvent_t ev0 = c->ev_t0, ev1 = c->ev_t1;
    HIP_TRY(hipEventRecord(ev0, st));
    int passes = 0, max_bounce = 0;
    int rc = RD_OK;
    std::vector<BatchSeg> segs;
    for (size_t g = 0; g < groups.size() && !rc; g++) {
        const std::vector<int>& G = groups[g];
        DevScene rep = dscenes[G[0]];   // the traversal variant must cover the deepest tree of the group
        for (int i : G) {
            rep.stack_needed = std::max(rep.stack_needed, dscenes[i].stack_needed);
            rep.num_insts = std::max(rep.num_insts, dscenes[i].num_insts);   // any instanced scene: two-level kernels
            rep.mat_mask |= dscenes[i].mat_mask;   // the shading kernels must hold every scene's materials
        }
        rep.stack_needed = std::max(rep.stack_needed, c->stack_min);   // rd_tuning.stack_min
        std::vector<int> next_sample(n, 0);
        size_t gi = 0;
        while (gi < G.size() && !rc) {
            if (cancel && *cancel) { rc = fail(RD_ERR_CANCELLED, "render cancelled"); break; }
            segs.clear();
            size_t used = 0;
            int pass_max_npix = 0;
            for (size_t j = gi; j < G.size() && (int)segs.size() < kMaxSegs; j++) {
                int i = G[j];
                const Crop& cr = crops[i];
                int cw = cr.x1 - cr.x0;
                size_t np = (size_t)cw * (cr.y1 - cr.y0);
                int left = reqs[i].params.spp - next_sample[i];
                if (left <= 0) continue;
                int k = (int)std::min<size_t>((size_t)left, (pass_cap - used) / np);
                if (k <= 0) break;
                BatchSeg sg{};
                sg.req = i;
                sg.path_offset = (int)used;
                sg.sample_base = next_sample[i];
                sg.nsamples = k;
                sg.x0 = cr.x0; sg.y0 = cr.y0; sg.crop_w = cw; sg.npix = (int)np;
                sg.width = reqs[i].scene->width;
                sg.height = reqs[i].scene->height;
                sg.seed = reqs[i].params.seed;
                sg.tile = c->tile;
                sg.sum_offset = (int)sum_off[i];
                segs.push_back(sg);
                used += np * (size_t)k;
                next_sample[i] += k;
                pass_max_npix = std::max(pass_max_npix, (int)np);
                if (k < left) break;   // this request fills the pass
            }
            while (gi < G.size() && next_sample[G[gi]] >= reqs[G[gi]].params.spp) gi++;
            if (segs.empty()) break;
            // the segment table of the previous pass may still be read by its kernels
            HIP_TRY(hipStreamSynchronize(st));
            std::memcpy(wb.host_segs, segs.data(), sizeof(BatchSeg) * segs.size());
            HIP_TRY(hipMemcpyAsync(wb.segs, wb.host_segs, sizeof(BatchSeg) * segs.size(), hipMemcpyHostToDevice, st));
            PassLaunch L{rep, d_scenes, nullptr, kinds[g].second, kinds[g].first == RD_INTEGRATOR_VOL_PATH, wb.segs,
                         (int)segs.size(), (int)used, (int)std::min(pool, used), &wb, !noalloc,
                         !(single && (reqs[0].params.flags & RD_RENDER_NO_GRAPH))};
            L.sss = need_sss;
            L.sss_lut = c->sss_lut;
            rc = run_pass(c, tm, L, &max_bounce, cancel);
            if (rc) break;
            HIP_TRY(tm.run("accumulate", [&] {
                return launch_accumulate_batch(wb.segs, (int)segs.size(), pass_max_npix, wb.ps, wb.sum, wb.sum2,
                                               c->counts, st);
            }));
            passes++;
            if (progress) *progress = passes;
        }
    }
    for (int i = 0; i < n && !rc; i++) {
        const Crop& cr = crops[i];
        WaveParams wp{};
        wp.x0 = cr.x0; wp.y0 = cr.y0; wp.crop_w = cr.x1 - cr.x0; wp.crop_h = cr.y1 - cr.y0;
        wp.npix = wp.crop_w * wp.crop_h;
        wp.width = reqs[i].scene->width;
        wp.tile = c->tile;
        HIP_TRY(tm.run("finalize", [&] {
            return launch_finalize(wp, reqs[i].scene->height, reqs[i].params.spp, wb.sum + sum_off[i],
                                   wb.sum2 + sum_off[i], reqs[i].out_rgb, reqs[i].out_sum2, single ? out64 : nullptr,
                                   st);
        }));
    }
    HIP_TRY(hipEventRecord(ev1, st));
    HIP_TRY(hipMemcpyAsync(c->host_counts, c->counts, kCntSlots * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (!rc && c->host_counts[kCntChainStuck]) rc = fail(RD_ERR_DEVICE, "medium chain did not terminate");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev0, ev1);
    tm.collect();
    collect_stats(c, total_paths, ms, passes, max_bounce);
    return rc;
}

// BidirectionalPathTracingRenderer (Renderers/BidirectionalPathTracingRenderer.cs:380-452)
// on the device: passes of whole sample slices of the crop, each pass's
// subpaths walked together, its strategies connected in parallel and their
// visibility traced as Transmittance chains; per-pixel float64 sums in
// sample order, then the mean.
int render_bdpt(rd_ctx* c, rd_scene* s, const rd_render_params* p, float* d_out, float* d_out_sum2, double* out64,
                volatile int32_t* cancel, volatile int32_t* progress) {
    if (int rc = check_stack(s)) return rc;
    if (p->spp <= 0 || p->spp >= (1 << 29)) return fail(RD_ERR_INVALID, "bdpt spp must be in [1, 2^29)");
    Crop cr;
    if (int rc = resolve_crop(s, p, &cr)) return rc;
    const int crop_w = cr.x1 - cr.x0, crop_h = cr.y1 - cr.y0;
    const int npix = crop_w * crop_h;
    const int walk_cap = s->dev.integ.max_depth - 1;
    const int vmax = walk_cap == 0 ? 1 : 1 + std::max(1, walk_cap - 1);
    if (vmax > 256) return fail(RD_ERR_UNSUPPORTED, "bdpt max_depth above 257");
    const int nstrat = (vmax - 1) * vmax;
    // bytes per sample: vertices, walkers and their queues, strategy slots
    // with their chain state, queues and hits
    const size_t per_sample = 2 * (size_t)vmax * 68 + 2 * (36 * 2 + 32 + 20 + 8) + (size_t)nstrat * (48 + 72 + 20) + 16;
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    const size_t budget = std::min<size_t>(free_b / 4, (size_t)6 << 30);
    size_t cap = std::max<size_t>(budget / per_sample, (size_t)npix);
    if (p->wave_paths > 0) cap = std::max<size_t>((size_t)p->wave_paths, (size_t)npix);
    const int ns_pass = (int)std::max<size_t>(1, std::min<size_t>(cap / npix, (size_t)p->spp));
    const size_t Pmax = (size_t)npix * ns_pass;
    if (Pmax > (size_t)INT32_MAX / 4 || Pmax * std::max(nstrat, 2) > (size_t)INT32_MAX)
        return fail(RD_ERR_INVALID, "bdpt pass too large");
    const size_t qw = 2 * Pmax, qc = std::max<size_t>(Pmax * nstrat, 1);
    const size_t nv = 2 * (size_t)vmax * Pmax;
    hipStream_t st = c->stream;
    ScopedAllocs A;
    BdptBuffers b{};
    b.vmax = vmax; b.nstrat = nstrat; b.walk_cap = walk_cap;
    b.v0 = A.get<float4>(nv); b.v1 = A.get<float4>(nv); b.v2 = A.get<float4>(nv); b.v3 = A.get<float4>(nv);
    b.v4 = A.get<float>(nv);
    b.nverts = A.get<int>(2 * Pmax);
    b.wcoeff = A.get<float4>(qw); b.wenv = A.get<int>(qw); b.wkey = A.get<uint64_t>(qw); b.wn = A.get<uint32_t>(qw);
    b.conn = A.get<float4>(qc); b.cto = A.get<float4>(qc); b.ctr = A.get<float4>(qc);
    const size_t qmax = std::max(qw, qc);
    RayQueue qa{nullptr, A.get<int>(qmax), A.get<float4>(qmax), A.get<float4>(qmax)};
    RayQueue qb{nullptr, A.get<int>(qmax), A.get<float4>(qmax), A.get<float4>(qmax)};
    float4* hits = A.get<float4>(qmax);
    int* hsub = s->dev.num_insts > 0 ? A.get<int>(qmax) : nullptr;
    float4* res = A.get<float4>(Pmax);
    double4* sum = A.get<double4>(npix);
    float4* sum2 = A.get<float4>(npix);
    if (!b.v0 || !b.v1 || !b.v2 || !b.v3 || !b.v4 || !b.nverts || !b.wcoeff || !b.wenv || !b.wkey || !b.wn ||
        !b.conn || !b.cto || !b.ctr || !qa.pid || !qa.org || !qa.dir || !qb.pid || !qb.org || !qb.dir || !hits ||
        (s->dev.num_insts > 0 && !hsub) || !res || !sum || !sum2)
        return fail(RD_ERR_OOM, "hipMalloc failed for the bidirectional render");
    PathState tps{};
    tps.hsub = hsub;
    Timer tm(c);
    hipEvent_t ev0 = c->ev_t0, ev1 = c->ev_t1;
    HIP_TRY(hipEventRecord(ev0, st));
    HIP_TRY(launch_zero_counts(c->counts, kCntSlots, st));
    HIP_TRY(launch_zero(sum, npix, st));
    HIP_TRY(launch_zero(reinterpret_cast<float4*>(sum2), npix, st));
    BdptPass w{};
    w.x0 = cr.x0; w.y0 = cr.y0; w.crop_w = crop_w; w.npix = npix; w.width = s->width; w.seed = p->seed;
    int passes = 0, max_bounce = 0;
    int rc = RD_OK;
    // runs a queue loop: trace the queue at cin, then `stage` into the other
    // queue, until a queue is empty
    auto queue_loop = [&](const char* stage, int sa, int sb, auto&& launch_stage) -> int {
        RayQueue qin = qa, qout = qb;
        int cin = sa, cout = sb;
        for (int step = 0;; step++) {
            HIP_TRY(hipMemcpyAsync(c->host_counts + cin, c->counts + cin, sizeof(int), hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            if (c->host_counts[cin] == 0) break;
            if (step >= kMaxChainSteps) return fail(RD_ERR_DEVICE, "bidirectional queue did not drain");
            if (sa == kCntExtA) max_bounce = std::max(max_bounce, step + 1);
            HIP_TRY(tm.run("trace_closest", [&] {
                return launch_trace_ext(s->dev, nullptr, qin, tps, hits, c->counts, cin, -1, -1, -1, nullptr, nullptr,
                                        0, st);
            }));
            HIP_TRY(launch_zero_slots(c->counts, cout, cout, st));
            HIP_TRY(tm.run(stage, [&] { return launch_stage(qin, cin, qout, cout); }));
            std::swap(qin, qout);
            std::swap(cin, cout);
        }
        return RD_OK;
    };
    for (int base = 0; base < p->spp && !rc; base += ns_pass) {
        const int ns = std::min(ns_pass, p->spp - base);
        b.P = npix * ns;
        w.sample_base = base;
        HIP_TRY(launch_zero_slots(c->counts, kCntExtA, kCntExtB, st));
        HIP_TRY(tm.run("bd_start", [&] { return launch_bd_start(s->dev, w, b, qa, c->counts, kCntExtA, st); }));
        rc = queue_loop("bd_walk", kCntExtA, kCntExtB, [&](RayQueue qi, int ci, RayQueue qo, int co) {
            return launch_bd_walk(s->dev, b, qi, ci, hits, hsub, qo, co, c->counts, st);
        });
        if (rc) break;
        if (nstrat > 0) {
            HIP_TRY(launch_zero_slots(c->counts, kCntShadow, kCntMis, st));
            HIP_TRY(tm.run("bd_connect", [&] {
                return launch_bd_connect(s->dev, s->lt_cam, w, b, qa, kCntShadow, c->counts, st);
            }));
            rc = queue_loop("bd_chain", kCntShadow, kCntMis, [&](RayQueue qi, int ci, RayQueue qo, int co) {
                return launch_bd_chain(s->dev, b, qi, ci, hits, hsub, qo, co, c->counts, st);
            });
            if (rc) break;
        } else {
            HIP_TRY(launch_zero(b.conn, 1, st));
        }
        HIP_TRY(tm.run("bd_sum", [&] { return launch_bd_sum(b, res, st); }));
        HIP_TRY(tm.run("accumulate", [&] { return launch_bd_accumulate(w, ns, res, sum, sum2, st); }));
        passes++;
        if (progress) *progress = passes;
        if (cancel && *cancel) rc = fail(RD_ERR_CANCELLED, "render cancelled");
    }
    if (!rc) {
        HIP_TRY(tm.run("finalize", [&] {
            return launch_bd_finalize(w, s->height, p->spp, sum, sum2, d_out, d_out_sum2, out64, st);
        }));
    }
    HIP_TRY(hipEventRecord(ev1, st));
    HIP_TRY(hipMemcpyAsync(c->host_counts, c->counts, kCntSlots * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev0, ev1);
    tm.collect();
    collect_stats(c, (uint64_t)npix * (uint64_t)p->spp, ms, passes, max_bounce);
    return rc;
}

// AdjointParticleRenderer (Renderers/AdjointParticleRenderer.cs:127-178) on
// the device: waves of particles (particle i draws from
// particle_key(seed, i), so the image does not depend on the wave size),
// each wave bounced until no particle is left, the splats summed in a
// float64 image; then RefRGB *= W*H / particles and the directly seen
// emission of each pixel corner is added.
int render_particles(rd_ctx* c, rd_scene* s, const rd_render_params* p, float* d_out, volatile int32_t* cancel,
                     volatile int32_t* progress) {
    if (int rc = check_stack(s)) return rc;
    if (p->x0 || p->y0 || p->x1 || p->y1) {
        Crop cr;
        if (int rc = resolve_crop(s, p, &cr)) return rc;
        if (cr.x0 != 0 || cr.y0 != 0 || cr.x1 != s->width || cr.y1 != s->height)
            return fail(RD_ERR_INVALID, "the particle renderer covers the whole image (no crop)");
    }
    if (p->particles < 0) return fail(RD_ERR_INVALID, "particle count must not be negative");
    const uint64_t total = p->particles > 0 ? (uint64_t)p->particles : 16384;   // "particle_count" default
    const int W = s->width, H = s->height;
    const size_t npix = (size_t)W * H;
    // particles in flight: the caller's wave_paths, else as many as a quarter
    // of the free device memory holds (about 180 bytes a particle: three ray
    // queues, hits, the particle state), at most 2^25; fewer waves mean fewer
    // of the long Russian-roulette tails, each of which runs nearly empty
    size_t wave_cap = (size_t)p->wave_paths;
    if (p->wave_paths <= 0) {
        size_t free_b = 0, total_b = 0;
        wave_cap = (size_t)1 << 22;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
            wave_cap = std::max(wave_cap, std::min((size_t)1 << 25, free_b / 4 / 192));
    }
    const size_t P = std::max<size_t>(1, std::min<size_t>(wave_cap, total));
    const size_t qcap = std::max(P, npix);
    hipStream_t st = c->stream;
    ScopedAllocs A;
    const size_t acc4 = (npix * 3 + 3) / 4;   // cleared as double4s
    double* acc = A.get<double>(acc4 * 4);
    RayQueue qa{nullptr, A.get<int>(qcap), A.get<float4>(qcap), A.get<float4>(qcap)};
    RayQueue qb{nullptr, A.get<int>(P), A.get<float4>(P), A.get<float4>(P)};
    RayQueue qs{nullptr, A.get<int>(P), A.get<float4>(P), A.get<float4>(P)};
    float4* hits = A.get<float4>(qcap);
    ParticleState ps{{A.get<float4>(P), A.get<float4>(P)}, {A.get<uint32_t>(P), A.get<uint32_t>(P)}, A.get<float4>(P),
                     A.get<uint8_t>(P),
                     s->dev.num_insts > 0 ? A.get<int>(qcap) : nullptr};
    if (!acc || !qa.pid || !qa.org || !qa.dir || !qb.pid || !qb.org || !qb.dir || !qs.pid || !qs.org || !qs.dir ||
        !hits || !ps.st[0] || !ps.st[1] || !ps.rng_n[0] || !ps.rng_n[1] || !ps.splat || !ps.vis ||
        (s->dev.num_insts > 0 && !ps.hsub))
        return fail(RD_ERR_OOM, "hipMalloc failed for the particle render");
    // the traversal kernels read queue counts, write hits / visibility and
    // instance triangles through a PathState
    PathState tps{};
    tps.vis = ps.vis;
    tps.hsub = ps.hsub;
    const int kA = kCntExtA, kB = kCntExtB, kS = kCntShadow;
    Timer tm(c);
    hipEvent_t ev0 = c->ev_t0, ev1 = c->ev_t1;
    HIP_TRY(hipEventRecord(ev0, st));
    HIP_TRY(launch_zero_counts(c->counts, kCntSlots, st));
    HIP_TRY(launch_zero(reinterpret_cast<double4*>(acc), (int)acc4, st));
    int waves = 0, max_bounce = 0;
    int rc = RD_OK;
    for (uint64_t base = 0; base < total && !rc; base += P) {
        const int n = (int)std::min<uint64_t>(P, total - base);
        RayQueue qin = qa, qout = qb;
        int cin = kA, cout = kB;
        HIP_TRY(launch_zero_slots(c->counts, kA, kB, st));
        HIP_TRY(tm.run("lt_emit", [&] { return launch_lt_emit(s->dev, p->seed, base, n, ps, qin, c->counts, cin, st); }));
        int side = 0;   // the state buffer beside qin
        // every kernel reads its queue size on the device, so once the queue
        // is small the bounces are queued kLtGroup at a time with one host
        // read per group (bounces past the end run on empty queues); a tick
        // counts the bounces that had particles
        constexpr int kLtGroup = 8;
        constexpr int kLtGroupBelow = 1 << 18;
        HIP_TRY(launch_zero_slots(c->counts, kCntTail, kCntTail, st));
        for (int bounce = 0;;) {
            HIP_TRY(hipMemcpyAsync(c->host_counts + cin, c->counts + cin, sizeof(int), hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            if (c->host_counts[cin] == 0) break;
            const int group = c->host_counts[cin] <= kLtGroupBelow ? kLtGroup : 1;
            for (int g = 0; g < group; g++, bounce++) {
                qin.count = qout.count = nullptr;
                HIP_TRY(launch_tail_tick(c->counts, cin, st));
                // grid-stride traversal (schedule 0): a particle's rays are
                // cheap, and the work-cursor schedules' atomics cost more than
                // they balance (2^24 particles in the Cornell box: 130 -> 106 ms)
                HIP_TRY(tm.run("trace_closest", [&] {
                    return launch_trace_ext(s->dev, nullptr, qin, tps, hits, c->counts, cin, -1, -1, -1, nullptr,
                                            nullptr, 0, st);
                }));
                HIP_TRY(launch_zero_slots(c->counts, cout, kS, st));
                HIP_TRY(tm.run("shade", [&] {
                    return launch_lt_shade(s->dev, s->lt_cam, p->seed, base, side, qin, cin, hits, ps, qout, cout, qs,
                                           kS, c->counts, st);
                }));
                HIP_TRY(tm.run("trace_any", [&] {
                    return launch_trace_shadow(s->dev, nullptr, qs, tps, c->counts, kS, 0, st);
                }));
                HIP_TRY(tm.run("splat", [&] { return launch_lt_splat(qs, kS, ps, acc, c->counts, st); }));
                std::swap(qin, qout);
                std::swap(cin, cout);
                side = 1 - side;
            }
            if (cancel && *cancel) { rc = fail(RD_ERR_CANCELLED, "render cancelled"); break; }
            if (bounce > (1 << 20)) { rc = fail(RD_ERR_DEVICE, "particle wave did not end"); break; }
        }
        HIP_TRY(hipMemcpyAsync(c->host_counts + kCntTail, c->counts + kCntTail, sizeof(int), hipMemcpyDeviceToHost,
                               st));
        HIP_TRY(hipStreamSynchronize(st));
        max_bounce = std::max(max_bounce, c->host_counts[kCntTail]);
        waves++;
        if (progress) *progress = waves;
        if (cancel && *cancel && !rc) rc = fail(RD_ERR_CANCELLED, "render cancelled");
    }
    if (!rc) {
        // AfterRender: v = (float)ScreenX * ScreenY / particles, then CameraTraceLight
        const float v = ((float)W * (float)H) / (float)total;
        HIP_TRY(launch_lt_corner_rays(s->dev, qa, c->counts, kA, st));
        HIP_TRY(tm.run("trace_closest", [&] {
            return launch_trace_ext(s->dev, nullptr, qa, tps, hits, c->counts, kA, -1, -1, -1, nullptr, nullptr,
                                    c->trace_schedule[0], st);
        }));
        HIP_TRY(tm.run("finalize", [&] { return launch_lt_final(s->dev, qa, hits, ps.hsub, acc, v, d_out, st); }));
    }
    HIP_TRY(hipEventRecord(ev1, st));
    HIP_TRY(hipMemcpyAsync(c->host_counts, c->counts, kCntSlots * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev0, ev1);
    tm.collect();
    collect_stats(c, total, ms, waves, max_bounce);
    return rc;
}

}  // namespace

int rd_render_device(rd_ctx* c, rd_scene* s, const rd_render_params* p, float* d_out_rgb, float* d_out_sum2,
                     volatile int32_t* cancel, volatile int32_t* progress) {
    if (!c || !s || !p || !d_out_rgb) return fail(RD_ERR_INVALID, "null argument");
    if (s->ctx != c) return fail(RD_ERR_INVALID, "scene belongs to another context");
    if (s->dev.integ.type == RD_INTEGRATOR_PARTICLE) {
        if (d_out_sum2) return fail(RD_ERR_INVALID, "the particle renderer has no per-pixel sample sums");
        (void)hipSetDevice(c->device);
        return render_particles(c, s, p, d_out_rgb, cancel, progress);
    }
    if (s->dev.integ.type == RD_INTEGRATOR_BDPT) {
        (void)hipSetDevice(c->device);
        return render_bdpt(c, s, p, d_out_rgb, d_out_sum2, nullptr, cancel, progress);
    }
    rd_request r{};
    r.scene = s;
    r.params = *p;
    r.out_rgb = d_out_rgb;
    r.out_sum2 = d_out_sum2;
    return render_requests(c, 1, &r, true, p->wave_paths, cancel, progress);
}

// rd_render_device with every work buffer carved from the context's fixed
// work area (no device allocation; see render_requests).
int rd_render_device_noalloc(rd_ctx* c, rd_scene* s, const rd_render_params* p, float* d_out_rgb,
                             float* d_out_sum2, volatile int32_t* cancel, volatile int32_t* progress) {
    if (!c || !s || !p || !d_out_rgb) return fail(RD_ERR_INVALID, "null argument");
    rd_request r{};
    r.scene = s;
    r.params = *p;
    r.out_rgb = d_out_rgb;
    r.out_sum2 = d_out_sum2;
    return render_requests(c, 1, &r, true, p->wave_paths, cancel, progress, true);
}

// Batched rendering: the requests' paths share passes and the slot pool; see
// render_requests.
int rd_render_batch_device(rd_ctx* c, int32_t n, const rd_request* reqs, int32_t wave_paths,
                           volatile int32_t* cancel, volatile int32_t* progress) {
    if (!c || n < 0 || (n > 0 && !reqs)) return fail(RD_ERR_INVALID, "bad argument");
    if (n == 0) return RD_OK;
    return render_requests(c, n, reqs, false, wave_paths, cancel, progress);
}

// Path-replay backpropagation of an image loss to the scene's texture
// parameters.  Every pass replays its paths twice with the primal's random
// numbers: first the primal (per-sample radiance into the result buffer),
// then the adjoint, whose resolve scatters dLoss/dtexel.
int rd_adjoint_device(rd_ctx* c, rd_scene* s, const rd_render_params* p, const float* d_dl_dimage,
                      float* const* d_tex_grads) {
    if (!c || !s || !p || !d_dl_dimage || !d_tex_grads) return fail(RD_ERR_INVALID, "null argument");
    if (p->spp <= 0) return fail(RD_ERR_INVALID, "spp must be positive");
    if (s->integ.type != RD_INTEGRATOR_PATH) return fail(RD_ERR_UNSUPPORTED, "the adjoint covers the path integrator");
    if (s->sss && s->integ.method == RD_METHOD_MIS)
        return fail(RD_ERR_UNSUPPORTED, "the adjoint covers scenes without a BSSRDF");
    if (int rc = check_stack(s)) return rc;
    Crop cr;
    if (int rc = resolve_crop(s, p, &cr)) return rc;
    CtxLock lock(c);
    if (!lock.ok()) return fail(RD_ERR_BUSY, "context in use by another thread");
    (void)hipSetDevice(c->device);
    const int cw = cr.x1 - cr.x0;
    const size_t npix = (size_t)cw * (cr.y1 - cr.y0);
    const size_t total = npix * (size_t)p->spp;
    const size_t pool = p->wave_paths > 0 ? std::min((size_t)p->wave_paths, std::max<size_t>(total, 1))
                                          : default_pool(total, c->pool_log2);
    const size_t pass_cap = std::max(std::min(pass_limit(), total), npix);
    if (int rc = ensure_wave_buffers(c, pool, pass_cap, npix, false)) return rc;
    if (int rc = ensure_adjoint_buffers(c, (size_t)s->num_textures)) return rc;
    WaveBuffers& wb = c->wb;
    wb.ps.sss_q = wb.ps.sss_q2 = nullptr;
    wb.ps.sss0 = wb.ps.sss1 = wb.ps.sss2 = nullptr;
    hipStream_t st = c->stream;
    // float64 accumulators for the requested textures
    std::vector<double*> acc(std::max(s->num_textures, 1), nullptr);
    std::vector<void*> acc_allocs;
    auto free_acc = [&] { for (void* q : acc_allocs) (void)hipFree(q); acc_allocs.clear(); };
    for (int t = 0; t < s->num_textures; t++) {
        if (!d_tex_grads[t]) continue;
        size_t bytes = s->tex_elems[t] * 3 * sizeof(double);
        void* q = nullptr;
        if (hipMalloc(&q, bytes) != hipSuccess) { free_acc(); return fail(RD_ERR_OOM, "hipMalloc gradient accumulator"); }
        acc_allocs.push_back(q);
        acc[t] = (double*)q;
        if (hipMemsetAsync(q, 0, bytes, st) != hipSuccess) { free_acc(); return fail(RD_ERR_DEVICE, "clear accumulator"); }
    }
    if (hipMemcpyAsync(wb.grad_ptrs, acc.data(), sizeof(double*) * acc.size(), hipMemcpyHostToDevice, st) != hipSuccess) {
        free_acc();
        return fail(RD_ERR_DEVICE, "upload gradient table");
    }
    AdjArgs aa{d_dl_dimage, (double* const*)wb.grad_ptrs, 1.0f / (float)p->spp, 0};
    Timer tm(c);
    std::memset(&c->stats, 0, sizeof(c->stats));
    HIP_TRY(launch_zero_counts(c->counts, kCntSlots, st));
    hipEvent_t ev0, ev1;
    HIP_TRY(hipEventCreate(&ev0));
    HIP_TRY(hipEventCreate(&ev1));
    HIP_TRY(hipEventRecord(ev0, st));
    const int per_pass = (int)std::max<size_t>(1, pass_cap / npix);
    int passes = 0, max_bounce = 0, rc = RD_OK;
    for (int sb = 0; sb < p->spp && !rc; sb += per_pass) {
        BatchSeg sg{};
        sg.req = 0;
        sg.path_offset = 0;
        sg.sample_base = sb;
        sg.nsamples = std::min(per_pass, p->spp - sb);
        sg.x0 = cr.x0; sg.y0 = cr.y0; sg.crop_w = cw; sg.npix = (int)npix;
        sg.width = s->width;
        sg.height = s->height;
        sg.seed = p->seed;
        sg.tile = c->tile;
        HIP_TRY(hipStreamSynchronize(st));
        *wb.host_segs = sg;
        HIP_TRY(hipMemcpyAsync(wb.segs, wb.host_segs, sizeof(BatchSeg), hipMemcpyHostToDevice, st));
        const int npass = (int)(npix * (size_t)sg.nsamples);
        PassLaunch primal{s->dev, nullptr, nullptr, s->integ.method, false, wb.segs, 1, npass,
                          (int)std::min(pool, (size_t)npass)};
        rc = run_pass(c, tm, primal, &max_bounce, nullptr);
        if (rc) break;
        PassLaunch adjoint = primal;
        adjoint.adj = &aa;
        rc = run_pass(c, tm, adjoint, &max_bounce, nullptr);
        passes++;
    }
    for (int t = 0; t < s->num_textures && !rc; t++)
        if (acc[t] && launch_add_f64_to_f32(acc[t], d_tex_grads[t], (int64_t)s->tex_elems[t] * 3, st) != hipSuccess)
            rc = fail(RD_ERR_DEVICE, "gradient conversion");
    HIP_TRY(hipEventRecord(ev1, st));
    HIP_TRY(hipMemcpyAsync(c->host_counts, c->counts, kCntSlots * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev0, ev1);
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    tm.collect();
    collect_stats(c, 2 * total, ms, passes, max_bounce);
    free_acc();
    return rc;
}

int rd_adjoint(rd_ctx* c, rd_scene* s, const rd_render_params* p, const float* dl_dimage, float* const* tex_grads) {
    if (!c || !s || !p || !dl_dimage || !tex_grads) return fail(RD_ERR_INVALID, "null argument");
    (void)hipSetDevice(c->device);
    const size_t nimg = (size_t)s->width * s->height * 3;
    std::vector<void*> bufs;
    std::vector<float*> dgrads(std::max(s->num_textures, 1), nullptr);
    int rc = RD_OK;
    auto dalloc = [&](size_t bytes) -> void* {
        void* q = nullptr;
        if (hipMalloc(&q, bytes) != hipSuccess) return nullptr;
        bufs.push_back(q);
        return q;
    };
    float* ddl = (float*)dalloc(nimg * 4);
    if (!ddl) rc = fail(RD_ERR_OOM, "hipMalloc dLoss/dImage");
    if (!rc && hipMemcpy(ddl, dl_dimage, nimg * 4, hipMemcpyHostToDevice) != hipSuccess) rc = fail(RD_ERR_DEVICE, "copy");
    for (int t = 0; t < s->num_textures && !rc; t++) {
        if (!tex_grads[t]) continue;
        size_t bytes = s->tex_elems[t] * 3 * sizeof(float);
        dgrads[t] = (float*)dalloc(bytes);
        if (!dgrads[t]) { rc = fail(RD_ERR_OOM, "hipMalloc gradient"); break; }
        if (hipMemcpy(dgrads[t], tex_grads[t], bytes, hipMemcpyHostToDevice) != hipSuccess) rc = fail(RD_ERR_DEVICE, "copy");
    }
    if (!rc) rc = rd_adjoint_device(c, s, p, ddl, dgrads.data());
    for (int t = 0; t < s->num_textures && !rc; t++) {
        if (!tex_grads[t]) continue;
        if (hipMemcpy(tex_grads[t], dgrads[t], s->tex_elems[t] * 3 * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(RD_ERR_DEVICE, "copy gradient");
    }
    for (void* q : bufs) (void)hipFree(q);
    return rc;
}

// Replaces a texture's texels (w*h*3 floats, the layout of rd_texture_desc;
// constants take 3 floats) — e.g. parameters updated by an optimiser step.
static int set_texture(rd_ctx* c, rd_scene* s, int32_t tex, const float* data, hipMemcpyKind kind) {
    if (!c || !s || !data || tex < 0 || tex >= s->num_textures) return fail(RD_ERR_INVALID, "bad argument");
    for (int e : s->env_textures)
        if (e == tex) return fail(RD_ERR_UNSUPPORTED, "environment-map textures are fixed (sampling tables)");
    CtxLock lock(c);
    if (!lock.ok()) return fail(RD_ERR_BUSY, "context in use by another thread");
    (void)hipSetDevice(c->device);
    const DevTexture& t = s->texs[tex];
    if (t.w == 0) {
        DevTexture nt = t;
        nt.v[0] = data[0]; nt.v[1] = data[1]; nt.v[2] = data[2];
        if (kind == hipMemcpyDeviceToDevice) {
            float v[3];
            HIP_TRY(hipMemcpy(v, data, sizeof(v), hipMemcpyDeviceToHost));
            nt.v[0] = v[0]; nt.v[1] = v[1]; nt.v[2] = v[2];
        }
        s->texs[tex] = nt;
        HIP_TRY(hipMemcpy((DevTexture*)s->dev.texs + tex, &nt, sizeof(DevTexture), hipMemcpyHostToDevice));
        return RD_OK;
    }
    HIP_TRY(hipMemcpy((float*)s->dev.texels + t.offset, data, (size_t)t.w * t.h * 3 * sizeof(float), kind));
    return RD_OK;
}
int rd_scene_set_texture(rd_ctx* c, rd_scene* s, int32_t tex, const float* data) {
    return set_texture(c, s, tex, data, hipMemcpyHostToDevice);
}
int rd_scene_set_texture_device(rd_ctx* c, rd_scene* s, int32_t tex, const float* d_data) {
    return set_texture(c, s, tex, d_data, hipMemcpyDeviceToDevice);
}

int rd_render_batch(rd_ctx* c, int32_t n, const rd_request* reqs, int32_t wave_paths, volatile int32_t* cancel,
                    volatile int32_t* progress) {
    if (!c || n < 0 || (n > 0 && !reqs)) return fail(RD_ERR_INVALID, "bad argument");
    CtxLock lock(c);
    if (!lock.ok()) return fail(RD_ERR_BUSY, "context in use by another thread");
    (void)hipSetDevice(c->device);
    // device outputs for every request; pixels outside a crop keep the caller's values
    std::vector<rd_request> dreq(reqs, reqs + n);
    std::vector<void*> bufs;
    int rc = RD_OK;
    for (int i = 0; i < n && !rc; i++) {
        if (!reqs[i].scene || !reqs[i].out_rgb) { rc = fail(RD_ERR_INVALID, "request without scene or output"); break; }
        size_t bytes = (size_t)reqs[i].scene->width * reqs[i].scene->height * 3 * 4;
        void* d1 = nullptr;
        void* d2 = nullptr;
        if (hipMalloc(&d1, bytes) != hipSuccess) { rc = fail(RD_ERR_OOM, "hipMalloc output"); break; }
        bufs.push_back(d1);
        if (reqs[i].out_sum2) {
            if (hipMalloc(&d2, bytes) != hipSuccess) { rc = fail(RD_ERR_OOM, "hipMalloc output"); break; }
            bufs.push_back(d2);
        }
        if (hipMemcpy(d1, reqs[i].out_rgb, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = fail(RD_ERR_DEVICE, "copy output");
        if (!rc && d2 && hipMemcpy(d2, reqs[i].out_sum2, bytes, hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(RD_ERR_DEVICE, "copy output");
        dreq[i].out_rgb = (float*)d1;
        dreq[i].out_sum2 = (float*)d2;
    }
    if (!rc) rc = rd_render_batch_device(c, n, dreq.data(), wave_paths, cancel, progress);
    for (int i = 0; i < n && !rc; i++) {
        size_t bytes = (size_t)reqs[i].scene->width * reqs[i].scene->height * 3 * 4;
        if (hipMemcpy(reqs[i].out_rgb, dreq[i].out_rgb, bytes, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(RD_ERR_DEVICE, "copy output");
        if (!rc && reqs[i].out_sum2 &&
            hipMemcpy(reqs[i].out_sum2, dreq[i].out_sum2, bytes, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(RD_ERR_DEVICE, "copy output");
    }
    for (void* p : bufs) (void)hipFree(p);
    return rc;
}

int rd_render(rd_ctx* c, rd_scene* s, const rd_render_params* p, float* out_rgb, float* out_sum2,
              volatile int32_t* cancel, volatile int32_t* progress) {
    if (!c || !s || !p || !out_rgb) return fail(RD_ERR_INVALID, "null argument");
    CtxLock lock(c);
    if (!lock.ok()) return fail(RD_ERR_BUSY, "context in use by another thread");
    (void)hipSetDevice(c->device);
    size_t n = (size_t)s->width * s->height * 3;
    float *d1 = nullptr, *d2 = nullptr;
    if (hipMalloc(&d1, n * 4) != hipSuccess) return fail(RD_ERR_OOM, "hipMalloc output");
    if (out_sum2 && hipMalloc(&d2, n * 4) != hipSuccess) { (void)hipFree(d1); return fail(RD_ERR_OOM, "hipMalloc output"); }
    int rc = RD_OK;
    // pixels outside a crop keep the caller's values
    if (hipMemcpy(d1, out_rgb, n * 4, hipMemcpyHostToDevice) != hipSuccess) rc = fail(RD_ERR_DEVICE, "copy output");
    if (!rc && d2 && hipMemcpy(d2, out_sum2, n * 4, hipMemcpyHostToDevice) != hipSuccess) rc = fail(RD_ERR_DEVICE, "copy output");
    if (!rc) rc = rd_render_device(c, s, p, d1, d2, cancel, progress);
    if (!rc && hipMemcpy(out_rgb, d1, n * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = fail(RD_ERR_DEVICE, "copy output");
    if (!rc && d2 && hipMemcpy(out_sum2, d2, n * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = fail(RD_ERR_DEVICE, "copy output");
    (void)hipFree(d1);
    if (d2) (void)hipFree(d2);
    return rc;
}

int rd_render_mean_f64(rd_ctx* c, rd_scene* s, const rd_render_params* p, double* out_mean) {
    if (!c || !s || !p || !out_mean) return fail(RD_ERR_INVALID, "null argument");
    CtxLock lock(c);
    if (!lock.ok()) return fail(RD_ERR_BUSY, "context in use by another thread");
    (void)hipSetDevice(c->device);
    size_t n = (size_t)s->width * s->height * 3;
    double* d = nullptr;
    if (hipMalloc(&d, n * sizeof(double)) != hipSuccess) return fail(RD_ERR_OOM, "hipMalloc output");
    int rc = RD_OK;
    // pixels outside a crop keep the caller's values
    if (hipMemcpy(d, out_mean, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(RD_ERR_DEVICE, "copy output");
    rd_request r{};
    r.scene = s;
    r.params = *p;
    if (!rc && s->dev.integ.type == RD_INTEGRATOR_BDPT) rc = render_bdpt(c, s, p, nullptr, nullptr, d, nullptr, nullptr);
    else if (!rc) rc = render_requests(c, 1, &r, true, p->wave_paths, nullptr, nullptr, false, d);
    if (!rc && hipMemcpy(out_mean, d, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(RD_ERR_DEVICE, "copy output");
    (void)hipFree(d);
    return rc;
}

int rd_get_render_stats(const rd_ctx* c, rd_render_stats* st) {
    if (!c || !st) return fail(RD_ERR_INVALID, "null argument");
    *st = c->stats;
    return RD_OK;
}

int rd_get_tuning(const rd_ctx* c, rd_tuning* t) {
    if (!c || !t) return fail(RD_ERR_INVALID, "null argument");
    *t = rd_tuning{};
    for (int k = 0; k < 3; k++) t->trace_schedule[k] = c->trace_schedule[k];
    t->ray_sort = c->ray_sort ? 1 : 0;
    t->ray_sort_bits = c->ray_sort_bits;
    t->ray_sort_min = c->ray_sort_min;
    t->ray_sort_shadow = c->ray_sort_shadow ? 1 : 0;
    t->overlap = c->overlap ? 1 : 0;
    t->device_loop = c->device_loop;
    t->tail_rays = c->tail_rays;
    t->chain_lockstep = c->chain_lockstep ? 1 : 0;
    t->pool_log2 = c->pool_log2;
    t->tile = c->tile;
    t->stack_min = c->stack_min;
    t->bvh_build = c->bvh_build;
    return RD_OK;
}

int rd_set_tuning(rd_ctx* c, const rd_tuning* t) {
    if (!c || !t) return fail(RD_ERR_INVALID, "null argument");
    for (int k = 0; k < 3; k++)
        if (t->trace_schedule[k] < -1 || t->trace_schedule[k] >= kNumSchedules)
            return fail(RD_ERR_INVALID, "trace_schedule: 0 to 3, or -1");
    if (t->ray_sort_bits != -1 && (t->ray_sort_bits < 3 || t->ray_sort_bits > 30))
        return fail(RD_ERR_INVALID, "ray_sort_bits: 3..30 or -1");
    if (t->ray_sort_min < -1) return fail(RD_ERR_INVALID, "ray_sort_min: >= 0 or -1");
    if (t->device_loop < -1 || t->device_loop > 2) return fail(RD_ERR_INVALID, "device_loop: 0, 1, 2 or -1");
    if (t->tail_rays < -1) return fail(RD_ERR_INVALID, "tail_rays: >= 0 or -1");
    if (t->pool_log2 != -1 && t->pool_log2 != 0 && (t->pool_log2 < 10 || t->pool_log2 > 27))
        return fail(RD_ERR_INVALID, "pool_log2: 0, 10..27 or -1");
    if (t->tile != -1 && (t->tile < 1 || t->tile > 64)) return fail(RD_ERR_INVALID, "tile: 1..64 or -1");
    if (t->stack_min < -1 || t->stack_min > 64) return fail(RD_ERR_INVALID, "stack_min: 0..64 or -1");
    if (t->bvh_build < -1 || t->bvh_build > 2) return fail(RD_ERR_INVALID, "bvh_build: 0, 1, 2 or -1");
    CtxLock lock(c);
    if (!lock.ok()) return fail(RD_ERR_BUSY, "context in use by another thread");
    for (int k = 0; k < 3; k++)
        if (t->trace_schedule[k] >= 0) c->trace_schedule[k] = t->trace_schedule[k];
    if (t->ray_sort >= 0) c->ray_sort = t->ray_sort != 0;
    if (t->ray_sort_bits >= 0) c->ray_sort_bits = t->ray_sort_bits;
    if (t->ray_sort_min >= 0) c->ray_sort_min = t->ray_sort_min;
    if (t->ray_sort_shadow >= 0) c->ray_sort_shadow = t->ray_sort_shadow != 0;
    if (t->overlap >= 0) c->overlap = t->overlap != 0;
    if (t->device_loop >= 0) c->device_loop = t->device_loop;
    if (t->tail_rays >= 0) c->tail_rays = t->tail_rays;
    if (t->chain_lockstep >= 0) c->chain_lockstep = t->chain_lockstep != 0;
    if (t->pool_log2 >= 0) c->pool_log2 = t->pool_log2;
    if (t->tile >= 0) c->tile = t->tile;
    if (t->stack_min >= 0) c->stack_min = t->stack_min;
    if (t->bvh_build >= 0) c->bvh_build = t->bvh_build;
    return RD_OK;
}

namespace {

// Tunings found by rd_autotune, per device and scene class, for the process.
struct TuneCache {
    std::mutex mu;
    std::map<std::string, rd_tuning> m;
};
TuneCache& tune_cache() {
    static TuneCache t;
    return t;
}

int ilog2(uint64_t v) {
    int l = 0;
    while (v > 1) { v >>= 1; l++; }
    return l;
}

// device: architecture and compute units; scene class: integrator and
// method, instancing, log2 of the primitives and of the render's paths
// (both rounded down to even)
std::string tune_key(const rd_ctx* c, const rd_scene* s, const rd_render_params* p) {
    hipDeviceProp_t prop{};
    std::string dev = "unknown";
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess)
        dev = std::string(prop.gcnArchName) + "/" + std::to_string(prop.multiProcessorCount);
    Crop cr{0, 0, s->width, s->height};
    (void)resolve_crop(s, p, &cr);
    const uint64_t paths = (uint64_t)(cr.x1 - cr.x0) * (uint64_t)(cr.y1 - cr.y0) * (uint64_t)std::max(p->spp, 1);
    const DevIntegrator& I = s->dev.integ;
    char buf[160];
    std::snprintf(buf, sizeof(buf), "%s|%s/%d|inst%d|prims2^%d|paths2^%d", dev.c_str(),
                  I.type == RD_INTEGRATOR_VOL_PATH ? "vol_path" : "path", I.type == RD_INTEGRATOR_VOL_PATH ? 0 : I.method,
                  s->dev.num_insts > 0 ? 1 : 0, ilog2((uint64_t)std::max(s->dev.num_prims, 1)) & ~1,
                  ilog2(std::max<uint64_t>(paths, 1)) & ~1);
    return buf;
}

const int kTuneFields = 15;
void tuning_to_ints(const rd_tuning& t, int* v) {
    const int32_t f[kTuneFields] = {t.trace_schedule[0], t.trace_schedule[1], t.trace_schedule[2], t.ray_sort,
                                    t.ray_sort_bits,     t.ray_sort_min,      t.ray_sort_shadow,   t.overlap,
                                    t.device_loop,       t.tail_rays,         t.chain_lockstep,    t.pool_log2,
                                    t.tile,              t.stack_min,         t.bvh_build};
    for (int k = 0; k < kTuneFields; k++) v[k] = f[k];
}
rd_tuning tuning_from_ints(const int* v) {
    rd_tuning t{};
    for (int k = 0; k < 3; k++) t.trace_schedule[k] = v[k];
    t.ray_sort = v[3]; t.ray_sort_bits = v[4]; t.ray_sort_min = v[5]; t.ray_sort_shadow = v[6]; t.overlap = v[7];
    t.device_loop = v[8]; t.tail_rays = v[9]; t.chain_lockstep = v[10]; t.pool_log2 = v[11]; t.tile = v[12];
    t.stack_min = v[13];
    t.bvh_build = v[14];
    return t;
}

}  // namespace

int rd_autotune(rd_ctx* c, rd_scene* s, const rd_render_params* p, rd_tuning* out) {
    if (!c || !s || !p) return fail(RD_ERR_INVALID, "null argument");
    if (p->spp <= 0) return fail(RD_ERR_INVALID, "spp must be positive");
    {
        CtxLock lock(c);
        if (!lock.ok()) return fail(RD_ERR_BUSY, "context in use by another thread");
        (void)hipSetDevice(c->device);
        float* d = nullptr;
        if (hipMalloc(&d, (size_t)s->width * s->height * 3 * sizeof(float)) != hipSuccess)
            return fail(RD_ERR_OOM, "hipMalloc output");
        rd_request r{};
        r.scene = s;
        r.params = *p;
        r.out_rgb = d;
        int rc = RD_OK;
        auto time_it = [&]() -> double {   // device ms, best of two renders
            double best = 1e300;
            for (int k = 0; k < 2 && !rc; k++) {
                rc = render_requests(c, 1, &r, true, p->wave_paths, nullptr, nullptr);
                best = std::min(best, (double)c->stats.render_ms);
            }
            return best;
        };
        // each knob in turn: its values are tried with the others fixed and
        // the fastest kept (a change must gain 1 %)
        struct Knob {
            int* field;
            std::vector<int> values;
        };
        Crop cr{0, 0, s->width, s->height};
        (void)resolve_crop(s, p, &cr);
        const uint64_t paths = (uint64_t)(cr.x1 - cr.x0) * (cr.y1 - cr.y0) * (uint64_t)p->spp;
        // slot pools below the render's paths (larger ones change nothing)
        std::vector<int> pools = {0};
        if (p->wave_paths <= 0)
            for (int l : {22, 24, 26})
                if (((uint64_t)1 << l) < paths) pools.push_back(l);
        // the next deeper LDS stack variant (from 40 entries: 128-thread blocks)
        std::vector<int> stacks = {0};
        {
            const int need = std::max(s->dev.stack_needed, 1);
            for (int v : {24, 26, 32, 40, 48, 64})
                if (v > need) { if (std::find(stacks.begin(), stacks.end(), v) == stacks.end()) stacks.push_back(v); break; }
            if (need <= 32) stacks.push_back(40);   // the 128-thread family
            std::sort(stacks.begin(), stacks.end());
            stacks.erase(std::unique(stacks.begin(), stacks.end()), stacks.end());
        }
        int sort = c->ray_sort ? 1 : 0, ovl = c->overlap ? 1 : 0, sort_sh = c->ray_sort_shadow ? 1 : 0;
        std::vector<Knob> knobs = {{&c->trace_schedule[0], {0, 2, 3}},
                                   {&c->trace_schedule[1], {0, 2, 3}},
                                   {&sort, {0, 1}},
                                   {&c->ray_sort_bits, {12, 15, 18}},
                                   {&sort_sh, {0, 1}},
                                   {&ovl, {0, 1}},
                                   {&c->tile, {4, 8, 16}},
                                   {&c->pool_log2, pools},
                                   {&c->stack_min, stacks},
                                   {&c->device_loop, {0, 2}},
                                   {&c->tail_rays, {1 << 16, 1 << 18, 1 << 20}}};
        rc = render_requests(c, 1, &r, true, p->wave_paths, nullptr, nullptr);   // buffers, code objects
        double best = rc ? 0.0 : time_it();
        for (Knob& k : knobs) {
            if (rc) break;
            if ((k.field == &c->ray_sort_bits || k.field == &sort_sh) && !c->ray_sort) continue;
            if (k.field == &c->tail_rays && c->device_loop == 0) continue;
            const int keep = *k.field;
            int chosen = keep;
            for (int v : k.values) {
                if (v == keep || rc) continue;
                *k.field = v;
                c->ray_sort = sort != 0;
                c->ray_sort_shadow = sort_sh != 0;
                c->overlap = ovl != 0;
                const double t = time_it();
                if (!rc && t < best * 0.99) { best = t; chosen = v; }
            }
            *k.field = chosen;
            c->ray_sort = sort != 0;
            c->ray_sort_shadow = sort_sh != 0;
            c->overlap = ovl != 0;
        }
        (void)hipFree(d);
        if (rc) return rc;
    }
    rd_tuning t{};
    (void)rd_get_tuning(c, &t);
    {
        TuneCache& tc = tune_cache();
        std::lock_guard<std::mutex> g(tc.mu);
        tc.m[tune_key(c, s, p)] = t;
    }
    if (out) *out = t;
    return RD_OK;
}

int rd_autotune_cached(rd_ctx* c, rd_scene* s, const rd_render_params* p, rd_tuning* out, int32_t* found) {
    if (!c || !s || !p) return fail(RD_ERR_INVALID, "null argument");
    if (found) *found = 0;
    rd_tuning t{};
    bool hit = false;
    {
        TuneCache& tc = tune_cache();
        std::lock_guard<std::mutex> g(tc.mu);
        auto it = tc.m.find(tune_key(c, s, p));
        if (it != tc.m.end()) { t = it->second; hit = true; }
    }
    if (!hit) return rd_autotune(c, s, p, out);
    if (int rc = rd_set_tuning(c, &t)) return rc;
    if (found) *found = 1;
    return out ? rd_get_tuning(c, out) : RD_OK;
}

int rd_tuning_cache_save(const char* path) {
    if (!path) return fail(RD_ERR_INVALID, "null path");
    TuneCache& tc = tune_cache();
    std::lock_guard<std::mutex> g(tc.mu);
    std::FILE* f = std::fopen(path, "w");
    if (!f) return fail(RD_ERR_IO, std::string("cannot open ") + path);
    for (const auto& kv : tc.m) {
        int v[kTuneFields];
        tuning_to_ints(kv.second, v);
        std::fprintf(f, "%s", kv.first.c_str());
        for (int k = 0; k < kTuneFields; k++) std::fprintf(f, "\t%d", v[k]);
        std::fprintf(f, "\n");
    }
    return std::fclose(f) == 0 ? RD_OK : fail(RD_ERR_IO, std::string("cannot write ") + path);
}

int rd_tuning_cache_load(const char* path) {
    if (!path) return fail(RD_ERR_INVALID, "null path");
    std::FILE* f = std::fopen(path, "r");
    if (!f) return fail(RD_ERR_IO, std::string("cannot open ") + path);
    std::map<std::string, rd_tuning> got;
    char line[1024];
    int rc = RD_OK;
    while (std::fgets(line, sizeof(line), f)) {
        std::string l(line);
        while (!l.empty() && (l.back() == '\n' || l.back() == '\r')) l.pop_back();
        if (l.empty()) continue;
        const size_t tab = l.find('\t');
        if (tab == std::string::npos) { rc = fail(RD_ERR_INVALID, "tuning cache: bad line"); break; }
        int v[kTuneFields];
        const char* q = l.c_str() + tab;
        int k = 0;
        for (; k < kTuneFields; k++) {
            char* e = nullptr;
            const long x = std::strtol(q, &e, 10);
            if (e == q) break;
            v[k] = (int)x;
            q = e;
        }
        if (k != kTuneFields) { rc = fail(RD_ERR_INVALID, "tuning cache: bad line"); break; }
        got[l.substr(0, tab)] = tuning_from_ints(v);
    }
    std::fclose(f);
    if (rc) return rc;
    TuneCache& tc = tune_cache();
    std::lock_guard<std::mutex> g(tc.mu);
    for (auto& kv : got) tc.m[kv.first] = kv.second;
    return RD_OK;
}

void rd_tuning_cache_clear(void) {
    TuneCache& tc = tune_cache();
    std::lock_guard<std::mutex> g(tc.mu);
    tc.m.clear();
}

void* rd_device_alloc(rd_ctx* c, uint64_t bytes) {
    if (!c) return nullptr;
    (void)hipSetDevice(c->device);
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) { g_err = "hipMalloc failed"; return nullptr; }
    return p;
}
int rd_device_free(rd_ctx* c, void* ptr) {
    if (!c) return fail(RD_ERR_INVALID, "null context");
    HIP_TRY(hipFree(ptr));
    return RD_OK;
}
int rd_memcpy_h2d(rd_ctx* c, void* dst, const void* src, uint64_t bytes) {
    if (!c) return fail(RD_ERR_INVALID, "null context");
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return RD_OK;
}
int rd_memcpy_d2h(rd_ctx* c, void* dst, const void* src, uint64_t bytes) {
    if (!c) return fail(RD_ERR_INVALID, "null context");
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return RD_OK;
}
int rd_synchronize(rd_ctx* c) {
    if (!c) return fail(RD_ERR_INVALID, "null context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipDeviceSynchronize());
    return RD_OK;
}
int rd_kernel_timing(const rd_ctx* c, const char* stage, double* avg_ms, int64_t* count) {
    if (!c || !stage) return fail(RD_ERR_INVALID, "null argument");
    auto it = c->timing.find(stage);
    if (it == c->timing.end()) {
        if (avg_ms) *avg_ms = 0;
        if (count) *count = 0;
        return RD_OK;
    }
    if (avg_ms) *avg_ms = it->second.count ? it->second.total_ms / it->second.count : 0;
    if (count) *count = it->second.count;
    return RD_OK;
}
int rd_reset_kernel_timing(rd_ctx* c) {
    if (!c) return fail(RD_ERR_INVALID, "null context");
    c->timing.clear();
    return RD_OK;
}

}  // extern "C"
