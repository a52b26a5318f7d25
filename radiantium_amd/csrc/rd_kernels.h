// rd_kernels.h — host-side launchers for the gfx950 kernels in rd_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "rd_types.h"

namespace rd {

// Ray queue, structure of arrays.  org = {o.xyz, tmin}, dir = {d.xyz, tmax}.
struct RayQueue {
    int* count;
    int* pid;
    float4* org;
    float4* dir;
};

// Per-path wavefront state (capacity = paths per wave).
struct PathState {
    float4* L;          // radiance
    float4* beta;       // path throughput (PathTracer "coeff")
    float4* nee_beta;   // throughput at the NEE vertex; w = light selection pdf
    float4* nee_le;     // light-sample contribution, gated by the shadow ray
    float4* mis_f;      // BSDF-sample f*|cos|; w = MIS weight
    float4* mis_li;     // radiance found by the MIS ray; w = BSDF-sample pdf
    int* flags;
    int* sel_light;
    uint64_t* key;      // RNG key of the path: path_key(seed, pixel, sample)
    int* req;           // request (scene) index of the path
    int* gid;           // the path's index in the current pass
    int* alive;         // 1: the slot's path continues with ray_o / ray_d next bounce
    float4* ray_o;      // next ray of the slot's path: {o.xyz, tmin}
    float4* ray_d;      //                              {d.xyz, tmax}
    float4* res;        // per pass path: final radiance (read in sample order by accumulate)
    int* pix;           // image element of the path's pixel: x * height + y
    // adjoint replay only: {R, material (int bits)}, {u, v, r_nee, r_mis}, {r_cont}
    float4* adj0;
    float4* adj1;
    float4* adj2;
    uint32_t* rng_n;
    int* depth;
    uint8_t* vis;
    // vol_path only (VolumetricPathTracer.cs:28-314)
    int* env;           // medium of the path's current ray (-1: vacuum)
    float4* nee_li;     // light sample Li; w = light pdf (nee_le then holds f*|cos|, w = MIS weight)
    float4* sh_to;      // Transmittance chain: light point; w = current medium (int bits)
    float4* sh_tr;      // Transmittance chain: accumulated Tr
    float4* mis_tr;     // IntersectTr chain: accumulated Tr; w = current medium (int bits)
    // instanced scenes: set triangle (DevPrim index) of the hit at each queue
    // position, for the extension / Transmittance hits and the IntersectTr hits
    int* hsub;
    int* hsub2;
    // path Mis over subsurface materials (PathTracer.cs:244-259): the paths
    // whose sample entered a BSSRDF this bounce (kCntSss), those whose probe
    // found a point (kCntSssB), and per path the point {P, material},
    // {shading normal, u}, {Wr, v}: first po, then the sampled pi.  Null
    // when the scene has no BSSRDF.
    int* sss_q;
    int* sss_q2;
    float4* sss0;
    float4* sss1;
    float4* sss2;
};

struct WaveParams {
    int x0, y0, crop_w, crop_h;
    int npix;           // crop_w * crop_h
    int sample_base;    // first sample index of this wave
    int nsamples;       // samples per pixel in this wave
    int width;          // full image width (pixel id = y * width + x)
    uint32_t seed;
    int tile;           // path-order tile edge (the accumulation buffers' order)
};

// Device counters used by the bounce loop.
// 64-bit statistics occupy two int slots at an even index.
enum : int {
    kCntExtA = 0, kCntExtB = 1, kCntShadow = 2, kCntMis = 3,
    kCntRaysClosest = 4, kCntRaysAny = 6, kCntVertices = 8, kCntInvalid = 10,
    kCntWork = 12,        // work-queue cursors (4 slots)
    kCntClassRays = 16,   // 3 x u64: extension, MIS, shadow
    kCntClassNodes = 22,  // 3 x u64
    kCntClassPrims = 28,  // 3 x u64
    kCntShadowB = 34, kCntMisB = 35,   // vol_path chain queues, second buffers
    kCntFree = 36,        // pass path index of the first path the last compaction started
    kCntNext = 37,        // next pass path to start
    kCntTail = 38,        // bounces the device-driven tail processed
    kCntChainStuck = 39,  // set when a medium chain reached kMaxChainSteps
    kCntSss = 40, kCntSssB = 41,   // BSSRDF queues
    kCntSlots = 42
};

// A Transmittance / IntersectTr chain takes one step per medium boundary it
// crosses; this bounds it for pathological geometry (the render then fails).
constexpr int kMaxChainSteps = 1 << 16;

// One segment of a batched wave: samples [sample_base, sample_base + nsamples)
// of every pixel of request `req`'s crop, stored at paths [path_offset, ...)
// sample-major, and accumulated into that request's sums at sum_offset.
struct BatchSeg {
    int req;
    int path_offset;
    int sample_base, nsamples;
    int x0, y0, crop_w, npix;
    int width, height;
    uint32_t seed;
    int sum_offset;
    int tile;   // path-order tile edge in pixels (tile_pixel)
};

// Adjoint replay inputs: dLoss/dImage (column-major W*H*3, device), the
// float64 gradient accumulator of every scene texture (device array of device
// pointers, null = no gradient; layout as the texture, 3 values per texel:
// many path terms of both signs meet in one texel, so they are summed in
// float64), 1/spp.
struct AdjArgs {
    const float* dl;
    double* const* grads;
    float inv_spp;
    int _pad;
};

// out[i] += (float)acc[i] for i < n.
hipError_t launch_add_f64_to_f32(const double* acc, float* out, int64_t n, hipStream_t s);

struct Timing;  // per-stage HIP event timing (rd_api.cpp)

// Traversal over caller-provided ray arrays (rd_ray / rd_hit layouts).
hipError_t launch_trace_rays(const DevScene& sc, const float4* rays, int n, float4* hits, uint8_t* occ, int* sub,
                             unsigned long long* counters, bool any, int* counts, hipStream_t s);

// Wavefront path tracing stages.  A pass is a list of segments (BatchSeg);
// its paths stream through a pool of `pool` slots.  init_pool frees every
// slot; compact builds the next extension queue in slot order from the slots
// whose path continues and starts the pass's next paths in the free slots
// (slot order, consecutive pass indices); resolve retires finished paths into
// ps.res.  blk: compact_scratch_int2(pool) int2 of scratch.
// compact_queue: the next queue from the n entries of the previous one whose
// path continues, same order; for bounces after every path of the pass has
// started (it starts none).
hipError_t launch_init_pool(PathState ps, int pool, int* counts, hipStream_t s);
hipError_t launch_compact(const DevScene& sc, const DevScene* scenes, const BatchSeg* segs, int nseg, int npass,
                          int pool, int2* blk, RayQueue q, int cnt_q, PathState ps, int* counts, hipStream_t s);
// n_dev (optional): the queue's count on the device, at most n, which then
// only sizes the grids (the device-driven tail does not read counts back).
hipError_t launch_compact_queue(RayQueue qin, int n, int npass, int2* blk, RayQueue q, int cnt_q, PathState ps,
                                int* counts, hipStream_t s, const int* n_dev = nullptr);
inline size_t compact_scratch_int2(size_t pool) {
    const size_t nblk = (pool + 1023) / 1024;
    return 2 * nblk + 2 * ((nblk + 1023) / 1024);
}
// perm: trace order (null = queue order); rank (with perm) receives each
// queue position's place in that order.
hipError_t launch_trace_ext(const DevScene& sc, const DevScene* scenes, RayQueue q, PathState ps, float4* hits,
                            int* counts, int cnt_slot, int za, int zb, int zc, const int* perm, int* rank, int sched,
                            hipStream_t s);
// Coherence order of the first n rays of q (direction octant, then the
// origin's Morton code in the root box of `nodes`, `bits` key bits in all,
// 3..30): perm receives the queue
// positions in traversal order.  keys / keys_alt / vals: n each; temp of
// ray_sort_temp_bytes(>= n) bytes (0: the size query failed).
size_t ray_sort_temp_bytes(int max_n);
hipError_t launch_ray_sort(const float4* nodes, RayQueue q, int n, int bits, uint32_t* keys, uint32_t* keys_alt,
                           int* vals, int* perm, void* temp, size_t temp_bytes, hipStream_t s);
// `scenes` (device array, may be null): batched mode, the path's scene is
// scenes[ps.req[pid]]; otherwise every path uses `sc`.
// rank (null = append): write each path's shadow ray at rank[queue
// position], holes pid -1, shadow queue size = the extension queue's; MIS
// rays are appended either way.
// rank_identity (rank null): ranked with every ray at its own queue position.
// PathTracer.Mis's BSSRDF branch for the paths k_shade queued (kCntSss):
// sample pi and weight the path (queue kCntSssB), then light pi through the
// adapter and continue (lut: the 2048 radii of the normalised-diffusion CDF).
hipError_t launch_sss_probe(const DevScene& sc, PathState ps, int* counts, const float* lut, hipStream_t s);
hipError_t launch_sss_shade(const DevScene& sc, RayQueue qshadow, RayQueue qmis, PathState ps, int* counts,
                            hipStream_t s);
hipError_t launch_shade(const DevScene& sc, const DevScene* scenes, RayQueue qin, const float4* hits,
                        RayQueue qshadow, RayQueue qmis, PathState ps, int* counts, int cnt_in, bool adjoint,
                        const int* rank, bool rank_identity, hipStream_t s);
// sum: float64 per-pixel sums {r, g, b, 0}; sum2: float32 sums of squares
hipError_t launch_accumulate_batch(const BatchSeg* segs, int nseg, int max_npix, PathState ps, double4* sum,
                                   float4* sum2, int* counts, hipStream_t s);
hipError_t launch_trace_shadow(const DevScene& sc, const DevScene* scenes, RayQueue q, PathState ps, int* counts,
                               int cnt_slot, int sched, hipStream_t s);
hipError_t launch_trace_mis(const DevScene& sc, const DevScene* scenes, RayQueue q, PathState ps, int* counts,
                            int sched, hipStream_t s);
// Both of the above in one launch where the schedules allow it.
hipError_t launch_trace_secondary(const DevScene& sc, const DevScene* scenes, RayQueue qshadow, int cnt_shadow,
                                  RayQueue qmis, PathState ps, int* counts, int sched_shadow, int sched_mis,
                                  hipStream_t s);
// The n entries of a ranked queue that hold a ray, in order, into q (count at
// cnt_q); blk as for launch_compact.
hipError_t launch_compact_holes(RayQueue qin, int n, int2* blk, RayQueue q, int cnt_q, int* counts, hipStream_t s,
                                const int* n_dev = nullptr);
// Counts one more tail bounce (kCntTail) when the queue at cnt_in is not empty.
hipError_t launch_tail_tick(int* counts, int cnt_in, hipStream_t s);
// adj: null for rendering; the adjoint replay's gradient inputs otherwise.
hipError_t launch_resolve(const DevScene& sc, int method, RayQueue qin, PathState ps, int* counts, int cnt_in,
                          const AdjArgs* adj, hipStream_t s);
// vol_path stages (VolumetricPathTracer.Li and its Scene.Transmittance /
// Scene.IntersectTr chains): shade_vol draws every random number of the
// bounce and starts one Transmittance chain (light sample) and one
// IntersectTr chain (BSDF / phase sample) per vertex; each chain step is a
// closest-hit query (trace_chain) followed by chain_advance, which either
// finishes the chain into the path state or appends its next segment.
hipError_t launch_shade_vol(const DevScene& sc, const DevScene* scenes, RayQueue qin, const float4* hits,
                            RayQueue qshadow, RayQueue qmis, PathState ps, int* counts, int cnt_in, hipStream_t s);
hipError_t launch_trace_chain(const DevScene& sc, const DevScene* scenes, RayQueue q, PathState ps, float4* hits,
                              bool mis, int* counts, int cnt_slot, int zero_slot, int sched, hipStream_t s);
hipError_t launch_chain_advance(const DevScene& sc, const DevScene* scenes, RayQueue qin, const float4* hits,
                                RayQueue qout, bool mis, PathState ps, int* counts, int cnt_in, int cnt_out,
                                hipStream_t s);
hipError_t launch_resolve_vol(RayQueue qin, PathState ps, int* counts, int cnt_in, hipStream_t s);
// Every chain of the queue (count at cnt_slot) walked to its end by one lane:
// its segments' closest-hit queries and chain steps, in the order the
// lockstep stages above take them; the work cursor must be zero on entry
// (shade_vol clears it).
hipError_t launch_chain_full(const DevScene& sc, const DevScene* scenes, RayQueue q, PathState ps, bool mis,
                             int* counts, int cnt_slot, hipStream_t s);
// The per-pixel mean into out_rgb (float32), or into out64 (float64) when it is not null.
hipError_t launch_finalize(const WaveParams& wp, int height, int spp, const double4* sum, const float4* sum2,
                           float* out_rgb, float* out_sum2, double* out64, hipStream_t s);
hipError_t launch_zero(float4* buf, int n, hipStream_t s);
hipError_t launch_zero(double4* buf, int n, hipStream_t s);
hipError_t launch_zero_counts(int* counts, int n, hipStream_t s);
hipError_t launch_zero_slots(int* counts, int a, int b, hipStream_t s);

// Traversal schedules (the `sched` argument of the queue launchers):
// 0 = straight (grid-stride, one ray per lane to completion),
// 1 = refill (idle lanes refilled from an atomic cursor),
// 2 = straight with 64-ray chunks per wavefront from an atomic cursor,
// 3 = sequences: 4 x 64-ray chunks per wavefront, each lane tracing its four
//     rays back to back.
// Batched queues take sequences for 3 and chunks otherwise; instanced scenes
// run straight (0, or chunks for 2 and 3).
constexpr int kNumSchedules = 4;

// Occupancy-sized persistent grid for the device.
void set_device_grid(int num_cus);
int persistent_blocks();

}  // namespace rd
