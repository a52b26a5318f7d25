// rd_light.h — light tracing (the reference's "adjoint_particle" renderer):
// camera state for connecting surface points to the camera, and the
// launchers of rd_light.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "rd_kernels.h"

namespace rd {

// What Camera.SampleWi and Camera.Importance read (Camera.cs:128-239):
// the clip transform and its inverse's image-plane bounds at z = 1, and the
// world/camera transforms (row-major M11..M44, row vectors).
struct LtCamera {
    float camera_to_clip[16];
    float world_to_camera[16];
    float camera_to_world[16];
    float min_x, min_y, max_x, max_y;
    float normalization;   // 1 / image-plane area at z = 1
    float near_z, far_z;
    int width, height;
    int _pad[3];
};

// Particle state of one wave (capacity = particles per wave).
// A live particle's state travels with its extension-queue entry (two
// buffers, one beside each of the ping-pong queues), so every stage reads
// and writes it in queue order; its random stream's key is
// particle_key(seed, wave base + the entry's pid), recomputed.
struct ParticleState {
    float4* st[2];      // per queue entry: throughput; w = the loop's bounce count (int bits)
    uint32_t* rng_n[2]; // per queue entry: draws taken so far
    float4* splat;      // per shadow-queue entry: colour; w = image element x * H + y (int bits)
    uint8_t* vis;       // per shadow-queue entry: 1 when the camera sees the vertex
    int* hsub;          // instanced scenes: set triangle of each extension hit
};

// Emits particles [base, base + n): light pick, Light.SampleEmit, the first
// loop head; live particles are appended to q (count at cnt_q).
hipError_t launch_lt_emit(const DevScene& sc, uint32_t seed, uint64_t base, int n, ParticleState ps, RayQueue q,
                          int* counts, int cnt_q, hipStream_t s);
// One bounce of every particle in qin (closest hits in hits): the camera
// connection (a shadow ray into qs with its splat) and the importance-mode
// BSDF sample (the next ray into qout).
// side: the state buffer beside qin (qout's is the other one).
hipError_t launch_lt_shade(const DevScene& sc, const LtCamera& cam, uint32_t seed, uint64_t base, int side,
                           RayQueue qin, int cnt_in, const float4* hits,
                           ParticleState ps, RayQueue qout, int cnt_out, RayQueue qs, int cnt_s, int* counts,
                           hipStream_t s);
// Adds the splat of every unoccluded shadow-queue entry into acc (float64,
// ColorBuffer layout).
hipError_t launch_lt_splat(RayQueue qs, int cnt_s, ParticleState ps, double* acc, int* counts, hipStream_t s);
// The camera ray through each pixel corner (x, y) of a W x H image into q,
// pixel x * H + y at queue position x * H + y.
hipError_t launch_lt_corner_rays(const DevScene& sc, RayQueue q, int* counts, int cnt_q, hipStream_t s);
// out = float(acc) * v + the emission the corner ray meets.
hipError_t launch_lt_final(const DevScene& sc, RayQueue q, const float4* hits, const int* hsub, const double* acc,
                           float v, float* out, hipStream_t s);

// ---------------------------------------------------------------- bidirectional (rd_bdpt.hip)
// One pass of bidirectional samples: every pixel of the crop (raster order)
// times samples [sample_base, sample_base + P / npix).
struct BdptPass {
    int x0, y0, crop_w, npix;
    int width;          // image width (pixel id = y * width + x)
    int sample_base;
    uint32_t seed;
    int _pad;
};

// Per-pass device storage.  Vertex slot k of subpath `side` (0 camera,
// 1 light) of sample i sits at (side * vmax + k) * P + i in five arrays:
//   v0 {p, type}  v1 {n, medium}  v2 {coeff, aux}  v3 {w, u}  v4 v
// (aux: a surface vertex's primitive, a light vertex's light or -1; w: a
// surface vertex's Wr, a medium vertex's Wo).  Strategy slot kk = (t - 1) *
// vmax + s of sample i at kk * P + i.
struct BdptBuffers {
    int P;              // samples in the pass
    int vmax;           // vertex slots per subpath
    int nstrat;         // (vmax - 1) * vmax strategy slots per sample
    int walk_cap;       // RandomWalk's maxDepth: the integrator's max_depth - 1
    float4 *v0, *v1, *v2, *v3;
    float* v4;
    int* nverts;        // [2][P]: nCamera, nLight
    float4* wcoeff;     // [2P] walkers (camera i, light P + i): coeff, w = bounces
    int* wenv;
    uint64_t* wkey;
    uint32_t* wn;
    float4* conn;       // strategy value without visibility; w = the factor g the transmittance is scaled by
    float4* cto;        // Transmittance chain: target, w = current medium
    float4* ctr;        // Transmittance chain: accumulated Tr
};

hipError_t launch_bd_start(const DevScene& sc, const BdptPass& w, const BdptBuffers& b, RayQueue q, int* counts,
                           int cnt_q, hipStream_t s);
hipError_t launch_bd_walk(const DevScene& sc, const BdptBuffers& b, RayQueue qin, int cnt_in, const float4* hits,
                          const int* hsub, RayQueue qout, int cnt_out, int* counts, hipStream_t s);
hipError_t launch_bd_connect(const DevScene& sc, const LtCamera& cam, const BdptPass& w, const BdptBuffers& b,
                             RayQueue qc, int cnt_c, int* counts, hipStream_t s);
hipError_t launch_bd_chain(const DevScene& sc, const BdptBuffers& b, RayQueue qin, int cnt_in, const float4* hits,
                           const int* hsub, RayQueue qout, int cnt_out, int* counts, hipStream_t s);
hipError_t launch_bd_sum(const BdptBuffers& b, float4* res, hipStream_t s);
hipError_t launch_bd_accumulate(const BdptPass& w, int nsamples, const float4* res, double4* sum, float4* sum2,
                                hipStream_t s);
hipError_t launch_bd_finalize(const BdptPass& w, int height, int spp, const double4* sum, const float4* sum2,
                              float* out, float* out_sum2, double* out64, hipStream_t s);

}  // namespace rd
