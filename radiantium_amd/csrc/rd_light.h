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
struct ParticleState {
    float4* coeff;      // throughput; w = the loop's bounce count (int bits)
    uint64_t* key;      // particle_key(seed, particle index)
    uint32_t* rng_n;    // draws taken so far
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
hipError_t launch_lt_shade(const DevScene& sc, const LtCamera& cam, RayQueue qin, int cnt_in, const float4* hits,
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

}  // namespace rd
