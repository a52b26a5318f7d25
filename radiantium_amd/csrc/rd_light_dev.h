// rd_light_dev.h — device helpers of the light-path integrators (light
// tracing, rd_light.hip; bidirectional path tracing, rd_bdpt.hip): the
// camera's SampleWi / Importance and the lights' SampleEmit.
#pragma once
#include "rd_light.h"
#include "rd_shading.h"

namespace rd {
namespace lightpath {

// Vector3.Transform with a row-major 4x4 (row vectors, no w divide).
__device__ __forceinline__ f3 xf_point(const float* M, f3 v) {
    return mk(v.x * M[0] + v.y * M[4] + v.z * M[8] + M[12], v.x * M[1] + v.y * M[5] + v.z * M[9] + M[13],
              v.x * M[2] + v.y * M[6] + v.z * M[10] + M[14]);
}
__device__ __forceinline__ f3 xf_vector(const float* M, f3 v) {
    return mk(v.x * M[0] + v.y * M[4] + v.z * M[8], v.x * M[1] + v.y * M[5] + v.z * M[9],
              v.x * M[2] + v.y * M[6] + v.z * M[10]);
}

// PerspectiveCamera.SampleRay (Camera.cs:182-191) at screen position (sx, sy).
__device__ __forceinline__ void camera_ray(const DevCamera& c, float sx, float sy, f3* o, f3* d, float* tmin,
                                           float* tmax) {
    f3 ndc = mk(sx * c.inv_res_x, sy * c.inv_res_y, 0.0f);
    f3 nearp = xf_point(c.clip_to_camera, ndc);
    f3 dir = normalize(nearp);
    float inv_z = 1.0f / dir.z;
    *o = xf_point(c.camera_to_world, mk(0.0f));
    *d = normalize(xf_vector(c.camera_to_world, dir));
    *tmin = c.near_z * inv_z;
    *tmax = c.far_z * inv_z;
}

// Camera.Importance (Camera.cs:193-204) of a camera-space unit direction.
__device__ __forceinline__ float cam_importance(const LtCamera& c, f3 d) {
    float ct = d.z;
    if (ct <= 0) return 0;
    float focus = 1 / ct;
    float px = d.x * focus, py = d.y * focus;
    if (px < c.min_x || px > c.max_x || py < c.min_y || py > c.max_y) return 0.0f;
    return c.normalization * focus * focus * focus;
}

struct CamWi {
    f3 pos, dir;
    float pdf, we, sx, sy;
};
// PerspectiveCamera.SampleWi (Camera.cs:221-239): the camera position, the
// direction to it, We = Importance / distance^2 and the screen position
// (clip x, y divided by clip z, scaled by ScreenX - 1 / ScreenY - 1); pdf 0
// outside [near, far] or the screen.
__device__ __forceinline__ CamWi cam_sample_wi(const LtCamera& c, f3 pos) {
    CamWi r;
    r.pos = mk(0.0f); r.dir = mk(0.0f); r.pdf = 0.0f; r.we = 0.0f; r.sx = 0.0f; r.sy = 0.0f;
    f3 lt = xf_point(c.world_to_camera, pos);
    if (lt.z < c.near_z || lt.z > c.far_z) return r;
    f3 screen = xf_point(c.camera_to_clip, lt);
    screen.x /= screen.z;
    screen.y /= screen.z;
    if (screen.x < 0 || screen.x > 1 || screen.y < 0 || screen.y > 1) return r;
    r.sx = screen.x * (float)(c.width - 1);
    r.sy = screen.y * (float)(c.height - 1);
    float dist = length(lt);
    float inv = 1 / dist;
    f3 local_dir = lt * inv;
    r.pos = xf_point(c.camera_to_world, mk(0.0f));
    r.dir = (r.pos - pos) / dist;
    r.pdf = 1;
    r.we = cam_importance(c, local_dir) * inv * inv;
    return r;
}

struct EmitS {
    f3 p, dir, n, le;
    float pdf_pos, pdf_dir;
};
// Light.SampleEmit: DiffuseAreaLight.SampleEmit (DiffuseAreaLight.cs:70-77),
// InfiniteAreaLight.SampleEmit (InfiniteAreaLight.cs:122-133).  False when
// the environment map's SampleLightDir gives pdf 0 (the reference would
// divide by it; the particle ends instead).
__device__ __forceinline__ bool sample_emit(const DevScene& sc, const DevLight& L, Rng& rng, EmitS* e) {
    if (L.type == 1) {
        const DevEnv& env = sc.envs[L.tex];
        float u0 = rng.next_float(), u1 = rng.next_float();
        f3 wdir;
        float pdf, x, y;
        if (!env_sample_dir(env, u0, u1, &wdir, &pdf, &x, &y)) return false;
        e->le = tex_sample(sc, env.tex, x, y);
        f3 normal = -wdir;
        Frame coord = make_frame(normal);
        float c0 = rng.next_float(), c1 = rng.next_float();
        // Probability.SquareToConcentricDisk (Probability.cs:157-174)
        float ox = 2.0f * c0 - 1, oy = 2.0f * c1 - 1;
        float cx = 0.0f, cy = 0.0f;
        if (!(ox == 0 && oy == 0)) {
            float theta, r;
            if (fabsf(ox) > fabsf(oy)) { r = ox; theta = (kPi / 4) * (oy / ox); }
            else { r = oy; theta = (kPi / 2) - (kPi / 4) * (ox / oy); }
            cx = r * cosf(theta);
            cy = r * sinf(theta);
        }
        const f3 center = mk(env.center[0], env.center[1], env.center[2]);
        f3 disk = center + env.radius * (cx * coord.X + cy * coord.Y);
        e->p = disk + wdir * env.radius;
        e->dir = -wdir;
        e->n = -wdir;
        e->pdf_pos = 1 / (kPi * sqr(env.radius));
        e->pdf_dir = pdf;
        return true;
    }
    float ux = rng.next_float(), uy = rng.next_float();
    f3 P, N;
    area_shape_sample(sc, L, ux, uy, &P, &N);
    // Probability.SquareToCosineHemisphere (:94-100) via SquareToUniformDisk
    float u0 = rng.next_float(), u1 = rng.next_float();
    float radius = sqrtf(u0);
    float angle = u1 * kPi * 2;
    float x = radius * cosf(angle), y = radius * sinf(angle);
    f3 lw = normalize(mk(x, y, sqrtf(1 - x * x - y * y)));
    e->pdf_dir = cos_theta(lw) / kPi;
    e->dir = normalize(to_world(make_frame(N), lw));
    e->p = P;
    e->n = N;
    e->le = mk(L.le[0], L.le[1], L.le[2]);
    e->pdf_pos = 1.0f / L.area;
    return true;
}

}  // namespace lightpath
}  // namespace rd
