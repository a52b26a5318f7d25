/* cpu_ref.h — TEST INFRASTRUCTURE: C API of the oracle ("cpu_ref"), a scalar
 * C++ restatement of the reference's path / vol_path hot path
 * (ksgfk/Radiantium, C#/.NET 6).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never links it.
 * It consumes the same rd_scene_desc as the product's C ABI. */
#ifndef CPU_REF_H
#define CPU_REF_H
#include <stdint.h>
#include "../include/radiantium_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cr_scene cr_scene;

const char* cr_last_error(void);
cr_scene* cr_scene_create(const rd_scene_desc* desc);
void cr_scene_destroy(cr_scene* s);
/* nodes32: num_nodes*32 bytes in LinearBVHNode layout; prim_order: BVH order
 * of primitive indices.  Returns the node count (or -1). */
int32_t cr_scene_num_nodes(const cr_scene* s);
int32_t cr_scene_max_depth(const cr_scene* s);
int cr_scene_export_bvh(const cr_scene* s, void* nodes32, int32_t* prim_order);
/* The loose octree of an RD_ACCEL_OCTREE scene (Accelerators/Octree.cs):
 * 32-byte LinearNode records (bound, NodeCount / LeafIndex, NextHead = -1
 * for leaves) and leaf i's primitives leaf_prims[leaf_first[i] ..
 * + leaf_count[i]); sizes3 = {nodes, leaves, leaf entries}; null buffers are
 * skipped, so a first call with nulls sizes them. */
int cr_scene_export_octree(const cr_scene* s, void* nodes32, int32_t* leaf_first, int32_t* leaf_count,
                           int32_t* leaf_prims, int32_t* sizes3);
/* The same tree over caller bounds (bmin / bmax: [n][3]); now_depth = the
 * deepest level built (root 0). */
int cr_octree_build(const float* bmin, const float* bmax, int32_t n, int32_t max_depth, int32_t max_count,
                    float out_bound, void* nodes32, int32_t* leaf_first, int32_t* leaf_count, int32_t* leaf_prims,
                    int32_t* sizes3, int32_t* now_depth);
/* counters (may be NULL): [0] nodes visited (AABB tests), [1] primitives tested */
int cr_trace_closest(const cr_scene* s, const rd_ray* rays, int32_t n, rd_hit* hits, uint64_t* counters);
int cr_trace_any(const cr_scene* s, const rd_ray* rays, int32_t n, uint8_t* occluded, uint64_t* counters);
/* Brute force over all primitives in input order (the Naive aggregate,
 * Accelerators/Naive.cs): smallest t, first index on ties. */
int cr_trace_brute(const cr_scene* s, const rd_ray* rays, int32_t n, rd_hit* hits);
/* As cr_trace_closest / cr_trace_brute, plus sub[i] = the hit triangle's index
 * within its instance set (-1 for other hits and misses). */
int cr_trace_closest_sub(const cr_scene* s, const rd_ray* rays, int32_t n, rd_hit* hits, int32_t* sub,
                         uint64_t* counters);
int cr_trace_brute_sub(const cr_scene* s, const rd_ray* rays, int32_t n, rd_hit* hits, int32_t* sub);
/* Camera.SampleRay for pixel sample positions (x,y) (already offset by U^2). */
int cr_camera_rays(const cr_scene* s, const float* xy, int32_t n, rd_ray* rays);
/* Renders like BlockBasedRenderer with `threads` workers over 32x32 blocks
 * of the crop in params.  out_rgb / out_sum2: full W*H*3 column-major
 * (pixels outside the crop untouched).  stats: [0] closest queries,
 * [1] any-hit queries, [2] paths, [3] vertices, [4] invalid samples. */
int cr_render(const cr_scene* s, const rd_render_params* params, int32_t threads,
              float* out_rgb, float* out_sum2, uint64_t* stats, double* seconds);
/* Single-sample radiance for (pixel x, y, sample index): returns the Li
 * value before the IsValid replacement; for debugging / per-sample tests. */
int cr_sample_li(const cr_scene* s, uint32_t seed, int32_t x, int32_t y, int32_t sample, float* rgb);

/* BSDF probes in the local shading frame (uv given): value, pdf, sampled
 * direction, with an explicit RNG key so sampling is reproducible. */
int cr_material_eval(const cr_scene* s, int32_t material, const float* wo, const float* wi,
                     const float* uv, float* fr_rgb, float* pdf);
int cr_material_sample(const cr_scene* s, int32_t material, const float* wo, const float* uv,
                       uint64_t key, float* wi, float* fr_rgb, float* pdf, int32_t* type);
/* The same over n direction pairs / n samples (keys key0 + j). */
int cr_material_eval_n(const cr_scene* s, int32_t material, int32_t n, const float* wo, const float* wi,
                       const float* uv, float* fr_rgb, float* pdf);
int cr_material_sample_n(const cr_scene* s, int32_t material, int32_t n, const float* wo, const float* uv,
                         uint64_t key0, float* wi, float* fr_rgb, float* pdf, int32_t* type);

/* Distribution1D / Distribution2D probes on a caller's function values.
 * dist1d: out = {continuous sample, pdf, discrete pdf of the sampled
 * interval} and *offset; dist2d_sample: out = {d0, d1, pdf}, offsets (vj, vi);
 * dist2d_pdf: ContinuousPdf(u, v).  f2 is nv rows of nu values. */
void cr_dist1d_sample(const float* f, int32_t n, float u, float* out3, int32_t* offset);
void cr_dist2d_sample(const float* f2, int32_t nu, int32_t nv, float u0, float u1, float* out3, int32_t* vj,
                      int32_t* vi);
float cr_dist2d_pdf(const float* f2, int32_t nu, int32_t nv, float u, float v);
/* InfiniteAreaLight probes for the scene's env-th infinite light: Le of a
 * world direction, PdfDir, SampleLightDir from two uniforms (returns 0 when
 * the reference returns pdf 0), world centre / radius after Preprocess. */
int cr_env_le(const cr_scene* s, int32_t env, const float* dir, float* rgb);
float cr_env_pdf(const cr_scene* s, int32_t env, const float* wi);
int cr_env_sample(const cr_scene* s, int32_t env, float u0, float u1, float* wdir, float* pdf, float* uv);
int cr_env_bounds(const cr_scene* s, int32_t env, float* center, float* radius);

/* Homogeneous medium and Henyey-Greenstein probes (Mediums/Homogeneous.cs,
 * IPhaseFunction.cs): Tr over [tmin, tmax]; free-flight Sample with an RNG
 * key (returns 1 for a medium vertex: p, weight tr, t; 0 when the segment
 * is passed: weight tr); HG value and SampleWi (returns p). */
void cr_medium_tr(const rd_medium_desc* m, float tmin, float tmax, float* rgb);
int cr_medium_sample(const rd_medium_desc* m, const float* o, const float* d, float tmin, float tmax, uint64_t key,
                     float* p, float* tr, float* t);
float cr_hg_p(float g, const float* wo, const float* wi);
float cr_hg_sample(float g, const float* wo, uint64_t key, float* wi);

/* Adjoint reference (the test reference for rd_adjoint): adds dLoss/dtexel
 * to tex_grads[t] (NULL = skip) for dl_dimage = dLoss/dImage of the render
 * with the same params; cr_rho_ratio probes dC/drho / C of a rough_plastic
 * term f(wo, wi) at uv. */
int cr_adjoint(const cr_scene* s, const rd_render_params* params, int32_t threads, const float* dl_dimage,
               float* const* tex_grads);
float cr_rho_ratio(const cr_scene* s, int32_t material, const float* uv, const float* wo, const float* wi);

/* Closed-form probes: Fresnel.DielectricFunc / ConductorFunc, and the
 * microfacet distribution (kind 0 Beckmann, 1 GGX): out = {D(wh), G(wo,wi),
 * Pdf(wo,wh)}; sample_wh draws with the given RNG key. */
float cr_fresnel_dielectric(float cos_i, float eta_i, float eta_t);
/* BSSRDF probes (Bssrdf.cs): the 2048-entry radius table; FresnelMoment;
 * the normalised-diffusion profile for n (channel, u) pairs: out[6j..] =
 * {SampleSr radius, PdfSr(channel, r), Sr(r) rgb, D[channel]}; the adapter
 * for n (wo, wi): out[4j..] = {Fr rgb, Pdf}; SamplePi at a point po of
 * primitive `prim` (shading normal N, Wr, UV) for n keys key0 + j:
 * out[9j..] = {ok, S rgb, pdf, pi.P xyz, closest-hit queries}. */
void cr_sss_lut(float* out2048);
float cr_fresnel_moment(float eta);
void cr_sss_profile(const float* a, const float* d, int32_t n, const int32_t* channel, const float* u, float* out);
void cr_sss_adapter(float eta, int32_t n, const float* wo, const float* wi, float* out);
int cr_sss_sample_pi(const cr_scene* s, int32_t prim, const float* P, const float* N, const float* wr, const float* uv,
                     uint64_t key0, int32_t n, float* out);
void cr_fresnel_conductor(float cos_i, const float* eta, const float* k, float* out_rgb);
void cr_mf_eval(int32_t kind, float roughness, float aniso, const float* wh, const float* wo, const float* wi, float* out3);
void cr_mf_sample(int32_t kind, float roughness, float aniso, const float* wo, uint64_t key, float* wh);
/* Pixel-sample camera ray for an explicit RNG-free sample position is
 * cr_camera_rays; this returns the oracle RNG's first n floats for a key. */
void cr_rng_floats(uint64_t key, int32_t n, float* out);
uint64_t cr_path_key(uint32_t seed, uint32_t pixel, uint32_t sample);

/* Light tracing (AdjointParticleIntegrator / AdjointParticleRenderer):
 * `particles` particles with keys particle_key(seed, i); out_rgb W*H*3
 * column-major; stats as cr_render ([4] unused). */
int cr_render_particles(const cr_scene* s, uint64_t particles, uint32_t seed, int32_t threads, float* out_rgb,
                        uint64_t* stats, double* seconds);
/* Bidirectional path tracing (BidirectionalPathTracer under
 * BidirectionalPathTracingRenderer): per pixel of the crop the mean of spp
 * samples; the camera subpath draws from path_key(seed, pixel, sample), the
 * light subpath from path_key(seed, pixel, sample | 2^30), the s = 1 light
 * samples from path_key(seed, pixel, sample | 2^29).  As cr_render. */
int cr_render_bdpt(const cr_scene* s, const rd_render_params* params, int32_t threads, float* out_rgb,
                   float* out_sum2, uint64_t* stats, double* seconds);
/* The same with strategy (s, t) of Connect alone (every other one 0). */
int cr_render_bdpt_strategy(const cr_scene* s, const rd_render_params* params, int32_t threads, int32_t strat_s,
                            int32_t strat_t, float* out_rgb);
/* Camera.SampleWi of a world point: out6 = {ok, screen x, screen y, We,
 * pdf, distance}; Camera.Importance of a camera-space unit direction. */
int cr_camera_sample_wi(const cr_scene* s, const float* pos, float* out6);
float cr_camera_importance(const cr_scene* s, const float* d);
/* Light.SampleEmit of light `light` drawn with RNG key `key`: out14 =
 * {p, dir, n, le, pdf_pos, pdf_dir}; 1 on success, 0 when the pdf is 0. */
int cr_light_sample_emit(const cr_scene* s, int32_t light, uint64_t key, float* out14);
/* Material probes under a transport mode: imp = 1 TransportMode.Importance,
 * 0 Radiance (cr_material_eval / cr_material_sample). */
int cr_material_eval_mode(const cr_scene* s, int32_t material, int32_t imp, const float* wo, const float* wi,
                          const float* uv, float* fr_rgb);
int cr_material_sample_mode(const cr_scene* s, int32_t material, int32_t imp, const float* wo, const float* uv,
                            uint64_t key, float* wi, float* fr_rgb, float* pdf, int32_t* type);

#ifdef __cplusplus
}
#endif
#endif
