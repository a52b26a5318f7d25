// rd_light.hip — gfx950 wavefront stages of light tracing: the reference's
// AdjointParticleIntegrator.EmitParticle
// (Radiantium.Offline/Integrators/AdjointParticleIntegrator.cs:24-80) under
// AdjointParticleRenderer (Renderers/AdjointParticleRenderer.cs:127-178).
//
// One particle is a light path; a wave holds up to the pool's particles and
// every bounce of the wave is
//   trace_ext   Scene.Intersect of the live particles (closest hit, the path
//               tracer's traversal kernels)
//   lt_shade    the camera connection (Camera.SampleWi + the splat's colour,
//               its IsOccluded ray appended to the shadow queue) and the
//               importance-mode BSDF sample, then the next loop head (depth
//               cap, Russian roulette) — every random draw of the bounce in
//               the reference's order
//   trace_any   the shadow rays (any hit)
//   lt_splat    the splats of the unoccluded rays, added into a float64
//               image with atomics (the reference adds under a lock, in
//               whatever order its threads arrive)
// and after the last wave one closest-hit query per pixel corner adds the
// emission the camera sees directly (CameraTraceLight).
#include "rd_light.h"
#include "rd_light_dev.h"
#include "rd_shading.h"
#include "rd_wave.h"

namespace rd {

namespace {

using namespace lightpath;

// The loop head of EmitParticle (AdjointParticleIntegrator.cs:35-44): the
// depth cap, then Russian roulette with the fixed threshold.
__device__ __forceinline__ bool loop_head(const DevIntegrator& I, int bounces, f3& coeff, Rng& rng) {
    if (I.max_depth != -1 && bounces >= I.max_depth) return false;
    if (bounces > I.min_depth) {
        if (rng.next_float() > I.rr_threshold) return false;
        coeff = coeff / I.rr_threshold;
    }
    return true;
}

__global__ __launch_bounds__(256) void k_lt_emit(DevScene sc, uint32_t seed, uint64_t base, int n, ParticleState ps,
                                                 RayQueue q, int* counts, int cnt_q) {
    __shared__ int sh[256 / 64 + 1];
    if (blockIdx.x == 0 && threadIdx.x < 2) counts[kCntWork + threadIdx.x] = 0;   // traversal work cursors
    for (int b0 = blockIdx.x * 256; b0 < n; b0 += gridDim.x * 256) {
        const int i = b0 + threadIdx.x;
        bool alive = false;
        f3 coeff = mk(0.0f);
        EmitS e;
        e.p = mk(0.0f); e.dir = mk(0.0f);
        Rng rng{particle_key(seed, base + (uint64_t)i), 0u};
        if (i < n && sc.num_lights > 0) {
            int l = rng.next_int(sc.num_lights);                  // Scene.SampleLight (Scene.cs:47-57)
            float select_pdf = 1.0f / sc.num_lights;
            if (sample_emit(sc, sc.lights[l], rng, &e) && !is_black(e.le)) {
                coeff = e.le * abs_dot(e.dir, e.n) / (select_pdf * e.pdf_pos * e.pdf_dir);
                alive = loop_head(sc.integ, 1, coeff, rng);
            }
        }
        int pos = block_append<256>(counts + cnt_q, alive, sh);
        if (alive) {
            q.pid[pos] = i;
            q.org[pos] = make_float4(e.p.x, e.p.y, e.p.z, 0.0001f);
            q.dir[pos] = make_float4(e.dir.x, e.dir.y, e.dir.z, kFloatMax);
            ps.st[0][pos] = make_float4(coeff.x, coeff.y, coeff.z, __int_as_float(1));
            ps.rng_n[0][pos] = rng.n;
        }
    }
}

template <unsigned kM>
__global__ __launch_bounds__(256) void k_lt_shade(DevScene sc, LtCamera cam, uint32_t seed, uint64_t base, int side,
                                                  RayQueue qin, int cnt_in,
                                                  const float4* __restrict__ hits, ParticleState ps, RayQueue qout,
                                                  int cnt_out, RayQueue qs, int cnt_s, int* counts) {
    __shared__ int sh[256 / 64 + 1];
    const int n = counts[cnt_in];
    if (blockIdx.x == 0 && threadIdx.x < 2) counts[kCntWork + threadIdx.x] = 0;
    uint64_t nvert = 0;
    for (int b0 = blockIdx.x * 256; b0 < n; b0 += gridDim.x * 256) {
        const int k = b0 + threadIdx.x;
        bool want_ext = false, want_shadow = false;
        f3 no = mk(0.f), nd = mk(0.f), so = mk(0.f), sd = mk(0.f), color = mk(0.f);
        float ntmin = 0.f, stmax = 0.f;
        int pid = 0, pixel = 0;
        f3 coeff = mk(0.f);
        int bounces = 0;
        Rng rng{0, 0u};
        if (k < n) {
            pid = qin.pid[k];
            const float4 h = hits[k];
            const int prim = __float_as_int(h.y);
            if (prim >= 0) {
                const float4 ro = qin.org[k], rd = qin.dir[k];
                const f3 o = mk(ro.x, ro.y, ro.z), d = mk(rd.x, rd.y, rd.z);
                const float4 cf = ps.st[side][k];
                coeff = mk(cf.x, cf.y, cf.z);
                bounces = __float_as_int(cf.w);
                rng = Rng{particle_key(seed, base + (uint64_t)pid), ps.rng_n[side][k]};
                const DevPrim pr = sc.prim_info[prim];
                const int sub = (pr.kind == kInst && ps.hsub) ? ps.hsub[k] : -1;
                Inct it = make_inct_any(sc, pr, sub, o, d, h.z, h.w, h.x);
                const int mat = pr.material;
                if (mat < 0) {
                    // a medium boundary: the same ray goes on from the hit,
                    // `bounces--; continue;` (the loop head runs again)
                    no = it.P; nd = d; ntmin = ro.w;
                    want_ext = loop_head(sc.integ, bounces, coeff, rng);
                } else {
                    nvert++;
                    MatCtx m = load_material<kM>(sc, mat, it.u, it.v);
                    const f3 wo_l = to_local(it.sh, it.wr);
                    CamWi cw = cam_sample_wi(cam, it.P);
                    if (cw.pdf > 0 && cw.we != 0.0f) {
                        // Scene.IsOccluded(camera, P) (Scene.cs:58-70)
                        f3 vec = it.P - cw.pos;
                        float dis = length(vec);
                        if (!(dis < 0.001f)) {
                            const f3 wi_l = to_local(it.sh, cw.dir);
                            f3 f = mat_fr<kM>(m, wo_l, wi_l, true);
                            float ac = abs_cos_theta(wi_l);
                            color = coeff * f * mk(cw.we) * ac / cw.pdf;
                            pixel = (int)floorf(cw.sx) * cam.height + (int)floorf(cw.sy);
                            so = cw.pos; sd = normalize(vec); stmax = dis - 0.0001f;
                            want_shadow = true;
                        }
                    }
                    BSample bs = mat_sample<kM>(m, wo_l, rng, true);
                    if (bs.pdf > 0.0f) {
                        coeff = coeff * (bs.fr * abs_cos_theta(bs.wi) / bs.pdf);
                        no = it.P; nd = to_world(it.sh, bs.wi); ntmin = 0.0001f;
                        bounces++;
                        want_ext = loop_head(sc.integ, bounces, coeff, rng);
                    }
                }
            }
        }
        // one atomic per block and queue: same-address atomics serialise,
        // and particles of a cheap scene are shaded faster than they can be
        // handed out one wavefront at a time
        int sp = block_append<256>(counts + cnt_s, want_shadow, sh);
        if (want_shadow) {
            qs.pid[sp] = sp;
            qs.org[sp] = make_float4(so.x, so.y, so.z, 0.0001f);
            qs.dir[sp] = make_float4(sd.x, sd.y, sd.z, stmax);
            ps.splat[sp] = make_float4(color.x, color.y, color.z, __int_as_float(pixel));
        }
        int ep = block_append<256>(counts + cnt_out, want_ext, sh);
        if (want_ext) {
            qout.pid[ep] = pid;
            qout.org[ep] = make_float4(no.x, no.y, no.z, ntmin);
            qout.dir[ep] = make_float4(nd.x, nd.y, nd.z, kFloatMax);
            ps.st[1 - side][ep] = make_float4(coeff.x, coeff.y, coeff.z, __int_as_float(bounces));
            ps.rng_n[1 - side][ep] = rng.n;
        }
    }
    add_counter((unsigned long long*)(counts + kCntVertices), nvert);
}

__global__ __launch_bounds__(256) void k_lt_splat(RayQueue qs, int cnt_s, ParticleState ps, double* acc,
                                                  const int* counts) {
    const int n = counts[cnt_s];
    for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
        if (!ps.vis[k]) continue;
        const float4 c = ps.splat[k];
        double* px = acc + 3 * (size_t)__float_as_int(c.w);
        atomicAdd(px + 0, (double)c.x);
        atomicAdd(px + 1, (double)c.y);
        atomicAdd(px + 2, (double)c.z);
    }
}

// Camera.SampleRay (Camera.cs:182-191) at pixel corner (x, y).
__global__ __launch_bounds__(256) void k_lt_corner_rays(DevScene sc, RayQueue q, int* counts, int cnt_q) {
    const DevCamera& c = sc.cam;
    const int n = c.width * c.height;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counts[cnt_q] = n;
        counts[kCntWork + 0] = 0;
    }
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const int x = i / c.height, y = i % c.height;
        f3 o, d;
        float tmin, tmax;
        camera_ray(c, (float)x, (float)y, &o, &d, &tmin, &tmax);
        q.pid[i] = i;
        q.org[i] = make_float4(o.x, o.y, o.z, tmin);
        q.dir[i] = make_float4(d.x, d.y, d.z, tmax);
    }
}

// AfterRender (AdjointParticleRenderer.cs:161-177): RefRGB *= v, then
// += CameraTraceLight (AdjointParticleIntegrator.cs:82-85).
__global__ __launch_bounds__(256) void k_lt_final(DevScene sc, RayQueue q, const float4* __restrict__ hits,
                                                  const int* __restrict__ hsub, const double* __restrict__ acc,
                                                  float v, float* out) {
    const int n = sc.cam.width * sc.cam.height;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        f3 c = mk((float)acc[3 * (size_t)i], (float)acc[3 * (size_t)i + 1], (float)acc[3 * (size_t)i + 2]);
        c = c * v;
        const float4 h = hits[i];
        const int prim = __float_as_int(h.y);
        const float4 rd = q.dir[i];
        const f3 d = mk(rd.x, rd.y, rd.z);
        f3 le = mk(0.0f);
        if (prim >= 0) {
            const DevPrim pr = sc.prim_info[prim];
            if (pr.kind != kInst && pr.light >= 0) {
                const float4 ro = q.org[i];
                Inct it = make_inct(sc, pr, mk(ro.x, ro.y, ro.z), d, h.z, h.w, h.x);
                const DevLight& L = sc.lights[pr.light];
                // DiffuseAreaLight.L (DiffuseAreaLight.cs:26-29)
                le = dot(it.sh.Z, -d) > 0 ? mk(L.le[0], L.le[1], L.le[2]) : mk(0.0f);
            }
        } else {
            le = eval_infinite(sc, d);
        }
        c = c + le;
        out[3 * (size_t)i] = c.x;
        out[3 * (size_t)i + 1] = c.y;
        out[3 * (size_t)i + 2] = c.z;
    }
}

int grid_for(int n) {
    int b = (n + 255) / 256;
    if (b < 1) b = 1;
    if (b > persistent_blocks()) b = persistent_blocks();
    return b;
}

}  // namespace

hipError_t launch_lt_emit(const DevScene& sc, uint32_t seed, uint64_t base, int n, ParticleState ps, RayQueue q,
                          int* counts, int cnt_q, hipStream_t s) {
    hipLaunchKernelGGL(k_lt_emit, dim3(grid_for(n)), dim3(256), 0, s, sc, seed, base, n, ps, q, counts, cnt_q);
    return hipGetLastError();
}

hipError_t launch_lt_shade(const DevScene& sc, const LtCamera& cam, uint32_t seed, uint64_t base, int side,
                           RayQueue qin, int cnt_in, const float4* hits,
                           ParticleState ps, RayQueue qout, int cnt_out, RayQueue qs, int cnt_s, int* counts,
                           hipStream_t s) {
    const dim3 g(persistent_blocks()), b(256);
    const unsigned set = mat_set(sc.mat_mask);
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, g, b, 0, s, sc, cam, seed, base, side, qin, cnt_in, hits, ps, qout, cnt_out, qs,
                           cnt_s, counts);
    };
    if (set == kMatsDelta) go(k_lt_shade<kMatsDelta>);
    else if (set == kMatsMicrofacet) go(k_lt_shade<kMatsMicrofacet>);
    else go(k_lt_shade<kMatsAll>);
    return hipGetLastError();
}

hipError_t launch_lt_splat(RayQueue qs, int cnt_s, ParticleState ps, double* acc, int* counts, hipStream_t s) {
    hipLaunchKernelGGL(k_lt_splat, dim3(persistent_blocks()), dim3(256), 0, s, qs, cnt_s, ps, acc, counts);
    return hipGetLastError();
}

hipError_t launch_lt_corner_rays(const DevScene& sc, RayQueue q, int* counts, int cnt_q, hipStream_t s) {
    hipLaunchKernelGGL(k_lt_corner_rays, dim3(grid_for(sc.cam.width * sc.cam.height)), dim3(256), 0, s, sc, q, counts,
                       cnt_q);
    return hipGetLastError();
}

hipError_t launch_lt_final(const DevScene& sc, RayQueue q, const float4* hits, const int* hsub, const double* acc,
                           float v, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_lt_final, dim3(grid_for(sc.cam.width * sc.cam.height)), dim3(256), 0, s, sc, q, hits, hsub,
                       acc, v, out);
    return hipGetLastError();
}

}  // namespace rd
