// rd_bdpt.hip — gfx950 wavefront stages of the reference's bidirectional
// path tracer: BidirectionalPathTracer
// (Radiantium.Offline/Integrators/BidirectionalPathTracer.cs:11-435) under
// BidirectionalPathTracingRenderer (Renderers/BidirectionalPathTracingRenderer.cs:278-452).
//
// A pass holds P bidirectional samples (every pixel of the crop, a run of
// sample indices).  Its stages:
//   bd_cam_start / bd_light_start   vertex 0 of both subpaths and their first
//                                    rays (GenerateCameraSubpath :42-57,
//                                    GenerateLightSubpath :59-82)
//   trace_ext + bd_walk, repeated    RandomWalk (:84-162) of the 2P walkers
//                                    together: medium free flight, a vertex
//                                    per scattering event, the next ray
//   bd_connect                       one thread per (sample, strategy) in the
//                                    bounds of Connect (:164-217): the
//                                    strategy's value without its visibility,
//                                    and a Transmittance chain when it needs one
//   trace_ext + bd_chain, repeated   Scene.Transmittance (Scene.cs:72-103):
//                                    closest-hit segments to the target
//   bd_sum                           the sample's value: the t = 1 strategies
//                                    in s order, then the others in loop order
//   bd_accumulate / bd_finalize      per-pixel float64 sums in sample order,
//                                    then the mean
// The reference's MisWeight returns 1 (its body is commented out), so the
// subpaths' pdf bookkeeping never reaches the image and is not kept.
//
// Random numbers: the camera subpath draws from path_key(seed, pixel,
// sample), the light subpath from path_key(seed, pixel, sample | 2^30), the
// light sample of strategy (s = 1, t) from path_key(seed, pixel,
// sample | 2^29) at draws 3 (t - 2) .. 3 (t - 2) + 2.
#include "rd_light.h"
#include "rd_light_dev.h"
#include "rd_shading.h"
#include "rd_wave.h"

namespace rd {

namespace {

using namespace lightpath;

enum : int { kBvCamera = 0, kBvLight = 1, kBvSurface = 2, kBvMedium = 3 };

struct BVert {
    int type;
    f3 p, n;       // n = 0 off surfaces (BdptVertex.IsOnSurface)
    f3 coeff;
    int env;       // the vertex's medium (-1: null; always -1 on surfaces)
    int aux;       // surface: primitive; light: light index (-1: the escaped-ray vertex)
    f3 w;          // surface: Wr; medium: Wo
    float u, v;    // surface: texture coordinates
};

__device__ __forceinline__ size_t vslot(const BdptBuffers& b, int side, int k, int i) {
    return ((size_t)side * b.vmax + k) * (size_t)b.P + i;
}
__device__ __forceinline__ void store_vertex(const BdptBuffers& b, int side, int k, int i, const BVert& v) {
    const size_t j = vslot(b, side, k, i);
    b.v0[j] = make_float4(v.p.x, v.p.y, v.p.z, __int_as_float(v.type));
    b.v1[j] = make_float4(v.n.x, v.n.y, v.n.z, __int_as_float(v.env));
    b.v2[j] = make_float4(v.coeff.x, v.coeff.y, v.coeff.z, __int_as_float(v.aux));
    b.v3[j] = make_float4(v.w.x, v.w.y, v.w.z, v.u);
    b.v4[j] = v.v;
}
__device__ __forceinline__ BVert load_vertex(const BdptBuffers& b, int side, int k, int i) {
    const size_t j = vslot(b, side, k, i);
    const float4 a = b.v0[j], c = b.v1[j], d = b.v2[j], e = b.v3[j];
    BVert v;
    v.p = mk(a.x, a.y, a.z); v.type = __float_as_int(a.w);
    v.n = mk(c.x, c.y, c.z); v.env = __float_as_int(c.w);
    v.coeff = mk(d.x, d.y, d.z); v.aux = __float_as_int(d.w);
    v.w = mk(e.x, e.y, e.z); v.u = e.w; v.v = b.v4[j];
    return v;
}
__device__ __forceinline__ bool on_surface(const BVert& v) { return !(v.n.x == 0 && v.n.y == 0 && v.n.z == 0); }

// The sample index (in the pass) -> pixel of the crop in raster order and
// the sample number.
__device__ __forceinline__ void sample_pixel(const BdptPass& w, int i, int* x, int* y, int* si) {
    const int q = i % w.npix;
    *si = w.sample_base + i / w.npix;
    *x = w.x0 + q % w.crop_w;
    *y = w.y0 + q / w.crop_w;
}

__global__ __launch_bounds__(256) void k_bd_cam_start(DevScene sc, BdptPass w, BdptBuffers b, RayQueue q, int* counts,
                                                      int cnt_q) {
    __shared__ int sh[256 / 64 + 1];   // block_append
    if (blockIdx.x == 0 && threadIdx.x == 0) counts[kCntWork + 0] = 0;
    for (int b0 = blockIdx.x * 256; b0 < b.P; b0 += gridDim.x * 256) {
        const int i = b0 + threadIdx.x;
        bool go = false;
        f3 o = mk(0.0f), d = mk(0.0f);
        float tmin = 0.0f, tmax = 0.0f;
        if (i < b.P) {
            int x, y, si;
            sample_pixel(w, i, &x, &y, &si);
            Rng rng{path_key(w.seed, (uint32_t)(y * w.width + x), (uint32_t)si), 0u};
            float ox = rng.next_float(), oy = rng.next_float();
            camera_ray(sc.cam, (float)x + ox, (float)y + oy, &o, &d, &tmin, &tmax);
            BVert v;
            v.type = kBvCamera; v.p = o; v.n = mk(0.0f); v.coeff = mk(1.0f); v.env = sc.global_medium; v.aux = -1;
            v.w = mk(0.0f); v.u = 0.0f; v.v = 0.0f;
            store_vertex(b, 0, 0, i, v);
            b.wcoeff[i] = make_float4(1.0f, 1.0f, 1.0f, __int_as_float(1));
            b.wenv[i] = sc.global_medium;
            b.wkey[i] = rng.key;
            b.wn[i] = rng.n;
            go = b.walk_cap != 0;
            if (!go) b.nverts[i] = 1;
        }
        int pos = block_append<256>(counts + cnt_q, go, sh);
        if (go) {
            q.pid[pos] = i;
            q.org[pos] = make_float4(o.x, o.y, o.z, tmin);
            q.dir[pos] = make_float4(d.x, d.y, d.z, tmax);
        }
    }
}

__global__ __launch_bounds__(256) void k_bd_light_start(DevScene sc, BdptPass w, BdptBuffers b, RayQueue q,
                                                        int* counts, int cnt_q) {
    __shared__ int sh[256 / 64 + 1];   // block_append
    for (int b0 = blockIdx.x * 256; b0 < b.P; b0 += gridDim.x * 256) {
        const int i = b0 + threadIdx.x;
        bool go = false;
        EmitS e;
        e.p = mk(0.0f); e.dir = mk(0.0f);
        if (i < b.P) {
            int x, y, si;
            sample_pixel(w, i, &x, &y, &si);
            Rng rng{path_key(w.seed, (uint32_t)(y * w.width + x), 0x40000000u | (uint32_t)si), 0u};
            int n = 0;
            if (sc.num_lights > 0) {
                const int l = rng.next_int(sc.num_lights);
                const float select_pdf = 1.0f / sc.num_lights;
                const DevLight& L = sc.lights[l];
                if (sample_emit(sc, L, rng, &e) && e.pdf_pos != 0 && e.pdf_dir != 0 && !is_black(e.le)) {
                    int lenv = -1;   // Light.GetMedium: the emitter's side (DiffuseAreaLight.cs:88-91); infinite: null
                    if (L.type == 0) {
                        const DevPrim pr = sc.prim_info[L.prim];
                        lenv = dot(e.n, e.dir) > 0 ? pr.med_out : pr.med_in;
                    }
                    BVert v;
                    v.type = kBvLight; v.p = e.p; v.n = e.n; v.coeff = e.le; v.env = lenv; v.aux = l;
                    v.w = mk(0.0f); v.u = 0.0f; v.v = 0.0f;
                    store_vertex(b, 1, 0, i, v);
                    f3 coeff = e.le * abs_dot(e.dir, e.n) / (select_pdf * e.pdf_pos * e.pdf_dir);
                    b.wcoeff[b.P + i] = make_float4(coeff.x, coeff.y, coeff.z, __int_as_float(1));
                    b.wenv[b.P + i] = lenv;
                    b.wkey[b.P + i] = rng.key;
                    b.wn[b.P + i] = rng.n;
                    n = 1;
                    go = b.walk_cap != 0;
                }
            }
            if (!go) b.nverts[b.P + i] = n;
        }
        int pos = block_append<256>(counts + cnt_q, go, sh);
        if (go) {
            q.pid[pos] = b.P + i;
            q.org[pos] = make_float4(e.p.x, e.p.y, e.p.z, 0.0001f);
            q.dir[pos] = make_float4(e.dir.x, e.dir.y, e.dir.z, kFloatMax);
        }
    }
}

// One RandomWalk iteration of every walker in qin (BidirectionalPathTracer.cs:94-160).
template <unsigned kM>
__global__ __launch_bounds__(256) void k_bd_walk(DevScene sc, BdptBuffers b, RayQueue qin, int cnt_in,
                                                 const float4* __restrict__ hits, const int* __restrict__ hsub,
                                                 RayQueue qout, int cnt_out, int* counts) {
    __shared__ int sh[256 / 64 + 1];   // block_append
    const int n = counts[cnt_in];
    if (blockIdx.x == 0 && threadIdx.x == 0) counts[kCntWork + 0] = 0;
    uint64_t nvert = 0;
    for (int b0 = blockIdx.x * 256; b0 < n; b0 += gridDim.x * 256) {
        const int k = b0 + threadIdx.x;
        bool go = false;
        int wid = 0;
        f3 no = mk(0.f), nd = mk(0.f);
        float ntmin = 0.f;
        if (k < n) {
            wid = qin.pid[k];
            const int side = wid >= b.P ? 1 : 0;
            const int i = wid - side * b.P;
            const float4 ro = qin.org[k], rd4 = qin.dir[k];
            const f3 o = mk(ro.x, ro.y, ro.z), d = mk(rd4.x, rd4.y, rd4.z);
            const float4 h = hits[k];
            const int prim = __float_as_int(h.y);
            const bool hit = prim >= 0;
            const float4 cf = b.wcoeff[wid];
            f3 coeff = mk(cf.x, cf.y, cf.z);
            int bounces = __float_as_int(cf.w);
            int env = b.wenv[wid];
            Rng rng{b.wkey[wid], b.wn[wid]};
            MSample ms;
            ms.medium = false;
            if (env >= 0) {
                ms = medium_sample(sc.media[env], o, d, ro.w, hit ? h.x : kFloatMax, rng);
                coeff = coeff * ms.tr;
            }
            bool end = is_black(coeff);
            if (!end && ms.medium) {
                BVert v;
                v.type = kBvMedium; v.p = ms.p; v.n = mk(0.0f); v.coeff = coeff; v.env = env; v.aux = -1;
                v.w = -d; v.u = 0.0f; v.v = 0.0f;
                store_vertex(b, side, bounces, i, v);
                ++bounces;
                if (bounces >= b.walk_cap) {
                    end = true;
                } else {
                    PSample ps = hg_sample(sc.media[env], -d, rng);
                    no = ms.p; nd = ps.wi; ntmin = ro.w;
                    go = true;
                }
            } else if (!end) {
                if (!hit) {
                    if (side == 0) {   // the escaped camera ray: a light vertex with no light
                        BVert v;
                        v.type = kBvLight; v.p = o + 10.0f * d; v.n = -d; v.coeff = coeff; v.env = env; v.aux = -1;
                        v.w = mk(0.0f); v.u = 0.0f; v.v = 0.0f;
                        store_vertex(b, side, bounces, i, v);
                        ++bounces;
                    }
                    end = true;
                } else {
                    const DevPrim pr = sc.prim_info[prim];
                    const int sub = (pr.kind == kInst && hsub) ? hsub[k] : -1;
                    Inct it = make_inct_any(sc, pr, sub, o, d, h.z, h.w, h.x);
                    if (pr.material < 0) {
                        no = it.P; nd = d; ntmin = ro.w;
                        env = get_medium(pr, it.sh.Z, d);
                        go = true;
                    } else {
                        nvert++;
                        BVert v;
                        v.type = kBvSurface; v.p = it.P; v.n = it.sh.Z; v.coeff = coeff; v.env = -1; v.aux = prim;
                        v.w = it.wr; v.u = it.u; v.v = it.v;
                        store_vertex(b, side, bounces, i, v);
                        ++bounces;
                        if (bounces >= b.walk_cap) {
                            end = true;
                        } else {
                            MatCtx m = load_material<kM>(sc, pr.material, it.u, it.v);
                            BSample bs = mat_sample<kM>(m, to_local(it.sh, it.wr), rng, side == 1);
                            if (bs.pdf > 0.0f) {
                                coeff = coeff * (bs.fr * abs_cos_theta(bs.wi) / bs.pdf);
                                nd = to_world(it.sh, bs.wi);
                                no = it.P; ntmin = ro.w;
                                env = get_medium(pr, it.sh.Z, nd);
                                go = true;
                            } else {
                                end = true;
                            }
                        }
                    }
                }
            }
            if (go) {
                b.wcoeff[wid] = make_float4(coeff.x, coeff.y, coeff.z, __int_as_float(bounces));
                b.wenv[wid] = env;
                b.wn[wid] = rng.n;
            } else {
                b.nverts[wid] = bounces;
            }
        }
        int pos = block_append<256>(counts + cnt_out, go, sh);
        if (go) {
            qout.pid[pos] = wid;
            qout.org[pos] = make_float4(no.x, no.y, no.z, ntmin);
            qout.dir[pos] = make_float4(nd.x, nd.y, nd.z, kFloatMax);
        }
    }
    add_counter((unsigned long long*)(counts + kCntVertices), nvert);
}

// BdptVertex.F (BidirectionalPathTracingRenderer.cs:94-106)
template <unsigned kM>
__device__ __forceinline__ f3 vertex_f(const DevScene& sc, const BVert& v, f3 next_p, bool imp) {
    f3 wi = next_p - v.p;
    if (length_sq(wi) == 0) return mk(0.0f);
    wi = normalize(wi);
    if (v.type == kBvSurface) {
        const DevPrim pr = sc.prim_info[v.aux];
        MatCtx m = load_material<kM>(sc, pr.material, v.u, v.v);
        Frame f = make_frame(v.n);
        return mat_fr<kM>(m, to_local(f, v.w), to_local(f, wi), imp);
    }
    if (v.type == kBvMedium) {
        const DevMedium& md = sc.media[v.env];
        return mk(hg_eval(dot(v.w, wi), hg_g(md)));
    }
    return mk(0.0f);
}

// BdptVertex.Le (:69-92) of a light vertex toward the previous vertex.
__device__ __forceinline__ f3 vertex_le(const DevScene& sc, const BVert& v, f3 prev_p) {
    const bool surf_light = v.type == kBvSurface && sc.prim_info[v.aux].light >= 0;
    if (!(v.type == kBvLight || surf_light)) return mk(0.0f);
    f3 w = prev_p - v.p;
    if (length_sq(w) == 0) return mk(0.0f);
    w = normalize(w);
    const bool infinite = v.type == kBvLight && (v.aux < 0 || sc.lights[v.aux].type == 1);
    if (infinite) {
        f3 le = mk(0.0f);
        for (int e = 0; e < sc.num_envs; e++) le = le + env_le(sc, sc.envs[e], -w);
        return le;
    }
    if (v.type == kBvLight) return mk(0.0f);   // an area light's own vertex is never evaluated here
    const DevLight& L = sc.lights[sc.prim_info[v.aux].light];
    return dot(v.n, w) > 0 ? mk(L.le[0], L.le[1], L.le[2]) : mk(0.0f);   // DiffuseAreaLight.L
}

// ConnectBdpt (:219-300) of strategy k of sample i; the value without its
// visibility goes to conn (w = the geometry factor the transmittance is
// scaled by), and a Transmittance chain starts when one is needed.
template <unsigned kM>
__global__ __launch_bounds__(256) void k_bd_connect(DevScene sc, LtCamera cam, BdptPass w, BdptBuffers b, RayQueue qc,
                                                    int cnt_c, int* counts) {
    __shared__ int sh[256 / 64 + 1];   // block_append
    if (blockIdx.x == 0 && threadIdx.x == 0) counts[kCntWork + 0] = 0;
    const int V = b.vmax;
    const long long total = (long long)b.P * b.nstrat;
    for (long long b0 = (long long)blockIdx.x * 256; b0 < total; b0 += (long long)gridDim.x * 256) {
        const long long idx = b0 + threadIdx.x;
        bool chain = false;
        f3 from = mk(0.f), to = mk(0.f);
        int env = -1;
        if (idx < total) {
            const int kk = (int)(idx / b.P), i = (int)(idx % b.P);
            const int t = 1 + kk / V, s = kk % V;
            const int nc = b.nverts[i], nl = b.nverts[b.P + i];
            const int depth = t + s - 2;
            f3 l = mk(0.0f);
            float g = 1.0f;
            if (t < nc && s < nl && !(s == 1 && t == 1) && depth >= 0 && depth <= sc.integ.max_depth) {
                if (s == 0) {
                    const BVert pt = load_vertex(b, 0, t - 1, i);
                    const BVert pv = load_vertex(b, 0, t - 2, i);
                    l = vertex_le(sc, pt, pv.p) * pt.coeff;
                } else if (t == 1) {
                    const BVert qs = load_vertex(b, 1, s - 1, i);
                    CamWi cw = cam_sample_wi(cam, qs.p);
                    if (cw.pdf > 0 && cw.we != 0.0f) {
                        const f3 cc = mk(cw.we) / cw.pdf;
                        l = qs.coeff * vertex_f<kM>(sc, qs, cw.pos, true) * cc;
                        if (on_surface(qs)) l = l * abs_dot(cw.dir, qs.n);
                        if (!is_black(l)) { chain = true; from = qs.p; to = cw.pos; env = qs.env; }
                    }
                } else {
                    const BVert pt = load_vertex(b, 0, t - 1, i);
                    if (pt.type != kBvLight) {
                        if (s == 1) {
                            int x, y, si;
                            sample_pixel(w, i, &x, &y, &si);
                            Rng rng{path_key(w.seed, (uint32_t)(y * w.width + x), 0x20000000u | (uint32_t)si),
                                    3u * (uint32_t)(t - 2)};
                            if (sc.num_lights > 0) {
                                const int li = rng.next_int(sc.num_lights);
                                const float select_pdf = 1.0f / sc.num_lights;
                                LSample ls = light_sample(sc, sc.lights[li], pt.p, rng);
                                if (ls.pdf > 0 && !is_black(ls.li)) {
                                    const f3 lc = ls.li / (ls.pdf * select_pdf);
                                    l = pt.coeff * vertex_f<kM>(sc, pt, ls.p, false) * lc;
                                    if (on_surface(pt)) l = l * abs_dot(ls.wi, pt.n);
                                    if (!is_black(l)) { chain = true; from = pt.p; to = ls.p; env = pt.env; }
                                }
                            }
                        } else {
                            const BVert qs = load_vertex(b, 1, s - 1, i);
                            l = qs.coeff * vertex_f<kM>(sc, qs, pt.p, true) * vertex_f<kM>(sc, pt, qs.p, false) *
                                pt.coeff;
                            if (!is_black(l)) {
                                // G (:311-320): the transmittance goes from qs to pt
                                f3 dd = qs.p - pt.p;
                                g = 1 / length_sq(dd);
                                dd = dd * sqrtf(g);
                                if (on_surface(qs)) g *= abs_dot(qs.n, dd);
                                if (on_surface(pt)) g *= abs_dot(pt.n, dd);
                                chain = true; from = qs.p; to = pt.p; env = qs.env;
                            }
                        }
                    }
                }
            }
            if (chain) {
                // the chain's first segment, or no query at all when the ends meet
                const f3 vec = to - from;
                const float dis = length(vec);
                if (dis < 0.0001f) {
                    l = l * (g * mk(1.0f));
                    chain = false;
                } else {
                    b.cto[idx] = make_float4(to.x, to.y, to.z, __int_as_float(env));
                    b.ctr[idx] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
                }
            }
            b.conn[idx] = make_float4(l.x, l.y, l.z, g);
        }
        int pos = block_append<256>(counts + cnt_c, chain, sh);
        if (chain) {
            const f3 vec = to - from;
            const float dis = length(vec);
            const f3 dir = normalize(to - from);
            qc.pid[pos] = (int)idx;
            qc.org[pos] = make_float4(from.x, from.y, from.z, 0.0001f);
            qc.dir[pos] = make_float4(dir.x, dir.y, dir.z, dis - 0.0001f);
        }
    }
}

// One Transmittance segment of every chain in qin (Scene.cs:79-101): a
// surface blocks; a medium multiplies its Tr over the segment; a boundary
// continues from the hit toward the target, in the medium behind it.  A
// finished chain scales its strategy: l * (g * tr).
__global__ __launch_bounds__(256) void k_bd_chain(DevScene sc, BdptBuffers b, RayQueue qin, int cnt_in,
                                                  const float4* __restrict__ hits, const int* __restrict__ hsub,
                                                  RayQueue qout, int cnt_out, int* counts) {
    __shared__ int sh[256 / 64 + 1];   // block_append
    const int n = counts[cnt_in];
    if (blockIdx.x == 0 && threadIdx.x == 0) counts[kCntWork + 0] = 0;
    for (int b0 = blockIdx.x * 256; b0 < n; b0 += gridDim.x * 256) {
        const int k = b0 + threadIdx.x;
        bool go = false;
        int id = 0;
        f3 no = mk(0.f), nd = mk(0.f);
        float ntmax = 0.f;
        if (k < n) {
            id = qin.pid[k];
            const float4 ro = qin.org[k], rd4 = qin.dir[k];
            const f3 o = mk(ro.x, ro.y, ro.z), d = mk(rd4.x, rd4.y, rd4.z);
            const float4 h = hits[k];
            const int prim = __float_as_int(h.y);
            const bool any = prim >= 0;
            const float4 to4 = b.cto[id];
            int env = __float_as_int(to4.w);
            f3 tr = ld3(b.ctr, id);
            bool done = true;
            DevPrim pr;
            if (any) pr = sc.prim_info[prim];
            if (any && pr.material >= 0) {
                tr = mk(0.0f);
            } else {
                if (env >= 0) tr = tr * medium_tr(sc.media[env], ro.w, rd4.w);
                if (any) {
                    const int sub = (pr.kind == kInst && hsub) ? hsub[k] : -1;
                    Inct it = make_inct_any(sc, pr, sub, o, d, h.z, h.w, h.x);
                    env = get_medium(pr, it.sh.Z, d);
                    const f3 to = mk(to4.x, to4.y, to4.z);
                    const float dis = length(to - it.P);
                    if (!(dis < 0.0001f)) {
                        done = false;
                        go = true;
                        no = it.P; nd = d; ntmax = dis - 0.0001f;
                        b.cto[id] = make_float4(to4.x, to4.y, to4.z, __int_as_float(env));
                        st3(b.ctr, id, tr);
                    }
                }
            }
            if (done) {
                const float4 c = b.conn[id];
                const f3 l = mk(c.x, c.y, c.z) * (c.w * tr);
                b.conn[id] = make_float4(l.x, l.y, l.z, c.w);
            }
        }
        int pos = block_append<256>(counts + cnt_out, go, sh);
        if (go) {
            qout.pid[pos] = id;
            qout.org[pos] = make_float4(no.x, no.y, no.z, 0.0001f);
            qout.dir[pos] = make_float4(nd.x, nd.y, nd.z, ntmax);
        }
    }
}

// The sample's value (Connect :168-216): the t = 1 strategies in s order
// (the reference's AtomicAddRGB splats, all on the sample's pixel), then
// the rest in loop order; a strategy that is not a valid colour counts 0.
__global__ __launch_bounds__(256) void k_bd_sum(BdptBuffers b, float4* __restrict__ res) {
    const int V = b.vmax;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < b.P; i += gridDim.x * 256) {
        f3 splat = mk(0.0f), l = mk(0.0f);
        for (int kk = 0; kk < b.nstrat; kk++) {
            const float4 c4 = b.conn[(size_t)kk * b.P + i];
            f3 c = mk(c4.x, c4.y, c4.z);
            if (!valid_color(c)) c = mk(0.0f);
            if (kk < V) splat = splat + c;   // t = 1
            else l = l + c;
        }
        const f3 v = splat + l;
        res[i] = make_float4(v.x, v.y, v.z, 0.0f);
    }
}

// Per pixel of the pass's crop: its samples' values in sample order into
// the float64 sums (and float32 sums of squares).
__global__ __launch_bounds__(256) void k_bd_accumulate(BdptPass w, int nsamples, const float4* __restrict__ res,
                                                       double4* sum, float4* sum2) {
    for (int q = blockIdx.x * 256 + threadIdx.x; q < w.npix; q += gridDim.x * 256) {
        double4 s = sum[q];
        float4 s2 = sum2[q];
        for (int j = 0; j < nsamples; j++) {
            const float4 c = res[(size_t)j * w.npix + q];
            s.x += (double)c.x; s.y += (double)c.y; s.z += (double)c.z;
            s2.x = s2.x + c.x * c.x; s2.y = s2.y + c.y * c.y; s2.z = s2.z + c.z * c.z;
        }
        sum[q] = s;
        sum2[q] = s2;
    }
}

__global__ __launch_bounds__(256) void k_bd_finalize(BdptPass w, int height, int spp, const double4* __restrict__ sum,
                                                     const float4* __restrict__ sum2, float* out, float* out_sum2,
                                                     double* out64) {
    for (int q = blockIdx.x * 256 + threadIdx.x; q < w.npix; q += gridDim.x * 256) {
        const int x = w.x0 + q % w.crop_w, y = w.y0 + q / w.crop_w;
        const size_t o = ((size_t)x * height + y) * 3;
        const double4 s = sum[q];
        const double n = (double)spp;
        if (out64) {
            out64[o] = s.x / n; out64[o + 1] = s.y / n; out64[o + 2] = s.z / n;
        }
        if (out) {
            out[o] = (float)(s.x / n); out[o + 1] = (float)(s.y / n); out[o + 2] = (float)(s.z / n);
        }
        if (out_sum2) {
            const float4 s2 = sum2[q];
            out_sum2[o] = s2.x; out_sum2[o + 1] = s2.y; out_sum2[o + 2] = s2.z;
        }
    }
}

int grid_for(long long n) {
    long long g = (n + 255) / 256;
    if (g < 1) g = 1;
    if (g > persistent_blocks()) g = persistent_blocks();
    return (int)g;
}

template <class F>
void by_mat_set(const DevScene& sc, F&& f) {
    const unsigned set = mat_set(sc.mat_mask);
    if (set == kMatsDelta) f(std::integral_constant<unsigned, kMatsDelta>{});
    else if (set == kMatsMicrofacet) f(std::integral_constant<unsigned, kMatsMicrofacet>{});
    else f(std::integral_constant<unsigned, kMatsAll>{});
}

}  // namespace

hipError_t launch_bd_start(const DevScene& sc, const BdptPass& w, const BdptBuffers& b, RayQueue q, int* counts,
                           int cnt_q, hipStream_t s) {
    hipLaunchKernelGGL(k_bd_cam_start, dim3(grid_for(b.P)), dim3(256), 0, s, sc, w, b, q, counts, cnt_q);
    hipLaunchKernelGGL(k_bd_light_start, dim3(grid_for(b.P)), dim3(256), 0, s, sc, w, b, q, counts, cnt_q);
    return hipGetLastError();
}

hipError_t launch_bd_walk(const DevScene& sc, const BdptBuffers& b, RayQueue qin, int cnt_in, const float4* hits,
                          const int* hsub, RayQueue qout, int cnt_out, int* counts, hipStream_t s) {
    by_mat_set(sc, [&](auto set) {
        hipLaunchKernelGGL(k_bd_walk<decltype(set)::value>, dim3(persistent_blocks()), dim3(256), 0, s, sc, b, qin,
                           cnt_in, hits, hsub, qout, cnt_out, counts);
    });
    return hipGetLastError();
}

hipError_t launch_bd_connect(const DevScene& sc, const LtCamera& cam, const BdptPass& w, const BdptBuffers& b,
                             RayQueue qc, int cnt_c, int* counts, hipStream_t s) {
    const long long total = (long long)b.P * b.nstrat;
    by_mat_set(sc, [&](auto set) {
        hipLaunchKernelGGL(k_bd_connect<decltype(set)::value>, dim3(grid_for(total)), dim3(256), 0, s, sc, cam, w, b,
                           qc, cnt_c, counts);
    });
    return hipGetLastError();
}

hipError_t launch_bd_chain(const DevScene& sc, const BdptBuffers& b, RayQueue qin, int cnt_in, const float4* hits,
                           const int* hsub, RayQueue qout, int cnt_out, int* counts, hipStream_t s) {
    hipLaunchKernelGGL(k_bd_chain, dim3(persistent_blocks()), dim3(256), 0, s, sc, b, qin, cnt_in, hits, hsub, qout,
                       cnt_out, counts);
    return hipGetLastError();
}

hipError_t launch_bd_sum(const BdptBuffers& b, float4* res, hipStream_t s) {
    hipLaunchKernelGGL(k_bd_sum, dim3(grid_for(b.P)), dim3(256), 0, s, b, res);
    return hipGetLastError();
}

hipError_t launch_bd_accumulate(const BdptPass& w, int nsamples, const float4* res, double4* sum, float4* sum2,
                                hipStream_t s) {
    hipLaunchKernelGGL(k_bd_accumulate, dim3(grid_for(w.npix)), dim3(256), 0, s, w, nsamples, res, sum, sum2);
    return hipGetLastError();
}

hipError_t launch_bd_finalize(const BdptPass& w, int height, int spp, const double4* sum, const float4* sum2,
                              float* out, float* out_sum2, double* out64, hipStream_t s) {
    hipLaunchKernelGGL(k_bd_finalize, dim3(grid_for(w.npix)), dim3(256), 0, s, w, height, spp, sum, sum2, out,
                       out_sum2, out64);
    return hipGetLastError();
}

}  // namespace rd
