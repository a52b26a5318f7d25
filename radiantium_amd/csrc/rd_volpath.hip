// rd_volpath.hip — gfx950 wavefront stages of VolumetricPathTracer.Li
// (Radiantium.Offline/Integrators/VolumetricPathTracer.cs:28-314).
//
// One reference bounce maps to:
//   trace_ext      Scene.Intersect of the path rays (closest hit, shared with path)
//   shade_vol      the medium's free-flight sample (Homogeneous.Sample), then
//                  either a medium vertex (light sample + HG phase sample +
//                  continuation phase sample) or the surface branch (emission,
//                  boundary crossing, light sample + BSDF samples), Russian
//                  roulette — every random draw of the bounce in the
//                  reference's order.  The direct-lighting visibility terms
//                  become two chains:
//                    Transmittance (Scene.cs:72-103) from the vertex to the
//                    light sample, and IntersectTr (Scene.cs:105-131) along
//                    the BSDF / phase sample.
//   trace_chain +  one closest-hit segment per chain and step; the advance
//   chain_advance  kernel multiplies the medium Tr of the segment in and
//                  either finishes the chain or appends its next segment.
//                  Neither chain draws random numbers (Homogeneous.Tr does not).
//   resolve_vol    radiance += coeff * (le_light + le_bsdf) / lightPdf.
#include "rd_chain.h"
#include "rd_kernels.h"
#include "rd_shading.h"
#include "rd_wave.h"

namespace rd {

template <bool kBatch, unsigned kM>
__global__ __launch_bounds__(256, 3) void k_shade_vol(DevScene sc0, const DevScene* __restrict__ scenes, RayQueue qin,
                                                   const float4* __restrict__ hits, RayQueue qshadow, RayQueue qmis,
                                                   PathState ps, int* counts, int cnt_in) {
    __shared__ int sh[256 / 64 + 1];   // block_append
    int n = counts[cnt_in];
    if (blockIdx.x == 0 && threadIdx.x < 3) counts[kCntWork + threadIdx.x] = 0;
    uint64_t nvert = 0;
    for (int base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {
        int k = base + threadIdx.x;
        bool want_ext = false, want_shadow = false, want_mis = false;
        f3 eo = mk(0.f), ed = mk(0.f), so = mk(0.f), sd = mk(0.f), mo = mk(0.f), md = mk(0.f);
        float etmin = 0.f, etmax = 0.f, stmax = 0.f, mtmin = 0.f;
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
            int env = ps.env[pid];
            int nflags = 0;
            f3 wo = -rd;
            DevPrim pr;
            Inct it;
            if (hit) {
                pr = sc.prim_info[prim];
                it = make_inct_any(sc, pr, pr.kind == kInst ? ps.hsub[k] : -1, ro, rd, h.z, h.w, h.x);
            }
            // starts the Transmittance chain toward the light point `to`
            auto start_shadow = [&](f3 from, f3 to, int medium) {
                f3 dir = normalize(to - from);
                float dis = length(to - from);
                ps.sh_tr[pid] = make_float4(1.f, 1.f, 1.f, 0.f);
                ps.sh_to[pid] = make_float4(to.x, to.y, to.z, __int_as_float(medium));
                if (!(dis < 0.0001f)) {
                    want_shadow = true;
                    so = from; sd = dir; stmax = dis - 0.0001f;
                }
            };
            // starts the IntersectTr chain of a BSDF / phase sample
            auto start_mis = [&](f3 o, f3 d, float tmin, f3 f, float w, float spdf, int l, int medium) {
                want_mis = true;
                mo = o; md = d; mtmin = tmin;
                st3(ps.mis_f, pid, f, w);
                ps.mis_li[pid] = make_float4(0.f, 0.f, 0.f, spdf);
                ps.mis_tr[pid] = make_float4(1.f, 1.f, 1.f, __int_as_float(medium));
                ps.sel_light[pid] = l;
            };
            bool alive = true, scattered = false;
            MSample ms;
            ms.medium = false;
            if (env >= 0) {
                // Homogeneous.Sample over [MinT, hit t or float.MaxValue] (VolumetricPathTracer.cs:43-48)
                ms = medium_sample(sc.media[env], ro, rd, rtmin, hit ? it.T : kFloatMax, rng);
                beta = beta * ms.tr;
            }
            if (is_black(beta)) {
                alive = false;
            } else if (ms.medium) {
                // medium vertex (:53-64); the depth test is written without the -1 escape
                if (depth >= in.max_depth) {
                    alive = false;
                } else {
                    nvert++;
                    const DevMedium mdm = sc.media[env];
                    if (sc.num_lights > 0) {
                        // EstimateDirectMedium (:252-314)
                        int l = rng.next_int(sc.num_lights);
                        float selpdf = 1.0f / (float)sc.num_lights;
                        const DevLight Lt = sc.lights[l];
                        LSample ls = light_sample(sc, Lt, ms.p, rng);
                        if (ls.pdf > 0.0f && !is_black(ls.li)) {
                            float p = hg_eval(dot(wo, ls.wi), hg_g(mdm));
                            if (p != 0.0f) {
                                start_shadow(ms.p, ls.p, env);
                                nflags |= kFlagShadowIssued;
                                st3(ps.nee_le, pid, mk(p), power_heuristic(ls.pdf, p));
                                st3(ps.nee_li, pid, ls.li, ls.pdf);
                            }
                        }
                        PSample ph = hg_sample(mdm, wo, rng);
                        if (ph.p > 0) {
                            float lp = light_pdf(sc, Lt, ms.p, ph.wi);
                            if (lp != 0.0f) {
                                // Ray3F(msr.P, wi): MinT = float.Epsilon
                                start_mis(ms.p, ph.wi, kFloatEps, mk(ph.p), power_heuristic(ph.p, lp), ph.p, l, env);
                                nflags |= kFlagMisIssued;
                            }
                        }
                        nflags |= kFlagNeePending;
                        st3(ps.nee_beta, pid, beta, selpdf);
                    }
                    PSample c = hg_sample(mdm, wo, rng);
                    eo = ms.p; ed = c.wi; etmin = rtmin; etmax = kFloatMax;   // Ray3F(msr.P, wi, ray.MinT)
                    spec = false;
                    scattered = true;
                }
            } else {
                if (depth == 0 || spec) {
                    if (hit) {
                        if (pr.light >= 0) {
                            const DevLight& Lt = sc.lights[pr.light];
                            if (dot(it.sh.Z, wo) > 0) L = L + beta * mk(Lt.le[0], Lt.le[1], Lt.le[2]);
                            else L = L + beta * mk(0.0f);
                        } else {
                            L = L + beta * mk(0.0f);
                        }
                    } else {
                        L = L + beta * eval_infinite(sc, rd);
                    }
                }
                if (!hit) {
                    alive = false;
                } else if (in.max_depth != -1 && depth >= in.max_depth) {
                    alive = false;
                } else if (pr.material < 0) {
                    // medium boundary without a surface: continue straight on in the medium behind it
                    eo = it.P; ed = rd; etmin = rtmin; etmax = kFloatMax;
                    env = get_medium(pr, it.sh.Z, rd);
                } else {
                    nvert++;
                    MatCtx m = load_material<kM>(sc, pr.material, it.u, it.v);
                    f3 wo_l = to_local(it.sh, wo);
                    f3 wo_d = to_local(it.sh, it.wr);   // EstimateDirectSurface: wo = inct.Wr (:156)
                    if (sc.num_lights > 0) {
                        // EstimateDirectSurface (:151-219)
                        int l = rng.next_int(sc.num_lights);
                        float selpdf = 1.0f / (float)sc.num_lights;
                        const DevLight Lt = sc.lights[l];
                        LSample ls = light_sample(sc, Lt, it.P, rng);
                        if (ls.pdf > 0.0f && !is_black(ls.li)) {
                            f3 wi_l = to_local(it.sh, ls.wi);
                            f3 fr = mat_fr<kM>(m, wo_d, wi_l);
                            float spdf = mat_pdf<kM>(m, wo_d, wi_l);
                            if (!is_black(fr)) {
                                start_shadow(it.P, ls.p, env);
                                nflags |= kFlagShadowIssued;
                                st3(ps.nee_le, pid, fr * abs_cos_theta(wi_l), power_heuristic(ls.pdf, spdf));
                                st3(ps.nee_li, pid, ls.li, ls.pdf);
                            }
                        }
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
                                start_mis(it.P, wi_w, 0.0001f, fr2, w2, sp2, l, env);   // Intersection.SpawnRay
                                nflags |= kFlagMisIssued;
                            }
                        }
                        nflags |= kFlagNeePending;
                        st3(ps.nee_beta, pid, beta, selpdf);
                    }
                    BSample bs = mat_sample<kM>(m, wo_l, rng);
                    spec = (bs.type & kSpecular) != 0;
                    if (bs.pdf > 0.0f) {
                        beta = beta * (bs.fr * abs_cos_theta(bs.wi) / bs.pdf);
                        f3 next = to_world(it.sh, bs.wi);
                        eo = it.P; ed = next; etmin = rtmin; etmax = kFloatMax;
                        env = get_medium(pr, it.sh.Z, next);
                        scattered = true;
                    } else {
                        alive = false;
                    }
                }
            }
            if (alive && scattered) {
                if (depth > in.min_depth) {
                    float qq = nmin(max_elem(beta), in.rr_threshold);
                    if (rng.next_float() > qq) alive = false;
                    else beta = beta / qq;
                }
                depth++;
            }
            // the next iteration starts with `if (ray.MaxT <= ray.MinT) break;` (:37-40)
            want_ext = alive && etmax > etmin;
            if (spec) nflags |= kFlagSpecular;
            if (!want_ext) nflags |= kFlagDone;
            st3(ps.L, pid, L);
            st3(ps.beta, pid, beta);
            ps.flags[pid] = nflags;
            ps.rng_n[pid] = rng.n;
            ps.depth[pid] = depth;
            ps.env[pid] = env;
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
        int slot = block_append<256>(&counts[kCntShadow], want_shadow, sh);
        if (want_shadow) {
            qshadow.pid[slot] = pid;
            qshadow.org[slot] = make_float4(so.x, so.y, so.z, 0.0001f);
            qshadow.dir[slot] = make_float4(sd.x, sd.y, sd.z, stmax);
        }
        slot = block_append<256>(&counts[kCntMis], want_mis, sh);
        if (want_mis) {
            qmis.pid[slot] = pid;
            qmis.org[slot] = make_float4(mo.x, mo.y, mo.z, mtmin);
            qmis.dir[slot] = make_float4(md.x, md.y, md.z, kFloatMax);
        }
    }
    add_counter((unsigned long long*)(counts + kCntVertices), nvert);
}

hipError_t launch_shade_vol(const DevScene& sc, const DevScene* scenes, RayQueue qin, const float4* hits,
                            RayQueue qshadow, RayQueue qmis, PathState ps, int* counts, int cnt_in, hipStream_t s) {
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(persistent_blocks()), dim3(256), 0, s, sc, scenes, qin, hits, qshadow, qmis,
                           ps, counts, cnt_in);
    };
    const unsigned set = mat_set(sc.mat_mask);
    if (scenes) {
        if (set == kMatsDelta) go(k_shade_vol<true, kMatsDelta>);
        else if (set == kMatsMicrofacet) go(k_shade_vol<true, kMatsMicrofacet>);
        else go(k_shade_vol<true, kMatsAll>);
    } else {
        if (set == kMatsDelta) go(k_shade_vol<false, kMatsDelta>);
        else if (set == kMatsMicrofacet) go(k_shade_vol<false, kMatsMicrofacet>);
        else go(k_shade_vol<false, kMatsAll>);
    }
    return hipGetLastError();
}

// One step of a chain after its closest-hit query (chain_step), for every
// chain ray of the queue; continuing chains append their next segment.
template <bool kBatch>
__global__ __launch_bounds__(256) void k_chain_advance(DevScene sc0, const DevScene* __restrict__ scenes, RayQueue qin,
                                                       const float4* __restrict__ hits, RayQueue qout, int mis,
                                                       PathState ps, int* counts, int cnt_in, int cnt_out) {
    __shared__ int sh[256 / 64 + 1];   // block_append
    int n = counts[cnt_in];
    // the traversal work cursor of this chain restarts at zero for its next step
    if (blockIdx.x == 0 && threadIdx.x == 0) counts[kCntWork + (mis ? 2 : 1)] = 0;
    for (int base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {
        int k = base + threadIdx.x;
        bool want = false;
        f3 no = mk(0.f), nd = mk(0.f);
        float ntmin = 0.f, ntmax = 0.f;
        int pid = 0;
        if (k < n) {
            pid = qin.pid[k];
            const DevScene& sc = kBatch ? scenes[ps.req[pid]] : sc0;
            float4 h = hits[k];
            int prim = __float_as_int(h.y);
            int hsub = (prim >= 0 && sc.prim_info[prim].kind == kInst) ? (mis ? ps.hsub2[k] : ps.hsub[k]) : -1;
            want = chain_step(sc, mis != 0, pid, qin.org[k], qin.dir[k], h, hsub, ps, &no, &nd, &ntmin, &ntmax);
        }
        int slot = block_append<256>(&counts[cnt_out], want, sh);
        if (want) {
            qout.pid[slot] = pid;
            qout.org[slot] = make_float4(no.x, no.y, no.z, ntmin);
            qout.dir[slot] = make_float4(nd.x, nd.y, nd.z, ntmax);
        }
    }
}

hipError_t launch_chain_advance(const DevScene& sc, const DevScene* scenes, RayQueue qin, const float4* hits,
                                RayQueue qout, bool mis, PathState ps, int* counts, int cnt_in, int cnt_out,
                                hipStream_t s) {
    if (scenes)
        hipLaunchKernelGGL(k_chain_advance<true>, dim3(persistent_blocks()), dim3(256), 0, s, sc, scenes, qin, hits,
                           qout, mis ? 1 : 0, ps, counts, cnt_in, cnt_out);
    else
        hipLaunchKernelGGL(k_chain_advance<false>, dim3(persistent_blocks()), dim3(256), 0, s, sc, scenes, qin, hits,
                           qout, mis ? 1 : 0, ps, counts, cnt_in, cnt_out);
    return hipGetLastError();
}

// SampleLightToEstimateDirect{Surface,Medium} result for every path of the
// bounce (VolumetricPathTracer.cs:121-147, 164-176, 199-215, 221-247, 264-311);
// a path that ended retires into the pass's per-sample result.
__global__ __launch_bounds__(256) void k_resolve_vol(RayQueue qin, PathState ps, int* counts, int cnt_in) {
    int n = counts[cnt_in];
    for (int base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {
        int k = base + threadIdx.x;
        bool done = false;
        int pid = 0;
        if (k < n) {
            pid = qin.pid[k];
            int flags = ps.flags[pid];
            done = (flags & kFlagDone) != 0;
            f3 L = ld3(ps.L, pid);
            if (flags & kFlagNeePending) {
                float4 nb = ps.nee_beta[pid];
                f3 beta = mk(nb.x, nb.y, nb.z);
                float selpdf = nb.w;
                f3 le = mk(0.0f);
                if (flags & kFlagShadowIssued) {
                    float4 li4 = ps.nee_li[pid];
                    f3 li = mk(li4.x, li4.y, li4.z) * ld3(ps.sh_tr, pid);
                    if (!is_black(li)) {
                        float4 f4 = ps.nee_le[pid];
                        le = le + mk(f4.x, f4.y, f4.z) * li * f4.w / li4.w;
                    }
                }
                if (flags & kFlagMisIssued) {
                    float4 li4 = ps.mis_li[pid];
                    f3 li = mk(li4.x, li4.y, li4.z);
                    if (!is_black(li)) {
                        float4 f4 = ps.mis_f[pid];
                        le = le + mk(f4.x, f4.y, f4.z) * li * ld3(ps.mis_tr, pid) * f4.w / li4.w;
                    }
                }
                f3 result = mk(0.0f) + le / selpdf;
                L = L + beta * result;
                st3(ps.L, pid, L);
            }
            if (done) ps.res[ps.gid[pid]] = make_float4(L.x, L.y, L.z, 0.f);
        }
    }
}

hipError_t launch_resolve_vol(RayQueue qin, PathState ps, int* counts, int cnt_in, hipStream_t s) {
    hipLaunchKernelGGL(k_resolve_vol, dim3(persistent_blocks()), dim3(256), 0, s, qin, ps, counts, cnt_in);
    return hipGetLastError();
}

}  // namespace rd
