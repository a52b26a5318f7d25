// rd_chain.h — one step of vol_path's Transmittance / IntersectTr chains after
// its closest-hit query, shared by the lockstep chain kernel (one step per
// launch over a queue) and the per-ray chain kernel (each lane walks its
// chain to the end).
#pragma once
#include "rd_kernels.h"
#include "rd_shading.h"
#include "rd_wave.h"

namespace rd {

// The chain ray (org = {o, tmin}, dir = {d, tmax}) of path pid hit h = {t,
// prim bits, u, v} (prim -1: miss); hsub = the set triangle of an instance hit.
//   Transmittance (Scene.cs:79-101): a surface blocks (Tr = 0); a medium
//   multiplies Tr over the whole segment [1e-4, dis - 1e-4]; a boundary hit
//   continues from the hit point toward the light, in the medium behind it.
//   IntersectTr (Scene.cs:108-130): a miss multiplies Tr to float.MaxValue and
//   takes the infinite lights; a hit multiplies Tr up to |ray.At(t)| (the
//   length of the hit's position vector, as written); a surface ends the chain
//   with the selected light's Le; a boundary continues with MinT 0.001.
// Updates the path's chain state; true (with the next segment) when the
// chain continues.
__device__ __forceinline__ bool chain_step(const DevScene& sc, bool mis, int pid, float4 a, float4 b, float4 h,
                                           int hsub, PathState ps, f3* no, f3* nd, float* ntmin, float* ntmax) {
    bool want = false;
    f3 o = mk(a.x, a.y, a.z), d = mk(b.x, b.y, b.z);
    float tmin = a.w, tmax = b.w;
    int prim = __float_as_int(h.y);
    bool any = prim >= 0;
    DevPrim pr;
    if (any) pr = sc.prim_info[prim];
    if (!mis) {
        float4 to4 = ps.sh_to[pid];
        int env = __float_as_int(to4.w);
        f3 tr = ld3(ps.sh_tr, pid);
        if (any && pr.material >= 0) {
            tr = mk(0.0f);
        } else {
            if (env >= 0) tr = tr * medium_tr(sc.media[env], tmin, tmax);
            if (any) {
                Inct it = make_inct_any(sc, pr, pr.kind == kInst ? hsub : -1, o, d, h.z, h.w, h.x);
                env = get_medium(pr, it.sh.Z, d);
                f3 to = mk(to4.x, to4.y, to4.z);
                float dis = length(to - it.P);
                if (!(dis < 0.0001f)) {
                    want = true;
                    *no = it.P; *nd = d; *ntmin = 0.0001f; *ntmax = dis - 0.0001f;
                }
            }
        }
        st3(ps.sh_tr, pid, tr);
        ps.sh_to[pid] = make_float4(to4.x, to4.y, to4.z, __int_as_float(env));
    } else {
        float4 tr4 = ps.mis_tr[pid];
        int env = __float_as_int(tr4.w);
        f3 tr = mk(tr4.x, tr4.y, tr4.z);
        if (!any) {
            if (env >= 0) tr = tr * medium_tr(sc.media[env], tmin, tmax);
            f3 li = eval_infinite(sc, d);
            float4 prev = ps.mis_li[pid];
            ps.mis_li[pid] = make_float4(li.x, li.y, li.z, prev.w);
        } else {
            if (env >= 0) tr = tr * medium_tr(sc.media[env], tmin, length(o + h.x * d));
            Inct it = make_inct_any(sc, pr, pr.kind == kInst ? hsub : -1, o, d, h.z, h.w, h.x);
            if (pr.material >= 0) {
                f3 li = mk(0.0f);
                int sel = ps.sel_light[pid];
                if (pr.light >= 0 && pr.light == sel) {
                    const DevLight& Lt = sc.lights[sel];
                    if (dot(it.sh.Z, -d) > 0) li = mk(Lt.le[0], Lt.le[1], Lt.le[2]);
                }
                float4 prev = ps.mis_li[pid];
                ps.mis_li[pid] = make_float4(li.x, li.y, li.z, prev.w);
            } else {
                env = get_medium(pr, it.sh.Z, d);
                want = true;
                *no = it.P; *nd = d; *ntmin = 0.001f; *ntmax = kFloatMax;
            }
        }
        ps.mis_tr[pid] = make_float4(tr.x, tr.y, tr.z, __int_as_float(env));
    }
    return want;
}

}  // namespace rd
