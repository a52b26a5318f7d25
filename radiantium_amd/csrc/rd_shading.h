// rd_shading.h — gfx950 device functions for intersection records, textures,
// BSDFs and area lights.  Each function restates the reference function it
// cites; texture parameters are fetched once per shading point into a
// MatCtx (the reference re-samples them in every Fr/Pdf/Sample call with the
// same uv, which gives the same values).
#pragma once
#include "rd_types.h"

namespace rd {

// ---------------------------------------------------------------- frames
// Coordinate (Radiantium.Offline/Coordinate.cs:14-36), Duff et al. basis.
struct Frame {
    f3 X, Y, Z;
};
__device__ __forceinline__ Frame make_frame(f3 n) {
    float sign = copysignf(1.0f, n.z);
    float a = -1 / (sign + n.z);
    float b = n.x * n.y * a;
    Frame f;
    f.X = mk(mulsign(sqr(n.x) * a, n.z) + 1.0f, mulsign(b, n.z), mulsign(n.x, -n.z));
    f.Y = mk(b, sign + sqr(n.y) * a, -n.y);
    f.Z = n;
    return f;
}
__device__ __forceinline__ f3 to_local(const Frame& f, f3 v) { return mk(dot(v, f.X), dot(v, f.Y), dot(v, f.Z)); }
__device__ __forceinline__ f3 to_world(const Frame& f, f3 v) { return f.X * v.x + f.Y * v.y + f.Z * v.z; }

// Coordinate helpers (Coordinate.cs:80-122); MathExt.Fma is a*b+c unfused.
__device__ __forceinline__ float cos_theta(f3 v) { return v.z; }
__device__ __forceinline__ float abs_cos_theta(f3 v) { return fabsf(v.z); }
__device__ __forceinline__ float cos2_theta(f3 v) { return sqr(v.z); }
__device__ __forceinline__ float sin2_theta(f3 v) { return v.x * v.x + sqr(v.y); }
__device__ __forceinline__ float tan_theta(f3 v) { return safe_sqrt(-v.z * v.z + 1.0f) / v.z; }
__device__ __forceinline__ float tan2_theta(f3 v) { return nmax(-v.z * v.z + 1.0f, 0.0f) / sqr(v.z); }
__device__ __forceinline__ float cos2_phi(f3 v) {
    float s2 = sin2_theta(v);
    return fabsf(s2) <= 4.0f * kFloatEps ? 1.0f : nclamp(sqr(v.x) / s2, -1.0f, 1.0f);
}
__device__ __forceinline__ float sin2_phi(f3 v) {
    float s2 = sin2_theta(v);
    return fabsf(s2) <= 4.0f * kFloatEps ? 0.0f : nclamp(sqr(v.y) / s2, -1.0f, 1.0f);
}
__device__ __forceinline__ bool same_hemisphere(f3 a, f3 b) { return a.z * b.z > 0; }
__device__ __forceinline__ f3 faceforward(f3 v, f3 v2) { return dot(v, v2) < 0.0f ? -v : v; }

// ---------------------------------------------------------------- geometry access
__device__ __forceinline__ f3 load_f3(const float* p, int i) { return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }

struct Inct {
    f3 P;
    float u, v;   // texture coordinates (Intersection.UV)
    float T;
    Frame sh;
    f3 wr;        // Intersection.Wr: -ray.D, or its image for an instance
};

// Vector3.Transform / TransformNormal with a 4x3 row-major matrix (rows 1-3
// linear, row 4 translation; row vectors).
__device__ __forceinline__ f3 xform_point12(const float* M, f3 v) {
    return mk(v.x * M[0] + v.y * M[3] + v.z * M[6] + M[9], v.x * M[1] + v.y * M[4] + v.z * M[7] + M[10],
              v.x * M[2] + v.y * M[5] + v.z * M[8] + M[11]);
}
__device__ __forceinline__ f3 xform_vector12(const float* M, f3 v) {
    return mk(v.x * M[0] + v.y * M[3] + v.z * M[6], v.x * M[1] + v.y * M[4] + v.z * M[7],
              v.x * M[2] + v.y * M[5] + v.z * M[8]);
}

// Triangle.GetIntersection (Shapes/Triangle.cs:96-120) / Sphere.GetIntersection
// (Shapes/Sphere.cs:34-48) for a hit (bu, bv, t) on original primitive `prim`.
__device__ __forceinline__ Inct make_inct(const DevScene& sc, const DevPrim& pr, f3 ro, f3 rd, float bu, float bv, float t) {
    Inct it;
    it.T = t;
    if (pr.kind == kTri) {
        int i0 = sc.tri_idx[3 * pr.index], i1 = sc.tri_idx[3 * pr.index + 1], i2 = sc.tri_idx[3 * pr.index + 2];
        f3 A = load_f3(sc.pos, i0), B = load_f3(sc.pos, i1), C = load_f3(sc.pos, i2);
        float bx = 1 - (bu + bv), by = bu, bz = bv;
        it.P = bx * A + by * B + bz * C;
        f3 n;
        if (pr.flags & 1) n = normalize(bx * load_f3(sc.nrm, i0) + by * load_f3(sc.nrm, i1) + bz * load_f3(sc.nrm, i2));
        else n = normalize(cross(B - A, C - A));
        if (pr.flags & 2) {
            it.u = bx * sc.uv[2 * i0] + by * sc.uv[2 * i1] + bz * sc.uv[2 * i2];
            it.v = bx * sc.uv[2 * i0 + 1] + by * sc.uv[2 * i1 + 1] + bz * sc.uv[2 * i2 + 1];
        } else {
            it.u = bx * 0.0f + by * 1.0f + bz * 1.0f;
            it.v = bx * 0.0f + by * 0.0f + bz * 1.0f;
        }
        it.sh = make_frame(n);
    } else {
        float4 s = sc.spheres[pr.index];
        f3 c = mk(s.x, s.y, s.z);
        f3 n = normalize((ro + t * rd) - c);
        it.P = n * s.w + c;
        it.sh = make_frame(n);
        // local = TransformNormal(n, WorldToModel) (Sphere.cs:40)
        f3 local = n;
        if (sc.sphere_w2m) {
            const float* M = sc.sphere_w2m + 9 * pr.index;
            local = mk(n.x * M[0] + n.y * M[3] + n.z * M[6], n.x * M[1] + n.y * M[4] + n.z * M[7],
                       n.x * M[2] + n.y * M[5] + n.z * M[8]);
        }
        float theta = acosf(-local.y);
        float phi = atan2f(-local.z, local.x) + kPi;
        it.u = phi / (2 * kPi);
        it.v = theta / kPi;
    }
    it.wr = -rd;
    return it;
}

// InstancedPrimitive.Intersect's hit record (Primitive.cs:170-183) for a hit
// on set triangle `sub` (a DevPrim index) with world distance T: the
// triangle's GetIntersection in model space (ray direction normalised into
// model space, Ray3F.Transform), then P and the shading normal to world
// space, a new frame from that normal, Wr = normalize(M2W * (-d_model)).
__device__ __forceinline__ Inct make_inct_inst(const DevScene& sc, const DevPrim& pr, int sub, f3 ro, f3 rd, float bu,
                                               float bv, float T) {
    const DevInstance& I = sc.insts[pr.index];
    f3 md = normalize(xform_vector12(I.w2m, rd));
    f3 mo = xform_point12(I.w2m, ro);
    Inct m = make_inct(sc, sc.prim_info[sub], mo, md, bu, bv, T);
    Inct w;
    w.P = xform_point12(I.m2w, m.P);
    w.sh = make_frame(normalize(xform_vector12(I.m2w, m.sh.Z)));
    w.wr = normalize(xform_vector12(I.m2w, m.wr));
    w.u = m.u;
    w.v = m.v;
    w.T = T;
    return w;
}

__device__ __forceinline__ Inct make_inct_any(const DevScene& sc, const DevPrim& pr, int sub, f3 ro, f3 rd, float bu,
                                              float bv, float t) {
    if (pr.kind == kInst) return make_inct_inst(sc, pr, sub, ro, rd, bu, bv, t);
    return make_inct(sc, pr, ro, rd, bu, bv, t);
}

// Geometric/shading normal only (for Le checks at a hit).
__device__ __forceinline__ f3 hit_normal(const DevScene& sc, const DevPrim& pr, f3 ro, f3 rd, float bu, float bv, float t) {
    if (pr.kind == kTri) {
        int i0 = sc.tri_idx[3 * pr.index], i1 = sc.tri_idx[3 * pr.index + 1], i2 = sc.tri_idx[3 * pr.index + 2];
        if (pr.flags & 1) {
            float bx = 1 - (bu + bv), by = bu, bz = bv;
            return normalize(bx * load_f3(sc.nrm, i0) + by * load_f3(sc.nrm, i1) + bz * load_f3(sc.nrm, i2));
        }
        f3 A = load_f3(sc.pos, i0), B = load_f3(sc.pos, i1), C = load_f3(sc.pos, i2);
        return normalize(cross(B - A, C - A));
    }
    float4 s = sc.spheres[pr.index];
    return normalize((ro + t * rd) - mk(s.x, s.y, s.z));
}

// Triangle.Intersect(ray, out SurfacePoint) (Triangle.cs:149-177) on vertices.
__device__ __forceinline__ bool tri_test(f3 A, f3 B, f3 C, f3 o, f3 d, float tmin, float tmax, float* tu, float* tv, float* tt) {
    f3 e1 = B - A, e2 = C - A;
    f3 pvec = cross(d, e2);
    float det = dot(e1, pvec);
    if (fabsf(det) < kFloatEps) return false;
    float inv_det = 1.0f / det;
    f3 tvec = o - A;
    float u = dot(tvec, pvec) * inv_det;
    if (u < 0.0f || u > 1.0f) return false;
    f3 qvec = cross(tvec, e1);
    float v = dot(d, qvec) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return false;
    float t = dot(e2, qvec) * inv_det;
    if (!(t >= tmin && t <= tmax)) return false;
    *tu = u; *tv = v; *tt = t;
    return true;
}

// Math.Min / Math.Max on doubles (NaN propagates, -0 < +0).
__device__ __forceinline__ double dmin_net(double a, double b) {
    if (a != b) { if (!isnan(a)) return a < b ? a : b; return a; }
    return signbit(a) ? a : b;
}
__device__ __forceinline__ double dmax_net(double a, double b) {
    if (a != b) { if (!isnan(a)) return b < a ? a : b; return a; }
    return signbit(b) ? a : b;
}

// Sphere.Intersect(ray, out SurfacePoint) (Sphere.cs:65-80) with
// MathExt.SolveQuadratic (MathExt.cs:119-138), double precision.
__device__ __forceinline__ bool sphere_test(float4 s, f3 o, f3 d, float tmin, float tmax, float* tt) {
    double mint = tmin, maxt = tmax;
    double ox = (double)o.x - (double)s.x, oy = (double)o.y - (double)s.y, oz = (double)o.z - (double)s.z;
    double dx = d.x, dy = d.y, dz = d.z;
    double a = dx * dx + dy * dy + dz * dz;
    double b = 2 * (ox * dx + oy * dy + oz * dz);
    double c = (ox * ox + oy * oy + oz * oz) - (double)(s.w * s.w);
    bool linear = a == 0.0;
    bool valid_linear = linear && b != 0.0;
    double x0 = -c / b, x1 = -c / b;
    double disc = b * b - 4.0 * a * c;
    bool valid_quad = !linear && (disc >= 0.0);
    if (valid_quad) {
        double rdd = sqrt(disc);
        double temp = -0.5 * (b + copysign(rdd, b));
        double x0p = temp / a, x1p = c / temp;
        double x0m = dmin_net(x0p, x1p);
        double x1m = dmax_net(x0p, x1p);
        x0 = x0m;
        x1 = x1m;
    }
    bool found = valid_linear || valid_quad;
    bool out_b = !(x0 <= maxt && x1 >= mint);
    bool in_b = x0 < mint && x1 > maxt;
    if (!(found && !out_b && !in_b)) return false;
    *tt = x0 < mint ? (float)x1 : (float)x0;
    return true;
}

// Intersects original primitive `prim` alone (Shape.Intersect).
__device__ __forceinline__ bool prim_test(const DevScene& sc, const DevPrim& pr, f3 o, f3 d, float tmin, float tmax,
                                          float* tu, float* tv, float* tt) {
    if (pr.kind == kTri) {
        int i0 = sc.tri_idx[3 * pr.index], i1 = sc.tri_idx[3 * pr.index + 1], i2 = sc.tri_idx[3 * pr.index + 2];
        return tri_test(load_f3(sc.pos, i0), load_f3(sc.pos, i1), load_f3(sc.pos, i2), o, d, tmin, tmax, tu, tv, tt);
    }
    *tu = 0; *tv = 0;
    return sphere_test(sc.spheres[pr.index], o, d, tmin, tmax, tt);
}

// ---------------------------------------------------------------- textures
// TextureSampler.Sample (TextureSampler.cs:28-72) over 3-channel column-major
// texels; constant textures return their colour.
__device__ __forceinline__ f3 tex_read(const DevScene& sc, const DevTexture& t, int x, int y) {
    const float* p = sc.texels + t.offset + ((int64_t)x * t.h + y) * 3;
    return mk(p[0], p[1], p[2]);
}
__device__ __forceinline__ f3 tex_sample(const DevScene& sc, int id, float su, float sv) {
    const DevTexture t = sc.texs[id];
    if (t.w == 0) return mk(t.v[0], t.v[1], t.v[2]);
    float u, v;
    if (t.wrap == 0) { u = nclamp(su, 0.0f, 1.0f); v = nclamp(sv, 0.0f, 1.0f); }
    else { u = nclamp(su - floorf(su), 0.0f, 1.0f); v = nclamp(sv - floorf(sv), 0.0f, 1.0f); }
    float fu = u * (t.w - 1), fv = v * (t.h - 1);
    int pu = (int)fu, pv = (int)fv;
    if (t.filter == 0) return tex_read(sc, t, pu, pv);
    int dpu = (fu > pu + 0.5f) ? 1 : -1;
    int dpv = (fv > pv + 0.5f) ? 1 : -1;
    int apu = iclamp(pu + dpu, 0, t.w - 1);
    int apv = iclamp(pv + dpv, 0, t.h - 1);
    float du = nmin(fabsf(fu - pu - 0.5f), 1.0f);
    float dv = nmin(fabsf(fv - pv - 0.5f), 1.0f);
    f3 a = tex_read(sc, t, pu, pv), b = tex_read(sc, t, apu, pv), c = tex_read(sc, t, pu, apv), e = tex_read(sc, t, apu, apv);
    return (a * (1 - du) + b * du) * (1 - dv) + (c * (1 - du) + e * du) * dv;
}

// ---------------------------------------------------------------- microfacets
// Bxdf/Microfacet.cs: Beckmann (kind 0) and GGX (kind 1).
struct Mf {
    int kind;
    float ax, ay;
};
__device__ __forceinline__ Mf make_mf(int kind, float roughness, float aniso) {   // RoughnessToAlpha :177-183
    Mf m;
    m.kind = kind;
    m.ax = nmax(roughness * roughness * (1 + aniso), 0.001f);
    m.ay = nmax(roughness * roughness * (1 - aniso), 0.001f);
    return m;
}
__device__ __forceinline__ float mf_D(const Mf& m, f3 wh) {
    float t2 = tan2_theta(wh);
    if (isinf(t2)) return 0.0f;
    float c4 = cos2_theta(wh) * cos2_theta(wh);
    if (m.kind == 0)
        return expf(-t2 * (cos2_phi(wh) / (m.ax * m.ax) + sin2_phi(wh) / (m.ay * m.ay))) / (kPi * m.ax * m.ay * c4);
    float e = (cos2_phi(wh) / (m.ax * m.ax) + sin2_phi(wh) / (m.ay * m.ay)) * t2;
    return 1 / (kPi * m.ax * m.ay * c4 * (1 + e) * (1 + e));
}
__device__ __forceinline__ float mf_lambda(const Mf& m, f3 w) {
    float at = fabsf(tan_theta(w));
    if (isinf(at)) return 0.0f;
    float alpha = sqrtf(cos2_phi(w) * m.ax * m.ax + sin2_phi(w) * m.ay * m.ay);
    if (m.kind == 0) {
        float a = 1 / (alpha * at);
        if (a >= 1.6f) return 0;
        return (1 - 1.259f * a + 0.396f * a * a) / (3.535f * a + 2.181f * a * a);
    }
    float a2t2 = (alpha * at) * (alpha * at);
    return (-1 + sqrtf(1.0f + a2t2)) / 2;
}
__device__ __forceinline__ float mf_G(const Mf& m, f3 wo, f3 wi) { return 1 / (1 + mf_lambda(m, wo) + mf_lambda(m, wi)); }
__device__ __forceinline__ float mf_pdf(const Mf& m, f3 wo, f3 wh) { return mf_D(m, wh) * abs_cos_theta(wh); }
__device__ __forceinline__ f3 spherical_dir(float st, float ct, float phi) {
    float sp = sinf(phi), cp = cosf(phi);
    return mk(st * cp, st * sp, ct);
}
__device__ __forceinline__ f3 mf_sample_wh(const Mf& m, f3 wo, Rng& rng) {
    float ux = rng.next_float(), uy = rng.next_float();
    float phi, ct;
    if (m.kind == 0) {   // Beckmann.SampleWh (Microfacet.cs:60-86)
        float t2;
        if (m.ax == m.ay) {
            float ls = logf(1 - ux);
            t2 = -m.ax * m.ax * ls;
            phi = uy * 2 * kPi;
        } else {
            float ls = logf(1 - ux);
            phi = atanf(m.ay / m.ax * tanf(2 * kPi * uy + 0.5f * kPi));
            if (uy > 0.5f) phi += kPi;
            float sp = sinf(phi), cp = cosf(phi);
            float ax2 = m.ax * m.ax, ay2 = m.ay * m.ay;
            t2 = -ls / (cp * cp / ax2 + sp * sp / ay2);
        }
        ct = 1 / sqrtf(1 + t2);
        float st = sqrtf(nmax(0.0f, 1 - ct * ct));
        f3 wh = spherical_dir(st, ct, phi);
        if (!same_hemisphere(wo, wh)) wh = -wh;
        return wh;
    }
    phi = (2 * kPi) * uy;   // SampleWhGGX (Microfacet.cs:185-209)
    if (m.ax == m.ay) {
        float t2 = m.ax * m.ax * ux / (1.0f - ux);
        ct = 1 / sqrtf(1 + t2);
    } else {
        phi = atanf(m.ay / m.ax * tanf(2 * kPi * uy + .5f * kPi));
        if (uy > 0.5f) phi += kPi;
        float sp = sinf(phi), cp = cosf(phi);
        float ax2 = m.ax * m.ax, ay2 = m.ay * m.ay;
        float a2 = 1 / (cp * cp / ax2 + sp * sp / ay2);
        float t2 = a2 * ux / (1 - ux);
        ct = 1 / sqrtf(1 + t2);
    }
    float st = sqrtf(nmax(0.0f, 1.0f - ct * ct));
    f3 wh = spherical_dir(st, ct, phi);
    if (!same_hemisphere(wo, wh)) wh = -wh;
    return wh;
}

// Fresnel.DielectricFunc (Bxdf/Fresnel.cs:50-68)
__device__ __forceinline__ float fr_dielectric(float ci, float ei, float et) {
    ci = nclamp(ci, -1, 1);
    if (!(ci > 0.0f)) { float t = ei; ei = et; et = t; ci = fabsf(ci); }
    float si = sqrtf(nmax(0, 1 - ci * ci));
    float st = ei / et * si;
    if (st >= 1) return 1;
    float ct = sqrtf(nmax(0, 1 - st * st));
    float rparl = ((et * ci) - (ei * ct)) / ((et * ci) + (ei * ct));
    float rperp = ((ei * ci) - (et * ct)) / ((ei * ci) + (et * ct));
    return (rparl * rparl + rperp * rperp) / 2;
}
// Fresnel.ConductorFunc (Bxdf/Fresnel.cs:72-95), etai = 1
__device__ __forceinline__ f3 fr_conductor(float ci, f3 etai, f3 etat, f3 k) {
    ci = nclamp(ci, -1, 1);
    f3 eta = etat / etai;
    f3 etak = k / etai;
    float ci2 = ci * ci;
    float si2 = 1.0f - ci2;
    f3 eta2 = eta * eta;
    f3 etak2 = etak * etak;
    f3 t0 = eta2 - etak2 - si2;
    f3 a2pb2 = sqrt3(t0 * t0 + 4 * eta2 * etak2);
    f3 t1 = a2pb2 + ci2;
    f3 a = sqrt3(0.5f * (a2pb2 + t0));
    f3 t2 = 2 * ci * a;
    f3 Rs = (t1 - t2) / (t1 + t2);
    f3 t3 = ci2 * a2pb2 + si2 * si2;
    f3 t4 = t2 * si2;
    f3 Rp = Rs * (t3 - t4) / (t3 + t4);
    return 0.5f * (Rp + Rs);
}
// MathExt.Refract (MathExt.cs:83-92)
__device__ __forceinline__ bool refract(f3 wi, f3 n, float eta, f3* wt) {
    float ci = dot(n, wi);
    float s2i = nmax(0, 1 - ci * ci);
    float s2t = eta * eta * s2i;
    if (s2t >= 1) return false;
    float ct = sqrtf(1 - s2t);
    *wt = normalize(eta * -wi + (eta * ci - ct) * n);
    return true;
}

// ---------------------------------------------------------------- materials
// Disney parameters (Materials/Disney.cs:65-85 CreateBsdf).
struct DisneyP {
    f3 color;
    float metallic, eta, roughness, spec_tint, aniso, sheen, sheen_tint, clearcoat, cc_gloss, spec_scale;
    float transmission, trans_rough, flatness;
    bool thin, sdist_black;
};

// Whether material kind k is compiled into a kernel for the set kM; the
// cases of the other kinds compile to nothing (callers pass only kinds in kM).
#define RD_MAT(k) ((kM & mat_bit(k)) != 0)

struct MatCtx {
    int type, two_sided;
    Mf mf;
    f3 r, t;        // diffuse kd / glass, plastic and mirror colours
    f3 eta, k;      // conductor
    float eta_a, eta_b, kd, ks;
    DisneyP dp;
    bool sss;       // subsurface: a non-black scatting distance (Subsurface.cs:50-54)
};

// kM: the material kinds compiled in (the others load nothing)
template <unsigned kM = kMatsAll>
__device__ __forceinline__ MatCtx load_material(const DevScene& sc, int id, float su, float sv) {
    const DevMaterial& m = sc.mats[id];
    MatCtx c;
    c.type = m.type;
    c.two_sided = m.two_sided;
    c.eta_a = m.eta_a;
    c.eta_b = m.eta_b;
    c.kd = m.kd;
    c.ks = m.ks;
    c.r = mk(0.0f); c.t = mk(0.0f); c.eta = mk(0.0f); c.k = mk(0.0f);
    c.mf.kind = 0; c.mf.ax = c.mf.ay = 0.001f;
    c.sss = false;
    switch (m.type) {
    case kMatDiffuse:
        if constexpr (RD_MAT(kMatDiffuse)) c.r = tex_sample(sc, m.tex[0], su, sv);
        break;
    case kMatRoughMetal:   // RoughMetal.GetParam (RoughMetal.cs:28-31)
        if constexpr (RD_MAT(kMatRoughMetal)) {
            c.mf = make_mf(m.dist, tex_sample(sc, m.tex[0], su, sv).x, tex_sample(sc, m.tex[1], su, sv).x);
            c.eta = mk(m.eta[0], m.eta[1], m.eta[2]);
            c.k = mk(m.k[0], m.k[1], m.k[2]);
        }
        break;
    case kMatRoughGlass:   // RoughGlass.GetParam (RoughGlass.cs:146-149)
        if constexpr (RD_MAT(kMatRoughGlass)) {
            c.mf = make_mf(m.dist, tex_sample(sc, m.tex[2], su, sv).x, tex_sample(sc, m.tex[3], su, sv).x);
            c.r = tex_sample(sc, m.tex[0], su, sv);
            c.t = tex_sample(sc, m.tex[1], su, sv);
        }
        break;
    case kMatRoughPlastic: // RoughPlastic.SampleParams / SampleRoughness (RoughPlastic.cs:37-45)
        if constexpr (RD_MAT(kMatRoughPlastic)) {
            c.mf = make_mf(m.dist, tex_sample(sc, m.tex[1], su, sv).x, tex_sample(sc, m.tex[2], su, sv).x);
            c.r = tex_sample(sc, m.tex[0], su, sv);
        }
        break;
    case kMatPerfectGlass:
        if constexpr (RD_MAT(kMatPerfectGlass)) {
            c.r = tex_sample(sc, m.tex[0], su, sv);
            c.t = tex_sample(sc, m.tex[1], su, sv);
        }
        break;
    case kMatPerfectMirror:
        if constexpr (RD_MAT(kMatPerfectMirror)) c.r = tex_sample(sc, m.tex[0], su, sv);
        break;
    case kMatSubsurface:   // Subsurface.Sample's FresnelSpecularBsdf(R, T, EtaA, EtaB)
        if constexpr (RD_MAT(kMatSubsurface)) {
            c.r = tex_sample(sc, m.tex[0], su, sv);
            c.t = tex_sample(sc, m.tex[1], su, sv);
            c.sss = !is_black(tex_sample(sc, m.tex[3], su, sv));
        }
        break;
    case kMatDisney:
        if constexpr (RD_MAT(kMatDisney)) {
            DisneyP& b = c.dp;
            b.color = tex_sample(sc, m.tex[0], su, sv);
            b.metallic = tex_sample(sc, m.tex[1], su, sv).x;
            b.roughness = tex_sample(sc, m.tex[2], su, sv).x;
            b.eta = tex_sample(sc, m.tex[3], su, sv).x;
            b.spec_scale = tex_sample(sc, m.tex[4], su, sv).x;
            b.spec_tint = tex_sample(sc, m.tex[5], su, sv).x;
            b.aniso = tex_sample(sc, m.tex[6], su, sv).x;
            b.sheen = tex_sample(sc, m.tex[7], su, sv).x;
            b.sheen_tint = tex_sample(sc, m.tex[8], su, sv).x;
            b.clearcoat = tex_sample(sc, m.tex[9], su, sv).x;
            b.cc_gloss = tex_sample(sc, m.tex[10], su, sv).x;
            b.sdist_black = is_black(tex_sample(sc, m.tex[11], su, sv));
            b.transmission = tex_sample(sc, m.tex[12], su, sv).x;
            b.trans_rough = tex_sample(sc, m.tex[13], su, sv).x;
            b.flatness = tex_sample(sc, m.tex[14], su, sv).x;
            b.thin = m.thin != 0;
        }
        break;
    default:
        break;
    }
    return c;
}

// MicrofacetReflectionBrdf<Conductor,D>.Fr with R = 1 (MicrofacetReflectionBrdf.cs:26-41)
__device__ __forceinline__ f3 conductor_fr(const MatCtx& c, f3 wo, f3 wi) {
    if (!same_hemisphere(wo, wi)) return mk(0.0f);
    float cto = abs_cos_theta(wo), cti = abs_cos_theta(wi);
    f3 wh = wi + wo;
    if (cti == 0 || cto == 0) return mk(0.0f);
    if (wh.x == 0 && wh.y == 0 && wh.z == 0) return mk(0.0f);
    wh = normalize(wh);
    f3 F = fr_conductor(dot(wi, faceforward(wh, mk(0, 0, 1))), mk(1.0f), c.eta, c.k);
    float D = mf_D(c.mf, wh);
    float G = mf_G(c.mf, wo, wi);
    return mk(1.0f) * fabs3((D * G * F) / (4 * cti * cto));
}

__device__ __forceinline__ float mf_reflect_pdf(const Mf& m, f3 wo, f3 wi) {   // MicrofacetReflectionBrdf.Pdf
    if (!same_hemisphere(wo, wi)) return 0;
    f3 wh = normalize(wo + wi);
    return mf_pdf(m, wo, wh) / (4 * dot(wo, wh));
}

// RoughPlastic FrImpl / PdfImpl (RoughPlastic.cs:47-110)
__device__ __forceinline__ f3 plastic_fr(const MatCtx& c, f3 wo, f3 wi) {
    if (cos_theta(wo) <= 0 || cos_theta(wi) <= 0) { if (!c.two_sided) return mk(0.0f); }
    f3 diffuse = same_hemisphere(wo, wi) ? c.r / kPi : mk(0.0f);
    f3 glossy = mk(0.0f);
    if (same_hemisphere(wo, wi)) {
        float cto = abs_cos_theta(wo), cti = abs_cos_theta(wi);
        f3 wh = wi + wo;
        if (!(cti == 0 || cto == 0) && !(wh.x == 0 && wh.y == 0 && wh.z == 0)) {
            wh = normalize(wh);
            f3 F = mk(fr_dielectric(dot(wi, faceforward(wh, mk(0, 0, 1))), c.eta_a, c.eta_b));
            float D = mf_D(c.mf, wh), G = mf_G(c.mf, wo, wi);
            glossy = c.r * fabs3((D * G * F) / (4 * cti * cto));
        }
    }
    return c.kd * diffuse + c.ks * glossy;
}
__device__ __forceinline__ float plastic_pdf(const MatCtx& c, f3 wo, f3 wi) {
    if (cos_theta(wo) <= 0 || cos_theta(wi) <= 0) { if (!c.two_sided) return 0.0f; }
    float p = c.ks / (c.kd + c.ks);
    float diffuse = same_hemisphere(wo, wi) ? abs_cos_theta(wi) / kPi : 0.0f;
    float glossy = mf_reflect_pdf(c.mf, wo, wi);
    return (1 - p) * diffuse + p * glossy;
}

// ---- Disney (Bxdf/DisneyBsdf.cs:13-512)
__device__ __forceinline__ float lerpf(float t, float a, float b) { return (1 - t) * a + t * b; }
__device__ __forceinline__ f3 lerp3(float w, f3 a, f3 b) { return (1 - w) * a + w * b; }
__device__ __forceinline__ float pow5f(float a) { return (a * a) * (a * a) * a; }
__device__ __forceinline__ float schlick_weight(float c) { return pow5f(nclamp(1 - c, 0, 1)); }
__device__ __forceinline__ f3 schlick3(f3 r0, float c) { return lerp3(schlick_weight(c), r0, mk(1.0f)); }
__device__ __forceinline__ float schlick1(float r0, float c) { return r0 + (1.0f - r0) * schlick_weight(c); }
__device__ __forceinline__ float luminance(f3 c) { return c.x * 0.212671f + c.y * 0.715160f + c.z * 0.072169f; }
__device__ __forceinline__ f3 color_tint(const DisneyP& b) {
    float l = luminance(b.color);
    return l > 0 ? b.color / l : mk(1.0f);
}
__device__ __forceinline__ Mf gtr2_dist(const DisneyP& b, float rough) {   // AnisAlpha
    float aspect = sqrtf(1 - b.aniso * 0.9f);
    Mf m;
    m.kind = 1;
    m.ax = nmax(0.001f, sqr(rough) / aspect);
    m.ay = nmax(0.001f, sqr(rough) * aspect);
    return m;
}
// DisneyDistributionGtr2.G: separable G1(wo) G1(wi)
__device__ __forceinline__ float gtr2_G(const Mf& m, f3 wo, f3 wi) {
    return (1 / (1 + mf_lambda(m, wo))) * (1 / (1 + mf_lambda(m, wi)));
}
__device__ __forceinline__ float gtr2_reflect_pdf(const Mf& m, f3 wo, f3 wi) {
    if (!same_hemisphere(wo, wi)) return 0;
    f3 wh = normalize(wo + wi);
    return mf_pdf(m, wo, wh) / (4 * dot(wo, wh));
}
__device__ __forceinline__ f3 disney_fresnel(const DisneyP& b, float ci) {   // DisneyFresnel.Eval
    f3 r0 = color_tint(b) * (sqr(b.eta - 1) / sqr(b.eta + 1));
    f3 fd = lerp3(b.spec_tint, mk(fr_dielectric(ci, 1, b.eta)), schlick3(r0, ci));
    return lerp3(b.metallic, fd * b.spec_scale, schlick3(b.color, ci));
}
// MicrofacetReflectionBrdf<F, Gtr2>.Fr: spec (DisneyFresnel, R=1) or inner (dielectric F, R=T*color)
__device__ __forceinline__ f3 gtr2_reflect_fr(const DisneyP& b, const Mf& m, bool inner, f3 wo, f3 wi) {
    if (!same_hemisphere(wo, wi)) return mk(0.0f);
    float cto = abs_cos_theta(wo), cti = abs_cos_theta(wi);
    f3 wh = wi + wo;
    if (cti == 0 || cto == 0) return mk(0.0f);
    if (wh.x == 0 && wh.y == 0 && wh.z == 0) return mk(0.0f);
    wh = normalize(wh);
    float ci = dot(wi, faceforward(wh, mk(0, 0, 1)));
    f3 F = inner ? mk(fr_dielectric(ci, 1, b.eta)) : disney_fresnel(b, ci);
    f3 R = inner ? b.transmission * b.color : mk(1.0f);
    float D = mf_D(m, wh), G = gtr2_G(m, wo, wi);
    return R * fabs3((D * G * F) / (4 * cti * cto));
}
// MicrofacetTransmissionBtdf<Gtr2> Fr / Pdf (MicrofacetTransmissionBtdf.cs:25-57), etaA = 1, etaB = Eta;
// imp: TransportMode.Importance (light paths), which drops the 1/eta radiance scaling (:40)
__device__ __forceinline__ f3 disney_btdf_fr(const DisneyP& b, f3 wo, f3 wi, bool imp = false) {
    if (same_hemisphere(wo, wi)) return mk(0.0f);
    float cto = cos_theta(wo), cti = cos_theta(wi);
    if (cti == 0 || cto == 0) return mk(0.0f);
    Mf m = gtr2_dist(b, b.trans_rough);
    bool entering = cos_theta(wo) > 0;
    float ei = entering ? 1.0f : b.eta, et = entering ? b.eta : 1.0f;
    float eta = et / ei;
    f3 wh = normalize(wo + wi * eta);
    if (cos_theta(wh) < 0) wh = -wh;
    if (dot(wo, wh) * dot(wi, wh) > 0) return mk(0.0f);
    f3 F = mk(fr_dielectric(dot(wo, wh), 1.0f, b.eta));
    float sd = dot(wo, wh) + eta * dot(wi, wh);
    float factor = imp ? 1.0f : 1 / eta;
    float D = mf_D(m, wh), G = gtr2_G(m, wo, wi);
    f3 T = (1 - b.metallic) * b.transmission * sqrt3(b.color);
    return (1 - F) * T * fabsf(D * G * eta * eta * abs_dot(wi, wh) * abs_dot(wo, wh) * factor * factor /
                              (cti * cto * sd * sd));
}
__device__ __forceinline__ float disney_btdf_pdf(const DisneyP& b, f3 wo, f3 wi) {
    if (same_hemisphere(wo, wi)) return 0.0f;
    Mf m = gtr2_dist(b, b.trans_rough);
    bool entering = cos_theta(wo) > 0;
    float ei = entering ? 1.0f : b.eta, et = entering ? b.eta : 1.0f;
    float eta = et / ei;
    f3 wh = normalize(wo + wi * eta);
    if (dot(wo, wh) * dot(wi, wh) > 0) return 0;
    float sd = dot(wo, wh) + eta * dot(wi, wh);
    float dwh = fabsf((eta * eta * dot(wi, wh)) / (sd * sd));
    return mf_pdf(m, wo, wh) * dwh;
}
__device__ __forceinline__ float gtr1(float c, float a) {
    float a2 = a * a;
    return (a2 - 1) / (kPi * logf(a2) * (1 + (a2 - 1) * c * c));
}
__device__ __forceinline__ float smith_g1_ggx(float c, float a) {
    float a2 = a * a, c2 = c * c;
    return 2 / (1 + sqrtf(1 + a2 * (1 - c2) / c2));
}
__device__ __forceinline__ f3 clearcoat_fr(const DisneyP& b, f3 wo, f3 wi) {   // DisneyClearcoatBrdf.Fr
    float gloss = lerpf(b.cc_gloss, 0.1f, 0.001f);
    f3 wh = wo + wi;
    if (wh.x == 0 && wh.y == 0 && wh.z == 0) return mk(0.0f);
    wh = normalize(wh);
    float D = gtr1(abs_cos_theta(wh), gloss);
    float F = schlick1(0.04f, dot(wo, wh));
    float G = smith_g1_ggx(abs_cos_theta(wo), 0.25f) * smith_g1_ggx(abs_cos_theta(wi), 0.25f);
    return mk(b.clearcoat * D * F * G / 4);
}
__device__ __forceinline__ float clearcoat_pdf(const DisneyP& b, f3 wo, f3 wi) {
    float gloss = lerpf(b.cc_gloss, 0.1f, 0.001f);
    if (!same_hemisphere(wo, wi)) return 0;
    f3 wh = wo + wi;
    if (wh.x == 0 && wh.y == 0 && wh.z == 0) return 0;
    wh = normalize(wh);
    return gtr1(abs_cos_theta(wh), gloss) * abs_cos_theta(wh) / (4 * dot(wo, wh));
}
__device__ __forceinline__ f3 disney_fr(const DisneyP& b, f3 wo, f3 wi, bool imp = false) {     // DisneyBsdf.FrImpl
    if (!same_hemisphere(wo, wi)) {
        if (b.transmission == 0) return mk(0.0f);
        return disney_btdf_fr(b, wo, wi, imp);
    }
    if (cos_theta(wo) < 0 && cos_theta(wi) < 0) return gtr2_reflect_fr(b, gtr2_dist(b, b.trans_rough), true, wo, wi);
    f3 diffuse = mk(0.0f);
    float dw = 1 - b.metallic;
    if (dw > 0) {
        float fo = schlick_weight(abs_cos_theta(wo)), fi = schlick_weight(abs_cos_theta(wi));
        f3 R = dw * b.color;
        f3 ddiff = mk(0.0f);
        f3 lambert = R * (1 / kPi) * (1 - fo / 2) * (1 - fi / 2);
        f3 wh = wo + wi;
        bool whz = wh.x == 0 && wh.y == 0 && wh.z == 0;
        f3 whn = whz ? wh : normalize(wh);
        if (b.thin) {
            f3 fake = mk(0.0f);
            if (!whz) {
                float cd = dot(wi, whn);
                float fss90 = cd * cd * b.roughness;
                float fss = lerpf(fo, 1.0f, fss90) * lerpf(fi, 1.0f, fss90);
                float ss = 1.25f * (fss * (1 / (abs_cos_theta(wo) + abs_cos_theta(wi)) - 0.5f) + 0.5f);
                fake = R * (1 / kPi) * ss;
            }
            ddiff = lambert * (1 - b.flatness) + fake * b.flatness;
        } else if (b.sdist_black) {
            ddiff = lambert;
        }
        f3 retro = mk(0.0f);
        if (!whz) {
            float cd = dot(wi, whn);
            float rr = 2 * b.roughness * cd * cd;
            retro = R * (1 / kPi) * rr * (fo + fi + fo * fi * (rr - 1));
        }
        f3 sheen = mk(0.0f);
        if (b.sheen > 0) {
            f3 scol = lerp3(b.sheen_tint, mk(1.0f), color_tint(b));
            if (!whz) sheen = (dw * b.sheen * scol) * schlick_weight(dot(wi, whn));
        }
        diffuse = (retro + sheen + ddiff) * (1 - b.transmission);
    }
    f3 spec = gtr2_reflect_fr(b, gtr2_dist(b, b.roughness), false, wo, wi);
    f3 cc = mk(0.0f);
    if (b.clearcoat > 0) cc = clearcoat_fr(b, wo, wi);
    return diffuse + spec + cc;
}
__device__ __forceinline__ void disney_weights(const DisneyP& b, float* dw, float* tw, float* sw, float* cw) {
    float a = nclamp(luminance(b.color) * (1 - b.metallic), 0.3f, 0.7f);
    float bb = 1 - a;
    *dw = a * (1 - b.transmission);
    *tw = a * b.transmission;
    *sw = bb * 2 / (2 + b.clearcoat);
    *cw = bb * b.clearcoat / (2 + b.clearcoat);
}
__device__ __forceinline__ float disney_pdf(const DisneyP& b, f3 wo, f3 wi) {  // DisneyBsdf.PdfImpl
    float dw, tw, sw, cw;
    disney_weights(b, &dw, &tw, &sw, &cw);
    if (cos_theta(wo) < 0) {
        if (b.transmission == 0) return 0;
        float fr = nclamp(fr_dielectric(cos_theta(wo), 1, b.eta), 0.1f, 0.9f);
        if (cos_theta(wi) > 0) return disney_btdf_pdf(b, wo, wi) * (1 - fr);
        return gtr2_reflect_pdf(gtr2_dist(b, b.trans_rough), wo, wi) * fr;
    }
    if (cos_theta(wi) < 0) return disney_btdf_pdf(b, wo, wi) * tw;
    float dp = 0;
    if (b.thin || b.sdist_black) dp = same_hemisphere(wo, wi) ? abs_cos_theta(wi) / kPi : 0.0f;
    float sp = gtr2_reflect_pdf(gtr2_dist(b, b.roughness), wo, wi);
    float cp = clearcoat_pdf(b, wo, wi);
    return dp * dw + sp * sw + cp * cw;
}

// Material.Fr (Diffuse.cs:21-27, RoughMetal.cs:33-58, RoughGlass.cs:151-195,
// RoughPlastic.cs:65-74, FresnelSpecularBsdf.cs:30, SpecularReflectionBrdf.cs:17, Disney.cs:87-90).
// imp: TransportMode.Importance (light paths): refraction without the radiance
// scaling (RoughGlass.cs:63, MicrofacetTransmissionBtdf.cs:40).
template <unsigned kM = kMatsAll>
__device__ __forceinline__ f3 mat_fr(const MatCtx& c, f3 wo, f3 wi, bool imp = false) {
    switch (c.type) {
    case kMatDiffuse:
        if constexpr (RD_MAT(kMatDiffuse)) {
            if (cos_theta(wo) <= 0 || cos_theta(wi) <= 0) { if (!c.two_sided) return mk(0.0f); }
            return same_hemisphere(wo, wi) ? c.r / kPi : mk(0.0f);
        }
        return mk(0.0f);
    case kMatRoughMetal:
        if constexpr (RD_MAT(kMatRoughMetal)) {
            if (cos_theta(wo) <= 0 || cos_theta(wi) <= 0) { if (!c.two_sided) return mk(0.0f); }
            return conductor_fr(c, wo, wi);
        }
        return mk(0.0f);
    case kMatRoughGlass: {
        if constexpr (RD_MAT(kMatRoughGlass)) {
            bool is_reflect = same_hemisphere(wo, wi);
            bool entering = cos_theta(wo) > 0;
            float ei = entering ? c.eta_a : c.eta_b;
            float et = entering ? c.eta_b : c.eta_a;
            float eta = is_reflect ? 1 : et / ei;
            f3 wh = normalize(wo + wi * eta);
            if (cos_theta(wh) < 0) wh = -wh;
            float sd = dot(wo, wh) + eta * dot(wi, wh);
            float D = mf_D(c.mf, wh);
            float G = mf_G(c.mf, wo, wi);
            float F = fr_dielectric(dot(wo, wh), ei, et);
            f3 bsdf;
            if (is_reflect) {
                bsdf = c.r * (D * G * F / (4.0f * cos_theta(wi) * cos_theta(wo)));
            } else {
                if (dot(wo, wh) * dot(wi, wh) > 0) return mk(0.0f);
                float factor = imp ? 1.0f : 1 / eta;
                bsdf = (1 - F) * c.t *
                       fabsf(D * G * eta * eta * abs_dot(wi, wh) * abs_dot(wo, wh) * factor * factor /
                             (cos_theta(wo) * cos_theta(wi) * sd * sd));
            }
            return valid_color(bsdf) ? bsdf : mk(0.0f);
        }
        return mk(0.0f);
    }
    case kMatRoughPlastic:
        if constexpr (RD_MAT(kMatRoughPlastic)) {
            return plastic_fr(c, wo, wi);
        }
        return mk(0.0f);
    case kMatDisney:
        if constexpr (RD_MAT(kMatDisney)) {
            return disney_fr(c.dp, wo, wi, imp);
        }
        return mk(0.0f);
    default:   // perfect glass / mirror: delta lobes
        return mk(0.0f);
    }
}

// Material.Pdf
template <unsigned kM = kMatsAll>
__device__ __forceinline__ float mat_pdf(const MatCtx& c, f3 wo, f3 wi) {
    switch (c.type) {
    case kMatDiffuse:
        if constexpr (RD_MAT(kMatDiffuse)) {
            if (cos_theta(wo) <= 0 || cos_theta(wi) <= 0) { if (!c.two_sided) return 0.0f; }
            return same_hemisphere(wo, wi) ? abs_cos_theta(wi) / kPi : 0.0f;
        }
        return 0.0f;
    case kMatRoughMetal: {
        if constexpr (RD_MAT(kMatRoughMetal)) {
            if (cos_theta(wo) <= 0 || cos_theta(wi) <= 0) { if (!c.two_sided) return 0.0f; }
            return mf_reflect_pdf(c.mf, wo, wi);
        }
        return 0.0f;
    }
    case kMatRoughGlass: {
        if constexpr (RD_MAT(kMatRoughGlass)) {
            bool is_reflect = same_hemisphere(wo, wi);
            bool entering = cos_theta(wo) > 0;
            float ei = entering ? c.eta_a : c.eta_b;
            float et = entering ? c.eta_b : c.eta_a;
            float pdf;
            if (is_reflect) {
                f3 wh = normalize(wo + wi);
                if (cos_theta(wh) < 0) wh = -wh;
                if (dot(wo, wh) * cos_theta(wo) <= 0 || dot(wi, wh) * cos_theta(wi) <= 0) return 0.0f;
                float jac = 1.0f / (4.0f * abs_dot(wh, wi));
                float F = fr_dielectric(dot(wo, wh), ei, et);
                pdf = mf_pdf(c.mf, wo, wh) * jac * F;
            } else {
                float eta = et / ei;
                f3 wh = normalize(wo + wi * eta);
                if (cos_theta(wh) < 0) wh = -wh;
                if (dot(wo, wh) * dot(wi, wh) > 0) return 0.0f;
                if (dot(wo, wh) * cos_theta(wo) <= 0 || dot(wi, wh) * cos_theta(wi) <= 0) return 0.0f;
                float sd = dot(wo, wh) + eta * dot(wi, wh);
                float dwh = fabsf((eta * eta * dot(wi, wh)) / (sd * sd));
                float F = fr_dielectric(dot(wo, wh), ei, et);
                pdf = mf_pdf(c.mf, wo, wh) * dwh * (1 - F);
            }
            return isinf(pdf) || isnan(pdf) ? 0.0f : pdf;
        }
        return 0.0f;
    }
    case kMatRoughPlastic:
        if constexpr (RD_MAT(kMatRoughPlastic)) {
            return plastic_pdf(c, wo, wi);
        }
        return 0.0f;
    case kMatDisney:
        if constexpr (RD_MAT(kMatDisney)) {
            return disney_pdf(c.dp, wo, wi);
        }
        return 0.0f;
    default:
        return 0.0f;
    }
}

struct BSample {
    f3 wi, fr;
    float pdf;
    int type;
};

// Cosine-hemisphere direction with LambertianReflectionBrdf.Sample's z flip
// (LambertianReflectionBrdf.cs:26-36, DisneyBsdf.DiffuseSample).
__device__ __forceinline__ bool cosine_dir(f3 wo, Rng& rng, f3* out) {
    float ux = rng.next_float(), uy = rng.next_float();
    float radius = sqrtf(ux);
    float angle = uy * kPi * 2;
    float x = radius * cosf(angle), y = radius * sinf(angle);
    f3 wi = normalize(mk(x, y, sqrtf(1 - x * x - y * y)));
    if (wo.z < 0) wi.z *= -1;
    if (!same_hemisphere(wo, wi)) return false;
    *out = wi;
    return true;
}
// Direction part of MicrofacetReflectionBrdf.Sample (MicrofacetReflectionBrdf.cs:50-59).
__device__ __forceinline__ bool mf_reflect_dir(const Mf& m, f3 wo, Rng& rng, f3* wi_out, f3* wh_out) {
    if (cos_theta(wo) == 0) return false;
    f3 wh = mf_sample_wh(m, wo, rng);
    if (dot(wo, wh) < 0) return false;
    f3 wi = reflect(-wo, wh);
    if (!same_hemisphere(wo, wi)) return false;
    *wi_out = wi;
    *wh_out = wh;
    return true;
}

__device__ __forceinline__ BSample disney_sample(const DisneyP& b, f3 wo, Rng& rng, bool imp = false) {  // DisneyBsdf.Sample
    BSample s;
    s.wi = mk(0.0f); s.fr = mk(0.0f); s.pdf = 0.0f; s.type = 0;
    Mf td = gtr2_dist(b, b.trans_rough);
    f3 wi, wh;
    int type = 0;
    bool ok = false;
    // MicrofacetTransmissionBtdf.Sample (MicrofacetTransmissionBtdf.cs:59-75)
    auto btdf = [&](f3* out) -> bool {
        if (wo.z == 0) return false;
        f3 h = mf_sample_wh(td, wo, rng);
        if (dot(wo, h) < 0) return false;
        bool entering = cos_theta(wo) > 0;
        float ei = entering ? 1.0f : b.eta, et = entering ? b.eta : 1.0f;
        f3 w;
        if (!refract(wo, h, ei / et, &w)) return false;
        f3 f = disney_btdf_fr(b, wo, w, imp);
        float p = disney_btdf_pdf(b, wo, w);
        if (!valid_color(f) || isinf(p) || p == 0) return false;
        *out = w;
        return true;
    };
    if (cos_theta(wo) < 0) {
        if (b.transmission == 0) return s;
        float fr = nclamp(fr_dielectric(cos_theta(wo), 1, b.eta), 0.1f, 0.9f);
        if (rng.next_float() >= fr) {
            ok = btdf(&wi);
            type = kGlossy | kTransmission;
        } else {
            ok = mf_reflect_dir(td, wo, rng, &wi, &wh) && (mf_pdf(td, wo, wh) / (4 * dot(wi, wh)) != 0);
            type = kReflection | kGlossy;
        }
    } else {
        float dw, tw, sw, cw;
        disney_weights(b, &dw, &tw, &sw, &cw);
        float u = rng.next_float();
        if (u < dw) {
            if (b.thin || b.sdist_black) {
                ok = cosine_dir(wo, rng, &wi) && ((same_hemisphere(wo, wi) ? abs_cos_theta(wi) / kPi : 0.0f) != 0);
                type = kReflection | kDiffuse;
            } else {
                // the BSSRDF's entry: SpecularTransmissionBtdf(1, 1, Eta).Sample
                // (SpecularTransmissionBtdf.cs:33-48), pdf times the diffuse
                // weight, flagged (DisneyBsdf.cs:357-364); no Fr / Pdf call
                const bool entering = cos_theta(wo) > 0;
                const float ei = entering ? 1.0f : b.eta, et = entering ? b.eta : 1.0f;
                const f3 n = entering ? mk(0, 0, 1) : mk(0, 0, -1);
                if (!refract(wo, n, ei / et, &wi) || cos_theta(wi) == 0.0f) return s;
                f3 ft = mk(1.0f) * (1.0f - fr_dielectric(cos_theta(wo), 1, b.eta));
                if (!imp) ft = ft * ((ei * ei) / (et * et));
                s.wi = wi;
                s.fr = ft / abs_cos_theta(wi);
                s.pdf = 1.0f * dw;
                s.type = kTransmission | kSpecular | kSubsurface;
                return s;
            }
        } else {
            u -= dw;
            if (u < tw) {
                ok = btdf(&wi);
                type = kGlossy | kTransmission;
            } else {
                u -= tw;
                if (u < sw) {
                    Mf sd = gtr2_dist(b, b.roughness);
                    ok = mf_reflect_dir(sd, wo, rng, &wi, &wh) && (mf_pdf(sd, wo, wh) / (4 * dot(wi, wh)) != 0);
                    type = kReflection | kGlossy;
                } else if (cos_theta(wo) != 0) {
                    // DisneyClearcoatBrdf.Sample
                    float gloss = lerpf(b.cc_gloss, 0.1f, 0.001f);
                    float ux = rng.next_float(), uy = rng.next_float();
                    float a2 = sqr(gloss);
                    float ct = sqrtf(nmax(0, (1 - powf(a2, 1 - ux)) / (1 - a2)));
                    float st = sqrtf(nmax(0, 1 - ct * ct));
                    float phi = 2 * kPi * uy;
                    wh = spherical_dir(st, ct, phi);
                    if (!same_hemisphere(wo, wh)) wh = -wh;
                    wi = reflect(-wo, wh);
                    ok = same_hemisphere(wo, wi) && clearcoat_pdf(b, wo, wi) != 0;
                    type = kReflection | kGlossy;
                }
            }
        }
    }
    if (!ok) return s;
    s.wi = wi;
    s.fr = disney_fr(b, wo, wi, imp);
    s.pdf = disney_pdf(b, wo, wi);
    s.type = type;
    return s;
}

// Material.Sample (Diffuse.cs:39-49, RoughMetal.cs:87-112, RoughGlass.cs:240-297,
// RoughPlastic.cs:112-166, FresnelSpecularBsdf.cs:34-65, PerfectMirror.cs:29-39, Disney.cs:97-100);
// imp as for mat_fr (RoughGlass.cs:161, FresnelSpecularBsdf.cs:58-61).
template <unsigned kM = kMatsAll>
__device__ __forceinline__ BSample mat_sample(const MatCtx& c, f3 wo, Rng& rng, bool imp = false) {
    BSample s;
    s.wi = mk(0.0f); s.fr = mk(0.0f); s.pdf = 0.0f; s.type = 0;
    switch (c.type) {
    case kMatDiffuse: {
        if constexpr (RD_MAT(kMatDiffuse)) {
            if (cos_theta(wo) <= 0) { if (!c.two_sided) return s; }
            f3 wi;
            if (!cosine_dir(wo, rng, &wi)) return s;
            s.wi = wi;
            s.pdf = same_hemisphere(wo, wi) ? abs_cos_theta(wi) / kPi : 0.0f;
            s.fr = c.r / kPi;
            s.type = kReflection | kDiffuse;
            return s;
        }
        return s;
    }
    case kMatRoughMetal: {
        if constexpr (RD_MAT(kMatRoughMetal)) {
            if (cos_theta(wo) <= 0) { if (!c.two_sided) return s; }
            f3 wi, wh;
            if (!mf_reflect_dir(c.mf, wo, rng, &wi, &wh)) return s;
            s.pdf = mf_pdf(c.mf, wo, wh) / (4 * dot(wi, wh));
            s.fr = conductor_fr(c, wo, wi);
            s.wi = wi;
            s.type = kReflection | kGlossy;
            return s;
        }
        return s;
    }
    case kMatRoughGlass: {
        if constexpr (RD_MAT(kMatRoughGlass)) {
            f3 wh = mf_sample_wh(c.mf, wo, rng);
            float F = fr_dielectric(dot(wo, wh), c.eta_a, c.eta_b);
            if (rng.next_float() < F) {
                f3 wi = reflect(-wo, wh);
                if (!same_hemisphere(wo, wi)) return s;
                float D = mf_D(c.mf, wh);
                float G = mf_G(c.mf, wo, wi);
                f3 brdf = c.r * (D * G * F / (4.0f * cos_theta(wi) * cos_theta(wo)));
                float jac = 1.0f / (4.0f * abs_dot(wh, wi));
                float pdf = mf_pdf(c.mf, wo, wh) * jac * F;
                if (isinf(pdf) || isnan(pdf) || !valid_color(brdf)) return s;
                s.wi = wi; s.fr = brdf; s.pdf = pdf; s.type = kReflection | kGlossy;
                return s;
            }
            bool entering = cos_theta(wo) > 0;
            float ei = entering ? c.eta_a : c.eta_b;
            float et = entering ? c.eta_b : c.eta_a;
            float eta = et / ei;
            f3 wi;
            if (refract(wo, wh, ei / et, &wi)) {
                if (dot(wo, wh) * dot(wi, wh) > 0) return s;
            } else {
                wi = reflect(-wo, wh);
                eta = 1;
            }
            float sd = dot(wo, wh) + eta * dot(wi, wh);
            float factor = imp ? 1.0f : 1 / eta;
            float D = mf_D(c.mf, wh);
            float G = mf_G(c.mf, wo, wi);
            f3 btdf = (1 - F) * c.t *
                      fabsf(D * G * eta * eta * abs_dot(wi, wh) * abs_dot(wo, wh) * factor * factor /
                            (cos_theta(wo) * cos_theta(wi) * sd * sd));
            float dwh = fabsf((eta * eta * dot(wi, wh)) / (sd * sd));
            float pdf = mf_pdf(c.mf, wo, wh) * dwh * (1 - F);
            if (isinf(pdf) || isnan(pdf) || !valid_color(btdf)) return s;
            s.wi = wi; s.fr = btdf; s.pdf = pdf; s.type = kTransmission | kGlossy;
            return s;
        }
        return s;
    }
    case kMatRoughPlastic: {
        if constexpr (RD_MAT(kMatRoughPlastic)) {
            if (cos_theta(wo) <= 0) { if (!c.two_sided) return s; }
            float p = c.ks / (c.kd + c.ks);
            f3 gwi = mk(0.0f), gwh, dwi = mk(0.0f);
            int gty = 0, dty = 0;
            if (mf_reflect_dir(c.mf, wo, rng, &gwi, &gwh)) gty = kReflection | kGlossy;
            else gwi = mk(0.0f);
            if (cosine_dir(wo, rng, &dwi)) dty = kReflection | kDiffuse;
            else dwi = mk(0.0f);
            bool glossy = rng.next_float() < p;
            f3 wi = glossy ? gwi : dwi;
            s.wi = wi;
            s.type = glossy ? gty : dty;
            s.pdf = plastic_pdf(c, wo, wi);
            s.fr = plastic_fr(c, wo, wi);
            return s;
        }
        return s;
    }
    case kMatSubsurface:   // Subsurface.Sample (Subsurface.cs:45-55): the glass's sample, flagged
    case kMatPerfectGlass: {
        if constexpr (RD_MAT(kMatPerfectGlass) || RD_MAT(kMatSubsurface)) {
            const int sss = (c.type == kMatSubsurface && c.sss) ? kSubsurface : 0;
            float f = fr_dielectric(cos_theta(wo), c.eta_a, c.eta_b);
            if (rng.next_float() < f) {
                if (cos_theta(wo) == 0) return s;
                f3 wi = mk(-wo.x, -wo.y, wo.z);
                s.wi = wi; s.fr = f * c.r / abs_cos_theta(wi); s.pdf = f; s.type = kReflection | kSpecular | sss;
                return s;
            }
            bool entering = cos_theta(wo) > 0;
            float ei = entering ? c.eta_a : c.eta_b, et = entering ? c.eta_b : c.eta_a;
            f3 n = entering ? mk(0, 0, 1) : mk(0, 0, -1);
            f3 wi;
            if (!refract(wo, n, ei / et, &wi) || cos_theta(wi) == 0.0f) return s;
            f3 ft = c.t * (1.0f - f);
            if (!imp) ft = ft * ((ei * ei) / (et * et));
            ft = ft / abs_cos_theta(wi);
            s.wi = wi; s.fr = ft; s.pdf = 1 - f; s.type = kTransmission | kSpecular | sss;
            return s;
        }
        return s;
    }
    case kMatPerfectMirror: {
        if constexpr (RD_MAT(kMatPerfectMirror)) {
            if (cos_theta(wo) <= 0) { if (!c.two_sided) return s; }
            if (abs_cos_theta(wo) == 0) return s;
            f3 wi = mk(-wo.x, -wo.y, wo.z);
            s.wi = wi; s.fr = c.r / abs_cos_theta(wi); s.pdf = 1.0f; s.type = kReflection | kSpecular;
            return s;
        }
        return s;
    }
    case kMatDisney:
        if constexpr (RD_MAT(kMatDisney)) {
            return disney_sample(c.dp, wo, rng, imp);
        }
        return s;
    default:
        return s;
    }
}

// ---------------------------------------------------------------- area lights
// ---------------------------------------------------------------- environment light
// Vector3.TransformNormal with a 3x3 row-major matrix: x*M11 + y*M21 + z*M31, ...
__device__ __forceinline__ f3 xform_normal3(const float* M, f3 v) {
    return mk(v.x * M[0] + v.y * M[3] + v.z * M[6],
              v.x * M[1] + v.y * M[4] + v.z * M[7],
              v.x * M[2] + v.y * M[5] + v.z * M[8]);
}

// Distribution1D.FindInterval (Distribution1D.cs:52-71) over cdf[0..n].
__device__ __forceinline__ int dist_find(const float* cdf, int n, float u) {
    int first = 0, len = n + 1;
    while (len > 0) {
        int half = len >> 1, middle = first + half;
        if (cdf[middle] <= u) { first = middle + 1; len -= half + 1; }
        else len = half;
    }
    return iclamp(first - 1, 0, n - 1);
}

// Equirectangular (u, v) of a world direction: phi = atan2(x, -z) in [0, 2 pi),
// theta = acos(y) (InfiniteAreaLight.Le / PdfDir, :56-85).
__device__ __forceinline__ void env_uv(const DevEnv& e, f3 d, float* u, float* v, float* theta_out) {
    f3 dir = normalize(xform_normal3(e.w2m, d));
    float phi = atan2f(dir.x, -dir.z);
    phi = (phi < 0) ? (phi + 2 * kPi) : phi;
    float theta = acosf(nclamp(dir.y, -1.0f, 1.0f));
    *u = phi / (2 * kPi);
    *v = theta / kPi;
    *theta_out = theta;
}

// InfiniteAreaLight.Le (:56-69)
__device__ __forceinline__ f3 env_le(const DevScene& sc, const DevEnv& e, f3 d) {
    float u, v, th;
    env_uv(e, d, &u, &v, &th);
    return tex_sample(sc, e.tex, u, v);
}

// Scene.EvalAllInfiniteLights (Scene.cs:37-45)
__device__ __forceinline__ f3 eval_infinite(const DevScene& sc, f3 d) {
    f3 l = mk(0.0f);
    for (int i = 0; i < sc.num_envs; i++) l = l + env_le(sc, sc.envs[i], d);
    return l;
}

// InfiniteAreaLight.PdfDir (:71-85) with Distribution2D.ContinuousPdf (Distribution2D.cs:55-60)
__device__ __forceinline__ float env_pdf_dir(const DevEnv& e, f3 wi) {
    float u, v, theta;
    env_uv(e, wi, &u, &v, &theta);
    float sin_theta = sinf(theta);
    if (sin_theta == 0.0f) return 0.0f;
    int iu = iclamp((int)(u * e.w), 0, e.w - 1);
    int iv = iclamp((int)(v * e.h), 0, e.h - 1);
    float pdf = e.cond_func[(size_t)iv * e.w + iu] / e.marg_int;
    return pdf / (2 * kPi * kPi * sin_theta);
}

struct LSample {
    f3 p, wi, li;
    float pdf;
};

// InfiniteAreaLight.SampleLi (:111-119) via SampleLightDir (:92-109): the
// Distribution2D picks a cell (marginal on u.Y, conditional on u.X); the
// direction is the cell's corner, the pdf the cell's continuous density over
// 2 pi^2 sin(theta); the shadow target lies 20 world radii out.
// SampleLightDir (:92-109) from two uniforms: false when the pdf is 0, else
// the world direction, its pdf and the cell corner (x, y) the map is read at.
__device__ __forceinline__ bool env_sample_dir(const DevEnv& e, float u0, float u1, f3* wdir_out, float* pdf_out,
                                               float* x_out, float* y_out) {
    int vj = dist_find(e.marg_cdf, e.h, u1);
    float pdf1 = e.marg_int > 0 ? e.marg_func[vj] / e.marg_int : 0;
    const float rint = e.marg_func[vj];
    int vi = dist_find(e.cond_cdf + (size_t)vj * (e.w + 1), e.w, u0);
    float pdf0 = rint > 0 ? e.cond_func[(size_t)vj * e.w + vi] / rint : 0;
    float pdf = pdf0 * pdf1;
    if (pdf <= 0) return false;
    float y = vj / (float)e.h;
    float x = vi / (float)e.w;
    float theta = y * kPi;
    float phi = x * 2 * kPi;
    float sin_theta = sinf(theta);
    if (sin_theta == 0) return false;
    pdf /= (2 * kPi * kPi * sin_theta);
    float st = sinf(theta), ct = cosf(theta), sp = sinf(phi), cp = cosf(phi);   // MathExt.SphericalCoordinates
    f3 sph = mk(cp * st, sp * st, ct);
    *wdir_out = normalize(xform_normal3(e.m2w, mk(sph.y, sph.z, -sph.x)));
    *pdf_out = pdf;
    *x_out = x;
    *y_out = y;
    return true;
}
__device__ __forceinline__ LSample env_sample_li(const DevScene& sc, const DevEnv& e, f3 refp, Rng& rng) {
    LSample out;
    out.p = mk(0.0f); out.wi = mk(0.0f); out.li = mk(0.0f); out.pdf = 0.0f;
    float u0 = rng.next_float(), u1 = rng.next_float();
    f3 wdir;
    float pdf, x, y;
    if (!env_sample_dir(e, u0, u1, &wdir, &pdf, &x, &y)) return out;
    out.li = tex_sample(sc, e.tex, x, y);
    out.p = wdir * e.radius * 20.0f + mk(e.center[0], e.center[1], e.center[2]);
    out.wi = normalize(out.p - refp);
    out.pdf = pdf;
    return out;
}

// DiffuseAreaLight.SampleLi (Lights/DiffuseAreaLight.cs:48-68) with
// Triangle.Sample (Triangle.cs:184-212) / Sphere.Sample (Sphere.cs:87-102).
__device__ __forceinline__ void area_shape_sample(const DevScene& sc, const DevLight& L, float ux, float uy, f3* P_out,
                                                  f3* N_out) {
    const DevPrim pr = sc.prim_info[L.prim];
    f3 P, N;
    if (pr.kind == kTri) {
        int i0 = sc.tri_idx[3 * pr.index], i1 = sc.tri_idx[3 * pr.index + 1], i2 = sc.tri_idx[3 * pr.index + 2];
        float alpha = 1 - sqrtf(1 - ux);
        float beta = uy * sqrtf(1 - ux);
        float gamma = (1 - alpha - beta);
        f3 A = load_f3(sc.pos, i0), B = load_f3(sc.pos, i1), C = load_f3(sc.pos, i2);
        P = alpha * A + beta * B + gamma * C;
        if (pr.flags & 1) N = normalize(alpha * load_f3(sc.nrm, i0) + beta * load_f3(sc.nrm, i1) + gamma * load_f3(sc.nrm, i2));
        else N = normalize(cross(B - A, C - A));
    } else {
        float4 s = sc.spheres[pr.index];
        float phi = ux * kPi * 2;   // Probability.SquareToUniformSphere (Probability.cs:62-71)
        float theta = acosf(1 - 2 * uy);
        f3 v = mk(sinf(theta) * cosf(phi), sinf(theta) * sinf(phi), cosf(theta));
        N = normalize(v);
        P = N * s.w + mk(s.x, s.y, s.z);
    }
    *P_out = P;
    *N_out = N;
}
__device__ __forceinline__ LSample light_sample(const DevScene& sc, const DevLight& L, f3 refp, Rng& rng) {
    if (L.type == 1) return env_sample_li(sc, sc.envs[L.tex], refp, rng);
    float ux = rng.next_float(), uy = rng.next_float();
    f3 P, N;
    area_shape_sample(sc, L, ux, uy, &P, &N);
    float spdf = 1.0f / L.area;
    LSample out;
    f3 wi = P - refp;
    f3 le = mk(L.le[0], L.le[1], L.le[2]);
    out.li = dot(N, -wi) > 0 ? le : mk(0.0f);
    if (length_sq(wi) == 0) {
        out.pdf = 0.0f;
    } else {
        wi = normalize(wi);
        out.pdf = spdf * length_sq(refp - P) / abs_dot(N, -wi);
        if (isinf(out.pdf)) out.pdf = 0.0f;
    }
    out.p = P;
    out.wi = wi;
    return out;
}

// DiffuseAreaLight.PdfLi (DiffuseAreaLight.cs:31-46): tests the emitter alone.
__device__ __forceinline__ float light_pdf(const DevScene& sc, const DevLight& L, f3 refp, f3 wi) {
    if (L.type == 1) return env_pdf_dir(sc.envs[L.tex], wi);   // InfiniteAreaLight.PdfLi (:87-90)
    const DevPrim pr = sc.prim_info[L.prim];
    float tu, tv, tt;
    if (!prim_test(sc, pr, refp, wi, kFloatEps, kFloatMax, &tu, &tv, &tt)) return 0.0f;
    Inct h = make_inct(sc, pr, refp, wi, tu, tv, tt);
    float pdf = length_sq(refp - h.P) / (abs_dot(h.sh.Z, -wi) * L.area);
    if (isinf(pdf)) pdf = 0.0f;
    return pdf;
}

// ---------------------------------------------------------------- parameter derivatives
// For the adjoint (path-replay backpropagation, Vicini et al. 2021).  Diffuse
// and RoughPlastic values are linear in the albedo texture R per channel, so a
// term C carrying f(R) has dC_c/dR_c = C_c / R_c.  RoughPlastic's roughness rho
// scales the distribution's alphas as (ax, ay) = s (a1, a2) with s = rho^2
// (RoughnessToAlpha, Microfacet.cs:177-183); only the glossy lobe's D * G
// depends on it, so dC_c/drho = C_c * (glossy / f) * (s dln(DG)/ds) * 2 / rho.
// s * dln(D)/ds for wh (Microfacet.cs:46-59 Beckmann, :93-107 GGX).
__device__ __forceinline__ float mf_sdlog_D(const Mf& m, f3 wh) {
    float t2 = tan2_theta(wh);
    if (isinf(t2)) return 0.0f;
    float e = (cos2_phi(wh) / (m.ax * m.ax) + sin2_phi(wh) / (m.ay * m.ay)) * t2;
    if (m.kind == 0) return 2 * e - 2;
    return -2 * (1 - e) / (1 + e);
}
// s * dLambda/ds for w (Microfacet.cs:117-132).
__device__ __forceinline__ float mf_sdlambda(const Mf& m, f3 w) {
    float at = fabsf(tan_theta(w));
    if (isinf(at)) return 0.0f;
    float alpha = sqrtf(cos2_phi(w) * m.ax * m.ax + sin2_phi(w) * m.ay * m.ay);
    if (m.kind == 0) {
        float a = 1 / (alpha * at);
        if (a >= 1.6f) return 0.0f;
        float num = 1 - 1.259f * a + 0.396f * a * a, den = 3.535f * a + 2.181f * a * a;
        float dnum = -1.259f + 0.792f * a, dden = 3.535f + 4.362f * a;
        return -a * (dnum * den - num * dden) / (den * den);
    }
    float k = (alpha * at) * (alpha * at);
    return k / (2 * sqrtf(1 + k));
}
// RoughPlastic split at (wo, wi) with f = R * (d + g): the diffuse and glossy
// weights and s * dln(D G)/ds of the glossy lobe (zero where plastic_fr is).
struct PlasticParts {
    float d, g, sdlog;
};
__device__ __forceinline__ PlasticParts plastic_parts(const MatCtx& c, f3 wo, f3 wi) {
    PlasticParts p{0.0f, 0.0f, 0.0f};
    if (cos_theta(wo) <= 0 || cos_theta(wi) <= 0) { if (!c.two_sided) return p; }
    if (!same_hemisphere(wo, wi)) return p;
    p.d = c.kd / kPi;
    float cto = abs_cos_theta(wo), cti = abs_cos_theta(wi);
    f3 wh = wi + wo;
    if (!(cti == 0 || cto == 0) && !(wh.x == 0 && wh.y == 0 && wh.z == 0)) {
        wh = normalize(wh);
        float F = fr_dielectric(dot(wi, faceforward(wh, mk(0, 0, 1))), c.eta_a, c.eta_b);
        float D = mf_D(c.mf, wh), G = mf_G(c.mf, wo, wi);
        p.g = c.ks * fabsf((D * G * F) / (4 * cti * cto));
        p.sdlog = mf_sdlog_D(c.mf, wh) - G * (mf_sdlambda(c.mf, wo) + mf_sdlambda(c.mf, wi));
    }
    return p;
}
// dC/drho / C for a term whose BSDF factor is f(wo, wi) of a RoughPlastic with roughness rho.
__device__ __forceinline__ float plastic_rho_ratio(const MatCtx& c, float rho, f3 wo, f3 wi) {
    if (c.type != kMatRoughPlastic || !(rho > 0)) return 0.0f;
    PlasticParts p = plastic_parts(c, wo, wi);
    float f = p.d + p.g;
    if (!(f > 0)) return 0.0f;
    return (p.g / f) * p.sdlog * 2 / rho;
}

// The texels a tex_sample(u, v) reads and their weights: one for Nearest,
// four for Linear (constant textures: element 0).  Element index is
// x * h + y like the texel layout.
struct TexFoot {
    int64_t idx[4];
    float w[4];
    int n;
};
__device__ __forceinline__ TexFoot tex_footprint(const DevScene& sc, int id, float su, float sv) {
    TexFoot f;
    const DevTexture t = sc.texs[id];
    if (t.w == 0) { f.n = 1; f.idx[0] = 0; f.w[0] = 1.0f; return f; }
    float u, v;
    if (t.wrap == 0) { u = nclamp(su, 0.0f, 1.0f); v = nclamp(sv, 0.0f, 1.0f); }
    else { u = nclamp(su - floorf(su), 0.0f, 1.0f); v = nclamp(sv - floorf(sv), 0.0f, 1.0f); }
    float fu = u * (t.w - 1), fv = v * (t.h - 1);
    int pu = (int)fu, pv = (int)fv;
    if (t.filter == 0) { f.n = 1; f.idx[0] = (int64_t)pu * t.h + pv; f.w[0] = 1.0f; return f; }
    int dpu = (fu > pu + 0.5f) ? 1 : -1;
    int dpv = (fv > pv + 0.5f) ? 1 : -1;
    int apu = iclamp(pu + dpu, 0, t.w - 1);
    int apv = iclamp(pv + dpv, 0, t.h - 1);
    float du = nmin(fabsf(fu - pu - 0.5f), 1.0f);
    float dv = nmin(fabsf(fv - pv - 0.5f), 1.0f);
    f.n = 4;
    f.idx[0] = (int64_t)pu * t.h + pv;   f.w[0] = (1 - du) * (1 - dv);
    f.idx[1] = (int64_t)apu * t.h + pv;  f.w[1] = du * (1 - dv);
    f.idx[2] = (int64_t)pu * t.h + apv;  f.w[2] = (1 - du) * dv;
    f.idx[3] = (int64_t)apu * t.h + apv; f.w[3] = du * dv;
    return f;
}

// ---------------------------------------------------------------- media
// Homogeneous.Tr (Mediums/Homogeneous.cs:22-30): exp(sigma_t * (MinT - MaxT)).
__device__ __forceinline__ f3 medium_tr(const DevMedium& m, float tmin, float tmax) {
    float neg = tmin - tmax;
    return mk(m.st[0] == 0 ? 1.0f : expf(m.st[0] * neg),
              m.st[1] == 0 ? 1.0f : expf(m.st[1] * neg),
              m.st[2] == 0 ? 1.0f : expf(m.st[2] * neg));
}

struct MSample {
    bool medium;
    f3 p, tr;
    float t;
};
// Homogeneous.Sample (:32-61): free flight on a uniformly chosen channel;
// a medium vertex carries tr * sigma_s / pdf, a pass tr / pdf.
__device__ __forceinline__ MSample medium_sample(const DevMedium& m, f3 o, f3 d, float tmin, float tmax, Rng& rng) {
    MSample out;
    int ch = rng.next_int(3);
    float dens = ch == 0 ? m.st[0] : (ch == 1 ? m.st[1] : m.st[2]);
    float sd = -logf(1 - rng.next_float()) / dens;
    float dist = tmax - tmin;
    if (sd < dist) {
        float t = tmin + sd;
        f3 tr = mk(expf(-m.st[0] * sd), expf(-m.st[1] * sd), expf(-m.st[2] * sd));
        float pdf = (tr.x * m.st[0] + tr.y * m.st[1] + tr.z * m.st[2]) / 3;
        if (pdf == 0) pdf = 0.0001f;
        out.medium = true;
        out.p = o + t * d;
        out.tr = tr * mk(m.ss[0], m.ss[1], m.ss[2]) / pdf;
        out.t = t;
    } else {
        f3 tr = mk(expf(-m.st[0] * dist), expf(-m.st[1] * dist), expf(-m.st[2] * dist));
        float pdf = (tr.x + tr.y + tr.z) / 3;
        if (pdf == 0) pdf = 0.0001f;
        out.medium = false;
        out.p = mk(0.0f);
        out.tr = tr / pdf;
        out.t = 0.0f;
    }
    return out;
}

// PhaseFunctionUtility.HenyeyGreenstein (IPhaseFunction.cs:39-44)
__device__ __forceinline__ float hg_eval(float cos_t, float g) {
    float denom = 1 + g * g + 2 * g * cos_t;
    return (1 / (4 * kPi)) * (1 - g * g) / (denom * sqrtf(denom));
}
__device__ __forceinline__ float hg_g(const DevMedium& m) { return nclamp(m.g, -0.999f, 0.999f); }   // :51-54

struct PSample {
    float p;
    f3 wi;
};
// HenyeyGreenstein.SampleWi (:62-81), directions around wo's Duff frame.
__device__ __forceinline__ PSample hg_sample(const DevMedium& m, f3 wo, Rng& rng) {
    float G = hg_g(m);
    float x = rng.next_float(), y = rng.next_float();
    float cos_t;
    if (fabsf(G) < 0.001f) {
        cos_t = 1 - 2 * x;
    } else {
        float sqr = (1 - G * G) / (1 + G - 2 * G * x);
        cos_t = -(1 + G * G - sqr * sqr) / (2 * G);
    }
    float sin_t = sqrtf(nmax(0.0f, 1 - cos_t * cos_t));
    float phi = 2 * kPi * y;
    Frame f = make_frame(wo);
    float sp = sinf(phi), cp = cosf(phi);
    PSample out;
    out.wi = sin_t * cp * f.X + sin_t * sp * f.Y + cos_t * f.Z;   // MathExt.SphericalDirection
    out.p = hg_eval(cos_t, G);
    return out;
}

// Intersection.GetMedium (Intersection.cs:68-71): the side of the shading normal.
__device__ __forceinline__ int get_medium(const DevPrim& pr, f3 n, f3 w) {
    return dot(w, n) > 0 ? pr.med_out : pr.med_in;
}

// PathUtility.PowerHeuristic (Integrators/PathUtility.cs:5-9), in double.
__device__ __forceinline__ float power_heuristic(float fp, float gp) {
    double f = (double)fp, g = (double)gp;
    return (float)((f * f) / (f * f + g * g));
}

}  // namespace rd
