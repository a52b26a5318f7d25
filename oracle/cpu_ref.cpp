// cpu_ref.cpp — TEST INFRASTRUCTURE: the oracle.  A scalar C++ restatement
// of the reference's path / vol_path hot path (ksgfk/Radiantium, C#/.NET 6),
// used only as the checker in tests/, by __graft_entry__.smoke() and as the
// CPU baseline in bench.py.  The product never links or calls it.
//
// Each function cites the reference file:line it restates (paths relative to
// the reference root).  Build: -O2 -ffp-contract=off, no fast-math, so float
// expressions round like .NET's (which does not contract to FMA).
#include "cpu_ref.h"
#include "cpu_ref_math.h"

#include <array>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

using namespace cref;

static thread_local std::string g_err;
static void set_err(const std::string& s) { g_err = s; }

// ---------------------------------------------------------------- AABB / ray
// BoundingBox3F (Radiantium.Core/BoundingBox3F.cs)
struct BBox {
    V3 mn, mx;
    BBox() : mn(FLT_MAX), mx(-FLT_MAX) {}                       // :37-41
    explicit BBox(V3 p) : mn(p), mx(p) {}                       // :43
    void expand(V3 p) { mn = vmin(mn, p); mx = vmax(mx, p); }   // ExpendBy :101-105
    void unite(const BBox& b) { mn = vmin(mn, b.mn); mx = vmax(mx, b.mx); }  // Union :107-111
    V3 diag() const { return mx - mn; }
    float area() const {                                        // SurfaceArea :24-31
        V3 d = diag();
        return 2 * (d.x * d.y + d.x * d.z + d.y * d.z);
    }
    int max_extent() const {                                    // MaximumExtent :216-225
        V3 d = diag();
        if (d.x > d.y && d.x > d.z) return 0;
        else if (d.y > d.z) return 1;
        return 2;
    }
    V3 offset(V3 p) const {                                     // Offset :140-147
        V3 o = p - mn;
        if (mx.x > mn.x) o.x /= mx.x - mn.x;
        if (mx.y > mn.y) o.y /= mx.y - mn.y;
        if (mx.z > mn.z) o.z /= mx.z - mn.z;
        return o;
    }
    V3 center() const { return (mx + mn) * 0.5f; }
};

// Ray3F (Radiantium.Core/Ray3F.cs:5-20)
struct Ray {
    V3 o, d;
    float tmin, tmax;
    V3 inv;
    Ray() : tmin(0), tmax(0) {}
    Ray(V3 o_, V3 d_, float mn = EPS_F, float mxv = FLT_MAX) : o(o_), d(d_), tmin(mn), tmax(mxv) {
        inv = V3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    }
    V3 at(float t) const { return o + t * d; }
};

// BoundingBox3F.Intersect(Ray3F) (BoundingBox3F.cs:155-176).  The factor
// 1+2*Gamma(3) is exactly 1.0f because Gamma uses float.Epsilon (:149-153),
// so it is omitted.
static inline bool box_hit(const BBox& b, const Ray& r) {
    bool nx = r.inv.x < 0, ny = r.inv.y < 0, nz = r.inv.z < 0;
    float tmin = ((nx ? b.mx : b.mn).x - r.o.x) * r.inv.x;
    float tmax = ((nx ? b.mn : b.mx).x - r.o.x) * r.inv.x;
    float tymin = ((ny ? b.mx : b.mn).y - r.o.y) * r.inv.y;
    float tymax = ((ny ? b.mn : b.mx).y - r.o.y) * r.inv.y;
    if (tmin > tymax || tymin > tmax) return false;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    float tzmin = ((nz ? b.mx : b.mn).z - r.o.z) * r.inv.z;
    float tzmax = ((nz ? b.mn : b.mx).z - r.o.z) * r.inv.z;
    if (tmin > tzmax || tzmin > tmax) return false;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    return (tmin < r.tmax) && (tmax > 0);
}

// Coordinate (Radiantium.Offline/Coordinate.cs:14-36): Duff et al. ONB.
struct Frame {
    V3 X, Y, Z;
    Frame() {}
    explicit Frame(V3 n) {
        float sign = std::copysign(1.0f, n.z);
        float a = -1 / (sign + n.z);
        float b = n.x * n.y * a;
        X = V3(mulsign(sqr(n.x) * a, n.z) + 1.0f, mulsign(b, n.z), mulsign(n.x, -n.z));
        Y = V3(b, sign + sqr(n.y) * a, -n.y);
        Z = n;
    }
    V3 to_local(V3 v) const { return V3(dot(v, X), dot(v, Y), dot(v, Z)); }   // :70-73
    V3 to_world(V3 v) const { return X * v.x + Y * v.y + Z * v.z; }            // :75-78
};
// Coordinate static helpers (:80-122).  MathExt.Fma is a*b+c (not fused).
static inline float cos_theta(V3 v) { return v.z; }
static inline float abs_cos_theta(V3 v) { return std::fabs(v.z); }
static inline float cos2_theta(V3 v) { return sqr(v.z); }
static inline float sin2_theta(V3 v) { return v.x * v.x + sqr(v.y); }
static inline float tan_theta(V3 v) { return safe_sqrt(-v.z * v.z + 1.0f) / v.z; }
static inline float tan2_theta(V3 v) { return nmax(-v.z * v.z + 1.0f, 0.0f) / sqr(v.z); }
static inline float cos2_phi(V3 v) {
    float s2 = sin2_theta(v);
    return std::fabs(s2) <= 4.0f * EPS_F ? 1.0f : nclamp(sqr(v.x) / s2, -1.0f, 1.0f);
}
static inline float sin2_phi(V3 v) {
    float s2 = sin2_theta(v);
    return std::fabs(s2) <= 4.0f * EPS_F ? 0.0f : nclamp(sqr(v.y) / s2, -1.0f, 1.0f);
}
static inline bool same_hemisphere(V3 a, V3 b) { return a.z * b.z > 0; }
static inline V3 faceforward(V3 v, V3 v2) { return dot(v, v2) < 0.0f ? -v : v; }

// Intersection (Radiantium.Offline/Intersection.cs:20-72)
struct Inct {
    V3 P;
    V2 UV;
    float T;
    int prim;   // primitive index (input order)
    Frame sh;
    V3 wr;
    V3 N() const { return sh.Z; }
};

// ---------------------------------------------------------------- scene data
struct Tex {
    int w = 0, h = 0, wrap = 0, filter = 0;
    C3 value;
    std::vector<float> data;
    // ColorBuffer.GetRGB (ColorBuffer.cs:44-60), 3 channels.
    C3 read(int x, int y) const {
        if (w == 0) return value;
        size_t i = ((size_t)x * h + y) * 3;
        return C3(data[i], data[i + 1], data[i + 2]);
    }
    // TextureSampler.Sample (TextureSampler.cs:28-72)
    C3 sample(V2 uv) const {
        if (w == 0) return value;   // ConstColorTexture2D.Read ignores (x, y)
        auto wrapf = [&](float a) {
            if (wrap == 0) return nclamp(a, 0.0f, 1.0f);
            return nclamp(a - std::floor(a), 0.0f, 1.0f);
        };
        float u = wrapf(uv.x), v = wrapf(uv.y);
        float fu = u * (w - 1), fv = v * (h - 1);
        int pu = (int)fu, pv = (int)fv;
        if (filter == 0) return read(pu, pv);
        int dpu = (fu > pu + 0.5f) ? 1 : -1;
        int dpv = (fv > pv + 0.5f) ? 1 : -1;
        int apu = iclamp(pu + dpu, 0, w - 1);
        int apv = iclamp(pv + dpv, 0, h - 1);
        float du = nmin(std::fabs(fu - pu - 0.5f), 1.0f);
        float dv = nmin(std::fabs(fv - pv - 0.5f), 1.0f);
        C3 u0v0 = read(pu, pv), u1v0 = read(apu, pv), u0v1 = read(pu, apv), u1v1 = read(apu, apv);
        return (u0v0 * (1 - du) + u1v0 * du) * (1 - dv) + (u0v1 * (1 - du) + u1v1 * du) * dv;
    }
};

struct Camera {
    // PerspectiveCamera (Radiantium.Offline/Camera.cs:101-191)
    float near_z, far_z;
    float inv_res_x, inv_res_y;
    M4 clip_to_camera, camera_to_world;
    M4 camera_to_clip, world_to_camera;
    float min_x, min_y, max_x, max_y, normalization;   // image plane at z = 1 (Camera.cs:172-179)
    int W, H;
    bool init(const rd_camera_desc& c) {
        W = c.width; H = c.height;
        near_z = c.near_z; far_z = c.far_z;
        M4 c2w;
        if (c.use_matrix) {
            for (int i = 0; i < 16; i++) c2w.m[i / 4][i % 4] = c.to_world[i];
        } else {
            // MathExt.LookAtLeftHand (MathExt.cs:149-161)
            V3 origin(c.origin[0], c.origin[1], c.origin[2]);
            V3 target(c.target[0], c.target[1], c.target[2]);
            V3 up(c.up[0], c.up[1], c.up[2]);
            V3 dir = normalize(target - origin);
            V3 left = normalize(cross(normalize(up), dir));
            V3 real_up = normalize(cross(dir, left));
            c2w = m4_identity();
            c2w.m[0][0] = left.x; c2w.m[0][1] = left.y; c2w.m[0][2] = left.z; c2w.m[0][3] = 0;
            c2w.m[1][0] = real_up.x; c2w.m[1][1] = real_up.y; c2w.m[1][2] = real_up.z; c2w.m[1][3] = 0;
            c2w.m[2][0] = dir.x; c2w.m[2][1] = dir.y; c2w.m[2][2] = dir.z; c2w.m[2][3] = 0;
            c2w.m[3][0] = origin.x; c2w.m[3][1] = origin.y; c2w.m[3][2] = origin.z; c2w.m[3][3] = 1;
        }
        // Update (Camera.cs:128-180)
        inv_res_x = 1.0f / (float)W;
        inv_res_y = 1.0f / (float)H;
        float aspect = W / (float)H;
        float recip = 1.0f / (far_z - near_z);
        float cot = 1.0f / std::tan(radian(c.fov / 2.0f));
        M4 persp{}; persp.m[0][0] = cot; persp.m[1][1] = cot; persp.m[2][2] = far_z * recip;
        persp.m[2][3] = 1; persp.m[3][2] = -near_z * far_z * recip; persp.m[3][3] = 0;
        M4 scale{}; scale.m[0][0] = -0.5f; scale.m[1][1] = -0.5f * aspect; scale.m[2][2] = 1; scale.m[3][3] = 1;
        M4 trans = m4_identity(); trans.m[3][0] = -1; trans.m[3][1] = -1 / aspect;
        M4 p = m4_mul(m4_mul(persp, trans), scale);
        if (!m4_invert(p, &clip_to_camera)) return false;
        camera_to_world = c2w;
        camera_to_clip = p;
        M4 w2c;
        if (!m4_invert(c2w, &w2c)) return false;
        world_to_camera = w2c;
        V3 pmin = xform_point(V3(0, 0, 0), clip_to_camera);
        V3 pmax = xform_point(V3(1, 1, 0), clip_to_camera);
        pmin = pmin / pmin.z;
        pmax = pmax / pmax.z;
        min_x = pmax.x; min_y = pmax.y;
        max_x = pmin.x; max_y = pmin.y;
        float area = std::fabs((max_x - min_x) * (max_y - min_y));
        normalization = 1 / area;
        return true;
    }
    // Importance (Camera.cs:193-204): d normalised, camera space
    float importance(V3 d) const {
        float ct = d.z;
        if (ct <= 0) return 0;
        float focus = 1 / ct;
        float px = d.x * focus, py = d.y * focus;
        if (px < min_x || px > max_x || py < min_y || py > max_y) return 0.0f;
        return normalization * focus * focus * focus;
    }
    // SampleWi (Camera.cs:221-239); ok = false for the default (pdf 0) result
    struct Wi { V3 pos, dir; float dist = 0, pdf = 0, we = 0, sx = 0, sy = 0; bool ok = false; };
    Wi sample_wi(V3 pos) const {
        Wi r;
        V3 lt = xform_point(pos, world_to_camera);
        if (lt.z < near_z || lt.z > far_z) return r;
        V3 screen = xform_point(lt, camera_to_clip);
        screen.x /= screen.z;
        screen.y /= screen.z;
        if (screen.x < 0 || screen.x > 1 || screen.y < 0 || screen.y > 1) return r;
        r.sx = screen.x * (float)(W - 1);
        r.sy = screen.y * (float)(H - 1);
        float dist = length(lt);
        float inv = 1 / dist;
        V3 local_dir = lt * inv;
        r.pos = xform_point(V3(0, 0, 0), camera_to_world);
        r.dir = (r.pos - pos) / dist;
        r.dist = dist;
        r.pdf = 1;
        r.we = importance(local_dir) * inv * inv;
        r.ok = true;
        return r;
    }
    // SampleRay (Camera.cs:182-191)
    Ray sample_ray(float sx, float sy) const {
        V3 ndc(sx * inv_res_x, sy * inv_res_y, 0.0f);
        V3 nearp = xform_point(ndc, clip_to_camera);
        V3 dir = normalize(nearp);
        float inv_z = 1.0f / dir.z;
        V3 o = xform_point(V3(0.0f), camera_to_world);
        V3 d = normalize(xform_vector(dir, camera_to_world));
        return Ray(o, d, near_z * inv_z, far_z * inv_z);
    }
};

struct Node {
    BBox b;
    int offset;          // PrimitivesIndex / SecondChildIndex
    uint16_t count;
    uint8_t axis;
};

// Distribution1D (Radiantium.Offline/Distribution1D.cs:19-111): pbrt's
// piecewise-constant 1D distribution with the reference's float operation order.
struct Dist1D {
    std::vector<float> func, cdf;
    float integ = 0;
    void build(const float* f, int n) {
        func.assign(f, f + n);
        cdf.assign(n + 1, 0.0f);
        for (int i = 1; i < n + 1; ++i) cdf[i] = cdf[i - 1] + func[i - 1] / n;
        integ = cdf[n];
        if (integ == 0) {
            for (int i = 1; i < n + 1; i++) cdf[i] = i / (float)n;
        } else {
            for (int i = 1; i < n + 1; i++) cdf[i] /= integ;
        }
    }
    int count() const { return (int)func.size(); }
    int find_interval(float u) const {                     // :52-71
        int first = 0, len = (int)cdf.size();
        while (len > 0) {
            int half = len >> 1, middle = first + half;
            if (cdf[middle] <= u) { first = middle + 1; len -= half + 1; }
            else len = half;
        }
        return iclamp(first - 1, 0, (int)cdf.size() - 2);
    }
    float continuous_pdf(int off) const { return integ > 0 ? func[off] / integ : 0; }   // :91-94
    float discrete_pdf(int i) const { return integ > 0 ? func[i] / (integ * count()) : 0; }  // :108-111
    float sample_continuous(float u, float* pdf, int* off) const {                     // :73-84
        *off = find_interval(u);
        float du = u - cdf[*off];
        if ((cdf[*off + 1] - cdf[*off]) > 0) du /= cdf[*off + 1] - cdf[*off];
        *pdf = continuous_pdf(*off);
        return (*off + du) / count();
    }
};

// Distribution2D (Radiantium.Offline/Distribution2D.cs:14-60).
struct Dist2D {
    std::vector<Dist1D> cond;
    Dist1D marg;
    void build(const float* data, int nu, int nv) {
        cond.resize(nv);
        std::vector<float> m(nv);
        for (int v = 0; v < nv; v++) { cond[v].build(data + (size_t)v * nu, nu); m[v] = cond[v].integ; }
        marg.build(m.data(), nv);
    }
    // SampleContinuous(u, out pdf, out (vj, vi)) (:33-40)
    V2 sample_continuous(float u0, float u1, float* pdf, int* vj, int* vi) const {
        float pdf1, pdf0;
        float d1 = marg.sample_continuous(u1, &pdf1, vj);
        float d0 = cond[*vj].sample_continuous(u0, &pdf0, vi);
        *pdf = pdf0 * pdf1;
        return V2(d0, d1);
    }
    float continuous_pdf(float u, float v) const {                                     // :55-60
        int iu = iclamp((int)(u * cond[0].count()), 0, cond[0].count() - 1);
        int iv = iclamp((int)(v * marg.count()), 0, marg.count() - 1);
        return cond[iv].func[iu] / marg.integ;
    }
};

// Homogeneous (Mediums/Homogeneous.cs:14-20): coefficients scaled at construction.
struct Homo {
    C3 sa, ss, st;
    float g;
};

// InfiniteAreaLight (Lights/InfiniteAreaLight.cs:21-119) after
// InfiniteLight.Preprocess (Light.cs:181-187).
struct EnvLight {
    int tex = -1, w = 1, h = 1;
    M4 m2w, w2m;
    Dist2D dist;
    V3 center;
    float radius = 0;
};

// An instanced mesh set (RendererBuilder's InstancedPrimitivesEntry,
// Config/RendererBuilder.cs:895-921): triangles [first, first+count) in the
// set's model space under the set's own Bvh.
struct InstSet {
    int first = 0, count = 0, max_prim = 1;
    std::vector<Node> nodes;
    std::vector<int> order;    // ordered position -> index within the set
    int leaf_count = 0, max_depth = 0;
};
// InstancedPrimitive's transform (Primitive.cs:94-108, 144-161)
struct Inst {
    int set = 0;
    M4 m2w, w2m;
};

struct cr_scene {
    std::vector<InstSet> sets;
    std::vector<Inst> insts;
    std::vector<V3> pos, nrm;
    std::vector<V2> uvs;
    std::vector<int32_t> idx;
    std::vector<uint8_t> tflags;
    std::vector<float> sph;
    std::vector<float> sph_w2m;         // 9 per sphere, empty = identity
    std::vector<rd_prim_desc> prims;
    std::vector<float> prim_area;
    std::vector<Tex> tex;
    std::vector<rd_material_desc> mats;
    std::vector<rd_light_desc> lights;
    std::vector<rd_medium_desc> media;
    std::vector<Homo> homo;
    std::vector<int> infinite;          // light index of each InfiniteLight, Scene.InfiniteLights order
    std::vector<int> env_of_light;      // light index -> EnvLight index (-1 for area lights)
    std::vector<EnvLight> env;
    int global_medium = -1;
    int max_prim = 1;
    std::vector<Node> nodes;
    std::vector<int> order;
    int leaf_count = 0;
    int max_depth = 0;
    int accel = RD_ACCEL_BVH;           // the scene aggregate: Bvh or Octree
    int oct_max_depth = 10, oct_max_count = 10;
    float oct_out_bound = 2.0f;
    struct OctLinear {                  // LinearNode (Octree.cs:28-35)
        BBox b;
        int count_or_leaf;              // NodeCount (interior) / LeafIndex (leaf)
        int next_head;                  // -1: leaf
    };
    std::vector<OctLinear> oct;
    std::vector<std::vector<int>> oct_leaves;
    int oct_depth = 0;                  // _nowDepth
    Camera cam;
    rd_integrator_desc integ;

    // Triangle accessors (Shapes/Triangle.cs:71-81)
    V3 pa(int t) const { return pos[idx[3 * t]]; }
    V3 pb(int t) const { return pos[idx[3 * t + 1]]; }
    V3 pc(int t) const { return pos[idx[3 * t + 2]]; }
    bool has_n(int t) const { return (tflags[t] & 1) != 0; }
    bool has_uv(int t) const { return (tflags[t] & 2) != 0; }
};

// ---------------------------------------------------------------- shapes
struct SurfPt { float u, v, t; };

// Triangle.Intersect(ray, out SurfacePoint) (Shapes/Triangle.cs:149-177)
static inline bool tri_hit(const cr_scene& s, int t, const Ray& r, SurfPt* sp) {
    V3 A = s.pa(t), B = s.pb(t), C = s.pc(t);
    V3 e1 = B - A, e2 = C - A;
    V3 pvec = cross(r.d, e2);
    float det = dot(e1, pvec);
    if (std::fabs(det) < EPS_F) return false;
    float inv_det = 1.0f / det;
    V3 tvec = r.o - A;
    float u = dot(tvec, pvec) * inv_det;
    if (u < 0.0f || u > 1.0f) return false;
    V3 qvec = cross(tvec, e1);
    float v = dot(r.d, qvec) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return false;
    float tt = dot(e2, qvec) * inv_det;
    bool hit = tt >= r.tmin && tt <= r.tmax;
    if (hit) { sp->u = u; sp->v = v; sp->t = tt; }
    return hit;
}

// Triangle.GetIntersection (Triangle.cs:96-120)
static Inct tri_inct(const cr_scene& s, int t, const Ray& r, const SurfPt& sp, int prim) {
    V3 bary(1 - (sp.u + sp.v), sp.u, sp.v);
    V3 A = s.pa(t), B = s.pb(t), C = s.pc(t);
    Inct it;
    it.P = bary.x * A + bary.y * B + bary.z * C;
    V3 n;
    if (s.has_n(t)) {
        n = normalize(bary.x * s.nrm[s.idx[3 * t]] + bary.y * s.nrm[s.idx[3 * t + 1]] + bary.z * s.nrm[s.idx[3 * t + 2]]);
    } else {
        n = normalize(cross(B - A, C - A));
    }
    if (s.has_uv(t)) {
        it.UV = bary.x * s.uvs[s.idx[3 * t]] + bary.y * s.uvs[s.idx[3 * t + 1]] + bary.z * s.uvs[s.idx[3 * t + 2]];
    } else {
        it.UV = bary.x * V2(0, 0) + bary.y * V2(1, 0) + bary.z * V2(1, 1);
    }
    it.sh = Frame(n);
    it.T = sp.t;
    it.wr = -r.d;
    it.prim = prim;
    return it;
}

// Math.Min / Math.Max on doubles (NaN propagates, -0 < +0).
static inline double dmin_net(double a, double b) {
    if (a != b) { if (!std::isnan(a)) return a < b ? a : b; return a; }
    return std::signbit(a) ? a : b;
}
static inline double dmax_net(double a, double b) {
    if (a != b) { if (!std::isnan(a)) return b < a ? a : b; return a; }
    return std::signbit(b) ? a : b;
}

// MathExt.SolveQuadratic (MathExt.cs:119-138), double precision
static bool solve_quadratic(double a, double b, double c, double* x0o, double* x1o) {
    bool linear = a == 0.0;
    bool valid_linear = linear && b != 0.0;
    double x0 = -c / b, x1 = -c / b;
    double disc = b * b - 4.0 * a * c;
    bool valid_quad = !linear && (disc >= 0.0);
    if (valid_quad) {
        double rd = std::sqrt(disc);
        double temp = -0.5 * (b + std::copysign(rd, b));
        double x0p = temp / a;
        double x1p = c / temp;
        double x0m = dmin_net(x0p, x1p);
        double x1m = dmax_net(x0p, x1p);
        x0 = linear ? x0 : x0m;
        x1 = linear ? x0 : x1m;
    }
    *x0o = x0; *x1o = x1;
    return valid_linear || valid_quad;
}

// Sphere.Intersect(ray, out SurfacePoint) (Shapes/Sphere.cs:65-80)
static inline bool sphere_hit(const cr_scene& s, int k, const Ray& r, SurfPt* sp) {
    const float* S = &s.sph[4 * k];
    double mint = r.tmin, maxt = r.tmax;
    double ox = (double)r.o.x - (double)S[0], oy = (double)r.o.y - (double)S[1], oz = (double)r.o.z - (double)S[2];
    double dx = r.d.x, dy = r.d.y, dz = r.d.z;
    double a = dx * dx + dy * dy + dz * dz;
    double b = 2 * (ox * dx + oy * dy + oz * dz);
    double c = (ox * ox + oy * oy + oz * oz) - (double)sqr(S[3]);
    double nt, ft;
    bool found = solve_quadratic(a, b, c, &nt, &ft);
    bool out_b = !(nt <= maxt && ft >= mint);
    bool in_b = nt < mint && ft > maxt;
    bool hit = found && !out_b && !in_b;
    if (hit) { sp->u = 0; sp->v = 0; sp->t = nt < mint ? (float)ft : (float)nt; }
    return hit;
}

// Sphere.GetIntersection (Sphere.cs:34-48): the centre is already in world
// space (the builder moves it, RendererBuilder.cs:49-54); WorldToModel only
// rotates the normal for the UV.
static Inct sphere_inct(const cr_scene& s, int k, const Ray& r, const SurfPt& sp, int prim) {
    const float* S = &s.sph[4 * k];
    V3 c(S[0], S[1], S[2]);
    Inct it;
    float t = sp.t;
    V3 n = normalize(r.at(t) - c);
    it.P = n * S[3] + c;
    it.sh = Frame(n);
    V3 local = n;                                            // TransformNormal(n, WorldToModel) (Sphere.cs:40)
    if (!s.sph_w2m.empty()) {
        const float* M = &s.sph_w2m[9 * k];
        local = V3(n.x * M[0] + n.y * M[3] + n.z * M[6], n.x * M[1] + n.y * M[4] + n.z * M[7],
                   n.x * M[2] + n.y * M[5] + n.z * M[8]);
    }
    float theta = std::acos(-local.y);
    float phi = std::atan2(-local.z, local.x) + PI_F;
    it.UV = V2(phi / (2 * PI_F), theta / PI_F);
    it.T = t;
    it.wr = -r.d;
    it.prim = prim;
    return it;
}

// A geometric primitive's test (GeometricPrimitive.Intersect, Primitive.cs:60-78).
static inline bool prim_hit(const cr_scene& s, int p, const Ray& r, SurfPt* sp) {
    const rd_prim_desc& pd = s.prims[p];
    return pd.kind == RD_PRIM_TRIANGLE ? tri_hit(s, pd.index, r, sp) : sphere_hit(s, pd.index, r, sp);
}

// Ray3F.Transform (Ray3F.cs:27-34): origin and normalised direction through
// m, MaxT = the distance to the image of ray.At(MaxT), MinT kept.
static Ray transform_ray(const Ray& r, const M4& m) {
    V3 o = xform_point(r.o, m);
    V3 d = normalize(xform_vector(r.d, m));
    V3 mx = xform_point(r.at(r.tmax), m);
    return Ray(o, d, r.tmin, length(o - mx));
}

// BoundingBox3F.Transform (BoundingBox3F.cs:232-243): the eight corners in
// the reference's order, through m.
static BBox bbox_transform(const BBox& b, const M4& m) {
    BBox r(xform_point(V3(b.mn.x, b.mn.y, b.mn.z), m));
    r.expand(xform_point(V3(b.mx.x, b.mn.y, b.mn.z), m));
    r.expand(xform_point(V3(b.mn.x, b.mx.y, b.mn.z), m));
    r.expand(xform_point(V3(b.mn.x, b.mn.y, b.mx.z), m));
    r.expand(xform_point(V3(b.mn.x, b.mx.y, b.mx.z), m));
    r.expand(xform_point(V3(b.mx.x, b.mx.y, b.mn.z), m));
    r.expand(xform_point(V3(b.mx.x, b.mn.y, b.mx.z), m));
    r.expand(xform_point(V3(b.mx.x, b.mx.y, b.mx.z), m));
    return r;
}

// One closest-hit candidate: world T, the scene primitive, the set triangle
// (index within its set) for an instance, and the SurfacePoint (barycentrics;
// t is the model-space t for an instance).
struct Hit {
    float T = 0;
    int prim = -1, sub = -1;
    SurfPt sp{0, 0, 0};
};

static inline Inct prim_inct(const cr_scene& s, const Hit& h, const Ray& r) {
    const rd_prim_desc& pd = s.prims[h.prim];
    if (pd.kind == RD_PRIM_INSTANCE) {
        // InstancedPrimitive.Intersect (Primitive.cs:165-187): the set
        // triangle's GetIntersection on the model ray, then P, N and Wr to world
        const Inst& I = s.insts[pd.index];
        Ray mr = transform_ray(r, I.w2m);
        Inct m = tri_inct(s, s.sets[I.set].first + h.sub, mr, h.sp, h.prim);
        Inct w;
        w.P = xform_point(m.P, I.m2w);
        w.sh = Frame(normalize(xform_vector(m.N(), I.m2w)));
        w.wr = normalize(xform_vector(m.wr, I.m2w));
        w.UV = m.UV;
        w.T = h.T;
        w.prim = h.prim;
        return w;
    }
    return pd.kind == RD_PRIM_TRIANGLE ? tri_inct(s, pd.index, r, h.sp, h.prim) : sphere_inct(s, pd.index, r, h.sp, h.prim);
}

static BBox prim_bound(const cr_scene& s, int p) {
    const rd_prim_desc& pd = s.prims[p];
    if (pd.kind == RD_PRIM_INSTANCE) {                     // InstancedPrimitive.WorldBound (Primitive.cs:160)
        const Inst& I = s.insts[pd.index];
        return bbox_transform(s.sets[I.set].nodes[0].b, I.m2w);
    }
    if (pd.kind == RD_PRIM_TRIANGLE) {                     // Triangle ctor (Triangle.cs:88-91)
        BBox b(s.pa(pd.index));
        b.expand(s.pb(pd.index));
        b.expand(s.pc(pd.index));
        return b;
    }
    const float* S = &s.sph[4 * pd.index];                 // Sphere ctor (Sphere.cs:30-31)
    V3 c(S[0], S[1], S[2]);
    BBox b; b.mn = c - V3(S[3]); b.mx = c + V3(S[3]);
    return b;
}

// ---------------------------------------------------------------- BVH build
// Accelerators/Bvh.cs:234-442 (SplitMethod.SAH).
struct PInfo { int num; BBox b; V3 c; };
struct BuildNode { BBox b; int left = -1, right = -1, axis = 0, first = 0, count = 0; };

// MathExt.Partition (MathExt.cs:170-201); `end` is an inclusive index.
template <class F>
static int partition_ref(std::vector<PInfo>& list, int begin, int end, F pred) {
    int first = begin, last = end;
    while (true) {
        while (true) {
            if (first == last) return first;
            if (!pred(list[first])) break;
            ++first;
        }
        while (!pred(list[last])) {
            if (first == last) return first;
            last--;
        }
        std::swap(list[first], list[last]);
        first++;
    }
}

struct Builder {
    const cr_scene& s;
    int max_prims;
    std::vector<PInfo> info;
    std::vector<BuildNode> bn;
    std::vector<int> ordered;
    int total = 0, leaves = 0;
    Builder(const cr_scene& sc, int mp) : s(sc), max_prims(mp) {}

    int make_leaf(int start, int end, const BBox& bounds) {
        BuildNode n; n.b = bounds; n.first = (int)ordered.size(); n.count = end - start;
        for (int i = start; i < end; i++) ordered.push_back(info[i].num);
        leaves++;
        bn.push_back(n);
        return (int)bn.size() - 1;
    }
    int build(int start, int end, int depth, int* max_depth) {
        total++;
        if (depth > *max_depth) *max_depth = depth;
        BBox bounds;
        for (int i = start; i < end; i++) bounds.unite(info[i].b);
        int n = end - start;
        if (n == 1) return make_leaf(start, end, bounds);
        BBox cb;
        for (int i = start; i < end; i++) cb.expand(info[i].c);
        int dim = cb.max_extent();
        int mid = (start + end) / 2;
        if (cb.mx[dim] == cb.mn[dim]) return make_leaf(start, end, bounds);
        if (n <= 2) {
            mid = (start + end) / 2;
            // List.Sort of two elements: swap when the comparer says l > r.
            float lv = info[start].c[dim], rv = info[start + 1].c[dim];
            if (!(lv == rv) && !(lv < rv)) std::swap(info[start], info[start + 1]);
        } else {
            const int nb = 12;
            int cnt[nb] = {0};
            BBox bb[nb];
            for (int i = start; i < end; i++) {
                V3 off = cb.offset(info[i].c);
                int b = (int)(nb * off[dim]);
                if (b == nb) b = nb - 1;
                cnt[b]++;
                bb[b].unite(info[i].b);
            }
            float cost[nb - 1];
            for (int i = 0; i < nb - 1; i++) {
                BBox b0, b1;
                int c0 = 0, c1 = 0;
                for (int j = 0; j <= i; j++) { b0.unite(bb[j]); c0 += cnt[j]; }
                for (int j = i + 1; j < nb; j++) { b1.unite(bb[j]); c1 += cnt[j]; }
                cost[i] = 1 + (c0 * b0.area() + c1 * b1.area()) / bounds.area();
            }
            float min_cost = cost[0];
            int split = 0;
            for (int i = 1; i < nb - 1; i++) if (cost[i] < min_cost) { min_cost = cost[i]; split = i; }
            float leaf_cost = (float)n;
            if (n > max_prims || min_cost < leaf_cost) {
                mid = partition_ref(info, start, end - 1, [&](const PInfo& pi) {
                    V3 off = cb.offset(pi.c);
                    int b = (int)(nb * off[dim]);
                    return b <= split;
                });
            } else {
                return make_leaf(start, end, bounds);
            }
        }
        int l = build(start, mid, depth + 1, max_depth);
        int r = build(mid, end, depth + 1, max_depth);
        BuildNode node;
        node.left = l; node.right = r; node.axis = dim;
        node.b = bn[l].b; node.b.unite(bn[r].b);           // BVHBuildNode(axis, l, r) :46-53
        bn.push_back(node);
        return (int)bn.size() - 1;
    }
    // FlattenBVHTree (Bvh.cs:417-442): depth-first, left child adjacent.
    int flatten(int id, std::vector<Node>& out, int* offset) {
        const BuildNode& n = bn[id];
        Node& ln = out[*offset];
        ln.b = n.b;
        int my = (*offset)++;
        if (n.count > 0) {
            out[my].offset = n.first;
            out[my].count = (uint16_t)n.count;
            out[my].axis = 0;
        } else {
            out[my].axis = (uint8_t)n.axis;
            out[my].count = 0;
            flatten(n.left, out, offset);
            out[my].offset = flatten(n.right, out, offset);
        }
        return my;
    }
};

// Bvh ctor over primitive bounds in primitive order.
static void build_tree(const cr_scene& s, const std::vector<BBox>& bounds, int max_prim, std::vector<Node>* nodes,
                       std::vector<int>* order, int* leaves, int* max_depth) {
    int np = (int)bounds.size();
    if (np == 0) return;
    Builder B(s, max_prim);
    B.info.resize(np);
    for (int i = 0; i < np; i++) {
        B.info[i].num = i;
        B.info[i].b = bounds[i];
        B.info[i].c = 0.5f * B.info[i].b.mn + 0.5f * B.info[i].b.mx;   // BVHPrimitiveInfo :61-66
    }
    B.bn.reserve(2 * np);
    B.ordered.reserve(np);
    int maxd = 0;
    int root = B.build(0, np, 0, &maxd);
    nodes->assign(B.total, Node());
    int off = 0;
    B.flatten(root, *nodes, &off);
    *order = B.ordered;
    *leaves = B.leaves;
    *max_depth = maxd;
}

static void build_bvh(cr_scene& s) {
    // every set's Bvh first (RendererBuilder.cs:895-921), then the scene's
    for (InstSet& st : s.sets) {
        std::vector<BBox> b(st.count);
        for (int j = 0; j < st.count; j++) {                // Triangle ctor (Triangle.cs:88-91)
            int t = st.first + j;
            b[j] = BBox(s.pa(t));
            b[j].expand(s.pb(t));
            b[j].expand(s.pc(t));
        }
        build_tree(s, b, st.max_prim, &st.nodes, &st.order, &st.leaf_count, &st.max_depth);
    }
    std::vector<BBox> b(s.prims.size());
    for (size_t i = 0; i < s.prims.size(); i++) b[i] = prim_bound(s, (int)i);
    build_tree(s, b, s.max_prim, &s.nodes, &s.order, &s.leaf_count, &s.max_depth);
}

// ---------------------------------------------------------------- traversal
// Bvh.Intersect(ray, out inct) (Bvh.cs:174-232) over one tree; `test` is the
// leaf primitive's Intersect and fills a candidate Hit (its T is what the
// loop compares).  nv / nt count nodes visited and primitives tested over
// every level.
template <class Test>
static bool dfs_closest(const std::vector<Node>& nodes, const std::vector<int>& order, Ray ray, Test test, Hit* out,
                        uint64_t* nv, uint64_t* nt) {
    if (nodes.empty()) return false;
    bool neg[3] = {ray.inv.x < 0, ray.inv.y < 0, ray.inv.z < 0};
    int stack[64];
    int sp = 0, cur = 0;
    float best = FLT_MAX;
    bool any = false;
    while (true) {
        const Node& node = nodes[cur];
        (*nv)++;
        if (box_hit(node.b, ray)) {
            if (node.count > 0) {
                for (int i = 0; i < node.count; ++i) {
                    Hit h;
                    (*nt)++;
                    if (test(order[node.offset + i], ray, &h)) {
                        any = true;
                        if (h.T < best) { best = h.T; *out = h; ray.tmax = h.T; }
                    }
                }
                if (sp == 0) break;
                cur = stack[--sp];
            } else {
                if (sp >= 64) { set_err("BVH stack overflow"); return false; }
                if (neg[node.axis]) { stack[sp++] = cur + 1; cur = node.offset; }
                else { stack[sp++] = node.offset; cur = cur + 1; }
            }
        } else {
            if (sp == 0) break;
            cur = stack[--sp];
        }
    }
    return any;
}

// Bvh.Intersect(ray) (Bvh.cs:126-172) over one tree.
template <class Test>
static bool dfs_any(const std::vector<Node>& nodes, const std::vector<int>& order, const Ray& ray, Test test,
                    uint64_t* nv, uint64_t* nt) {
    if (nodes.empty()) return false;
    bool neg[3] = {ray.inv.x < 0, ray.inv.y < 0, ray.inv.z < 0};
    int stack[64];
    int sp = 0, cur = 0;
    while (true) {
        const Node& node = nodes[cur];
        (*nv)++;
        if (box_hit(node.b, ray)) {
            if (node.count > 0) {
                for (int i = 0; i < node.count; ++i) {
                    (*nt)++;
                    if (test(order[node.offset + i], ray)) return true;
                }
                if (sp == 0) break;
                cur = stack[--sp];
            } else {
                if (neg[node.axis]) { stack[sp++] = cur + 1; cur = node.offset; }
                else { stack[sp++] = node.offset; cur = cur + 1; }
            }
        } else {
            if (sp == 0) break;
            cur = stack[--sp];
        }
    }
    return false;
}

// Primitive.Intersect(ray, out inct) of scene primitive p.  An
// InstancedPrimitive (Primitive.cs:165-187) runs its set's Bvh on the model
// ray; its T is the world distance from the ray origin to the transformed
// hit point (Triangle.GetIntersection's barycentric P, Triangle.cs:98-99).
static bool prim_closest(const cr_scene& s, int p, const Ray& r, Hit* h, uint64_t* nv, uint64_t* nt) {
    const rd_prim_desc& pd = s.prims[p];
    if (pd.kind != RD_PRIM_INSTANCE) {
        SurfPt sp;
        if (!prim_hit(s, p, r, &sp)) return false;
        h->T = sp.t; h->prim = p; h->sub = -1; h->sp = sp;
        return true;
    }
    const Inst& I = s.insts[pd.index];
    const InstSet& st = s.sets[I.set];
    Ray mr = transform_ray(r, I.w2m);
    Hit mh;
    auto tri = [&](int j, const Ray& rr, Hit* o) {
        SurfPt sp;
        if (!tri_hit(s, st.first + j, rr, &sp)) return false;
        o->T = sp.t; o->sub = j; o->sp = sp;
        return true;
    };
    if (!dfs_closest(st.nodes, st.order, mr, tri, &mh, nv, nt)) return false;
    int t = st.first + mh.sub;
    V3 bary(1 - (mh.sp.u + mh.sp.v), mh.sp.u, mh.sp.v);
    V3 pm = bary.x * s.pa(t) + bary.y * s.pb(t) + bary.z * s.pc(t);
    V3 pw = xform_point(pm, I.m2w);
    h->T = length(r.o - pw);            // Vector3.Distance(ray.O, p)
    h->prim = p; h->sub = mh.sub; h->sp = mh.sp;
    return true;
}

static bool prim_any(const cr_scene& s, int p, const Ray& r, uint64_t* nv, uint64_t* nt) {
    const rd_prim_desc& pd = s.prims[p];
    if (pd.kind != RD_PRIM_INSTANCE) {
        SurfPt sp;
        return prim_hit(s, p, r, &sp);
    }
    const Inst& I = s.insts[pd.index];
    const InstSet& st = s.sets[I.set];
    auto tri = [&](int j, const Ray& rr) { SurfPt sp; return tri_hit(s, st.first + j, rr, &sp); };
    return dfs_any(st.nodes, st.order, transform_ray(r, I.w2m), tri, nv, nt);
}

static bool bvh_closest(const cr_scene& s, const Ray& ray, Hit* out, uint64_t* cnt) {
    uint64_t nv = 0, nt = 0;
    auto test = [&](int p, const Ray& r, Hit* h) { return prim_closest(s, p, r, h, &nv, &nt); };
    bool any = dfs_closest(s.nodes, s.order, ray, test, out, &nv, &nt);
    if (cnt) { cnt[0] += nv; cnt[1] += nt; }
    return any;
}

static bool bvh_any(const cr_scene& s, const Ray& ray, uint64_t* cnt) {
    uint64_t nv = 0, nt = 0;
    auto test = [&](int p, const Ray& r) { return prim_any(s, p, r, &nv, &nt); };
    bool hit = dfs_any(s.nodes, s.order, ray, test, &nv, &nt);
    if (cnt) { cnt[0] += nv; cnt[1] += nt; }
    return hit;
}

// ---------------------------------------------------------------- octree
// Octree (Accelerators/Octree.cs), the config's "octree" accel: a sparse,
// loose octree.  BoundingBox3F's strict tests (Core/BoundingBox3F.cs:47-73)
// and GetCorner (:131-138) decide the build.
static bool oct_contains(const BBox& b, V3 p) {
    return p.x > b.mn.x && p.x < b.mx.x && p.y > b.mn.y && p.y < b.mx.y && p.z > b.mn.z && p.z < b.mx.z;
}
static bool oct_contains(const BBox& b, const BBox& o) {
    return o.mn.x > b.mn.x && o.mx.x < b.mx.x && o.mn.y > b.mn.y && o.mx.y < b.mx.y && o.mn.z > b.mn.z &&
           o.mx.z < b.mx.z;
}
static bool oct_overlaps(const BBox& b, const BBox& o) {
    return o.mn.x < b.mx.x && o.mx.x > b.mn.x && o.mn.y < b.mx.y && o.mx.y > b.mn.y && o.mn.z < b.mx.z &&
           o.mx.z > b.mn.z;
}
static V3 oct_corner(const BBox& b, int i) {
    return V3((i & 1) == 0 ? b.mn.x : b.mx.x, (i & 2) == 0 ? b.mn.y : b.mx.y, (i & 4) == 0 ? b.mn.z : b.mx.z);
}

namespace {
struct OctBuild {
    // build-time Node (:12-17): a leaf holds primitives, an interior node 8
    // child slots (null = no primitive went there)
    struct N {
        BBox box;
        std::vector<int> prims;
        std::unique_ptr<N> kids[8];
        bool interior = false;
    };
    const std::vector<BBox>& bound;
    int max_depth, max_count;
    float out_bound;
    int now_depth = 0, leaf_count = 0, node_count = 0;
    int64_t entries = 0;   // child-list entries so far (a guard: copies multiply per level)
    struct TooLarge {};

    OctBuild(const std::vector<BBox>& b, int md, int mc, float ob) : bound(b), max_depth(md), max_count(mc), out_bound(ob) {}

    // RecursionBuild (:227-309)
    std::unique_ptr<N> build(const std::vector<int>& prim_index, const BBox& inner_bound, const BBox& out_bnd, int depth) {
        const int cnt = (int)prim_index.size();
        if (cnt == 0) return nullptr;
        now_depth = std::max(now_depth, depth);
        if (cnt <= max_count || depth >= max_depth) {
            auto leaf = std::make_unique<N>();
            leaf->box = out_bnd;
            leaf->prims = prim_index;
            leaf_count++;
            node_count++;
            return leaf;
        }
        const V3 center = inner_bound.center();
        BBox inner[8], outer[8];
        std::vector<int> child_data[8];
        const float extend_len = (out_bound - 1) * 0.5f;
        for (int i = 0; i < 8; i++) {
            inner[i] = BBox(center);
            inner[i].expand(oct_corner(inner_bound, i));
            outer[i].mn = inner[i].mn - V3(extend_len);
            outer[i].mx = inner[i].mx + V3(extend_len);
        }
        for (int i : prim_index) {
            const BBox& ob = bound[i];
            const V3 cp = ob.center();
            bool inserted = false;
            for (int j = 0; j < 8 && !inserted; j++) {
                if (oct_contains(inner[j], cp) && oct_contains(outer[j], ob)) {
                    child_data[j].push_back(i);
                    inserted = true;
                }
            }
            if (!inserted)
                for (int j = 0; j < 8; j++)
                    if (oct_overlaps(inner[j], ob)) child_data[j].push_back(i);
        }
        for (int j = 0; j < 8; j++) entries += (int64_t)child_data[j].size();
        if (entries > ((int64_t)1 << 27)) throw TooLarge{};
        for (int i = 0; i < 8; i++) {
            BBox real;
            for (int j : child_data[i]) real.unite(bound[j]);
            outer[i] = real;
        }
        auto node = std::make_unique<N>();
        node->box = out_bnd;
        node->interior = true;
        for (int i = 0; i < 8; i++) node->kids[i] = build(child_data[i], inner[i], outer[i], depth + 1);
        node_count++;
        return node;
    }

    // FlattenNode (:311-346): a node's non-null children get consecutive
    // slots from `all`, then each child's subtree is flattened in turn
    void flatten(const N& node, int this_index, int* all, cr_scene& s) {
        s.oct[this_index].b = node.box;
        if (!node.interior) {
            s.oct_leaves.push_back(node.prims);
            s.oct[this_index].count_or_leaf = (int)s.oct_leaves.size() - 1;
            s.oct[this_index].next_head = -1;
            return;
        }
        int not_null = 0;
        for (const auto& k : node.kids) not_null += k ? 1 : 0;
        const int head = *all;
        s.oct[this_index].count_or_leaf = not_null;
        s.oct[this_index].next_head = head;
        *all += not_null;
        int j = 0;
        for (const auto& k : node.kids)
            if (k) flatten(*k, head + j++, all, s);
    }
};
}  // namespace

// Octree ctor (:48-93) over the scene primitives' WorldBound
static bool build_octree(cr_scene& s) {
    std::vector<BBox> b(s.prims.size());
    BBox root;
    std::vector<int> idx(s.prims.size());
    for (size_t i = 0; i < s.prims.size(); i++) {
        b[i] = prim_bound(s, (int)i);
        root.unite(b[i]);
        idx[i] = (int)i;
    }
    OctBuild B(b, s.oct_max_depth, s.oct_max_count, s.oct_out_bound);
    std::unique_ptr<OctBuild::N> r;
    try {
        r = B.build(idx, root, root, 0);
    } catch (const OctBuild::TooLarge&) {
        set_err("octree over 2^27 entries");
        return false;
    }
    s.oct.clear();
    s.oct_leaves.clear();
    s.oct_depth = B.now_depth;
    if (!r) return true;   // no primitive: an empty tree
    s.oct.resize(B.node_count);
    int all = 1;
    B.flatten(*r, 0, &all, s);
    if (all != B.node_count) { set_err("octree flatten count"); return false; }
    return true;
}

// Octree.Intersect(ray, out inct) (:133-225): the root box first, then a
// LIFO of the nodes whose box the ray met when their parent was expanded
// (children pushed in slot order); a popped node is not tested again.  nv
// counts box tests, nt primitive tests.
static bool oct_closest(const cr_scene& s, Ray ray, Hit* out, uint64_t* nv, uint64_t* nt) {
    if (s.oct.empty()) return false;
    (*nv)++;
    if (!box_hit(s.oct[0].b, ray)) return false;
    int stack[256];   // Octree.StackSize
    int sp = 0;
    stack[sp++] = 0;
    float best = FLT_MAX;
    bool any = false;
    while (sp > 0) {
        const cr_scene::OctLinear& n = s.oct[stack[--sp]];
        if (n.next_head == -1) {
            for (int p : s.oct_leaves[n.count_or_leaf]) {
                Hit h;
                (*nt)++;
                if (prim_closest(s, p, ray, &h, nv, nt)) {
                    any = true;
                    if (h.T < best) { best = h.T; *out = h; ray.tmax = h.T; }
                }
            }
        } else {
            for (int i = 0; i < n.count_or_leaf; i++) {
                (*nv)++;
                if (box_hit(s.oct[n.next_head + i].b, ray)) {
                    if (sp >= 256) { set_err("octree stack overflow"); return false; }
                    stack[sp++] = n.next_head + i;
                }
            }
        }
    }
    return any;
}

// Octree.Intersect(ray) (:95-131)
static bool oct_any(const cr_scene& s, const Ray& ray, uint64_t* nv, uint64_t* nt) {
    if (s.oct.empty()) return false;
    (*nv)++;
    if (!box_hit(s.oct[0].b, ray)) return false;
    int stack[256];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const cr_scene::OctLinear& n = s.oct[stack[--sp]];
        if (n.next_head == -1) {
            for (int p : s.oct_leaves[n.count_or_leaf]) {
                (*nt)++;
                if (prim_any(s, p, ray, nv, nt)) return true;
            }
        } else {
            for (int i = 0; i < n.count_or_leaf; i++) {
                (*nv)++;
                if (box_hit(s.oct[n.next_head + i].b, ray)) {
                    if (sp >= 256) { set_err("octree stack overflow"); return false; }
                    stack[sp++] = n.next_head + i;
                }
            }
        }
    }
    return false;
}

// The scene aggregate's closest / any query (Scene.cs:32-35, 58-70 call it).
static bool agg_closest(const cr_scene& s, const Ray& ray, Hit* out, uint64_t* cnt) {
    if (s.accel != RD_ACCEL_OCTREE) return bvh_closest(s, ray, out, cnt);
    uint64_t nv = 0, nt = 0;
    bool any = oct_closest(s, ray, out, &nv, &nt);
    if (cnt) { cnt[0] += nv; cnt[1] += nt; }
    return any;
}
static bool agg_any(const cr_scene& s, const Ray& ray, uint64_t* cnt) {
    if (s.accel != RD_ACCEL_OCTREE) return bvh_any(s, ray, cnt);
    uint64_t nv = 0, nt = 0;
    bool hit = oct_any(s, ray, &nv, &nt);
    if (cnt) { cnt[0] += nv; cnt[1] += nt; }
    return hit;
}

// Scene.Intersect(ray, out inct) (Scene.cs:32-35)
static bool scene_intersect(const cr_scene& s, const Ray& r, Inct* it, uint64_t* st) {
    Hit h;
    if (st) st[0]++;
    if (!agg_closest(s, r, &h, nullptr)) return false;
    *it = prim_inct(s, h, r);
    return true;
}
static bool scene_occluded_ray(const cr_scene& s, const Ray& r, uint64_t* st) {
    if (st) st[1]++;
    return agg_any(s, r, nullptr);
}
// Scene.IsOccluded (Scene.cs:58-70)
static bool scene_is_occluded(const cr_scene& s, V3 from, V3 to, uint64_t* st) {
    V3 vec = to - from;
    float dis = length(vec);
    if (dis < 0.001f) return true;
    V3 dir = normalize(vec);
    Ray sr(from, dir, 0.0001f, dis - 0.0001f);
    return scene_occluded_ray(s, sr, st);
}

// ---------------------------------------------------------------- sampling
static inline V3 square_to_cosine_hemisphere(float ux, float uy) {
    // Probability.SquareToUniformDisk + SquareToCosineHemisphere (Probability.cs:53-58,94-100)
    float radius = std::sqrt(ux);
    float angle = uy * PI_F * 2;
    float x = radius * std::cos(angle), y = radius * std::sin(angle);
    return V3(x, y, std::sqrt(1 - x * x - y * y));
}
static inline V3 spherical_direction(float st, float ct, float phi) {   // MathExt.cs:111-115
    float sp = std::sin(phi), cp = std::cos(phi);
    return V3(st * cp, st * sp, ct);
}

// ---------------------------------------------------------------- microfacet
// Bxdf/Microfacet.cs
struct MfDist {
    int kind;   // 0 Beckmann 1 GGX
    float ax, ay;
    MfDist(int k, float roughness, float aniso) : kind(k) {          // RoughnessToAlpha :177-183
        ax = nmax(roughness * roughness * (1 + aniso), 0.001f);
        ay = nmax(roughness * roughness * (1 - aniso), 0.001f);
    }
    float D(V3 wh) const {
        float t2 = tan2_theta(wh);
        if (std::isinf(t2)) return 0.0f;
        float c4 = cos2_theta(wh) * cos2_theta(wh);
        if (kind == 0)                                               // :117-123
            return std::exp(-t2 * (cos2_phi(wh) / (ax * ax) + sin2_phi(wh) / (ay * ay))) / (PI_F * ax * ay * c4);
        float e = (cos2_phi(wh) / (ax * ax) + sin2_phi(wh) / (ay * ay)) * t2;   // :125-132
        return 1 / (PI_F * ax * ay * c4 * (1 + e) * (1 + e));
    }
    float lambda(V3 w) const {
        float at = std::fabs(tan_theta(w));
        if (std::isinf(at)) return 0.0f;
        float alpha = std::sqrt(cos2_phi(w) * ax * ax + sin2_phi(w) * ay * ay);
        if (kind == 0) {                                             // :134-142
            float a = 1 / (alpha * at);
            if (a >= 1.6f) return 0;
            return (1 - 1.259f * a + 0.396f * a * a) / (3.535f * a + 2.181f * a * a);
        }
        float a2t2 = (alpha * at) * (alpha * at);                    // :144-151
        return (-1 + std::sqrt(1.0f + a2t2)) / 2;
    }
    float G(V3 wo, V3 wi) const { return 1 / (1 + lambda(wo) + lambda(wi)); }
    float pdf(V3 wo, V3 wh) const { return D(wh) * abs_cos_theta(wh); }
    V3 sample_wh(V3 wo, Rng& rng) const {
        float ux = rng.next_float(), uy = rng.next_float();
        if (kind == 0) {                                             // Beckmann.SampleWh :60-86
            float t2, phi;
            if (ax == ay) {
                float ls = std::log(1 - ux);
                t2 = -ax * ax * ls;
                phi = uy * 2 * PI_F;
            } else {
                float ls = std::log(1 - ux);
                phi = std::atan(ay / ax * std::tan(2 * PI_F * uy + 0.5f * PI_F));
                if (uy > 0.5f) phi += PI_F;
                float sp = std::sin(phi), cp = std::cos(phi);
                float ax2 = ax * ax, ay2 = ay * ay;
                t2 = -ls / (cp * cp / ax2 + sp * sp / ay2);
            }
            float ct = 1 / std::sqrt(1 + t2);
            float st = std::sqrt(nmax(0, 1 - ct * ct));
            V3 wh = spherical_direction(st, ct, phi);
            if (!same_hemisphere(wo, wh)) wh = -wh;
            return wh;
        }
        float phi = (2 * PI_F) * uy;                                 // SampleWhGGX :185-209
        float ct;
        if (ax == ay) {
            float t2 = ax * ax * ux / (1.0f - ux);
            ct = 1 / std::sqrt(1 + t2);
        } else {
            phi = std::atan(ay / ax * std::tan(2 * PI_F * uy + .5f * PI_F));
            if (uy > 0.5f) phi += PI_F;
            float sp = std::sin(phi), cp = std::cos(phi);
            float ax2 = ax * ax, ay2 = ay * ay;
            float a2 = 1 / (cp * cp / ax2 + sp * sp / ay2);
            float t2 = a2 * ux / (1 - ux);
            ct = 1 / std::sqrt(1 + t2);
        }
        float st = std::sqrt(nmax(0.0f, 1.0f - ct * ct));
        V3 wh = spherical_direction(st, ct, phi);
        if (!same_hemisphere(wo, wh)) wh = -wh;
        return wh;
    }
};

// Fresnel.DielectricFunc (Bxdf/Fresnel.cs:50-68)
static float fresnel_dielectric(float ci, float ei, float et) {
    ci = nclamp(ci, -1, 1);
    bool entering = ci > 0.0f;
    if (!entering) { std::swap(ei, et); ci = std::fabs(ci); }
    float si = std::sqrt(nmax(0, 1 - ci * ci));
    float st = ei / et * si;
    if (st >= 1) return 1;
    float ct = std::sqrt(nmax(0, 1 - st * st));
    float rparl = ((et * ci) - (ei * ct)) / ((et * ci) + (ei * ct));
    float rperp = ((ei * ci) - (et * ct)) / ((ei * ci) + (et * ct));
    return (rparl * rparl + rperp * rperp) / 2;
}
// Fresnel.ConductorFunc (Bxdf/Fresnel.cs:72-95)
static C3 fresnel_conductor(float ci, C3 etai, C3 etat, C3 k) {
    ci = nclamp(ci, -1, 1);
    C3 eta = etat / etai;
    C3 etak = k / etai;
    float ci2 = ci * ci;
    float si2 = 1.0f - ci2;
    C3 eta2 = eta * eta;
    C3 etak2 = etak * etak;
    C3 t0 = eta2 - etak2 - si2;
    C3 a2pb2 = csqrt(t0 * t0 + 4 * eta2 * etak2);
    C3 t1 = a2pb2 + ci2;
    C3 a = csqrt(0.5f * (a2pb2 + t0));
    C3 t2 = 2 * ci * a;
    C3 Rs = (t1 - t2) / (t1 + t2);
    C3 t3 = ci2 * a2pb2 + si2 * si2;
    C3 t4 = t2 * si2;
    C3 Rp = Rs * (t3 - t4) / (t3 + t4);
    return 0.5f * (Rp + Rs);
}
// MathExt.Refract (MathExt.cs:83-92)
static bool refract(V3 wi, V3 n, float eta, V3* wt) {
    float ci = dot(n, wi);
    float s2i = nmax(0, 1 - ci * ci);
    float s2t = eta * eta * s2i;
    if (s2t >= 1) return false;
    float ct = std::sqrt(1 - s2t);
    *wt = normalize(eta * -wi + (eta * ci - ct) * n);
    return true;
}

// ---------------------------------------------------------------- materials
enum { T_REFLECTION = 1, T_TRANSMISSION = 2, T_DIFFUSE = 4, T_GLOSSY = 8, T_SPECULAR = 16,
       T_SUBSURFACE = 32 };  // IBxdf.cs:7-15
struct BSample { V3 wi; C3 fr; float pdf = 0; int type = 0; };

static inline C3 tex_at(const cr_scene& s, int t, V2 uv) { return s.tex[t].sample(uv); }

// MicrofacetReflectionBrdf<Conductor, D>.Fr (Bxdf/MicrofacetReflectionBrdf.cs:26-41), R = 1
static C3 mf_reflect_fr(const MfDist& d, C3 eta, C3 k, V3 wo, V3 wi) {
    if (!same_hemisphere(wo, wi)) return C3(0.0f);
    float cto = abs_cos_theta(wo), cti = abs_cos_theta(wi);
    V3 wh = wi + wo;
    if (cti == 0 || cto == 0) return C3(0);
    if (wh.x == 0 && wh.y == 0 && wh.z == 0) return C3(0);
    wh = normalize(wh);
    C3 f = fresnel_conductor(dot(wi, faceforward(wh, V3(0, 0, 1))), C3(1.0f), eta, k);
    float D = d.D(wh);
    float G = d.G(wo, wi);
    return C3(1.0f) * cabs((D * G * f) / (4 * cti * cto));
}
static float mf_reflect_pdf(const MfDist& d, V3 wo, V3 wi) {        // :43-48
    if (!same_hemisphere(wo, wi)) return 0;
    V3 wh = normalize(wo + wi);
    return d.pdf(wo, wh) / (4 * dot(wo, wh));
}

// ---------------------------------------------------------------- plastic
// RoughPlastic (Materials/RoughPlastic.cs:37-166): kd*Lambert + ks*microfacet
// reflection with dielectric Fresnel and tint R; lobe probability ks/(kd+ks).
static C3 plastic_fr(const MfDist& d, C3 r, float kd, float ks, float ei, float et, bool two, V3 wo, V3 wi) {
    if (cos_theta(wo) <= 0 || cos_theta(wi) <= 0) { if (!two) return C3(0.0f); }
    C3 diffuse = same_hemisphere(wo, wi) ? r / PI_F : C3(0.0f);
    C3 glossy(0.0f);
    // MicrofacetReflectionBrdf<Fresnel.Dielectric, D>.Fr with R = r
    if (same_hemisphere(wo, wi)) {
        float cto = abs_cos_theta(wo), cti = abs_cos_theta(wi);
        V3 wh = wi + wo;
        if (!(cti == 0 || cto == 0) && !(wh.x == 0 && wh.y == 0 && wh.z == 0)) {
            wh = normalize(wh);
            C3 f(fresnel_dielectric(dot(wi, faceforward(wh, V3(0, 0, 1))), ei, et));
            float D = d.D(wh), G = d.G(wo, wi);
            glossy = r * cabs((D * G * f) / (4 * cti * cto));
        }
    }
    return kd * diffuse + ks * glossy;
}
static float plastic_pdf(const MfDist& d, float kd, float ks, bool two, V3 wo, V3 wi) {
    if (cos_theta(wo) <= 0 || cos_theta(wi) <= 0) { if (!two) return 0.0f; }
    float p = ks / (kd + ks);
    float diffuse = same_hemisphere(wo, wi) ? abs_cos_theta(wi) / PI_F : 0.0f;
    float glossy = mf_reflect_pdf(d, wo, wi);
    return (1 - p) * diffuse + p * glossy;
}

// ---------------------------------------------------------------- Disney
// DisneyBsdf and its lobes (Bxdf/DisneyBsdf.cs:13-512).  The reference
// author marks it abandoned (:11); this restates that code as written.
struct Disney {
    C3 color, sdist;
    float metallic, eta, roughness, spec_tint, aniso, sheen, sheen_tint, clearcoat, cc_gloss, spec_scale;
    float transmission, trans_rough, flatness;
    bool thin;
};
static inline float lerpf(float t, float a, float b) { return (1 - t) * a + t * b; }        // MathExt.Lerp
static inline C3 lerpc(float w, C3 a, C3 b) { return (1 - w) * a + w * b; }                // Color3F.Lerp
static inline float pow5(float a) { return (a * a) * (a * a) * a; }
static inline float schlick_weight(float c) { return pow5(nclamp(1 - c, 0, 1)); }
static inline C3 schlick_c(C3 r0, float c) { return lerpc(schlick_weight(c), r0, C3(1.0f)); }
static inline float schlick_f(float r0, float c) { return r0 + (1.0f - r0) * schlick_weight(c); }
static inline float gtr1(float c, float a) {
    float a2 = a * a;
    return (a2 - 1) / (PI_F * std::log(a2) * (1 + (a2 - 1) * c * c));
}
static inline float smith_g1_ggx(float c, float a) {
    float a2 = a * a, c2 = c * c;
    return 2 / (1 + std::sqrt(1 + a2 * (1 - c2) / c2));
}
static inline C3 color_tint(const Disney& b) {
    float l = b.color.lum();
    return l > 0 ? b.color / l : C3(1.0f);
}
static inline void aniso_alpha(const Disney& b, float rough, float* ax, float* ay) {
    float aspect = std::sqrt(1 - b.aniso * 0.9f);
    *ax = nmax(0.001f, sqr(rough) / aspect);
    *ay = nmax(0.001f, sqr(rough) * aspect);
}
// DisneyDistributionGtr2: GGX D, separable G = G1(wo) G1(wi)
struct Gtr2 {
    MfDist m;
    explicit Gtr2(float ax, float ay) : m(1, 0.0f, 0.0f) { m.ax = ax; m.ay = ay; }
    float D(V3 wh) const { return m.D(wh); }
    float G(V3 wo, V3 wi) const { return (1 / (1 + m.lambda(wo))) * (1 / (1 + m.lambda(wi))); }
    float pdf(V3 wo, V3 wh) const { return D(wh) * abs_cos_theta(wh); }
};
static C3 disney_fresnel(const Disney& b, float ci) {                                     // DisneyFresnel.Eval
    C3 r0 = color_tint(b) * (sqr(b.eta - 1) / sqr(b.eta + 1));
    C3 fd = lerpc(b.spec_tint, C3(fresnel_dielectric(ci, 1, b.eta)), schlick_c(r0, ci));
    return lerpc(b.metallic, fd * b.spec_scale, schlick_c(b.color, ci));
}
// MicrofacetReflectionBrdf with R, a Fresnel functor and Gtr2
template <class F>
static C3 mf_reflect_generic(const Gtr2& d, C3 R, F fres, V3 wo, V3 wi) {
    if (!same_hemisphere(wo, wi)) return C3(0.0f);
    float cto = abs_cos_theta(wo), cti = abs_cos_theta(wi);
    V3 wh = wi + wo;
    if (cti == 0 || cto == 0) return C3(0);
    if (wh.x == 0 && wh.y == 0 && wh.z == 0) return C3(0);
    wh = normalize(wh);
    C3 f = fres(dot(wi, faceforward(wh, V3(0, 0, 1))));
    float D = d.D(wh), G = d.G(wo, wi);
    return R * cabs((D * G * f) / (4 * cti * cto));
}
static float mf_reflect_pdf_g(const Gtr2& d, V3 wo, V3 wi) {
    if (!same_hemisphere(wo, wi)) return 0;
    V3 wh = normalize(wo + wi);
    return d.pdf(wo, wh) / (4 * dot(wo, wh));
}
static Gtr2 spec_dist(const Disney& b) { float ax, ay; aniso_alpha(b, b.roughness, &ax, &ay); return Gtr2(ax, ay); }
static Gtr2 trans_dist(const Disney& b) { float ax, ay; aniso_alpha(b, b.trans_rough, &ax, &ay); return Gtr2(ax, ay); }
static C3 disney_spec_brdf(const Disney& b, V3 wo, V3 wi) {
    return mf_reflect_generic(spec_dist(b), C3(1.0f), [&](float c) { return disney_fresnel(b, c); }, wo, wi);
}
static C3 disney_inner_brdf(const Disney& b, V3 wo, V3 wi) {
    return mf_reflect_generic(trans_dist(b), b.transmission * b.color,
                              [&](float c) { return C3(fresnel_dielectric(c, 1, b.eta)); }, wo, wi);
}
// MicrofacetTransmissionBtdf<Gtr2> (Bxdf/MicrofacetTransmissionBtdf.cs:25-77), etaA = 1, etaB = Eta
// imp: TransportMode.Importance (no 1/eta radiance scaling, MicrofacetTransmissionBtdf.cs:40)
static C3 disney_btdf_fr(const Disney& b, V3 wo, V3 wi, bool imp = false) {
    if (same_hemisphere(wo, wi)) return C3();
    float cto = cos_theta(wo), cti = cos_theta(wi);
    if (cti == 0 || cto == 0) return C3();
    Gtr2 d = trans_dist(b);
    bool entering = cos_theta(wo) > 0;
    float ei = entering ? 1.0f : b.eta, et = entering ? b.eta : 1.0f;
    float eta = et / ei;
    V3 wh = normalize(wo + wi * eta);
    if (cos_theta(wh) < 0) wh = -wh;
    if (dot(wo, wh) * dot(wi, wh) > 0) return C3();
    C3 f(fresnel_dielectric(dot(wo, wh), 1.0f, b.eta));
    float sd = dot(wo, wh) + eta * dot(wi, wh);
    float factor = imp ? 1.0f : 1 / eta;
    float D = d.D(wh), G = d.G(wo, wi);
    C3 T = (1 - b.metallic) * b.transmission * csqrt(b.color);
    return (1 - f) * T * std::fabs(D * G * eta * eta * abs_dot(wi, wh) * abs_dot(wo, wh) * factor * factor /
                                   (cti * cto * sd * sd));
}
static float disney_btdf_pdf(const Disney& b, V3 wo, V3 wi) {
    if (same_hemisphere(wo, wi)) return 0.0f;
    Gtr2 d = trans_dist(b);
    bool entering = cos_theta(wo) > 0;
    float ei = entering ? 1.0f : b.eta, et = entering ? b.eta : 1.0f;
    float eta = et / ei;
    V3 wh = normalize(wo + wi * eta);
    if (dot(wo, wh) * dot(wi, wh) > 0) return 0;
    float sd = dot(wo, wh) + eta * dot(wi, wh);
    float dwh = std::fabs((eta * eta * dot(wi, wh)) / (sd * sd));
    return d.pdf(wo, wh) * dwh;
}
static C3 disney_clearcoat_fr(const Disney& b, V3 wo, V3 wi) {                          // DisneyClearcoatBrdf
    float gloss = lerpf(b.cc_gloss, 0.1f, 0.001f);
    V3 wh = wo + wi;
    if (wh.x == 0 && wh.y == 0 && wh.z == 0) return C3(0);
    wh = normalize(wh);
    float D = gtr1(abs_cos_theta(wh), gloss);
    float F = schlick_f(0.04f, dot(wo, wh));
    float G = smith_g1_ggx(abs_cos_theta(wo), 0.25f) * smith_g1_ggx(abs_cos_theta(wi), 0.25f);
    return C3(b.clearcoat * D * F * G / 4);
}
static float disney_clearcoat_pdf(const Disney& b, V3 wo, V3 wi) {
    float gloss = lerpf(b.cc_gloss, 0.1f, 0.001f);
    if (!same_hemisphere(wo, wi)) return 0;
    V3 wh = wo + wi;
    if (wh.x == 0 && wh.y == 0 && wh.z == 0) return 0;
    wh = normalize(wh);
    float D = gtr1(abs_cos_theta(wh), gloss);
    return D * abs_cos_theta(wh) / (4 * dot(wo, wh));
}
static C3 disney_fr(const Disney& b, V3 wo, V3 wi, bool imp = false) {                                      // DisneyBsdf.FrImpl
    if (!same_hemisphere(wo, wi)) {
        if (b.transmission == 0) return C3(0.0f);
        return disney_btdf_fr(b, wo, wi, imp);
    }
    if (cos_theta(wo) < 0 && cos_theta(wi) < 0) return disney_inner_brdf(b, wo, wi);
    C3 diffuse(0.0f);
    float dw = 1 - b.metallic;
    if (dw > 0) {
        float fo = schlick_weight(abs_cos_theta(wo)), fi = schlick_weight(abs_cos_theta(wi));
        C3 R = dw * b.color;
        C3 ddiff(0.0f);
        C3 lambert = R * (1 / PI_F) * (1 - fo / 2) * (1 - fi / 2);                      // DisneyDiffuseBrdf
        if (b.thin) {
            C3 fake(0.0f);                                                               // DisneyFakeSubsurfaceBrdf
            V3 wh = wo + wi;
            if (!(wh.x == 0 && wh.y == 0 && wh.z == 0)) {
                wh = normalize(wh);
                float cd = dot(wi, wh);
                float fss90 = cd * cd * b.roughness;
                float fss = lerpf(fo, 1.0f, fss90) * lerpf(fi, 1.0f, fss90);
                float ss = 1.25f * (fss * (1 / (abs_cos_theta(wo) + abs_cos_theta(wi)) - 0.5f) + 0.5f);
                fake = R * (1 / PI_F) * ss;
            }
            ddiff = lambert * (1 - b.flatness) + fake * b.flatness;
        } else if (b.sdist.black()) {
            ddiff = lambert;
        }
        C3 retro(0.0f);                                                                  // DisneyRetroBrdf
        {
            V3 wh = wo + wi;
            if (!(wh.x == 0 && wh.y == 0 && wh.z == 0)) {
                wh = normalize(wh);
                float cd = dot(wi, wh);
                float rr = 2 * b.roughness * cd * cd;
                retro = R * (1 / PI_F) * rr * (fo + fi + fo * fi * (rr - 1));
            }
        }
        C3 sheen(0.0f);
        if (b.sheen > 0) {                                                               // DisneySheenBrdf
            C3 sc = lerpc(b.sheen_tint, C3(1.0f), color_tint(b));
            V3 wh = wo + wi;
            if (!(wh.x == 0 && wh.y == 0 && wh.z == 0)) {
                wh = normalize(wh);
                sheen = (dw * b.sheen * sc) * schlick_weight(dot(wi, wh));
            }
        }
        diffuse = (retro + sheen + ddiff) * (1 - b.transmission);
    }
    C3 spec = disney_spec_brdf(b, wo, wi);
    C3 cc(0.0f);
    if (b.clearcoat > 0) cc = disney_clearcoat_fr(b, wo, wi);
    return diffuse + spec + cc;
}
static void disney_weights(const Disney& b, float* dw, float* tw, float* sw, float* cw) {     // GetSampleWeight
    float a = nclamp(b.color.lum() * (1 - b.metallic), 0.3f, 0.7f);
    float bb = 1 - a;
    *dw = a * (1 - b.transmission);
    *tw = a * b.transmission;
    *sw = bb * 2 / (2 + b.clearcoat);
    *cw = bb * b.clearcoat / (2 + b.clearcoat);
}
static float disney_pdf(const Disney& b, V3 wo, V3 wi) {                                  // DisneyBsdf.PdfImpl
    float dw, tw, sw, cw;
    disney_weights(b, &dw, &tw, &sw, &cw);
    if (cos_theta(wo) < 0) {
        if (b.transmission == 0) return 0;
        float fr = nclamp(fresnel_dielectric(cos_theta(wo), 1, b.eta), 0.1f, 0.9f);
        if (cos_theta(wi) > 0) return disney_btdf_pdf(b, wo, wi) * (1 - fr);
        return mf_reflect_pdf_g(trans_dist(b), wo, wi) * fr;
    }
    if (cos_theta(wi) < 0) return disney_btdf_pdf(b, wo, wi) * tw;
    float dp = 0;
    if (b.thin || b.sdist.black()) dp = same_hemisphere(wo, wi) ? abs_cos_theta(wi) / PI_F : 0.0f;
    float sp = mf_reflect_pdf_g(spec_dist(b), wo, wi);
    float cp = disney_clearcoat_pdf(b, wo, wi);
    return dp * dw + sp * sw + cp * cw;
}

static Disney load_disney(const cr_scene& s, const rd_material_desc& m, V2 uv) {         // Disney.CreateBsdf
    Disney b;
    b.color = tex_at(s, m.tex[0], uv);
    b.metallic = tex_at(s, m.tex[1], uv).r;
    b.roughness = tex_at(s, m.tex[2], uv).r;
    b.eta = tex_at(s, m.tex[3], uv).r;
    b.spec_scale = tex_at(s, m.tex[4], uv).r;
    b.spec_tint = tex_at(s, m.tex[5], uv).r;
    b.aniso = tex_at(s, m.tex[6], uv).r;
    b.sheen = tex_at(s, m.tex[7], uv).r;
    b.sheen_tint = tex_at(s, m.tex[8], uv).r;
    b.clearcoat = tex_at(s, m.tex[9], uv).r;
    b.cc_gloss = tex_at(s, m.tex[10], uv).r;
    b.sdist = tex_at(s, m.tex[11], uv);
    b.transmission = tex_at(s, m.tex[12], uv).r;
    b.trans_rough = tex_at(s, m.tex[13], uv).r;
    b.flatness = tex_at(s, m.tex[14], uv).r;
    b.thin = m.thin != 0;
    return b;
}

// imp: TransportMode.Importance (light paths: RoughGlass.cs:63, FresnelSpecularBsdf.cs:58,
// MicrofacetTransmissionBtdf.cs:40 drop the radiance scaling of refraction)
static C3 material_fr(const cr_scene& s, const rd_material_desc& m, V3 wo, V3 wi, V2 uv, bool imp = false) {
    switch (m.type) {
    case RD_MAT_ROUGH_PLASTIC: {
        MfDist d(m.dist, tex_at(s, m.tex[1], uv).r, tex_at(s, m.tex[2], uv).r);
        return plastic_fr(d, tex_at(s, m.tex[0], uv), m.kd, m.ks, m.eta_a, m.eta_b, m.two_sided != 0, wo, wi);
    }
    case RD_MAT_PERFECT_GLASS:     // FresnelSpecularBsdf.Fr (FresnelSpecularBsdf.cs:30-33)
    case RD_MAT_SUBSURFACE:        // Subsurface.Fr: the same FresnelSpecularBsdf (Subsurface.cs:31-36)
    case RD_MAT_PERFECT_MIRROR:    // SpecularReflectionBrdf.Fr (SpecularReflectionBrdf.cs:17-20)
        return C3(0.0f);
    case RD_MAT_DISNEY:
        return disney_fr(load_disney(s, m, uv), wo, wi, imp);
    case RD_MAT_DIFFUSE: {                                           // Materials/Diffuse.cs:21-27
        if (cos_theta(wo) <= 0 || cos_theta(wi) <= 0) { if (!m.two_sided) return C3(0.0f); }
        C3 R = tex_at(s, m.tex[0], uv);                              // LambertianReflectionBrdf.Fr :16-19
        return same_hemisphere(wo, wi) ? R / PI_F : C3(0.0f);
    }
    case RD_MAT_ROUGH_METAL: {                                       // Materials/RoughMetal.cs:33-58
        float rough = tex_at(s, m.tex[0], uv).r, an = tex_at(s, m.tex[1], uv).r;
        MfDist d(m.dist, rough, an);
        if (cos_theta(wo) <= 0 || cos_theta(wi) <= 0) { if (!m.two_sided) return C3(0.0f); }
        return mf_reflect_fr(d, C3(m.eta[0], m.eta[1], m.eta[2]), C3(m.k[0], m.k[1], m.k[2]), wo, wi);
    }
    case RD_MAT_ROUGH_GLASS: {                                       // Materials/RoughGlass.cs:151-195
        float rough = tex_at(s, m.tex[2], uv).r, an = tex_at(s, m.tex[3], uv).r;
        MfDist d(m.dist, rough, an);
        C3 r = tex_at(s, m.tex[0], uv), t = tex_at(s, m.tex[1], uv);
        bool is_reflect = same_hemisphere(wo, wi);
        bool entering = cos_theta(wo) > 0;
        float ei = entering ? m.eta_a : m.eta_b;
        float et = entering ? m.eta_b : m.eta_a;
        float eta = is_reflect ? 1 : et / ei;
        V3 wh = normalize(wo + wi * eta);
        if (cos_theta(wh) < 0) wh = -wh;
        float sqrt_denom = dot(wo, wh) + eta * dot(wi, wh);
        float D = d.D(wh);
        float G = d.G(wo, wi);
        float F = fresnel_dielectric(dot(wo, wh), ei, et);
        C3 bsdf;
        if (is_reflect) {
            bsdf = r * (D * G * F / (4.0f * cos_theta(wi) * cos_theta(wo)));
        } else {
            if (dot(wo, wh) * dot(wi, wh) > 0) return C3();
            float factor = imp ? 1.0f : 1 / eta;   // RoughGlass.cs:63
            bsdf = (1 - F) * t *
                   std::fabs(D * G * eta * eta * abs_dot(wi, wh) * abs_dot(wo, wh) * factor * factor /
                             (cos_theta(wo) * cos_theta(wi) * sqrt_denom * sqrt_denom));
        }
        return bsdf.valid() ? bsdf : C3(0.0f);
    }
    default:
        return C3(0.0f);
    }
}

static float material_pdf(const cr_scene& s, const rd_material_desc& m, V3 wo, V3 wi, V2 uv) {
    switch (m.type) {
    case RD_MAT_ROUGH_PLASTIC: {
        MfDist d(m.dist, tex_at(s, m.tex[1], uv).r, tex_at(s, m.tex[2], uv).r);
        return plastic_pdf(d, m.kd, m.ks, m.two_sided != 0, wo, wi);
    }
    case RD_MAT_PERFECT_GLASS:
    case RD_MAT_SUBSURFACE:
    case RD_MAT_PERFECT_MIRROR:
        return 0.0f;
    case RD_MAT_DISNEY:
        return disney_pdf(load_disney(s, m, uv), wo, wi);
    case RD_MAT_DIFFUSE: {                                           // Diffuse.cs:29-37
        if (cos_theta(wo) <= 0 || cos_theta(wi) <= 0) { if (!m.two_sided) return 0.0f; }
        return same_hemisphere(wo, wi) ? abs_cos_theta(wi) / PI_F : 0.0f;   // Lambert.Pdf :21-24
    }
    case RD_MAT_ROUGH_METAL: {                                       // RoughMetal.cs:60-85
        float rough = tex_at(s, m.tex[0], uv).r, an = tex_at(s, m.tex[1], uv).r;
        MfDist d(m.dist, rough, an);
        if (cos_theta(wo) <= 0 || cos_theta(wi) <= 0) { if (!m.two_sided) return 0.0f; }
        return mf_reflect_pdf(d, wo, wi);
    }
    case RD_MAT_ROUGH_GLASS: {                                       // RoughGlass.cs:197-238
        float rough = tex_at(s, m.tex[2], uv).r, an = tex_at(s, m.tex[3], uv).r;
        MfDist d(m.dist, rough, an);
        bool is_reflect = same_hemisphere(wo, wi);
        bool entering = cos_theta(wo) > 0;
        float ei = entering ? m.eta_a : m.eta_b;
        float et = entering ? m.eta_b : m.eta_a;
        float pdf;
        if (is_reflect) {
            V3 wh = normalize(wo + wi);
            if (cos_theta(wh) < 0) wh = -wh;
            if (dot(wo, wh) * cos_theta(wo) <= 0 || dot(wi, wh) * cos_theta(wi) <= 0) return 0.0f;
            float jac = 1.0f / (4.0f * abs_dot(wh, wi));
            float F = fresnel_dielectric(dot(wo, wh), ei, et);
            pdf = d.pdf(wo, wh) * jac * F;
        } else {
            float eta = et / ei;
            V3 wh = normalize(wo + wi * eta);
            if (cos_theta(wh) < 0) wh = -wh;
            if (dot(wo, wh) * dot(wi, wh) > 0) return 0.0f;
            if (dot(wo, wh) * cos_theta(wo) <= 0 || dot(wi, wh) * cos_theta(wi) <= 0) return 0.0f;
            float sd = dot(wo, wh) + eta * dot(wi, wh);
            float dwh = std::fabs((eta * eta * dot(wi, wh)) / (sd * sd));
            float F = fresnel_dielectric(dot(wo, wh), ei, et);
            pdf = d.pdf(wo, wh) * dwh * (1 - F);
        }
        return std::isinf(pdf) || std::isnan(pdf) ? 0.0f : pdf;
    }
    default:
        return 0.0f;
    }
}

// Cosine-hemisphere sample with the z-flip of LambertianReflectionBrdf.Sample
// (LambertianReflectionBrdf.cs:26-36) / DisneyBsdf.DiffuseSample.
static bool cosine_sample(V3 wo, Rng& rng, V3* wi_out) {
    float ux = rng.next_float(), uy = rng.next_float();
    V3 wi = normalize(square_to_cosine_hemisphere(ux, uy));
    if (wo.z < 0) wi.z *= -1;
    if (!same_hemisphere(wo, wi)) return false;
    *wi_out = wi;
    return true;
}

// MicrofacetReflectionBrdf.Sample's direction part (MicrofacetReflectionBrdf.cs:50-59)
template <class SampleWh>
static bool mf_reflect_sample_dir(V3 wo, Rng& rng, SampleWh swh, V3* wi_out, V3* wh_out) {
    if (cos_theta(wo) == 0) return false;
    V3 wh = swh(wo, rng);
    if (dot(wo, wh) < 0) return false;
    V3 wi = reflect(-wo, wh);
    if (!same_hemisphere(wo, wi)) return false;
    *wi_out = wi;
    *wh_out = wh;
    return true;
}

static BSample disney_sample(const Disney& b, V3 wo, Rng& rng, bool imp = false) {                         // DisneyBsdf.Sample
    BSample out;
    Gtr2 sd = spec_dist(b), td = trans_dist(b);
    auto wh_spec = [&](V3 w, Rng& r) { return sd.m.sample_wh(w, r); };
    auto wh_trans = [&](V3 w, Rng& r) { return td.m.sample_wh(w, r); };
    auto finish = [&](V3 wi, int type) {
        out.wi = wi;
        out.fr = disney_fr(b, wo, wi, imp);
        out.pdf = disney_pdf(b, wo, wi);
        out.type = type;
        return out;
    };
    // MicrofacetTransmissionBtdf.Sample (MicrofacetTransmissionBtdf.cs:59-75), etaA = 1, etaB = Eta
    auto btdf_sample = [&](V3* wi_out) -> bool {
        if (wo.z == 0) return false;
        V3 wh = td.m.sample_wh(wo, rng);
        if (dot(wo, wh) < 0) return false;
        bool entering = cos_theta(wo) > 0;
        float ei = entering ? 1.0f : b.eta, et = entering ? b.eta : 1.0f;
        V3 wi;
        if (!refract(wo, wh, ei / et, &wi)) return false;
        C3 f = disney_btdf_fr(b, wo, wi, imp);
        float p = disney_btdf_pdf(b, wo, wi);
        if (!f.valid() || std::isinf(p)) return false;
        if (p == 0) return false;
        *wi_out = wi;
        return true;
    };
    V3 wi, wh;
    if (cos_theta(wo) < 0) {
        if (b.transmission == 0) return out;
        float fr = nclamp(fresnel_dielectric(cos_theta(wo), 1, b.eta), 0.1f, 0.9f);
        if (rng.next_float() >= fr) {
            if (!btdf_sample(&wi)) return BSample();
            return finish(wi, T_GLOSSY | T_TRANSMISSION);
        }
        // inner MicrofacetReflectionBrdf<Dielectric, Gtr2>: pdf 0 when its sample fails
        if (!mf_reflect_sample_dir(wo, rng, wh_trans, &wi, &wh)) return BSample();
        if (td.pdf(wo, wh) / (4 * dot(wi, wh)) == 0) return BSample();
        return finish(wi, T_REFLECTION | T_GLOSSY);
    }
    float dw, tw, sw, cw;
    disney_weights(b, &dw, &tw, &sw, &cw);
    float u = rng.next_float();
    if (u < dw) {
        if (b.thin || b.sdist.black()) {
            if (!cosine_sample(wo, rng, &wi)) return BSample();
            if ((same_hemisphere(wo, wi) ? abs_cos_theta(wi) / PI_F : 0.0f) == 0) return BSample();
            return finish(wi, T_REFLECTION | T_DIFFUSE);
        }
        // scatting_distance > 0: SpecularTransmissionBtdf(1, 1, Eta).Sample
        // (SpecularTransmissionBtdf.cs:33-48), pdf scaled by the diffuse
        // weight, flagged for the BSSRDF (DisneyBsdf.cs:357-364)
        bool entering = cos_theta(wo) > 0;
        float ei = entering ? 1.0f : b.eta, et = entering ? b.eta : 1.0f;
        V3 n = entering ? V3(0, 0, 1) : V3(0, 0, -1);
        if (!refract(wo, n, ei / et, &wi) || cos_theta(wi) == 0.0f) return BSample();
        C3 ft = C3(1.0f) * (1.0f - fresnel_dielectric(cos_theta(wo), 1, b.eta));
        if (!imp) ft = ft * ((ei * ei) / (et * et));
        ft = ft / abs_cos_theta(wi);
        out.wi = wi; out.fr = ft; out.pdf = 1.0f * dw; out.type = T_TRANSMISSION | T_SPECULAR | T_SUBSURFACE;
        return out;
    }
    u -= dw;
    if (u < tw) {
        if (!btdf_sample(&wi)) return BSample();
        return finish(wi, T_GLOSSY | T_TRANSMISSION);
    }
    u -= tw;
    if (u < sw) {
        if (!mf_reflect_sample_dir(wo, rng, wh_spec, &wi, &wh)) return BSample();
        if (sd.pdf(wo, wh) / (4 * dot(wi, wh)) == 0) return BSample();
        return finish(wi, T_REFLECTION | T_GLOSSY);
    }
    // DisneyClearcoatBrdf.Sample
    if (cos_theta(wo) == 0) return BSample();
    float gloss = lerpf(b.cc_gloss, 0.1f, 0.001f);
    float ux = rng.next_float(), uy = rng.next_float();
    float a2 = sqr(gloss);
    float ct = std::sqrt(nmax(0, (1 - std::pow(a2, 1 - ux)) / (1 - a2)));
    float st = std::sqrt(nmax(0, 1 - ct * ct));
    float phi = 2 * PI_F * uy;
    wh = spherical_direction(st, ct, phi);
    if (!same_hemisphere(wo, wh)) wh = -wh;
    wi = reflect(-wo, wh);
    if (!same_hemisphere(wo, wi)) return BSample();
    if (disney_clearcoat_pdf(b, wo, wi) == 0) return BSample();
    return finish(wi, T_REFLECTION | T_GLOSSY);
}

static BSample material_sample(const cr_scene& s, const rd_material_desc& m, V3 wo, V2 uv, Rng& rng, bool imp = false) {
    BSample out;
    switch (m.type) {
    case RD_MAT_ROUGH_PLASTIC: {                                     // RoughPlastic.SampleImpl (:134-151)
        if (cos_theta(wo) <= 0) { if (!m.two_sided) return out; }
        MfDist d(m.dist, tex_at(s, m.tex[1], uv).r, tex_at(s, m.tex[2], uv).r);
        C3 r = tex_at(s, m.tex[0], uv);
        float p = m.ks / (m.kd + m.ks);
        V3 gwi(0.0f), gwh, dwi(0.0f);
        int gty = 0, dty = 0;
        auto swh = [&](V3 w, Rng& rr) { return d.sample_wh(w, rr); };
        if (mf_reflect_sample_dir(wo, rng, swh, &gwi, &gwh)) gty = T_REFLECTION | T_GLOSSY;
        else gwi = V3(0.0f);
        if (cosine_sample(wo, rng, &dwi)) dty = T_REFLECTION | T_DIFFUSE;
        else dwi = V3(0.0f);
        V3 wi;
        int ty;
        if (rng.next_float() < p) { wi = gwi; ty = gty; }
        else { wi = dwi; ty = dty; }
        out.wi = wi;
        out.pdf = plastic_pdf(d, m.kd, m.ks, m.two_sided != 0, wo, wi);
        out.fr = plastic_fr(d, r, m.kd, m.ks, m.eta_a, m.eta_b, m.two_sided != 0, wo, wi);
        out.type = ty;
        return out;
    }
    case RD_MAT_SUBSURFACE:                                          // Subsurface.Sample (Subsurface.cs:45-55)
    case RD_MAT_PERFECT_GLASS: {                                     // FresnelSpecularBsdf.Sample (:34-65)
        C3 R = tex_at(s, m.tex[0], uv), T = tex_at(s, m.tex[1], uv);
        float f = fresnel_dielectric(cos_theta(wo), m.eta_a, m.eta_b);
        // a subsurface material with a non-black scatting distance adds the
        // SubsurfaceScatting flag to whatever FresnelSpecularBsdf sampled
        const int sss = (m.type == RD_MAT_SUBSURFACE && !tex_at(s, m.tex[3], uv).black()) ? T_SUBSURFACE : 0;
        if (rng.next_float() < f) {
            if (cos_theta(wo) == 0) return out;
            V3 wi(-wo.x, -wo.y, wo.z);
            out.wi = wi; out.fr = f * R / abs_cos_theta(wi); out.pdf = f;
            out.type = T_REFLECTION | T_SPECULAR | sss;
            return out;
        }
        bool entering = cos_theta(wo) > 0;
        float ei = entering ? m.eta_a : m.eta_b, et = entering ? m.eta_b : m.eta_a;
        V3 n = entering ? V3(0, 0, 1) : V3(0, 0, -1);
        V3 wi;
        if (!refract(wo, n, ei / et, &wi) || cos_theta(wi) == 0.0f) return out;
        C3 ft = T * (1.0f - f);
        if (!imp) ft = ft * ((ei * ei) / (et * et));   // TransportMode.Radiance only (:58-61)
        ft = ft / abs_cos_theta(wi);
        out.wi = wi; out.fr = ft; out.pdf = 1 - f; out.type = T_TRANSMISSION | T_SPECULAR | sss;
        return out;
    }
    case RD_MAT_PERFECT_MIRROR: {                                    // PerfectMirror.Sample + SpecularReflectionBrdf
        if (cos_theta(wo) <= 0) { if (!m.two_sided) return out; }
        if (abs_cos_theta(wo) == 0) return out;
        V3 wi(-wo.x, -wo.y, wo.z);
        out.wi = wi; out.fr = tex_at(s, m.tex[0], uv) / abs_cos_theta(wi); out.pdf = 1.0f;
        out.type = T_REFLECTION | T_SPECULAR;
        return out;
    }
    case RD_MAT_DISNEY:
        return disney_sample(load_disney(s, m, uv), wo, rng, imp);
    case RD_MAT_DIFFUSE: {                                           // Diffuse.cs:39-49
        if (cos_theta(wo) <= 0) { if (!m.two_sided) return out; }
        C3 R = tex_at(s, m.tex[0], uv);
        float ux = rng.next_float(), uy = rng.next_float();         // Lambert.Sample :26-36
        V3 wi = normalize(square_to_cosine_hemisphere(ux, uy));
        if (wo.z < 0) wi.z *= -1;
        if (!same_hemisphere(wo, wi)) return out;
        out.wi = wi;
        out.pdf = same_hemisphere(wo, wi) ? abs_cos_theta(wi) / PI_F : 0.0f;
        out.fr = R / PI_F;
        out.type = T_REFLECTION | T_DIFFUSE;
        return out;
    }
    case RD_MAT_ROUGH_METAL: {                                       // RoughMetal.cs:87-112
        float rough = tex_at(s, m.tex[0], uv).r, an = tex_at(s, m.tex[1], uv).r;
        MfDist d(m.dist, rough, an);
        if (cos_theta(wo) <= 0) { if (!m.two_sided) return out; }
        // MicrofacetReflectionBrdf.Sample (MicrofacetReflectionBrdf.cs:50-59)
        if (cos_theta(wo) == 0) return out;
        V3 wh = d.sample_wh(wo, rng);
        if (dot(wo, wh) < 0) return out;
        V3 wi = reflect(-wo, wh);
        if (!same_hemisphere(wo, wi)) return out;
        out.pdf = d.pdf(wo, wh) / (4 * dot(wi, wh));
        out.fr = mf_reflect_fr(d, C3(m.eta[0], m.eta[1], m.eta[2]), C3(m.k[0], m.k[1], m.k[2]), wo, wi);
        out.wi = wi;
        out.type = T_REFLECTION | T_GLOSSY;
        return out;
    }
    case RD_MAT_ROUGH_GLASS: {                                       // RoughGlass.cs:240-297
        float rough = tex_at(s, m.tex[2], uv).r, an = tex_at(s, m.tex[3], uv).r;
        MfDist d(m.dist, rough, an);
        C3 r = tex_at(s, m.tex[0], uv), t = tex_at(s, m.tex[1], uv);
        V3 wh = d.sample_wh(wo, rng);
        float F = fresnel_dielectric(dot(wo, wh), m.eta_a, m.eta_b);
        if (rng.next_float() < F) {
            V3 wi = reflect(-wo, wh);
            if (!same_hemisphere(wo, wi)) return out;
            float D = d.D(wh);
            float G = d.G(wo, wi);
            C3 brdf = r * (D * G * F / (4.0f * cos_theta(wi) * cos_theta(wo)));
            float jac = 1.0f / (4.0f * abs_dot(wh, wi));
            float pdf = d.pdf(wo, wh) * jac * F;
            if (std::isinf(pdf) || std::isnan(pdf) || !brdf.valid()) return out;
            out.wi = wi; out.fr = brdf; out.pdf = pdf; out.type = T_REFLECTION | T_GLOSSY;
            return out;
        }
        bool entering = cos_theta(wo) > 0;
        float ei = entering ? m.eta_a : m.eta_b;
        float et = entering ? m.eta_b : m.eta_a;
        float eta = et / ei;
        V3 wi;
        if (refract(wo, wh, ei / et, &wi)) {
            if (dot(wo, wh) * dot(wi, wh) > 0) return out;
        } else {
            wi = reflect(-wo, wh);
            eta = 1;
        }
        float sd = dot(wo, wh) + eta * dot(wi, wh);
        float factor = imp ? 1.0f : 1 / eta;   // RoughGlass.cs:161
        float D = d.D(wh);
        float G = d.G(wo, wi);
        C3 btdf = (1 - F) * t *
                  std::fabs(D * G * eta * eta * abs_dot(wi, wh) * abs_dot(wo, wh) * factor * factor /
                            (cos_theta(wo) * cos_theta(wi) * sd * sd));
        float dwh = std::fabs((eta * eta * dot(wi, wh)) / (sd * sd));
        float pdf = d.pdf(wo, wh) * dwh * (1 - F);
        if (std::isinf(pdf) || std::isnan(pdf) || !btdf.valid()) return out;
        out.wi = wi; out.fr = btdf; out.pdf = pdf; out.type = T_TRANSMISSION | T_GLOSSY;
        return out;
    }
    default:
        return out;
    }
}

// ---------------------------------------------------------------- lights
struct LSample { V3 p, wi; float pdf = 0; C3 li; };

// Shape.Sample for the emitting primitive: Triangle.Sample (Triangle.cs:184-212),
// Sphere.Sample (Sphere.cs:87-102).
static void shape_sample(const cr_scene& s, int prim, Rng& rng, V3* P, V3* N, float* pdf) {
    const rd_prim_desc& pd = s.prims[prim];
    float ux = rng.next_float(), uy = rng.next_float();
    if (pd.kind == RD_PRIM_TRIANGLE) {
        int t = pd.index;
        float alpha = 1 - std::sqrt(1 - ux);
        float beta = uy * std::sqrt(1 - ux);
        V3 bary(alpha, beta, (1 - alpha - beta));
        V3 A = s.pa(t), B = s.pb(t), C = s.pc(t);
        *P = bary.x * A + bary.y * B + bary.z * C;
        if (s.has_n(t)) *N = normalize(bary.x * s.nrm[s.idx[3 * t]] + bary.y * s.nrm[s.idx[3 * t + 1]] + bary.z * s.nrm[s.idx[3 * t + 2]]);
        else *N = normalize(cross(B - A, C - A));
        *pdf = 1.0f / s.prim_area[prim];
    } else {
        const float* S = &s.sph[4 * pd.index];
        // Probability.SquareToUniformSphere (Probability.cs:62-71)
        float phi = ux * PI_F * 2;
        float theta = std::acos(1 - 2 * uy);
        V3 v(std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta));
        V3 n = normalize(v);
        *P = n * S[3] + V3(S[0], S[1], S[2]);
        *N = n;
        *pdf = 1 / s.prim_area[prim];
    }
}

// ---- InfiniteAreaLight (Lights/InfiniteAreaLight.cs)
// Le (:56-69): the world direction in model space, then equirectangular
// (phi = atan2(x, -z), theta = acos y) through the map's sampler.
static C3 env_le(const cr_scene& s, const EnvLight& e, V3 d) {
    V3 dir = normalize(xform_vector(d, e.w2m));
    float phi = std::atan2(dir.x, -dir.z);
    phi = (phi < 0) ? (phi + 2 * PI_F) : phi;
    float theta = std::acos(nclamp(dir.y, -1.0f, 1.0f));
    float u = phi / (2 * PI_F);
    float v = theta / PI_F;
    return s.tex[e.tex].sample(V2(u, v));
}
// PdfDir (:71-85)
static float env_pdf_dir(const EnvLight& e, V3 wi) {
    V3 dir = normalize(xform_vector(wi, e.w2m));
    float phi = std::atan2(dir.x, -dir.z);
    phi = (phi < 0) ? (phi + 2 * PI_F) : phi;
    float theta = std::acos(nclamp(dir.y, -1.0f, 1.0f));
    float sin_theta = std::sin(theta);
    if (sin_theta == 0.0f) return 0.0f;
    float u = phi / (2 * PI_F);
    float v = theta / PI_F;
    return e.dist.continuous_pdf(u, v) / (2 * PI_F * PI_F * sin_theta);
}
// SampleLightDir (:92-109): the direction is that of the sampled cell's
// corner (offset / size), the pdf the continuous one of the cell.
static bool env_sample_dir(const EnvLight& e, float u0, float u1, V3* wdir, float* pdf_out, V2* uv) {
    float pdf;
    int vj, vi;
    e.dist.sample_continuous(u0, u1, &pdf, &vj, &vi);
    if (pdf <= 0) return false;
    float y = vj / (float)e.h;
    float x = vi / (float)e.w;
    *uv = V2(x, y);
    float theta = y * PI_F;
    float phi = x * 2 * PI_F;
    float sin_theta = std::sin(theta);
    if (sin_theta == 0) return false;
    pdf /= (2 * PI_F * PI_F * sin_theta);
    // MathExt.SphericalCoordinates (MathExt.cs:94-102)
    float st = std::sin(theta), ct = std::cos(theta), sp = std::sin(phi), cp = std::cos(phi);
    V3 sph(cp * st, sp * st, ct);
    V3 dir(sph.y, sph.z, -sph.x);
    *wdir = normalize(xform_vector(dir, e.m2w));
    *pdf_out = pdf;
    return true;
}
// SampleLi (:111-119): the shadow target sits 20 world radii out.
static LSample env_sample_li(const cr_scene& s, const EnvLight& e, V3 refp, Rng& rng) {
    LSample out;
    float u0 = rng.next_float(), u1 = rng.next_float();     // Random.NextVec2: X then Y
    V3 wdir;
    float pdf;
    V2 uv;
    if (!env_sample_dir(e, u0, u1, &wdir, &pdf, &uv)) return out;
    out.li = s.tex[e.tex].sample(uv);
    out.p = wdir * e.radius * 20.0f + e.center;
    out.wi = normalize(out.p - refp);
    out.pdf = pdf;
    return out;
}

// Light.SampleLi dispatch
static LSample area_sample_li(const cr_scene& s, const rd_light_desc& L, V3 refp, Rng& rng);
static LSample light_sample_li(const cr_scene& s, int light, V3 refp, Rng& rng) {
    const rd_light_desc& L = s.lights[light];
    if (L.type == RD_LIGHT_INFINITE) return env_sample_li(s, s.env[s.env_of_light[light]], refp, rng);
    return area_sample_li(s, L, refp, rng);
}

// DiffuseAreaLight.SampleLi (Lights/DiffuseAreaLight.cs:48-68)
static LSample area_sample_li(const cr_scene& s, const rd_light_desc& L, V3 refp, Rng& rng) {
    LSample out;
    V3 P, N;
    float spdf;
    shape_sample(s, L.prim, rng, &P, &N, &spdf);
    V3 wi = P - refp;
    C3 le(L.le[0], L.le[1], L.le[2]);
    out.li = dot(N, -wi) > 0 ? le : C3(0.0f);
    if (length_sq(wi) == 0) {
        out.pdf = 0.0f;
    } else {
        wi = normalize(wi);
        out.pdf = spdf * length_sq(refp - P) / abs_dot(N, -wi);
        if (std::isinf(out.pdf)) out.pdf = 0.0f;
    }
    out.p = P;
    out.wi = wi;
    return out;
}
// DiffuseAreaLight.PdfLi (DiffuseAreaLight.cs:31-46): intersects the light's
// own shape only.
static float light_pdf_li(const cr_scene& s, int light, V3 refp, V3 wi) {
    const rd_light_desc& L = s.lights[light];
    if (L.type == RD_LIGHT_INFINITE) return env_pdf_dir(s.env[s.env_of_light[light]], wi);   // PdfLi :87-90
    Ray ray(refp, wi);
    Hit h;
    if (!prim_hit(s, L.prim, ray, &h.sp)) return 0.0f;
    h.T = h.sp.t; h.prim = L.prim;
    Inct hit = prim_inct(s, h, ray);
    float pdf = length_sq(refp - hit.P) / (abs_dot(hit.N(), -wi) * s.prim_area[L.prim]);
    if (std::isinf(pdf)) pdf = 0.0f;
    return pdf;
}
// Intersection.Le -> DiffuseAreaLight.L (DiffuseAreaLight.cs:26-29)
static C3 inct_le(const cr_scene& s, const Inct& it, V3 w) {
    int l = s.prims[it.prim].light;
    if (l < 0) return C3(0.0f);
    const rd_light_desc& L = s.lights[l];
    return dot(it.N(), w) > 0 ? C3(L.le[0], L.le[1], L.le[2]) : C3(0.0f);
}
// Scene.EvalAllInfiniteLights (Scene.cs:37-45)
static C3 eval_infinite(const cr_scene& s, const Ray& r) {
    C3 l(0.0f);
    for (size_t i = 0; i < s.env.size(); i++) l = l + env_le(s, s.env[i], r.d);
    return l;
}

// PathUtility.PowerHeuristic (Integrators/PathUtility.cs:5-9), in double
static inline float power_heuristic(int nf, float fp, int ng, float gp) {
    double f = nf * (double)fp, g = ng * (double)gp;
    return (float)((f * f) / (f * f + g * g));
}

// ---------------------------------------------------------------- path
static inline Ray spawn_ray(const Inct& it, V3 d) { return Ray(it.P, d, 0.0001f); }  // Intersection.cs:48-51

// PathTracer.Mis.EstimateDirect (Integrators/PathTracer.cs:309-399)
// BssrdfAdapter (Bssrdf.cs:203-243): the exit lobe at a BSSRDF's sampled
// point, (1 - Fr(cos wi)) / (c pi) with c = 1 - 2 FresnelMoment(1 / eta),
// times eta^2 in radiance mode; no hemisphere test in Fr, a cosine lobe on
// wo's side for Pdf / Sample.
static float fresnel_moment(float eta) {                       // BssrdfUtility.FresnelMoment (:247-267)
    float eta2 = eta * eta, eta3 = eta2 * eta, eta4 = eta2 * eta2, eta5 = eta2 * eta3;
    if (eta < 1)
        return 0.45966f - 1.73965f * eta + 3.37668f * eta2 - 3.904945f * eta3 + 2.49277f * eta4 - 0.68441f * eta5;
    return -4.61686f + 11.1136f * eta - 10.4646f * eta2 + 5.11455f * eta3 - 1.27198f * eta4 + 0.12746f * eta5;
}
static C3 adapter_fr(float eta, V3 wi) {
    float ci = 1 - 2 * fresnel_moment(1 / eta);
    float fr = fresnel_dielectric(cos_theta(wi), 1, eta);
    C3 f = C3((1 - fr) / (ci * PI_F));
    return f * (eta * eta);
}
static float adapter_pdf(V3 wo, V3 wi) { return same_hemisphere(wo, wi) ? abs_cos_theta(wi) / PI_F : 0.0f; }
static BSample adapter_sample(float eta, V3 wo, Rng& rng) {
    BSample out;
    float ux = rng.next_float(), uy = rng.next_float();
    V3 wi = normalize(square_to_cosine_hemisphere(ux, uy));
    if (wo.z < 0) wi.z *= -1;
    if (!same_hemisphere(wo, wi)) return out;
    out.wi = wi;
    out.pdf = adapter_pdf(wo, wi);
    out.fr = adapter_fr(eta, wi);
    out.type = T_REFLECTION | T_DIFFUSE;
    return out;
}

// adapter_eta: the BSSRDF adapter's eta in place of the surface material
// (EstimateDirect's bssrdfAdapter argument).
static C3 estimate_direct(const cr_scene& s, Rng& rng, int light, const Inct& it, uint64_t* st,
                          const float* adapter_eta = nullptr) {
    const rd_material_desc& m = s.mats[s.prims[it.prim].material];
    V3 wo = it.wr;
    C3 le(0.0f);
    LSample ls = light_sample_li(s, light, it.P, rng);
    if (ls.pdf > 0.0f && !ls.li.black()) {
        C3 fr;
        float spdf;
        if (adapter_eta) {
            fr = adapter_fr(*adapter_eta, it.sh.to_local(ls.wi));
            spdf = adapter_pdf(it.sh.to_local(wo), it.sh.to_local(ls.wi));
        } else {
            fr = material_fr(s, m, it.sh.to_local(wo), it.sh.to_local(ls.wi), it.UV);
            spdf = material_pdf(s, m, it.sh.to_local(wo), it.sh.to_local(ls.wi), it.UV);
        }
        if (!fr.black()) {
            if (!scene_is_occluded(s, it.P, ls.p, st)) {
                fr = fr * abs_cos_theta(it.sh.to_local(ls.wi));
                float w = power_heuristic(1, ls.pdf, 1, spdf);        // area lights are not delta
                le = le + fr * ls.li * w / ls.pdf;
            }
        }
    }
    BSample bs = adapter_eta ? adapter_sample(*adapter_eta, it.sh.to_local(wo), rng)
                             : material_sample(s, m, it.sh.to_local(wo), it.UV, rng);
    C3 fr = bs.fr * abs_cos_theta(bs.wi);
    float spdf = bs.pdf;
    bool sampled_specular = (bs.type & T_SPECULAR) != 0;
    if (!fr.black() && spdf > 0) {
        float w = 1;
        if (!sampled_specular) {
            float lpdf = light_pdf_li(s, light, it.P, it.sh.to_world(bs.wi));
            if (lpdf == 0.0f) return le;
            w = power_heuristic(1, spdf, 1, lpdf);
        }
        Ray to_light = spawn_ray(it, it.sh.to_world(bs.wi));
        Inct li_it;
        bool hit = scene_intersect(s, to_light, &li_it, st);
        C3 li(0);
        if (hit) {
            if (s.prims[li_it.prim].light == light) li = inct_le(s, li_it, it.sh.to_world(-bs.wi));
        } else {
            li = eval_infinite(s, to_light);
        }
        if (!li.black()) le = le + fr * li * w / spdf;
    }
    return le;
}

// PathTracer.Mis.SampleLightToEstimateDirect (PathTracer.cs:273-307)
static C3 sample_light_direct(const cr_scene& s, Rng& rng, const Inct& it, uint64_t* st,
                              const float* adapter_eta = nullptr) {
    C3 result;
    if (s.integ.strategy == RD_STRATEGY_UNIFORM) {
        if (s.lights.empty()) return result;                       // Scene.SampleLight (Scene.cs:47-57)
        int l = rng.next_int((int)s.lights.size());
        float lpdf = 1.0f / s.lights.size();
        if (lpdf > 0.0f) result = result + estimate_direct(s, rng, l, it, st, adapter_eta) / lpdf;
    } else {
        for (int l = 0; l < (int)s.lights.size(); l++)
            result = result + estimate_direct(s, rng, l, it, st, adapter_eta);
    }
    return result;
}

// ---------------------------------------------------------------- BSSRDF
// BssrdfUtility.ScattingLUT (Bssrdf.cs:289-): 2048 radii r_i with
// 1 - e^(-r)/4 - 3 e^(-r/3)/4 = i / 2047, the inverse CDF of the
// normalised-diffusion radial pdf in units of d; solved here by bisection in
// double.  i = 2047 has no root; the reference stores 111.68667351 there, a
// value no accepted sample reaches (rMax is the 0.996 quantile).
static const std::vector<float>& scatter_lut() {
    static const std::vector<float> lut = [] {
        const int n = 2048;
        std::vector<float> t(n);
        for (int i = 0; i < n - 1; i++) {
            const double target = (double)i / (n - 1);
            double lo = 0, hi = 64;
            for (int it = 0; it < 200; it++) {
                const double mid = 0.5 * (lo + hi);
                const double f = 1 - 0.25 * std::exp(-mid) - 0.75 * std::exp(-mid / 3);
                if (f < target) lo = mid;
                else hi = mid;
            }
            t[i] = (float)(0.5 * (lo + hi));
        }
        t[n - 1] = 111.68667351f;
        return t;
    }();
    return lut;
}

// NormalizedDiffusionRadialProfile (Bssrdf.cs:64-124): A and d per
// channel, D = d / (100 (A - 0.33)^4 + 3.5).
struct Profile {
    C3 A, D;
    Profile(C3 a, C3 d) : A(a) {
        C3 t = a - C3(0.33f);
        C3 sc = t * t * t * t * 100.0f + C3(3.5f);
        D = C3(d.r / sc.r, d.g / sc.g, d.b / sc.b);
    }
    float dch(int c) const { return c == 0 ? D.r : (c == 1 ? D.g : D.b); }
    float pdf_sr(int c, float distance) const {                 // PdfSr
        float chd = dch(c);
        distance /= chd;
        return 0.25f * (std::exp(-distance) + std::exp(-distance / 3)) / chd;
    }
    C3 sr(float distance) const {                               // Sr
        auto one = [&](float a, float d) {
            return a * (std::exp(-distance / d) + std::exp(-distance / (3.0f * d))) / (8 * PI_F * d);
        };
        return C3(one(A.r, D.r), one(A.g, D.g), one(A.b, D.b));
    }
    float sample_sr(int c, float u) const {                     // SampleSr: the radius
        const std::vector<float>& lut = scatter_lut();
        float global = u * (float)(lut.size() - 1);
        int low = std::min((int)global, (int)lut.size() - 2);
        float local = global - low;
        float r = lerpf(local, lut[low], lut[low + 1]);
        return r * dch(c);
    }
};

struct SssSample {
    bool ok = false;
    C3 S;
    float pdf = 0;
    Inct pi;
};

// SeparableBssrdf.SamplePi / PdfPi (Bssrdf.cs:139-200) at po, for the
// material `mat` (Scene.Intersect's surfaces compared by material index).
static SssSample sss_sample_pi(const cr_scene& s, const Inct& po, int mat, const Profile& prof, float eta, Rng& rng,
                               uint64_t* st) {
    SssSample out;
    int channel = rng.next_int(3);
    Frame proj = po.sh;
    float u = rng.next_float();
    if (u < 0.5f) {
        proj = po.sh;
    } else if (u < 0.75f) {
        proj.X = po.sh.X; proj.Y = po.sh.Z; proj.Z = po.sh.Y;     // Coordinate(X, Z, Y)
    } else {
        proj.X = po.sh.Z; proj.Y = po.sh.Y; proj.Z = po.sh.X;     // Coordinate(Z, Y, X)
    }
    float r = prof.sample_sr(channel, rng.next_float());
    float r_max = prof.sample_sr(channel, 0.996f);
    if (r <= 0 || r > r_max) return out;
    float phi = 2 * PI_F * rng.next_float();
    float hl = std::sqrt(nmax(0, sqr(r_max) - sqr(r)));
    float sp = std::sin(phi), cp = std::cos(phi);
    V3 start(r * cp, r * sp, hl);
    float len = 2 * hl;
    Ray ray(po.P + proj.to_world(start), -proj.Z, 0.0001f, len);
    std::vector<Inct> hits;
    while (true) {
        if (ray.tmin >= ray.tmax) break;
        Inct it;
        if (!scene_intersect(s, ray, &it, st)) break;
        if (s.prims[it.prim].material == mat) hits.push_back(it);
        float adv = it.T + 0.0001f;
        ray = Ray(ray.at(adv), ray.d, ray.tmin, ray.tmax - adv);
    }
    if (hits.empty()) return out;
    int sel = rng.next_int((int)hits.size());
    const Inct& e = hits[sel];
    // PdfPi
    V3 ld = po.sh.to_local(po.P - e.P);
    V3 ln = po.sh.to_local(e.sh.Z);
    float rproj[3] = {std::sqrt(ld.y * ld.y + ld.z * ld.z), std::sqrt(ld.z * ld.z + ld.x * ld.x),
                      std::sqrt(ld.x * ld.x + ld.y * ld.y)};
    const float axis_pdf[3] = {0.25f, 0.25f, 0.5f};
    float pdf = 0;
    for (int axis = 0; axis < 3; axis++) {
        float lna = std::fabs(axis == 0 ? ln.x : (axis == 1 ? ln.y : ln.z));
        for (int c = 0; c < 3; c++) pdf += lna * prof.pdf_sr(c, rproj[axis]) * axis_pdf[axis];
    }
    pdf = pdf * (1.0f / 3.0f);
    out.pdf = pdf / (2 * PI_F) / (float)hits.size();
    float fro = 1 - fresnel_dielectric(cos_theta(po.sh.to_local(po.wr)), 1, eta);
    out.S = prof.sr(length(po.P - e.P)) * fro;
    out.pi = e;
    out.ok = true;
    return out;
}

// The BSSRDF of the material at po: a Disney material's profile is
// (BaseColor * DiffuseWeight, ScattingDistance) with its Eta
// (DisneyBsdf.GetBssrdf, DisneyBsdf.cs:514-518); a subsurface material's
// (A, ScattingDistance) with EtaB / EtaA (Subsurface.cs:57-64).
static Profile sss_profile(const cr_scene& s, const rd_material_desc& m, V2 uv, float* eta) {
    if (m.type == RD_MAT_DISNEY) {
        Disney b = load_disney(s, m, uv);
        *eta = b.eta;
        return Profile(b.color * (1 - b.metallic), b.sdist);
    }
    *eta = m.eta_b / m.eta_a;
    return Profile(tex_at(s, m.tex[2], uv), tex_at(s, m.tex[3], uv));
}

// PathTracer.Mis (PathTracer.cs:192-270)
static C3 li_mis(const cr_scene& s, Ray ray, Rng& rng, uint64_t* st) {
    C3 radiance(0.0f), coeff(1.0f);
    bool specular_path = true;
    const rd_integrator_desc& I = s.integ;
    for (int bounces = 0;; bounces++) {
        Inct it;
        bool hit = scene_intersect(s, ray, &it, st);
        V3 wo = -ray.d;
        if (bounces == 0 || specular_path) {
            if (hit) radiance = radiance + coeff * inct_le(s, it, wo);
            else radiance = radiance + coeff * eval_infinite(s, ray);
        }
        if (!hit) break;
        if (I.max_depth != -1 && bounces >= I.max_depth) break;
        int mat = s.prims[it.prim].material;
        if (mat < 0) {
            ray = Ray(it.P, ray.d, ray.tmin);
            bounces--;
            continue;
        }
        if (st) st[3]++;
        radiance = radiance + coeff * sample_light_direct(s, rng, it, st);
        BSample bs = material_sample(s, s.mats[mat], it.sh.to_local(wo), it.UV, rng);
        specular_path = (bs.type & T_SPECULAR) != 0;
        if (bs.pdf > 0.0f) coeff = coeff * (bs.fr * abs_cos_theta(bs.wi) / bs.pdf);
        else break;
        ray = spawn_ray(it, it.sh.to_world(bs.wi));
        if ((bs.type & T_SUBSURFACE) && (bs.type & T_TRANSMISSION)) {
            // the BSSRDF branch (PathTracer.cs:244-259): a point pi on the same
            // material, its direct light through the adapter, the adapter's
            // exit direction
            float eta;
            Profile prof = sss_profile(s, s.mats[mat], it.UV, &eta);
            SssSample sss = sss_sample_pi(s, it, mat, prof, eta, rng, st);
            if (!sss.ok || sss.pdf == 0 || sss.S.black()) break;
            coeff = coeff * (sss.S / sss.pdf);
            Inct pi = sss.pi;
            pi.T = length(it.P - pi.P);
            pi.prim = it.prim;                                  // Intersection(.., inct.Shape, ..)
            radiance = radiance + coeff * sample_light_direct(s, rng, pi, st, &eta);
            BSample ns = adapter_sample(eta, pi.sh.to_local(pi.wr), rng);
            if (ns.pdf == 0 || ns.fr.black()) break;
            coeff = coeff * (ns.fr * abs_cos_theta(ns.wi) / ns.pdf);
            specular_path = (ns.type & T_SPECULAR) != 0;
            ray = spawn_ray(pi, pi.sh.to_world(ns.wi));
        }
        if (bounces > I.min_depth) {
            float q = nmin(coeff.max_elem(), I.rr_threshold);
            if (rng.next_float() > q) break;
            coeff = coeff / q;
        }
    }
    return radiance;
}

// PathTracer.NextEventEstimation (PathTracer.cs:95-190)
static C3 li_nee(const cr_scene& s, Ray ray, Rng& rng, uint64_t* st) {
    C3 radiance(0.0f), coeff(1.0f);
    bool specular_path = true;
    const rd_integrator_desc& I = s.integ;
    for (int bounces = 0;; bounces++) {
        Inct it;
        bool hit = scene_intersect(s, ray, &it, st);
        V3 wo = -ray.d;
        if (bounces == 0 || specular_path) {
            if (hit) radiance = radiance + coeff * inct_le(s, it, wo);
            else radiance = radiance + coeff * eval_infinite(s, ray);
        }
        if (!hit) break;
        if (I.max_depth != -1 && bounces >= I.max_depth) break;
        int mat = s.prims[it.prim].material;
        if (mat < 0) { ray = Ray(it.P, ray.d, ray.tmin); bounces--; continue; }
        if (st) st[3]++;
        const rd_material_desc& m = s.mats[mat];
        auto estimate_light = [&](int l) -> C3 {                   // EstimateLight :117-129
            LSample ls = light_sample_li(s, l, it.P, rng);
            if (ls.pdf <= 0.0f) return C3(0);
            if (scene_is_occluded(s, it.P, ls.p, st)) return C3(0.0f);
            C3 fr = material_fr(s, m, it.sh.to_local(wo), it.sh.to_local(ls.wi), it.UV);
            return fr * ls.li * abs_cos_theta(it.sh.to_local(ls.wi)) / ls.pdf;
        };
        if (I.strategy == RD_STRATEGY_UNIFORM) {
            if (!s.lights.empty()) {
                int l = rng.next_int((int)s.lights.size());
                float lpdf = 1.0f / s.lights.size();
                if (lpdf > 0.0f) radiance = radiance + coeff * estimate_light(l) / lpdf;
            }
        } else {
            for (int l = 0; l < (int)s.lights.size(); l++) radiance = radiance + coeff * estimate_light(l);
        }
        BSample bs = material_sample(s, m, it.sh.to_local(wo), it.UV, rng);
        specular_path = (bs.type & T_SPECULAR) != 0;
        if (bs.pdf > 0.0f) coeff = coeff * (bs.fr * abs_cos_theta(bs.wi) / bs.pdf);
        else break;
        ray = spawn_ray(it, it.sh.to_world(bs.wi));
        if (bounces > I.min_depth) {
            float q = nmin(coeff.max_elem(), I.rr_threshold);
            if (rng.next_float() > q) break;
            coeff = coeff / q;
        }
    }
    return radiance;
}

// PathTracer.OnlySampleBxdf (PathTracer.cs:43-93)
static C3 li_bxdf(const cr_scene& s, Ray ray, Rng& rng, uint64_t* st) {
    C3 radiance(0.0f), coeff(1.0f);
    const rd_integrator_desc& I = s.integ;
    for (int bounces = 0;; bounces++) {
        Inct it;
        if (!scene_intersect(s, ray, &it, st)) {
            radiance = radiance + coeff * eval_infinite(s, ray);
            break;
        }
        if (I.max_depth != -1 && bounces >= I.max_depth) break;
        int mat = s.prims[it.prim].material;
        if (mat < 0) { ray = Ray(it.P, ray.d, ray.tmin); bounces--; continue; }
        if (st) st[3]++;
        V3 wo = -ray.d;
        if (s.prims[it.prim].light >= 0) radiance = radiance + coeff * inct_le(s, it, wo);
        BSample bs = material_sample(s, s.mats[mat], it.sh.to_local(-ray.d), it.UV, rng);
        if (bs.pdf > 0.0f) coeff = coeff * (bs.fr * abs_cos_theta(bs.wi) / bs.pdf);
        else break;
        ray = spawn_ray(it, it.sh.to_world(bs.wi));
        if (bounces > I.min_depth) {
            float q = nmin(coeff.max_elem(), I.rr_threshold);
            if (rng.next_float() > q) break;
            coeff = coeff / q;
        }
    }
    return radiance;
}

// ---------------------------------------------------------------- media
static Homo make_homo(const rd_medium_desc& m) {
    Homo h;
    h.sa = C3(m.sigma_a[0], m.sigma_a[1], m.sigma_a[2]) * m.scale;
    h.ss = C3(m.sigma_s[0], m.sigma_s[1], m.sigma_s[2]) * m.scale;
    h.st = h.sa + h.ss;
    h.g = m.g;
    return h;
}
// Homogeneous.Tr (:22-30): exp(sigma_t * (MinT - MaxT)) per channel
static C3 medium_tr(const Homo& m, float tmin, float tmax) {
    float neg = tmin - tmax;
    C3 tr;
    tr.r = m.st.r == 0 ? 1.0f : std::exp(m.st.r * neg);
    tr.g = m.st.g == 0 ? 1.0f : std::exp(m.st.g * neg);
    tr.b = m.st.b == 0 ? 1.0f : std::exp(m.st.b * neg);
    return tr;
}
struct MSample {
    bool medium = false;
    V3 p, wo;
    C3 tr;
    float t = 0;
};
// Homogeneous.Sample (:32-61): spectral free flight, channel chosen uniformly
static MSample medium_sample(const Homo& m, const Ray& r, Rng& rng) {
    MSample out;
    int ch = rng.next_int(3);
    float dens = ch == 0 ? m.st.r : (ch == 1 ? m.st.g : m.st.b);
    float sd = -std::log(1 - rng.next_float()) / dens;
    float dist = r.tmax - r.tmin;
    if (sd < dist) {
        float t = r.tmin + sd;
        out.p = r.o + t * r.d;
        C3 tr(std::exp(-m.st.r * sd), std::exp(-m.st.g * sd), std::exp(-m.st.b * sd));
        float pdf = (tr.r * m.st.r + tr.g * m.st.g + tr.b * m.st.b) / 3;
        if (pdf == 0) pdf = 0.0001f;
        out.medium = true;
        out.wo = -r.d;
        out.tr = tr * m.ss / pdf;
        out.t = t;
    } else {
        C3 tr(std::exp(-m.st.r * dist), std::exp(-m.st.g * dist), std::exp(-m.st.b * dist));
        float pdf = (tr.r + tr.g + tr.b) / 3;
        if (pdf == 0) pdf = 0.0001f;
        out.tr = tr / pdf;
    }
    return out;
}
// PhaseFunctionUtility.HenyeyGreenstein (IPhaseFunction.cs:39-44)
static inline float hg_eval(float cos_t, float g) {
    float denom = 1 + g * g + 2 * g * cos_t;
    return (1 / (4 * PI_F)) * (1 - g * g) / (denom * std::sqrt(denom));
}
static inline float hg_g(const Homo& m) { return nclamp(m.g, -0.999f, 0.999f); }     // :51-54
static inline float hg_p(const Homo& m, V3 wo, V3 wi) { return hg_eval(dot(wo, wi), hg_g(m)); }
struct PSample { float p; V3 wi; };
// HenyeyGreenstein.SampleWi (:62-81)
static PSample hg_sample(const Homo& m, V3 wo, Rng& rng) {
    float G = hg_g(m);
    float x = rng.next_float(), y = rng.next_float();
    float cos_t;
    if (std::fabs(G) < 0.001f) {
        cos_t = 1 - 2 * x;
    } else {
        float sqr = (1 - G * G) / (1 + G - 2 * G * x);
        cos_t = -(1 + G * G - sqr * sqr) / (2 * G);
    }
    float sin_t = std::sqrt(nmax(0.0f, 1 - cos_t * cos_t));
    float phi = 2 * PI_F * y;
    Frame f(wo);
    float sp = std::sin(phi), cp = std::cos(phi);
    PSample out;
    out.wi = sin_t * cp * f.X + sin_t * sp * f.Y + cos_t * f.Z;     // MathExt.SphericalDirection
    out.p = hg_eval(cos_t, G);
    return out;
}
// Intersection.GetMedium (Intersection.cs:68-71)
static inline int inct_medium(const cr_scene& s, const Inct& it, V3 w) {
    const rd_prim_desc& pd = s.prims[it.prim];
    return dot(w, it.N()) > 0 ? pd.medium_outside : pd.medium_inside;
}
// Scene.Transmittance (Scene.cs:72-103): closest-hit segments toward `to`;
// a surface blocks, medium boundaries are crossed.  Tr covers each whole
// segment [1e-4, dis - 1e-4].
static C3 scene_transmittance(const cr_scene& s, V3 from, V3 to, int env, uint64_t* st) {
    C3 tr(1.0f);
    V3 start = from;
    V3 dir = normalize(to - start);
    while (true) {
        V3 vec = to - start;
        float dis = length(vec);
        if (dis < 0.0001f) break;
        Ray ray(start, dir, 0.0001f, dis - 0.0001f);
        Inct it;
        bool any = scene_intersect(s, ray, &it, st);
        if (any && s.prims[it.prim].material >= 0) return C3(0.0f);
        if (env >= 0) tr = tr * medium_tr(s.homo[env], ray.tmin, ray.tmax);
        if (!any) break;
        start = it.P;
        env = inct_medium(s, it, dir);
    }
    return tr;
}
// Scene.IntersectTr (Scene.cs:105-131).  The medium segment before a hit
// ends at |ray.At(t)| (the length of the hit's position vector), as written.
static bool scene_intersect_tr(const cr_scene& s, Ray ray, int env, Inct* out, C3* tr_out, uint64_t* st) {
    C3 tr(1.0f);
    while (true) {
        bool any = scene_intersect(s, ray, out, st);
        if (!any) {
            if (env >= 0) tr = tr * medium_tr(s.homo[env], ray.tmin, ray.tmax);
            *tr_out = tr;
            return false;
        }
        if (env >= 0) tr = tr * medium_tr(s.homo[env], ray.tmin, length(ray.o + out->T * ray.d));
        if (s.prims[out->prim].material >= 0) {
            *tr_out = tr;
            return true;
        }
        ray = Ray(out->P, ray.d, 0.001f);
        env = inct_medium(s, *out, ray.d);
    }
}

// VolumetricPathTracer.EstimateDirectSurface (Integrators/VolumetricPathTracer.cs:151-219)
static C3 estimate_direct_surface(const cr_scene& s, Rng& rng, int light, const Inct& it, int env, uint64_t* st) {
    const rd_material_desc& m = s.mats[s.prims[it.prim].material];
    C3 le(0.0f);
    V3 wo = it.wr;
    LSample ls = light_sample_li(s, light, it.P, rng);
    if (ls.pdf > 0.0f && !ls.li.black()) {
        C3 fr = material_fr(s, m, it.sh.to_local(wo), it.sh.to_local(ls.wi), it.UV);
        float spdf = material_pdf(s, m, it.sh.to_local(wo), it.sh.to_local(ls.wi), it.UV);
        if (!fr.black()) {
            C3 li = ls.li * scene_transmittance(s, it.P, ls.p, env, st);
            if (!li.black()) {
                fr = fr * abs_cos_theta(it.sh.to_local(ls.wi));
                float w = power_heuristic(1, ls.pdf, 1, spdf);
                le = le + fr * li * w / ls.pdf;
            }
        }
    }
    BSample bs = material_sample(s, m, it.sh.to_local(wo), it.UV, rng);
    C3 fr = bs.fr * abs_cos_theta(bs.wi);
    float spdf = bs.pdf;
    bool sampled_specular = (bs.type & T_SPECULAR) != 0;
    if (!fr.black() && spdf > 0) {
        float w = 1;
        if (!sampled_specular) {
            float lpdf = light_pdf_li(s, light, it.P, it.sh.to_world(bs.wi));
            if (lpdf == 0.0f) return le;
            w = power_heuristic(1, spdf, 1, lpdf);
        }
        Ray to_light = spawn_ray(it, it.sh.to_world(bs.wi));
        Inct li_it;
        C3 tr;
        bool hit = scene_intersect_tr(s, to_light, env, &li_it, &tr, st);
        C3 li(0);
        if (hit) {
            if (s.prims[li_it.prim].light == light) li = inct_le(s, li_it, it.sh.to_world(-bs.wi));
        } else {
            li = eval_infinite(s, to_light);
        }
        if (!li.black()) le = le + fr * li * tr * w / spdf;
    }
    return le;
}

// VolumetricPathTracer.EstimateDirectMedium (VolumetricPathTracer.cs:252-314)
static C3 estimate_direct_medium(const cr_scene& s, Rng& rng, int light, const MSample& ms, int env, uint64_t* st) {
    const Homo& md = s.homo[env];
    C3 le(0.0f);
    LSample ls = light_sample_li(s, light, ms.p, rng);
    if (ls.pdf > 0.0f && !ls.li.black()) {
        float p = hg_p(md, ms.wo, ls.wi);
        float spdf = p;
        if (p != 0.0f) {
            C3 li = ls.li * scene_transmittance(s, ms.p, ls.p, env, st);
            C3 f(p);
            if (!li.black() && !f.black()) {
                float w = power_heuristic(1, ls.pdf, 1, spdf);
                le = le + f * li * w / ls.pdf;
            }
        }
    }
    PSample ps = hg_sample(md, ms.wo, rng);
    C3 f(ps.p);
    float spdf = ps.p;
    if (spdf > 0) {
        float lpdf = light_pdf_li(s, light, ms.p, ps.wi);
        if (lpdf == 0.0f) return le;
        float w = power_heuristic(1, spdf, 1, lpdf);
        Ray to_light(ms.p, ps.wi);                       // Ray3F default MinT = float.Epsilon
        Inct li_it;
        C3 tr;
        bool hit = scene_intersect_tr(s, to_light, env, &li_it, &tr, st);
        C3 li(0);
        if (hit) {
            if (s.prims[li_it.prim].light == light) li = inct_le(s, li_it, -ps.wi);
        } else {
            li = eval_infinite(s, to_light);
        }
        if (!li.black()) le = le + f * li * tr * w / spdf;
    }
    return le;
}

// SampleLightToEstimateDirectSurface / ...Medium (VolumetricPathTracer.cs:121-147, 221-247)
template <class F>
static C3 sample_light_vol(const cr_scene& s, Rng& rng, F&& estimate) {
    C3 result(0.0f);
    if (s.integ.strategy == RD_STRATEGY_UNIFORM) {
        if (s.lights.empty()) return result;
        int l = rng.next_int((int)s.lights.size());
        float lpdf = 1.0f / s.lights.size();
        if (lpdf > 0.0f) result = result + estimate(l) / lpdf;
    } else {
        for (int l = 0; l < (int)s.lights.size(); l++) result = result + estimate(l);
    }
    return result;
}

// VolumetricPathTracer.Li (VolumetricPathTracer.cs:28-119)
static C3 li_vol(const cr_scene& s, Ray ray, Rng& rng, uint64_t* st) {
    C3 radiance(0.0f), coeff(1.0f);
    bool specular_path = true;
    int env = s.global_medium;
    const rd_integrator_desc& I = s.integ;
    for (int bounces = 0;; bounces++) {
        V3 wo = -ray.d;
        if (ray.tmax <= ray.tmin) break;
        Inct it;
        bool hit = scene_intersect(s, ray, &it, st);
        MSample ms;
        if (env >= 0) {
            Ray real(ray.o, ray.d, ray.tmin, hit ? it.T : FLT_MAX);
            ms = medium_sample(s.homo[env], real, rng);
            coeff = coeff * ms.tr;
        }
        if (coeff.black()) break;
        if (ms.medium) {
            if (bounces >= I.max_depth) break;          // as written: also when max_depth is -1
            if (st) st[3]++;
            int menv = env;
            radiance = radiance + coeff * sample_light_vol(s, rng, [&](int l) {
                return estimate_direct_medium(s, rng, l, ms, menv, st);
            });
            PSample ps = hg_sample(s.homo[menv], wo, rng);
            ray = Ray(ms.p, ps.wi, ray.tmin);
            specular_path = false;
        } else {
            if (bounces == 0 || specular_path) {
                if (hit) radiance = radiance + coeff * inct_le(s, it, wo);
                else radiance = radiance + coeff * eval_infinite(s, ray);
            }
            if (!hit) break;
            if (I.max_depth != -1 && bounces >= I.max_depth) break;
            int mat = s.prims[it.prim].material;
            if (mat < 0) {
                ray = Ray(it.P, ray.d, ray.tmin);
                env = inct_medium(s, it, ray.d);
                bounces--;
                continue;
            }
            if (st) st[3]++;
            int senv = env;
            radiance = radiance + coeff * sample_light_vol(s, rng, [&](int l) {
                return estimate_direct_surface(s, rng, l, it, senv, st);
            });
            BSample bs = material_sample(s, s.mats[mat], it.sh.to_local(wo), it.UV, rng);
            specular_path = (bs.type & T_SPECULAR) != 0;
            if (bs.pdf > 0.0f) coeff = coeff * (bs.fr * abs_cos_theta(bs.wi) / bs.pdf);
            else break;
            V3 next = it.sh.to_world(bs.wi);
            ray = Ray(it.P, next, ray.tmin);
            env = inct_medium(s, it, next);
        }
        if (bounces > I.min_depth) {
            float q = nmin(coeff.max_elem(), I.rr_threshold);
            if (rng.next_float() > q) break;
            coeff = coeff / q;
        }
    }
    return radiance;
}

// ---------------------------------------------------------------- adjoint
// CPU reference for rd_adjoint: path-replay backpropagation (Vicini, Speierer,
// Jakob 2021) of dLoss/dradiance to the albedo texture of "diffuse" /
// "rough_plastic" and the roughness texture of "rough_plastic", with sampling,
// MIS weights and Russian roulette held fixed.  Diffuse and plastic values are
// linear in R per channel, so a term C with f(R) has dC_c/dR_c = C_c / R_c;
// plastic roughness rho scales both alphas by s = rho^2 and only the glossy
// lobe's D * G depends on it: dC/drho = C * (glossy / f) * s dln(DG)/ds * 2 / rho.
static float sdlog_D(const MfDist& m, V3 wh) {                 // s dln(D)/ds
    float t2 = tan2_theta(wh);
    if (std::isinf(t2)) return 0.0f;
    float e = (cos2_phi(wh) / (m.ax * m.ax) + sin2_phi(wh) / (m.ay * m.ay)) * t2;
    if (m.kind == 0) return 2 * e - 2;                          // D ~ exp(-e) / (ax ay)
    return -2 * (1 - e) / (1 + e);                              // D ~ 1 / (ax ay (1 + e)^2)
}
static float sdlambda(const MfDist& m, V3 w) {                  // s dLambda/ds
    float at = std::fabs(tan_theta(w));
    if (std::isinf(at)) return 0.0f;
    float alpha = std::sqrt(cos2_phi(w) * m.ax * m.ax + sin2_phi(w) * m.ay * m.ay);
    if (m.kind == 0) {
        float a = 1 / (alpha * at);
        if (a >= 1.6f) return 0.0f;
        float num = 1 - 1.259f * a + 0.396f * a * a, den = 3.535f * a + 2.181f * a * a;
        float dnum = -1.259f + 0.792f * a, dden = 3.535f + 4.362f * a;
        return -a * (dnum * den - num * dden) / (den * den);   // a ~ 1/s
    }
    float k = (alpha * at) * (alpha * at);                      // k ~ s^2
    return k / (2 * std::sqrt(1 + k));
}
// dC/drho / C for a term with f(wo, wi) of material m at uv (0 unless rough_plastic).
static float rho_ratio(const cr_scene& s, const rd_material_desc& m, V2 uv, V3 wo, V3 wi) {
    if (m.type != RD_MAT_ROUGH_PLASTIC) return 0.0f;
    float rho = tex_at(s, m.tex[1], uv).r;
    if (!(rho > 0)) return 0.0f;
    bool two = m.two_sided != 0;
    if (cos_theta(wo) <= 0 || cos_theta(wi) <= 0) { if (!two) return 0.0f; }
    if (!same_hemisphere(wo, wi)) return 0.0f;
    MfDist d(m.dist, rho, tex_at(s, m.tex[2], uv).r);
    float dw = m.kd / PI_F, gw = 0.0f, sd = 0.0f;
    float cto = abs_cos_theta(wo), cti = abs_cos_theta(wi);
    V3 wh = wi + wo;
    if (!(cti == 0 || cto == 0) && !(wh.x == 0 && wh.y == 0 && wh.z == 0)) {
        wh = normalize(wh);
        float F = fresnel_dielectric(dot(wi, faceforward(wh, V3(0, 0, 1))), m.eta_a, m.eta_b);
        float D = d.D(wh), G = d.G(wo, wi);
        gw = m.ks * std::fabs((D * G * F) / (4 * cti * cto));
        sd = sdlog_D(d, wh) - G * (sdlambda(d, wo) + sdlambda(d, wi));
    }
    float f = dw + gw;
    if (!(f > 0)) return 0.0f;
    return (gw / f) * sd * 2 / rho;
}

struct AdjGrads {
    std::vector<std::vector<double>> g;   // per texture: w*h*3 (3 for a constant); empty = not requested
};
// Adds dLoss/dR (RGB) and dLoss/drho at material `mi`'s lookups at uv.
static void adj_scatter(const cr_scene& s, int mi, V2 uv, C3 gr, float grho, AdjGrads& G) {
    const rd_material_desc& m = s.mats[mi];
    auto add = [&](int t, int ch, double v) {
        if (t < 0 || G.g[t].empty()) return;
        const Tex& tx = s.tex[t];
        if (tx.w == 0) { G.g[t][ch] += v; return; }
        // TextureSampler.Sample's texel reads (TextureSampler.cs:28-72)
        auto wrapf = [&](float a) { return tx.wrap == 0 ? nclamp(a, 0.0f, 1.0f) : nclamp(a - std::floor(a), 0.0f, 1.0f); };
        float u = wrapf(uv.x), v2 = wrapf(uv.y);
        float fu = u * (tx.w - 1), fv = v2 * (tx.h - 1);
        int pu = (int)fu, pv = (int)fv;
        if (tx.filter == 0) { G.g[t][((size_t)pu * tx.h + pv) * 3 + ch] += v; return; }
        int apu = iclamp(pu + ((fu > pu + 0.5f) ? 1 : -1), 0, tx.w - 1);
        int apv = iclamp(pv + ((fv > pv + 0.5f) ? 1 : -1), 0, tx.h - 1);
        float du = nmin(std::fabs(fu - pu - 0.5f), 1.0f), dv = nmin(std::fabs(fv - pv - 0.5f), 1.0f);
        G.g[t][((size_t)pu * tx.h + pv) * 3 + ch] += v * ((1 - du) * (1 - dv));
        G.g[t][((size_t)apu * tx.h + pv) * 3 + ch] += v * (du * (1 - dv));
        G.g[t][((size_t)pu * tx.h + apv) * 3 + ch] += v * ((1 - du) * dv);
        G.g[t][((size_t)apu * tx.h + apv) * 3 + ch] += v * (du * dv);
    };
    add(m.tex[0], 0, gr.r);
    add(m.tex[0], 1, gr.g);
    add(m.tex[0], 2, gr.b);
    if (m.type == RD_MAT_ROUGH_PLASTIC && grho != 0.0f) add(m.tex[1], 0, grho);
}

// One replay of PathTracer.Li (PathTracer.cs:43-270) with the same random
// numbers as pixel_sample, propagating delta = dLoss/dLi to the textures.
static void li_adjoint(const cr_scene& s, Ray ray, Rng& rng, C3 ltot, C3 delta, AdjGrads& G) {
    C3 radiance(0.0f), coeff(1.0f);
    bool specular_path = true;
    const rd_integrator_desc& I = s.integ;
    const int method = I.method;
    for (int bounces = 0;; bounces++) {
        Inct it;
        bool hit = scene_intersect(s, ray, &it, nullptr);
        V3 wo = -ray.d;
        if (method == RD_METHOD_BXDF) {
            if (!hit) { radiance = radiance + coeff * eval_infinite(s, ray); break; }
        } else {
            if (bounces == 0 || specular_path) {
                if (hit) radiance = radiance + coeff * inct_le(s, it, wo);
                else radiance = radiance + coeff * eval_infinite(s, ray);
            }
            if (!hit) break;
        }
        if (I.max_depth != -1 && bounces >= I.max_depth) break;
        int mat = s.prims[it.prim].material;
        if (mat < 0) { ray = Ray(it.P, ray.d, ray.tmin); bounces--; continue; }
        const rd_material_desc& m = s.mats[mat];
        V3 wo_l = it.sh.to_local(wo);
        C3 c_nee(0.0f), c_mis(0.0f);
        float r_nee = 0, r_mis = 0, r_cont = 0;
        if (method == RD_METHOD_BXDF) {
            if (s.prims[it.prim].light >= 0) radiance = radiance + coeff * inct_le(s, it, wo);
        } else if (!s.lights.empty()) {
            int l = rng.next_int((int)s.lights.size());
            float lpdf = 1.0f / s.lights.size();
            LSample ls = light_sample_li(s, l, it.P, rng);
            C3 le_nee(0.0f), le_mis(0.0f);
            if (method == RD_METHOD_MIS) {                          // EstimateDirect (PathTracer.cs:309-399)
                if (ls.pdf > 0.0f && !ls.li.black()) {
                    C3 fr = material_fr(s, m, wo_l, it.sh.to_local(ls.wi), it.UV);
                    float spdf = material_pdf(s, m, wo_l, it.sh.to_local(ls.wi), it.UV);
                    if (!fr.black() && !scene_is_occluded(s, it.P, ls.p, nullptr)) {
                        fr = fr * abs_cos_theta(it.sh.to_local(ls.wi));
                        le_nee = fr * ls.li * power_heuristic(1, ls.pdf, 1, spdf) / ls.pdf;
                        r_nee = rho_ratio(s, m, it.UV, wo_l, it.sh.to_local(ls.wi));
                    }
                }
                BSample bs = material_sample(s, m, wo_l, it.UV, rng);
                C3 fr = bs.fr * abs_cos_theta(bs.wi);
                if (!fr.black() && bs.pdf > 0) {
                    float w = 1;
                    bool ok = true;
                    if (!(bs.type & T_SPECULAR)) {
                        float lp = light_pdf_li(s, l, it.P, it.sh.to_world(bs.wi));
                        if (lp == 0.0f) ok = false;
                        else w = power_heuristic(1, bs.pdf, 1, lp);
                    }
                    if (ok) {
                        Ray to_light = spawn_ray(it, it.sh.to_world(bs.wi));
                        Inct li_it;
                        C3 li(0);
                        if (scene_intersect(s, to_light, &li_it, nullptr)) {
                            if (s.prims[li_it.prim].light == l) li = inct_le(s, li_it, it.sh.to_world(-bs.wi));
                        } else {
                            li = eval_infinite(s, to_light);
                        }
                        if (!li.black()) {
                            le_mis = fr * li * w / bs.pdf;
                            r_mis = rho_ratio(s, m, it.UV, wo_l, bs.wi);
                        }
                    }
                }
                C3 result = C3(0.0f) + (le_nee + le_mis) / lpdf;
                radiance = radiance + coeff * result;
            } else {                                                // NextEventEstimation.EstimateLight
                if (ls.pdf > 0.0f && !scene_is_occluded(s, it.P, ls.p, nullptr)) {
                    C3 fr = material_fr(s, m, wo_l, it.sh.to_local(ls.wi), it.UV);
                    le_nee = fr * ls.li * abs_cos_theta(it.sh.to_local(ls.wi)) / ls.pdf;
                    r_nee = rho_ratio(s, m, it.UV, wo_l, it.sh.to_local(ls.wi));
                }
                radiance = radiance + coeff * le_nee / lpdf;
            }
            c_nee = coeff * (le_nee / lpdf);
            c_mis = coeff * (le_mis / lpdf);
        }
        BSample bs = material_sample(s, m, wo_l, it.UV, rng);
        specular_path = (bs.type & T_SPECULAR) != 0;
        bool cont = bs.pdf > 0.0f;
        if (cont) {
            r_cont = rho_ratio(s, m, it.UV, wo_l, bs.wi);
            coeff = coeff * (bs.fr * abs_cos_theta(bs.wi) / bs.pdf);
            ray = spawn_ray(it, it.sh.to_world(bs.wi));
            if (bounces > I.min_depth) {
                float q = nmin(coeff.max_elem(), I.rr_threshold);
                if (rng.next_float() > q) cont = false;
                else coeff = coeff / q;
            }
        }
        if (m.type == RD_MAT_DIFFUSE || m.type == RD_MAT_ROUGH_PLASTIC) {
            C3 R = tex_at(s, m.tex[0], it.UV);
            C3 lr = ltot - radiance;                                // what the rest of the path adds
            C3 t = c_nee + c_mis + lr;
            C3 gr(R.r != 0 ? delta.r * t.r / R.r : 0.0f, R.g != 0 ? delta.g * t.g / R.g : 0.0f,
                  R.b != 0 ? delta.b * t.b / R.b : 0.0f);
            C3 q3 = c_nee * r_nee + c_mis * r_mis + lr * r_cont;
            float grho = delta.r * q3.r + delta.g * q3.g + delta.b * q3.b;
            adj_scatter(s, mat, it.UV, gr, grho, G);
        }
        if (!cont) break;
    }
}

static C3 li(const cr_scene& s, const Ray& ray, Rng& rng, uint64_t* st) {
    if (s.integ.type == RD_INTEGRATOR_VOL_PATH) return li_vol(s, ray, rng, st);
    switch (s.integ.method) {
    case RD_METHOD_BXDF: return li_bxdf(s, ray, rng, st);
    case RD_METHOD_NEE: return li_nee(s, ray, rng, st);
    default: return li_mis(s, ray, rng, st);
    }
}

// One pixel sample: BlockBasedRenderer.RenderTask body (BlockBasedRenderer.cs:136-146).
static C3 pixel_sample(const cr_scene& s, uint32_t seed, int x, int y, int t, uint64_t* st) {
    Rng rng;
    rng.key = path_key(seed, (uint32_t)(y * s.cam.W + x), (uint32_t)t);
    rng.n = 0;
    float ox = rng.next_float(), oy = rng.next_float();
    Ray ray = s.cam.sample_ray((float)x + ox, (float)y + oy);
    if (st) st[2]++;
    return li(s, ray, rng, st);
}

// ---------------------------------------------------------------- light tracing
// Light.SampleEmit: DiffuseAreaLight.SampleEmit (DiffuseAreaLight.cs:70-77) and
// InfiniteAreaLight.SampleEmit (InfiniteAreaLight.cs:122-133).  Returns false
// when the emission pdf is 0 (the environment map's SampleLightDir failing):
// the reference then divides by that pdf; the particle ends here instead.
struct Emit { V3 p, dir, n; C3 le; float pdf_pos = 0, pdf_dir = 0; };
// Probability.SquareToConcentricDisk (Probability.cs:157-174)
static inline void concentric_disk(float u0, float u1, float* dx, float* dy) {
    float ox = 2.0f * u0 - 1, oy = 2.0f * u1 - 1;
    if (ox == 0 && oy == 0) { *dx = 0; *dy = 0; return; }
    float theta, r;
    if (std::fabs(ox) > std::fabs(oy)) { r = ox; theta = (PI_F / 4) * (oy / ox); }
    else { r = oy; theta = (PI_F / 2) - (PI_F / 4) * (ox / oy); }
    *dx = r * std::cos(theta);
    *dy = r * std::sin(theta);
}
static bool light_sample_emit(const cr_scene& s, int light, Rng& rng, Emit* e) {
    const rd_light_desc& L = s.lights[light];
    if (L.type == RD_LIGHT_INFINITE) {
        const EnvLight& env = s.env[s.env_of_light[light]];
        float u0 = rng.next_float(), u1 = rng.next_float();
        V3 wdir;
        float pdf;
        V2 uv;
        if (!env_sample_dir(env, u0, u1, &wdir, &pdf, &uv)) return false;
        e->le = s.tex[env.tex].sample(uv);
        V3 normal = -wdir;
        Frame coord(normal);
        float c0 = rng.next_float(), c1 = rng.next_float();
        float cx, cy;
        concentric_disk(c0, c1, &cx, &cy);
        V3 disk = env.center + env.radius * (cx * coord.X + cy * coord.Y);
        e->p = disk + wdir * env.radius;
        e->dir = -wdir;
        e->n = -wdir;
        e->pdf_pos = 1 / (PI_F * sqr(env.radius));
        e->pdf_dir = pdf;
        return true;
    }
    V3 P, N;
    float pdf_pos;
    shape_sample(s, L.prim, rng, &P, &N, &pdf_pos);
    float u0 = rng.next_float(), u1 = rng.next_float();
    V3 lw = normalize(square_to_cosine_hemisphere(u0, u1));
    float pdf_dir = cos_theta(lw) / PI_F;
    V3 w = Frame(N).to_world(lw);
    e->p = P;
    e->dir = normalize(w);
    e->n = N;
    e->le = C3(L.le[0], L.le[1], L.le[2]);
    e->pdf_pos = pdf_pos;
    e->pdf_dir = pdf_dir;
    return true;
}

// AdjointParticleIntegrator.EmitParticle (AdjointParticleIntegrator.cs:24-80):
// splats into acc (float64, ColorBuffer layout) at the pixel the camera sees
// each surface vertex in.
static void emit_particle(const cr_scene& s, uint32_t seed, uint64_t index, double* acc, uint64_t* st) {
    const rd_integrator_desc& I = s.integ;
    Rng rng;
    rng.key = particle_key(seed, index);
    rng.n = 0;
    if (st) st[2]++;
    if (s.lights.empty()) return;                                   // Scene.SampleLight: pdf 0
    int light = rng.next_int((int)s.lights.size());
    float select_pdf = 1.0f / s.lights.size();
    Emit e;
    if (!light_sample_emit(s, light, rng, &e)) return;
    if (e.le.black()) return;
    C3 coeff = e.le * abs_dot(e.dir, e.n) / (select_pdf * e.pdf_pos * e.pdf_dir);
    Ray ray(e.p, e.dir, 0.0001f);
    const int H = s.cam.H;
    for (int bounces = 1;; bounces++) {
        if (I.max_depth != -1 && bounces >= I.max_depth) break;
        if (bounces > I.min_depth) {
            if (rng.next_float() > I.rr_threshold) break;
            coeff = coeff / I.rr_threshold;
        }
        Inct it;
        if (!scene_intersect(s, ray, &it, st)) break;
        int mat = s.prims[it.prim].material;
        if (mat < 0) {
            ray = Ray(it.P, ray.d, ray.tmin);
            bounces--;
            continue;
        }
        if (st) st[3]++;
        const rd_material_desc& m = s.mats[mat];
        Camera::Wi cw = s.cam.sample_wi(it.P);
        if (cw.pdf > 0 && cw.we != 0) {
            if (!scene_is_occluded(s, cw.pos, it.P, st)) {
                V3 wi_l = it.sh.to_local(cw.dir);
                C3 f = material_fr(s, m, it.sh.to_local(it.wr), wi_l, it.UV, true);
                float ac = abs_cos_theta(wi_l);
                C3 color = coeff * f * C3(cw.we) * ac / cw.pdf;
                int x = (int)std::floor(cw.sx), y = (int)std::floor(cw.sy);
                double* px = acc + ((size_t)x * H + y) * 3;
                px[0] += (double)color.r;
                px[1] += (double)color.g;
                px[2] += (double)color.b;
            }
        }
        BSample bs = material_sample(s, m, it.sh.to_local(it.wr), it.UV, rng, true);
        if (bs.pdf > 0.0f) coeff = coeff * (bs.fr * abs_cos_theta(bs.wi) / bs.pdf);
        else break;
        ray = spawn_ray(it, it.sh.to_world(bs.wi));
    }
}

// AdjointParticleIntegrator.CameraTraceLight (:82-85) for the ray through
// pixel corner (x, y) (AdjointParticleRenderer.AfterRender, :170-176).
static C3 camera_trace_light(const cr_scene& s, int x, int y, uint64_t* st) {
    Ray ray = s.cam.sample_ray((float)x, (float)y);
    Inct it;
    if (scene_intersect(s, ray, &it, st)) return inct_le(s, it, -ray.d);
    return eval_infinite(s, ray);
}

// ---------------------------------------------------------------- bidirectional
// BidirectionalPathTracer (Integrators/BidirectionalPathTracer.cs:11-435) and
// BdptVertex / BdptPath (Renderers/BidirectionalPathTracingRenderer.cs:13-277).
// The reference's MisWeight returns 1 for every strategy (its body is
// commented out, :305-414), so the vertices' forward / reverse densities
// never reach the image and are not kept here.
enum { BV_CAMERA = 0, BV_LIGHT = 1, BV_SURFACE = 2, BV_MEDIUM = 3 };
struct BVert {
    int type = BV_CAMERA;
    V3 p, n;            // n = 0 off surfaces (IsOnSurface => N != 0)
    C3 coeff;
    int env = -1;       // Msr.Medium: the medium a medium vertex sampled in;
                        // the camera's / light's medium; -1 (null) on surfaces
    Inct it;            // surface vertices
    V3 wo;              // medium vertices: Msr.Wo
    int light = -1;     // light vertices: the light (-1: the escaped-ray vertex, Light null)
};
static inline bool bv_on_surface(const BVert& v) { return !(v.n.x == 0 && v.n.y == 0 && v.n.z == 0); }
static inline bool bv_is_light(const cr_scene& s, const BVert& v) {
    return v.type == BV_LIGHT || (v.type == BV_SURFACE && s.prims[v.it.prim].light >= 0);
}
static inline bool bv_is_infinite(const cr_scene& s, const BVert& v) {
    return v.type == BV_LIGHT && (v.light < 0 || s.lights[v.light].type == RD_LIGHT_INFINITE);
}
// BdptVertex.F (:94-106)
static C3 bv_f(const cr_scene& s, const BVert& v, const BVert& next, bool imp) {
    V3 wi = next.p - v.p;
    if (length_sq(wi) == 0) return C3(0.0f);
    wi = normalize(wi);
    if (v.type == BV_SURFACE) {
        const rd_material_desc& m = s.mats[s.prims[v.it.prim].material];
        return material_fr(s, m, v.it.sh.to_local(v.it.wr), v.it.sh.to_local(wi), v.it.UV, imp);
    }
    if (v.type == BV_MEDIUM) return C3(hg_p(s.homo[v.env], v.wo, wi));
    return C3(0.0f);
}
// BdptVertex.Le (:69-92)
static C3 bv_le(const cr_scene& s, const BVert& v, const BVert& prev) {
    if (!bv_is_light(s, v)) return C3(0.0f);
    V3 w = prev.p - v.p;
    if (length_sq(w) == 0) return C3(0.0f);
    w = normalize(w);
    if (bv_is_infinite(s, v)) {
        C3 le(0.0f);
        for (size_t i = 0; i < s.env.size(); i++) le = le + env_le(s, s.env[i], -w);
        return le;
    }
    return inct_le(s, v.it, w);
}

// BidirectionalPathTracer.RandomWalk (:84-162): appends the walk's vertices
// after path[0]; returns the walk count.  Rays after the first keep the
// first ray's MinT, as written.
static int bd_random_walk(const cr_scene& s, Ray ray, int env, Rng& rng, C3 coeff, std::vector<BVert>& path,
                          int max_depth, bool imp, uint64_t* st) {
    if (max_depth == 0) return 0;
    int bounces = 1;
    while (true) {
        Inct it;
        bool hit = scene_intersect(s, ray, &it, st);
        MSample ms;
        ms.medium = false;
        if (env >= 0) {
            Ray real(ray.o, ray.d, ray.tmin, hit ? it.T : FLT_MAX);
            ms = medium_sample(s.homo[env], real, rng);
            coeff = coeff * ms.tr;
        }
        if (coeff.black()) break;
        if (ms.medium) {
            BVert v;
            v.type = BV_MEDIUM; v.p = ms.p; v.n = V3(0.0f); v.coeff = coeff; v.env = env; v.wo = ms.wo;
            path.push_back(v);
            if (++bounces >= max_depth) break;
            PSample ps = hg_sample(s.homo[env], ms.wo, rng);
            ray = Ray(ms.p, ps.wi, ray.tmin);
        } else {
            if (!hit) {
                if (!imp) {
                    BVert v;
                    v.type = BV_LIGHT; v.light = -1; v.p = ray.at(10); v.n = -ray.d; v.coeff = coeff; v.env = env;
                    path.push_back(v);
                    ++bounces;
                }
                break;
            }
            int mat = s.prims[it.prim].material;
            if (mat < 0) {
                ray = Ray(it.P, ray.d, ray.tmin);
                env = inct_medium(s, it, ray.d);
                continue;
            }
            if (st) st[3]++;
            BVert v;
            v.type = BV_SURFACE; v.p = it.P; v.n = it.N(); v.coeff = coeff; v.env = -1; v.it = it;
            path.push_back(v);
            if (++bounces >= max_depth) break;
            BSample bs = material_sample(s, s.mats[mat], it.sh.to_local(it.wr), it.UV, rng, imp);
            if (bs.pdf > 0.0f) coeff = coeff * (bs.fr * abs_cos_theta(bs.wi) / bs.pdf);
            else break;
            V3 nd = it.sh.to_world(bs.wi);
            ray = Ray(it.P, nd, ray.tmin);
            env = inct_medium(s, it, nd);
        }
    }
    return bounces - 1;
}

// The three random streams of bidirectional sample (pixel, sample): the
// camera subpath (the sample position, then its walk), the light subpath,
// and the light samples of the s = 1 strategies (three draws each, strategy
// t at draws 3 (t - 2) ...).  Sample indices stay below 2^29.
static inline uint64_t bd_light_key(uint32_t seed, uint32_t pixel, uint32_t sample) {
    return path_key(seed, pixel, 0x40000000u | sample);
}
static inline uint64_t bd_nee_key(uint32_t seed, uint32_t pixel, uint32_t sample) {
    return path_key(seed, pixel, 0x20000000u | sample);
}

// ConnectBdpt (:219-300) for strategy (s, t); MisWeight is 1.
static C3 bd_connect(const cr_scene& s, const std::vector<BVert>& lp, int sn, const std::vector<BVert>& cp, int t,
                     uint32_t seed, uint32_t pixel, uint32_t sample, uint64_t* st) {
    C3 l(0.0f);
    if (t > 1 && sn != 0 && cp[t - 1].type == BV_LIGHT) return C3(0.0f);
    if (sn == 0) {
        const BVert& pt = cp[t - 1];
        if (bv_is_light(s, pt)) l = bv_le(s, pt, cp[t - 2]) * pt.coeff;
    } else if (t == 1) {
        const BVert& qs = lp[sn - 1];   // every vertex type past the light's own is connectible
        Camera::Wi cw = s.cam.sample_wi(qs.p);
        if (cw.pdf > 0 && cw.we != 0) {
            BVert cam;
            cam.type = BV_CAMERA; cam.p = cw.pos; cam.coeff = C3(cw.we) / cw.pdf;
            l = qs.coeff * bv_f(s, qs, cam, true) * cam.coeff;
            if (bv_on_surface(qs)) l = l * abs_dot(cw.dir, qs.n);
            if (!l.black()) l = l * scene_transmittance(s, qs.p, cw.pos, qs.env, st);
        }
    } else if (sn == 1) {
        const BVert& pt = cp[t - 1];
        Rng rng;
        rng.key = bd_nee_key(seed, pixel, sample);
        rng.n = 3 * (uint32_t)(t - 2);
        if (!s.lights.empty()) {
            int light = rng.next_int((int)s.lights.size());
            float select_pdf = 1.0f / s.lights.size();
            LSample ls = light_sample_li(s, light, pt.p, rng);
            if (ls.pdf > 0 && !ls.li.black()) {
                BVert lv;
                lv.type = BV_LIGHT; lv.light = light; lv.p = ls.p; lv.coeff = ls.li / (ls.pdf * select_pdf);
                l = pt.coeff * bv_f(s, pt, lv, false) * lv.coeff;
                if (bv_on_surface(pt)) l = l * abs_dot(ls.wi, pt.n);
                if (!l.black()) l = l * scene_transmittance(s, pt.p, ls.p, pt.env, st);
            }
        }
    } else {
        const BVert& qs = lp[sn - 1];
        const BVert& pt = cp[t - 1];
        l = qs.coeff * bv_f(s, qs, pt, true) * bv_f(s, pt, qs, false) * pt.coeff;
        if (!l.black()) {
            // G (:311-320)
            V3 d = qs.p - pt.p;
            float g = 1 / length_sq(d);
            d = d * std::sqrt(g);
            if (bv_on_surface(qs)) g *= abs_dot(qs.n, d);
            if (bv_on_surface(pt)) g *= abs_dot(pt.n, d);
            C3 tr = scene_transmittance(s, qs.p, pt.p, qs.env, st);
            l = l * (g * tr);
        }
    }
    return l;
}

// One bidirectional sample (BidirectionalPathTracingRenderer.RenderTask
// body, :396-407): the camera subpath, the light subpath, then every
// strategy of Connect (:164-217) in its loop order.  All of them land on the
// sample's own pixel (the reference's t = 1 splat takes floor(samplePoint)
// whenever SampleWi set a screen position), so the sample's value is
// (sum of the t = 1 strategies in s order) + (sum of the others in loop
// order), each strategy that is not a valid colour counted as 0.
// only_s / only_t >= 0: that strategy alone (a test probe of the estimator).
static C3 bd_sample(const cr_scene& s, uint32_t seed, int x, int y, int sample, uint64_t* st, int only_s = -1,
                    int only_t = -1) {
    const rd_integrator_desc& I = s.integ;
    const uint32_t pixel = (uint32_t)(y * s.cam.W + x);
    Rng rng;
    rng.key = path_key(seed, pixel, (uint32_t)sample);
    rng.n = 0;
    float ox = rng.next_float(), oy = rng.next_float();
    if (st) st[2]++;
    // GenerateCameraSubpath (:42-57)
    std::vector<BVert> cp, lp;
    cp.reserve(16); lp.reserve(16);
    Ray ray = s.cam.sample_ray((float)x + ox, (float)y + oy);
    BVert c0;
    c0.type = BV_CAMERA; c0.p = ray.o; c0.n = V3(0.0f); c0.coeff = C3(1.0f); c0.env = s.global_medium;
    cp.push_back(c0);
    int n_cam = bd_random_walk(s, ray, s.global_medium, rng, C3(1.0f), cp, I.max_depth - 1, false, st) + 1;
    // GenerateLightSubpath (:59-82)
    int n_light = 0;
    Rng lr;
    lr.key = bd_light_key(seed, pixel, (uint32_t)sample);
    lr.n = 0;
    if (!s.lights.empty()) {
        int light = lr.next_int((int)s.lights.size());
        float select_pdf = 1.0f / s.lights.size();
        Emit e;
        if (light_sample_emit(s, light, lr, &e) && e.pdf_pos != 0 && e.pdf_dir != 0 && !e.le.black()) {
            const rd_light_desc& L = s.lights[light];
            int lenv = -1;                       // Light.GetMedium (DiffuseAreaLight.cs:88-91; infinite: null)
            if (L.type == RD_LIGHT_AREA) {
                const rd_prim_desc& pd = s.prims[L.prim];
                lenv = dot(e.n, e.dir) > 0 ? pd.medium_outside : pd.medium_inside;
            }
            BVert l0;
            l0.type = BV_LIGHT; l0.light = light; l0.p = e.p; l0.n = e.n; l0.coeff = e.le; l0.env = lenv;
            lp.push_back(l0);
            C3 coeff = e.le * abs_dot(e.dir, e.n) / (select_pdf * e.pdf_pos * e.pdf_dir);
            Ray lray(e.p, e.dir, 0.0001f);
            n_light = bd_random_walk(s, lray, lenv, lr, coeff, lp, I.max_depth - 1, true, st) + 1;
        }
    }
    // Connect
    C3 splat(0.0f), l(0.0f);
    for (int t = 1; t < n_cam; t++) {
        for (int sn = 0; sn < n_light; sn++) {
            int depth = t + sn - 2;
            if ((sn == 1 && t == 1) || depth < 0 || depth > I.max_depth) continue;
            if (only_s >= 0 && (sn != only_s || t != only_t)) continue;
            C3 c = bd_connect(s, lp, sn, cp, t, seed, pixel, (uint32_t)sample, st);
            if (!c.valid()) c = C3(0.0f);
            if (t != 1) l = l + c;
            else splat = splat + c;
        }
    }
    return splat + l;
}

// ---------------------------------------------------------------- C API
extern "C" {

const char* cr_last_error(void) { return g_err.c_str(); }

cr_scene* cr_scene_create(const rd_scene_desc* d) {
    if (!d) { set_err("null desc"); return nullptr; }
    cr_scene* s = new cr_scene();
    s->pos.resize(d->num_vertices);
    for (int i = 0; i < d->num_vertices; i++) s->pos[i] = V3(d->positions[3 * i], d->positions[3 * i + 1], d->positions[3 * i + 2]);
    if (d->normals) {
        s->nrm.resize(d->num_vertices);
        for (int i = 0; i < d->num_vertices; i++) s->nrm[i] = V3(d->normals[3 * i], d->normals[3 * i + 1], d->normals[3 * i + 2]);
    }
    if (d->uvs) {
        s->uvs.resize(d->num_vertices);
        for (int i = 0; i < d->num_vertices; i++) s->uvs[i] = V2(d->uvs[2 * i], d->uvs[2 * i + 1]);
    }
    s->idx.assign(d->indices, d->indices + 3 * (size_t)d->num_triangles);
    s->tflags.assign(d->num_triangles, 0);
    if (d->tri_flags) s->tflags.assign(d->tri_flags, d->tri_flags + d->num_triangles);
    for (int t = 0; t < d->num_triangles; t++) {
        if ((s->tflags[t] & 1) && s->nrm.empty()) { set_err("triangle flags normals but no normals"); delete s; return nullptr; }
        if ((s->tflags[t] & 2) && s->uvs.empty()) { set_err("triangle flags uvs but no uvs"); delete s; return nullptr; }
        for (int k = 0; k < 3; k++) {
            int v = s->idx[3 * t + k];
            if (v < 0 || v >= d->num_vertices) { set_err("index out of range"); delete s; return nullptr; }
        }
    }
    if (d->num_spheres) s->sph.assign(d->spheres, d->spheres + 4 * (size_t)d->num_spheres);
    if (d->num_spheres && d->sphere_w2m) s->sph_w2m.assign(d->sphere_w2m, d->sphere_w2m + 9 * (size_t)d->num_spheres);
    s->prims.assign(d->prims, d->prims + d->num_prims);
    s->mats.assign(d->materials, d->materials + d->num_materials);
    s->lights.assign(d->lights, d->lights + d->num_lights);
    if (d->num_media) s->media.assign(d->media, d->media + d->num_media);
    s->global_medium = d->global_medium;
    s->max_prim = d->bvh_max_prim > 0 ? d->bvh_max_prim : 1;
    s->accel = d->accel;
    if (s->accel != RD_ACCEL_BVH && s->accel != RD_ACCEL_OCTREE) { set_err("unknown accelerator"); delete s; return nullptr; }
    s->oct_max_depth = d->octree_max_depth;
    s->oct_max_count = d->octree_max_count;
    s->oct_out_bound = d->octree_out_bound;
    s->integ = d->integrator;
    s->tex.resize(d->num_textures);
    for (int i = 0; i < d->num_textures; i++) {
        const rd_texture_desc& td = d->textures[i];
        Tex& t = s->tex[i];
        t.w = td.width; t.h = td.height; t.wrap = td.wrap; t.filter = td.filter;
        t.value = C3(td.value[0], td.value[1], td.value[2]);
        if (t.w > 0) {
            if (!td.data || t.h <= 0) { set_err("image texture without data"); delete s; return nullptr; }
            t.data.assign(td.data, td.data + (size_t)t.w * t.h * 3);
        }
    }
    if (d->num_instance_sets < 0 || d->num_instances < 0) { set_err("negative count"); delete s; return nullptr; }
    for (int i = 0; i < d->num_instance_sets; i++) {
        const rd_instance_set_desc& sd = d->instance_sets[i];
        if (sd.first_triangle < 0 || sd.num_triangles <= 0 || sd.num_triangles > d->num_triangles - sd.first_triangle) {
            set_err("instance set range"); delete s; return nullptr;
        }
        InstSet st;
        st.first = sd.first_triangle; st.count = sd.num_triangles; st.max_prim = sd.bvh_max_prim > 0 ? sd.bvh_max_prim : 1;
        s->sets.push_back(st);
    }
    for (int i = 0; i < d->num_instances; i++) {
        const rd_instance_desc& id = d->instances[i];
        if (id.set < 0 || id.set >= d->num_instance_sets) { set_err("instance set index"); delete s; return nullptr; }
        Inst in;
        in.set = id.set;
        for (int k = 0; k < 16; k++) in.m2w.m[k / 4][k % 4] = id.to_world[k];
        // InstancedTransform ctor: Matrix4x4.Invert (Primitive.cs:99-106)
        if (!m4_invert(in.m2w, &in.w2m)) { set_err("instance: invalid transform matrix"); delete s; return nullptr; }
        s->insts.push_back(in);
    }
    s->prim_area.resize(s->prims.size());
    for (size_t p = 0; p < s->prims.size(); p++) {
        const rd_prim_desc& pd = s->prims[p];
        if (pd.kind == RD_PRIM_INSTANCE) {
            if (pd.index < 0 || pd.index >= d->num_instances) { set_err("prim instance index"); delete s; return nullptr; }
            if (pd.light != -1) { set_err("an instanced primitive cannot be a light"); delete s; return nullptr; }
            s->prim_area[p] = 0.0f;
            if (pd.material >= d->num_materials) { set_err("prim material/light index"); delete s; return nullptr; }
            continue;
        }
        if (pd.kind == RD_PRIM_TRIANGLE) {
            if (pd.index < 0 || pd.index >= d->num_triangles) { set_err("prim triangle index"); delete s; return nullptr; }
            int t = pd.index;
            s->prim_area[p] = 0.5f * length(cross(s->pb(t) - s->pa(t), s->pc(t) - s->pa(t)));   // Triangle.cs:93
        } else if (pd.kind == RD_PRIM_SPHERE) {
            if (pd.index < 0 || pd.index >= d->num_spheres) { set_err("prim sphere index"); delete s; return nullptr; }
            float r = s->sph[4 * pd.index + 3];
            s->prim_area[p] = 4 * r * r * PI_F;                                                   // Sphere.cs:18
        } else { set_err("prim kind"); delete s; return nullptr; }
        if (pd.material >= d->num_materials || pd.light >= d->num_lights) { set_err("prim material/light index"); delete s; return nullptr; }
    }
    for (int i = 0; i < d->num_materials; i++) {
        int ty = s->mats[i].type;
        if (ty < RD_MAT_DIFFUSE || ty > RD_MAT_SUBSURFACE) { set_err("unknown material type"); delete s; return nullptr; }
    }
    s->env_of_light.assign(d->num_lights, -1);
    for (int i = 0; i < d->num_lights; i++) {
        const rd_light_desc& L = s->lights[i];
        if (L.type == RD_LIGHT_INFINITE) {
            if (L.tex < 0 || L.tex >= d->num_textures) { set_err("infinite light texture index"); delete s; return nullptr; }
            EnvLight e;
            e.tex = L.tex;
            for (int k = 0; k < 16; k++) e.m2w.m[k / 4][k % 4] = L.to_world[k];
            if (!m4_invert(e.m2w, &e.w2m)) { set_err("infinite light: invalid matrix"); delete s; return nullptr; }
            s->env_of_light[i] = (int)s->env.size();
            s->infinite.push_back(i);
            s->env.push_back(e);
            continue;
        }
        if (L.type != RD_LIGHT_AREA) { set_err("unknown light type"); delete s; return nullptr; }
        int p = L.prim;
        if (p < 0 || p >= d->num_prims) { set_err("light prim index"); delete s; return nullptr; }
    }
    if (d->integrator.type < RD_INTEGRATOR_PATH || d->integrator.type > RD_INTEGRATOR_BDPT) {
        set_err("unknown integrator type"); delete s; return nullptr;
    }
    for (int i = 0; i < d->num_media; i++) s->homo.push_back(make_homo(d->media[i]));
    if (s->global_medium < -1 || s->global_medium >= d->num_media) { set_err("global medium index"); delete s; return nullptr; }
    for (const rd_prim_desc& pd : s->prims)
        if (pd.medium_inside < -1 || pd.medium_inside >= d->num_media || pd.medium_outside < -1 ||
            pd.medium_outside >= d->num_media) { set_err("prim medium index"); delete s; return nullptr; }
    if (!s->cam.init(d->camera)) { set_err("camera matrix not invertible"); delete s; return nullptr; }
    build_bvh(*s);
    if (s->accel == RD_ACCEL_OCTREE && !build_octree(*s)) { delete s; return nullptr; }
    for (EnvLight& e : s->env) {
        // InfiniteAreaLight ctor (InfiniteAreaLight.cs:31-53): luminance * sin(theta)
        // of the map sampled at texel corners (x/W, y/H), rows of the 2D distribution
        const Tex& t = s->tex[e.tex];
        e.w = t.w > 0 ? t.w : 1;     // ConstColorTexture2D is 1x1
        e.h = t.w > 0 ? t.h : 1;
        std::vector<float> lum((size_t)e.w * e.h);
        for (int y = 0; y < e.h; y++) {
            float sin_theta = std::sin((y + 0.5f) / e.h * PI_F);
            float v = y / (float)e.h;
            for (int x = 0; x < e.w; x++) {
                float u = x / (float)e.w;
                C3 c = t.sample(V2(u, v));
                lum[(size_t)y * e.w + x] = (c.r * 0.212671f + c.g * 0.715160f + c.b * 0.072169f) * sin_theta;
            }
        }
        e.dist.build(lum.data(), e.w, e.h);
        // InfiniteLight.Preprocess (Light.cs:181-187): Aggregate.WorldBound = root node bounds
        BBox b = s->nodes.empty() ? BBox() : s->nodes[0].b;
        if (s->nodes.empty()) { b.mn = V3(0.0f); b.mx = V3(0.0f); }
        e.center = (b.mx + b.mn) * 0.5f;
        e.radius = length(b.mx - b.mn) / 2.0f;
    }
    return s;
}

void cr_scene_destroy(cr_scene* s) { delete s; }
int32_t cr_scene_num_nodes(const cr_scene* s) { return s ? (int32_t)s->nodes.size() : -1; }
int32_t cr_scene_max_depth(const cr_scene* s) { return s ? s->max_depth : -1; }

int cr_scene_export_bvh(const cr_scene* s, void* nodes32, int32_t* prim_order) {
    if (!s) return RD_ERR_INVALID;
    unsigned char* out = (unsigned char*)nodes32;
    for (size_t i = 0; i < s->nodes.size(); i++) {
        const Node& n = s->nodes[i];
        float b[6] = {n.b.mn.x, n.b.mn.y, n.b.mn.z, n.b.mx.x, n.b.mx.y, n.b.mx.z};
        unsigned char rec[32];
        memset(rec, 0, 32);
        memcpy(rec, b, 24);
        memcpy(rec + 24, &n.offset, 4);
        memcpy(rec + 28, &n.count, 2);
        memcpy(rec + 30, &n.axis, 1);
        memcpy(out + 32 * i, rec, 32);
    }
    if (prim_order) memcpy(prim_order, s->order.data(), s->order.size() * sizeof(int32_t));
    return RD_OK;
}

static inline Ray to_ray(const rd_ray& r) {
    return Ray(V3(r.o[0], r.o[1], r.o[2]), V3(r.d[0], r.d[1], r.d[2]), r.tmin, r.tmax);
}

// The octree as 32-byte LinearNode records plus the leaves' primitive lists.
static void export_octree(const cr_scene& s, void* nodes32, int32_t* leaf_first, int32_t* leaf_count,
                          int32_t* leaf_prims, int32_t* sizes3) {
    int entries = 0;
    for (const auto& l : s.oct_leaves) entries += (int)l.size();
    if (sizes3) { sizes3[0] = (int32_t)s.oct.size(); sizes3[1] = (int32_t)s.oct_leaves.size(); sizes3[2] = entries; }
    if (nodes32) {
        unsigned char* out = (unsigned char*)nodes32;
        for (size_t i = 0; i < s.oct.size(); i++) {
            const cr_scene::OctLinear& n = s.oct[i];
            float b[6] = {n.b.mn.x, n.b.mn.y, n.b.mn.z, n.b.mx.x, n.b.mx.y, n.b.mx.z};
            memcpy(out + 32 * i, b, 24);
            memcpy(out + 32 * i + 24, &n.count_or_leaf, 4);
            memcpy(out + 32 * i + 28, &n.next_head, 4);
        }
    }
    int e = 0;
    for (size_t l = 0; l < s.oct_leaves.size(); l++) {
        if (leaf_first) leaf_first[l] = e;
        if (leaf_count) leaf_count[l] = (int32_t)s.oct_leaves[l].size();
        for (int p : s.oct_leaves[l]) {
            if (leaf_prims) leaf_prims[e] = p;
            e++;
        }
    }
}

int cr_scene_export_octree(const cr_scene* s, void* nodes32, int32_t* leaf_first, int32_t* leaf_count,
                           int32_t* leaf_prims, int32_t* sizes3) {
    if (!s || s->accel != RD_ACCEL_OCTREE) return RD_ERR_INVALID;
    export_octree(*s, nodes32, leaf_first, leaf_count, leaf_prims, sizes3);
    return RD_OK;
}

int cr_octree_build(const float* bmin, const float* bmax, int32_t n, int32_t max_depth, int32_t max_count,
                    float out_bound, void* nodes32, int32_t* leaf_first, int32_t* leaf_count, int32_t* leaf_prims,
                    int32_t* sizes3, int32_t* now_depth) {
    if (n < 0) return RD_ERR_INVALID;
    cr_scene s;
    s.oct_max_depth = max_depth; s.oct_max_count = max_count; s.oct_out_bound = out_bound;
    std::vector<BBox> b(n);
    BBox root;
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) {
        b[i].mn = V3(bmin[3 * i], bmin[3 * i + 1], bmin[3 * i + 2]);
        b[i].mx = V3(bmax[3 * i], bmax[3 * i + 1], bmax[3 * i + 2]);
        root.unite(b[i]);
        idx[i] = i;
    }
    OctBuild B(b, max_depth, max_count, out_bound);
    std::unique_ptr<OctBuild::N> r;
    try {
        r = B.build(idx, root, root, 0);
    } catch (const OctBuild::TooLarge&) {
        set_err("octree over 2^27 entries");
        return RD_ERR_INVALID;
    }
    if (r) {
        s.oct.resize(B.node_count);
        int all = 1;
        B.flatten(*r, 0, &all, s);
    }
    if (now_depth) *now_depth = B.now_depth;
    export_octree(s, nodes32, leaf_first, leaf_count, leaf_prims, sizes3);
    return RD_OK;
}

int cr_trace_closest(const cr_scene* s, const rd_ray* rays, int32_t n, rd_hit* hits, uint64_t* counters) {
    return cr_trace_closest_sub(s, rays, n, hits, nullptr, counters);
}

int cr_trace_closest_sub(const cr_scene* s, const rd_ray* rays, int32_t n, rd_hit* hits, int32_t* sub,
                         uint64_t* counters) {
    if (!s) return RD_ERR_INVALID;
    for (int i = 0; i < n; i++) {
        Ray r = to_ray(rays[i]);
        Hit h;
        if (agg_closest(*s, r, &h, counters)) { hits[i].t = h.T; hits[i].prim = h.prim; hits[i].u = h.sp.u; hits[i].v = h.sp.v; }
        else { hits[i].t = 0; hits[i].prim = -1; hits[i].u = 0; hits[i].v = 0; }
        if (sub) sub[i] = h.prim >= 0 ? h.sub : -1;
    }
    return RD_OK;
}

int cr_trace_any(const cr_scene* s, const rd_ray* rays, int32_t n, uint8_t* occ, uint64_t* counters) {
    if (!s) return RD_ERR_INVALID;
    for (int i = 0; i < n; i++) occ[i] = agg_any(*s, to_ray(rays[i]), counters) ? 1 : 0;
    return RD_OK;
}

// Every primitive (and every triangle of every instance, on the model ray
// with the original MaxT) tested; the smallest world T wins.
int cr_trace_brute(const cr_scene* s, const rd_ray* rays, int32_t n, rd_hit* hits) {
    return cr_trace_brute_sub(s, rays, n, hits, nullptr);
}

int cr_trace_brute_sub(const cr_scene* s, const rd_ray* rays, int32_t n, rd_hit* hits, int32_t* sub) {
    if (!s) return RD_ERR_INVALID;
    for (int i = 0; i < n; i++) {
        Ray r = to_ray(rays[i]);
        hits[i].prim = -1; hits[i].t = 0; hits[i].u = 0; hits[i].v = 0;
        if (sub) sub[i] = -1;
        float best = FLT_MAX;
        for (int p = 0; p < (int)s->prims.size(); p++) {
            const rd_prim_desc& pd = s->prims[p];
            if (pd.kind != RD_PRIM_INSTANCE) {
                SurfPt sp;
                if (prim_hit(*s, p, r, &sp) && sp.t < best) {
                    best = sp.t; hits[i].t = sp.t; hits[i].prim = p; hits[i].u = sp.u; hits[i].v = sp.v;
                    if (sub) sub[i] = -1;
                }
                continue;
            }
            const Inst& I = s->insts[pd.index];
            const InstSet& st = s->sets[I.set];
            Ray mr = transform_ray(r, I.w2m);
            for (int j = 0; j < st.count; j++) {
                SurfPt sp;
                int t = st.first + j;
                if (!tri_hit(*s, t, mr, &sp)) continue;
                V3 bary(1 - (sp.u + sp.v), sp.u, sp.v);
                V3 pw = xform_point(bary.x * s->pa(t) + bary.y * s->pb(t) + bary.z * s->pc(t), I.m2w);
                float T = length(r.o - pw);
                if (T < best) {
                    best = T; hits[i].t = T; hits[i].prim = p; hits[i].u = sp.u; hits[i].v = sp.v;
                    if (sub) sub[i] = j;
                }
            }
        }
    }
    return RD_OK;
}

int cr_camera_rays(const cr_scene* s, const float* xy, int32_t n, rd_ray* rays) {
    if (!s) return RD_ERR_INVALID;
    for (int i = 0; i < n; i++) {
        Ray r = s->cam.sample_ray(xy[2 * i], xy[2 * i + 1]);
        rays[i].o[0] = r.o.x; rays[i].o[1] = r.o.y; rays[i].o[2] = r.o.z;
        rays[i].d[0] = r.d.x; rays[i].d[1] = r.d.y; rays[i].d[2] = r.d.z;
        rays[i].tmin = r.tmin; rays[i].tmax = r.tmax;
    }
    return RD_OK;
}

int cr_render(const cr_scene* s, const rd_render_params* p, int32_t threads, float* out_rgb,
              float* out_sum2, uint64_t* stats, double* seconds) {
    if (!s || !p || !out_rgb) return RD_ERR_INVALID;
    int W = s->cam.W, H = s->cam.H;
    int x0 = p->x0, y0 = p->y0, x1 = p->x1, y1 = p->y1;
    if (x0 == 0 && y0 == 0 && x1 == 0 && y1 == 0) { x1 = W; y1 = H; }
    if (x0 < 0 || y0 < 0 || x1 > W || y1 > H || x0 >= x1 || y0 >= y1 || p->spp <= 0) { set_err("bad crop/spp"); return RD_ERR_INVALID; }
    // 32x32 blocks over the crop (RenderBlockGenerator.cs; scheduling only:
    // results do not depend on block order because the RNG is keyed per sample).
    const int bs = 32;
    int nbx = (x1 - x0 + bs - 1) / bs, nby = (y1 - y0 + bs - 1) / bs;
    int nblocks = nbx * nby;
    std::atomic<int> next(0);
    std::mutex mu;
    uint64_t tot[5] = {0, 0, 0, 0, 0};
    if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
    auto t0 = std::chrono::steady_clock::now();
    auto worker = [&]() {
        uint64_t st[5] = {0, 0, 0, 0, 0};
        while (true) {
            int b = next.fetch_add(1);
            if (b >= nblocks) break;
            int bx = x0 + (b % nbx) * bs, by = y0 + (b / nbx) * bs;
            int ex = std::min(bx + bs, x1), ey = std::min(by + bs, y1);
            for (int x = bx; x < ex; x++) {
                for (int y = by; y < ey; y++) {
                    // the per-pixel sum and mean in double over float32
                    // samples (the reference's ColorBuffer sums in float32,
                    // BlockBasedRenderer.cs:140-178); the sum of squares float
                    double sr = 0, sg = 0, sb = 0;
                    C3 sum2(0.0f);
                    for (int t = 0; t < p->spp; t++) {
                        C3 c = pixel_sample(*s, p->seed, x, y, t, st);
                        if (!c.valid()) { c = C3(0.0f); st[4]++; }
                        sr += (double)c.r;
                        sg += (double)c.g;
                        sb += (double)c.b;
                        sum2 = sum2 + c * c;
                    }
                    size_t o = ((size_t)x * H + y) * 3;
                    const double n = (double)p->spp;
                    out_rgb[o] = (float)(sr / n); out_rgb[o + 1] = (float)(sg / n); out_rgb[o + 2] = (float)(sb / n);
                    if (out_sum2) { out_sum2[o] = sum2.r; out_sum2[o + 1] = sum2.g; out_sum2[o + 2] = sum2.b; }
                }
            }
        }
        std::lock_guard<std::mutex> g(mu);
        for (int i = 0; i < 5; i++) tot[i] += st[i];
    };
    std::vector<std::thread> pool;
    for (int i = 0; i < threads; i++) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (stats) for (int i = 0; i < 5; i++) stats[i] = tot[i];
    return RD_OK;
}

// AdjointParticleRenderer (Renderers/AdjointParticleRenderer.cs:127-178):
// `particles` particles split over the workers, each worker summing its
// splats in float64; the sums are added in worker order, rounded to float,
// scaled by W*H / particles and the directly seen emission added.
int cr_render_particles(const cr_scene* s, uint64_t particles, uint32_t seed, int32_t threads, float* out_rgb,
                        uint64_t* stats, double* seconds) {
    if (!s || !out_rgb || particles == 0) { set_err("bad particle render arguments"); return RD_ERR_INVALID; }
    const int W = s->cam.W, H = s->cam.H;
    const size_t n = (size_t)W * H * 3;
    if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
    std::vector<std::vector<double>> acc(threads, std::vector<double>(n, 0.0));
    std::vector<std::array<uint64_t, 5>> sts(threads);
    auto t0 = std::chrono::steady_clock::now();
    const uint64_t chunk = (particles + threads - 1) / threads;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) {
        pool.emplace_back([&, t]() {
            uint64_t st[5] = {0, 0, 0, 0, 0};
            uint64_t b = (uint64_t)t * chunk, e = std::min(particles, b + chunk);
            for (uint64_t i = b; i < e; i++) emit_particle(*s, seed, i, acc[t].data(), st);
            for (int k = 0; k < 5; k++) sts[t][k] = st[k];
        });
    }
    for (auto& th : pool) th.join();
    uint64_t st[5] = {0, 0, 0, 0, 0};
    for (int t = 0; t < threads; t++) for (int k = 0; k < 5; k++) st[k] += sts[t][k];
    const float v = ((float)W * (float)H) / (float)particles;
    for (int x = 0; x < W; x++) {
        for (int y = 0; y < H; y++) {
            size_t o = ((size_t)x * H + y) * 3;
            double sr = 0, sg = 0, sb = 0;
            for (int t = 0; t < threads; t++) { sr += acc[t][o]; sg += acc[t][o + 1]; sb += acc[t][o + 2]; }
            C3 c((float)sr, (float)sg, (float)sb);
            c = c * v;
            c = c + camera_trace_light(*s, x, y, st);
            out_rgb[o] = c.r; out_rgb[o + 1] = c.g; out_rgb[o + 2] = c.b;
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (stats) for (int k = 0; k < 5; k++) stats[k] = st[k];
    return RD_OK;
}

// BidirectionalPathTracingRenderer (:380-452): per pixel of the crop the
// mean of its spp bidirectional samples (float64 sums over float32 sample
// values), 32x32 blocks over `threads` workers; out_sum2 the float32 sums of
// squares.  stats as cr_render.
static int render_bdpt(const cr_scene* s, const rd_render_params* p, int32_t threads, float* out_rgb, float* out_sum2,
                       uint64_t* stats, double* seconds, int only_s, int only_t);
int cr_render_bdpt(const cr_scene* s, const rd_render_params* p, int32_t threads, float* out_rgb, float* out_sum2,
                   uint64_t* stats, double* seconds) {
    return render_bdpt(s, p, threads, out_rgb, out_sum2, stats, seconds, -1, -1);
}
// One strategy (s, t) of the bidirectional estimator alone.
int cr_render_bdpt_strategy(const cr_scene* s, const rd_render_params* p, int32_t threads, int32_t strat_s,
                            int32_t strat_t, float* out_rgb) {
    if (strat_s < 0 || strat_t < 1) return RD_ERR_INVALID;
    return render_bdpt(s, p, threads, out_rgb, nullptr, nullptr, nullptr, strat_s, strat_t);
}
static int render_bdpt(const cr_scene* s, const rd_render_params* p, int32_t threads, float* out_rgb, float* out_sum2,
                       uint64_t* stats, double* seconds, int only_s, int only_t) {
    if (!s || !p || !out_rgb) return RD_ERR_INVALID;
    int W = s->cam.W, H = s->cam.H;
    int x0 = p->x0, y0 = p->y0, x1 = p->x1, y1 = p->y1;
    if (x0 == 0 && y0 == 0 && x1 == 0 && y1 == 0) { x1 = W; y1 = H; }
    if (x0 < 0 || y0 < 0 || x1 > W || y1 > H || x0 >= x1 || y0 >= y1 || p->spp <= 0 || p->spp >= (1 << 29)) {
        set_err("bad crop/spp");
        return RD_ERR_INVALID;
    }
    const int bs = 32;
    int nbx = (x1 - x0 + bs - 1) / bs, nby = (y1 - y0 + bs - 1) / bs;
    int nblocks = nbx * nby;
    std::atomic<int> next(0);
    std::mutex mu;
    uint64_t tot[5] = {0, 0, 0, 0, 0};
    if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
    auto t0 = std::chrono::steady_clock::now();
    auto worker = [&]() {
        uint64_t st[5] = {0, 0, 0, 0, 0};
        while (true) {
            int b = next.fetch_add(1);
            if (b >= nblocks) break;
            int bx = x0 + (b % nbx) * bs, by = y0 + (b / nbx) * bs;
            int ex = std::min(bx + bs, x1), ey = std::min(by + bs, y1);
            for (int x = bx; x < ex; x++) {
                for (int y = by; y < ey; y++) {
                    double sr = 0, sg = 0, sb = 0;
                    C3 sum2(0.0f);
                    for (int t = 0; t < p->spp; t++) {
                        C3 c = bd_sample(*s, p->seed, x, y, t, st, only_s, only_t);
                        sr += (double)c.r; sg += (double)c.g; sb += (double)c.b;
                        sum2 = sum2 + c * c;
                    }
                    size_t o = ((size_t)x * H + y) * 3;
                    const double n = (double)p->spp;
                    out_rgb[o] = (float)(sr / n); out_rgb[o + 1] = (float)(sg / n); out_rgb[o + 2] = (float)(sb / n);
                    if (out_sum2) { out_sum2[o] = sum2.r; out_sum2[o + 1] = sum2.g; out_sum2[o + 2] = sum2.b; }
                }
            }
        }
        std::lock_guard<std::mutex> g(mu);
        for (int i = 0; i < 5; i++) tot[i] += st[i];
    };
    std::vector<std::thread> pool;
    for (int i = 0; i < threads; i++) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (stats) for (int i = 0; i < 5; i++) stats[i] = tot[i];
    return RD_OK;
}

// Camera probes: SampleWi of a world point {ok, sx, sy, we, pdf, dist} and
// Importance of a camera-space direction.
int cr_camera_sample_wi(const cr_scene* s, const float* pos, float* out6) {
    if (!s || !pos || !out6) return RD_ERR_INVALID;
    Camera::Wi w = s->cam.sample_wi(V3(pos[0], pos[1], pos[2]));
    out6[0] = w.ok ? 1.0f : 0.0f; out6[1] = w.sx; out6[2] = w.sy; out6[3] = w.we; out6[4] = w.pdf; out6[5] = w.dist;
    return RD_OK;
}
float cr_camera_importance(const cr_scene* s, const float* d) {
    return s ? s->cam.importance(V3(d[0], d[1], d[2])) : 0.0f;
}
// Light.SampleEmit probe for light `light` with RNG key `key`: out =
// {p.xyz, dir.xyz, n.xyz, le.rgb, pdf_pos, pdf_dir}; returns 0 when the
// emission pdf is 0.
int cr_light_sample_emit(const cr_scene* s, int32_t light, uint64_t key, float* out14) {
    if (!s || light < 0 || light >= (int)s->lights.size()) return -1;
    Rng rng;
    rng.key = key;
    rng.n = 0;
    Emit e;
    if (!light_sample_emit(*s, light, rng, &e)) return 0;
    const float v[14] = {e.p.x, e.p.y, e.p.z, e.dir.x, e.dir.y, e.dir.z, e.n.x, e.n.y, e.n.z,
                         e.le.r, e.le.g, e.le.b, e.pdf_pos, e.pdf_dir};
    for (int i = 0; i < 14; i++) out14[i] = v[i];
    return 1;
}

int cr_adjoint(const cr_scene* s, const rd_render_params* p, int32_t threads, const float* dl_dimage,
               float* const* tex_grads) {
    if (!s || !p || !dl_dimage || !tex_grads || p->spp <= 0) return RD_ERR_INVALID;
    if (s->integ.type != RD_INTEGRATOR_PATH) { set_err("the adjoint covers the path integrator"); return RD_ERR_UNSUPPORTED; }
    int W = s->cam.W, H = s->cam.H;
    int x0 = p->x0, y0 = p->y0, x1 = p->x1, y1 = p->y1;
    if (x0 == 0 && y0 == 0 && x1 == 0 && y1 == 0) { x1 = W; y1 = H; }
    if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
    std::atomic<int> next(x0);
    std::mutex mu;
    AdjGrads total;
    total.g.resize(s->tex.size());
    for (size_t t = 0; t < s->tex.size(); t++)
        if (tex_grads[t]) total.g[t].assign(s->tex[t].w > 0 ? (size_t)s->tex[t].w * s->tex[t].h * 3 : 3, 0.0);
    auto worker = [&]() {
        AdjGrads G;
        G.g.resize(total.g.size());
        for (size_t t = 0; t < total.g.size(); t++) if (!total.g[t].empty()) G.g[t].assign(total.g[t].size(), 0.0);
        while (true) {
            int x = next.fetch_add(1);
            if (x >= x1) break;
            for (int y = y0; y < y1; y++) {
                size_t o = ((size_t)x * H + y) * 3;
                C3 delta = C3(dl_dimage[o], dl_dimage[o + 1], dl_dimage[o + 2]) / (float)p->spp;
                for (int t = 0; t < p->spp; t++) {
                    C3 ltot = pixel_sample(*s, p->seed, x, y, t, nullptr);   // the primal replay
                    if (!ltot.valid()) continue;
                    Rng rng;
                    rng.key = path_key(p->seed, (uint32_t)(y * s->cam.W + x), (uint32_t)t);
                    rng.n = 0;
                    float ox = rng.next_float(), oy = rng.next_float();
                    Ray ray = s->cam.sample_ray((float)x + ox, (float)y + oy);
                    li_adjoint(*s, ray, rng, ltot, delta, G);
                }
            }
        }
        std::lock_guard<std::mutex> g(mu);
        for (size_t t = 0; t < G.g.size(); t++)
            for (size_t i = 0; i < G.g[t].size(); i++) total.g[t][i] += G.g[t][i];
    };
    std::vector<std::thread> pool;
    for (int i = 0; i < threads; i++) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
    for (size_t t = 0; t < total.g.size(); t++)
        for (size_t i = 0; i < total.g[t].size(); i++) tex_grads[t][i] += (float)total.g[t][i];
    return RD_OK;
}

float cr_rho_ratio(const cr_scene* s, int32_t material, const float* uv, const float* wo, const float* wi) {
    if (!s || material < 0 || material >= (int)s->mats.size()) return 0.0f;
    return rho_ratio(*s, s->mats[material], V2(uv[0], uv[1]), V3(wo[0], wo[1], wo[2]), V3(wi[0], wi[1], wi[2]));
}

int cr_sample_li(const cr_scene* s, uint32_t seed, int32_t x, int32_t y, int32_t sample, float* rgb) {
    if (!s) return RD_ERR_INVALID;
    C3 c = pixel_sample(*s, seed, x, y, sample, nullptr);
    rgb[0] = c.r; rgb[1] = c.g; rgb[2] = c.b;
    return RD_OK;
}

void cr_dist1d_sample(const float* f, int32_t n, float u, float* out3, int32_t* offset) {
    Dist1D d;
    d.build(f, n);
    float pdf;
    int off;
    out3[0] = d.sample_continuous(u, &pdf, &off);
    out3[1] = pdf;
    out3[2] = d.discrete_pdf(off);
    *offset = off;
}
void cr_dist2d_sample(const float* f2, int32_t nu, int32_t nv, float u0, float u1, float* out3, int32_t* vj,
                      int32_t* vi) {
    Dist2D d;
    d.build(f2, nu, nv);
    float pdf;
    int j, i;
    V2 r = d.sample_continuous(u0, u1, &pdf, &j, &i);
    out3[0] = r.x; out3[1] = r.y; out3[2] = pdf;
    *vj = j; *vi = i;
}
float cr_dist2d_pdf(const float* f2, int32_t nu, int32_t nv, float u, float v) {
    Dist2D d;
    d.build(f2, nu, nv);
    return d.continuous_pdf(u, v);
}
int cr_env_le(const cr_scene* s, int32_t env, const float* dir, float* rgb) {
    if (!s || env < 0 || env >= (int)s->env.size()) return RD_ERR_INVALID;
    C3 c = env_le(*s, s->env[env], V3(dir[0], dir[1], dir[2]));
    rgb[0] = c.r; rgb[1] = c.g; rgb[2] = c.b;
    return RD_OK;
}
float cr_env_pdf(const cr_scene* s, int32_t env, const float* wi) {
    if (!s || env < 0 || env >= (int)s->env.size()) return -1.0f;
    return env_pdf_dir(s->env[env], V3(wi[0], wi[1], wi[2]));
}
int cr_env_sample(const cr_scene* s, int32_t env, float u0, float u1, float* wdir, float* pdf, float* uv) {
    if (!s || env < 0 || env >= (int)s->env.size()) return -1;
    V3 w;
    V2 t;
    float p = 0;
    if (!env_sample_dir(s->env[env], u0, u1, &w, &p, &t)) { *pdf = 0; return 0; }
    wdir[0] = w.x; wdir[1] = w.y; wdir[2] = w.z;
    *pdf = p; uv[0] = t.x; uv[1] = t.y;
    return 1;
}
int cr_env_bounds(const cr_scene* s, int32_t env, float* center, float* radius) {
    if (!s || env < 0 || env >= (int)s->env.size()) return RD_ERR_INVALID;
    const EnvLight& e = s->env[env];
    center[0] = e.center.x; center[1] = e.center.y; center[2] = e.center.z;
    *radius = e.radius;
    return RD_OK;
}

void cr_medium_tr(const rd_medium_desc* m, float tmin, float tmax, float* rgb) {
    C3 t = medium_tr(make_homo(*m), tmin, tmax);
    rgb[0] = t.r; rgb[1] = t.g; rgb[2] = t.b;
}
int cr_medium_sample(const rd_medium_desc* m, const float* o, const float* d, float tmin, float tmax, uint64_t key,
                     float* p, float* tr, float* t) {
    Rng rng;
    rng.key = key;
    rng.n = 0;
    Ray r(V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]), tmin, tmax);
    MSample ms = medium_sample(make_homo(*m), r, rng);
    p[0] = ms.p.x; p[1] = ms.p.y; p[2] = ms.p.z;
    tr[0] = ms.tr.r; tr[1] = ms.tr.g; tr[2] = ms.tr.b;
    *t = ms.t;
    return ms.medium ? 1 : 0;
}
float cr_hg_p(float g, const float* wo, const float* wi) {
    Homo h{};
    h.g = g;
    return hg_p(h, V3(wo[0], wo[1], wo[2]), V3(wi[0], wi[1], wi[2]));
}
float cr_hg_sample(float g, const float* wo, uint64_t key, float* wi) {
    Homo h{};
    h.g = g;
    Rng rng;
    rng.key = key;
    rng.n = 0;
    PSample ps = hg_sample(h, V3(wo[0], wo[1], wo[2]), rng);
    wi[0] = ps.wi.x; wi[1] = ps.wi.y; wi[2] = ps.wi.z;
    return ps.p;
}

int cr_material_eval(const cr_scene* s, int32_t material, const float* wo, const float* wi, const float* uv,
                     float* fr_rgb, float* pdf) {
    if (!s || material < 0 || material >= (int)s->mats.size()) return RD_ERR_INVALID;
    const rd_material_desc& m = s->mats[material];
    V3 o(wo[0], wo[1], wo[2]), i(wi[0], wi[1], wi[2]);
    V2 u(uv[0], uv[1]);
    C3 f = material_fr(*s, m, o, i, u);
    fr_rgb[0] = f.r; fr_rgb[1] = f.g; fr_rgb[2] = f.b;
    *pdf = material_pdf(*s, m, o, i, u);
    return RD_OK;
}

int cr_material_sample(const cr_scene* s, int32_t material, const float* wo, const float* uv, uint64_t key,
                       float* wi, float* fr_rgb, float* pdf, int32_t* type) {
    if (!s || material < 0 || material >= (int)s->mats.size()) return RD_ERR_INVALID;
    Rng rng; rng.key = key; rng.n = 0;
    BSample b = material_sample(*s, s->mats[material], V3(wo[0], wo[1], wo[2]), V2(uv[0], uv[1]), rng);
    wi[0] = b.wi.x; wi[1] = b.wi.y; wi[2] = b.wi.z;
    fr_rgb[0] = b.fr.r; fr_rgb[1] = b.fr.g; fr_rgb[2] = b.fr.b;
    *pdf = b.pdf;
    *type = b.type;
    return RD_OK;
}

// The two probes above under TransportMode.Importance (imp = 1) or Radiance (0).
int cr_material_eval_mode(const cr_scene* s, int32_t material, int32_t imp, const float* wo, const float* wi,
                          const float* uv, float* fr_rgb) {
    if (!s || material < 0 || material >= (int)s->mats.size()) return RD_ERR_INVALID;
    C3 f = material_fr(*s, s->mats[material], V3(wo[0], wo[1], wo[2]), V3(wi[0], wi[1], wi[2]), V2(uv[0], uv[1]),
                       imp != 0);
    fr_rgb[0] = f.r; fr_rgb[1] = f.g; fr_rgb[2] = f.b;
    return RD_OK;
}
int cr_material_sample_mode(const cr_scene* s, int32_t material, int32_t imp, const float* wo, const float* uv,
                            uint64_t key, float* wi, float* fr_rgb, float* pdf, int32_t* type) {
    if (!s || material < 0 || material >= (int)s->mats.size()) return RD_ERR_INVALID;
    Rng rng; rng.key = key; rng.n = 0;
    BSample b = material_sample(*s, s->mats[material], V3(wo[0], wo[1], wo[2]), V2(uv[0], uv[1]), rng, imp != 0);
    wi[0] = b.wi.x; wi[1] = b.wi.y; wi[2] = b.wi.z;
    fr_rgb[0] = b.fr.r; fr_rgb[1] = b.fr.g; fr_rgb[2] = b.fr.b;
    *pdf = b.pdf;
    *type = b.type;
    return RD_OK;
}

// Batched forms of the two above for quadrature and sampling statistics:
// n direction pairs (eval) or n samples with keys key0 + j (sample).
int cr_material_eval_n(const cr_scene* s, int32_t material, int32_t n, const float* wo, const float* wi,
                       const float* uv, float* fr_rgb, float* pdf) {
    for (int32_t j = 0; j < n; j++) {
        int rc = cr_material_eval(s, material, wo + 3 * j, wi + 3 * j, uv, fr_rgb + 3 * j, pdf + j);
        if (rc) return rc;
    }
    return RD_OK;
}

int cr_material_sample_n(const cr_scene* s, int32_t material, int32_t n, const float* wo, const float* uv,
                         uint64_t key0, float* wi, float* fr_rgb, float* pdf, int32_t* type) {
    for (int32_t j = 0; j < n; j++) {
        int rc = cr_material_sample(s, material, wo + 3 * j, uv, key0 + (uint64_t)j, wi + 3 * j, fr_rgb + 3 * j,
                                    pdf + j, type + j);
        if (rc) return rc;
    }
    return RD_OK;
}

float cr_fresnel_dielectric(float ci, float ei, float et) { return fresnel_dielectric(ci, ei, et); }

// BSSRDF probes
void cr_sss_lut(float* out) {
    const std::vector<float>& l = scatter_lut();
    for (size_t i = 0; i < l.size(); i++) out[i] = l[i];
}

float cr_fresnel_moment(float eta) { return fresnel_moment(eta); }

void cr_sss_profile(const float* a, const float* d, int32_t n, const int32_t* channel, const float* u, float* out) {
    Profile p(C3(a[0], a[1], a[2]), C3(d[0], d[1], d[2]));
    for (int32_t j = 0; j < n; j++) {
        float r = p.sample_sr(channel[j], u[j]);
        C3 sr = p.sr(r);
        float* o = out + 6 * j;
        o[0] = r; o[1] = p.pdf_sr(channel[j], r); o[2] = sr.r; o[3] = sr.g; o[4] = sr.b; o[5] = p.dch(channel[j]);
    }
}

void cr_sss_adapter(float eta, int32_t n, const float* wo, const float* wi, float* out) {
    for (int32_t j = 0; j < n; j++) {
        V3 o(wo[3 * j], wo[3 * j + 1], wo[3 * j + 2]), i(wi[3 * j], wi[3 * j + 1], wi[3 * j + 2]);
        C3 f = adapter_fr(eta, i);
        out[4 * j] = f.r; out[4 * j + 1] = f.g; out[4 * j + 2] = f.b; out[4 * j + 3] = adapter_pdf(o, i);
    }
}

int cr_sss_sample_pi(const cr_scene* s, int32_t prim, const float* P, const float* N, const float* wr, const float* uv,
                     uint64_t key0, int32_t n, float* out) {
    if (!s || prim < 0 || prim >= (int)s->prims.size()) return RD_ERR_INVALID;
    const int mat = s->prims[prim].material;
    if (mat < 0) return RD_ERR_INVALID;
    Inct po;
    po.P = V3(P[0], P[1], P[2]);
    po.UV = V2(uv[0], uv[1]);
    po.T = 0;
    po.prim = prim;
    po.sh = Frame(normalize(V3(N[0], N[1], N[2])));
    po.wr = V3(wr[0], wr[1], wr[2]);
    float eta;
    Profile prof = sss_profile(*s, s->mats[mat], po.UV, &eta);
    for (int32_t j = 0; j < n; j++) {
        Rng rng; rng.key = key0 + (uint64_t)j; rng.n = 0;
        uint64_t st[5] = {0, 0, 0, 0, 0};
        SssSample r = sss_sample_pi(*s, po, mat, prof, eta, rng, st);
        float* o = out + 9 * j;
        o[0] = r.ok ? 1.0f : 0.0f;
        o[1] = r.S.r; o[2] = r.S.g; o[3] = r.S.b; o[4] = r.pdf;
        o[5] = r.pi.P.x; o[6] = r.pi.P.y; o[7] = r.pi.P.z;
        o[8] = (float)st[0];
    }
    return RD_OK;
}

void cr_fresnel_conductor(float ci, const float* eta, const float* k, float* out) {
    C3 f = fresnel_conductor(ci, C3(1.0f), C3(eta[0], eta[1], eta[2]), C3(k[0], k[1], k[2]));
    out[0] = f.r; out[1] = f.g; out[2] = f.b;
}

void cr_mf_eval(int32_t kind, float roughness, float aniso, const float* wh, const float* wo, const float* wi, float* out) {
    MfDist d(kind, roughness, aniso);
    V3 h(wh[0], wh[1], wh[2]), o(wo[0], wo[1], wo[2]), i(wi[0], wi[1], wi[2]);
    out[0] = d.D(h);
    out[1] = d.G(o, i);
    out[2] = d.pdf(o, h);
}

void cr_mf_sample(int32_t kind, float roughness, float aniso, const float* wo, uint64_t key, float* wh) {
    MfDist d(kind, roughness, aniso);
    Rng rng; rng.key = key; rng.n = 0;
    V3 h = d.sample_wh(V3(wo[0], wo[1], wo[2]), rng);
    wh[0] = h.x; wh[1] = h.y; wh[2] = h.z;
}

void cr_rng_floats(uint64_t key, int32_t n, float* out) {
    Rng rng; rng.key = key; rng.n = 0;
    for (int i = 0; i < n; i++) out[i] = rng.next_float();
}

uint64_t cr_path_key(uint32_t seed, uint32_t pixel, uint32_t sample) { return path_key(seed, pixel, sample); }

}  // extern "C"
