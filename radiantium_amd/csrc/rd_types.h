// rd_types.h — the scene and wavefront layouts held in HBM.
//
// BVH nodes keep the reference's 32-byte LinearBVHNode (Accelerators/Bvh.cs:18-26)
// so one node is two 16-byte loads:
//   float4 lo = {min.x, min.y, min.z, max.x}
//   float4 hi = {max.y, max.z, offset (int bits), count | axis << 16 (int bits)}
// offset is PrimitivesIndex for leaves and SecondChildIndex for interior nodes.
// Primitives are stored in BVH leaf order, 48 bytes each (three float4):
//   triangle: {a.xyz, orig}, {b.xyz, kind=0}, {c.xyz, 0}
//   sphere:   {center.xyz, orig}, {radius, 0, 0, kind=1}, {0,0,0,0}
//   instance: {instance (int bits), 0, 0, orig}, {0, 0, 0, kind=2}, {0,0,0,0}
// so a leaf's primitives are contiguous and each test reads 36 (triangle) or
// 16 (sphere) payload bytes.
// Instanced scenes keep every set's BVH in the same two arrays after the
// top-level tree (child and primitive offsets absolute); a set triangle's
// "orig" is its DevPrim index (after the scene primitives), and its vertices
// are in the set's model space.
#pragma once
#include <stdint.h>
#include "rd_math.h"

namespace rd {

enum : int { kTri = 0, kSphere = 1, kInst = 2 };
enum : int {
    kMatDiffuse = 0, kMatRoughMetal = 1, kMatRoughGlass = 2, kMatRoughPlastic = 3,
    kMatPerfectGlass = 4, kMatPerfectMirror = 5, kMatDisney = 6, kMatSubsurface = 7
};
// Sets of material kinds a shading kernel is compiled for (bit k: kind k).  A
// scene's set is known on the host, and the kernels compiled for the smallest
// set that holds it keep only those kinds' code: the same arithmetic for
// every material they meet, fewer registers.
constexpr unsigned mat_bit(int k) { return 1u << k; }
constexpr unsigned kMatsAll = 0xffu;
constexpr unsigned kMatsDelta = mat_bit(kMatDiffuse) | mat_bit(kMatPerfectGlass) | mat_bit(kMatPerfectMirror);
constexpr unsigned kMatsMicrofacet = kMatsDelta | mat_bit(kMatRoughMetal) | mat_bit(kMatRoughGlass);
// the smallest compiled set holding `mask`
constexpr unsigned mat_set(unsigned mask) {
    return (mask & ~kMatsDelta) == 0 ? kMatsDelta : ((mask & ~kMatsMicrofacet) == 0 ? kMatsMicrofacet : kMatsAll);
}
// BxdfType flags (IBxdf.cs:7-15)
enum : int { kReflection = 1, kTransmission = 2, kDiffuse = 4, kGlossy = 8, kSpecular = 16, kSubsurface = 32 };

struct DevTexture {
    int w, h, wrap, filter;
    float v[3];
    int _pad;
    int64_t offset;  // float offset into texel pool (3 floats per texel)
};

struct DevMaterial {
    int type, two_sided, dist, thin;
    int tex[16];
    float eta_a, eta_b, kd, ks;
    float eta[3], k[3];
    float _pad[2];
};

struct DevLight {
    int type;     // 0 area, 1 infinite
    int prim;     // area: original primitive index of the emitter shape
    float le[3];
    int tex;      // infinite: index into DevScene::envs
    float area;   // Shape.SurfaceArea of the emitter
    int _pad;
};

// InfiniteAreaLight (Lights/InfiniteAreaLight.cs) after Preprocess
// (Light.cs:181-187): the radiance map, the linear part of the model/world
// transforms (row-major M11..M33, row vectors as Vector3.TransformNormal),
// the scene's bounding sphere, and the Distribution2D tables (host-built).
struct DevEnv {
    int tex, w, h, _pad;
    float m2w[9], w2m[9];
    float center[3], radius;
    const float* cond_func;  // [h][w]
    const float* cond_cdf;   // [h][w+1]
    const float* marg_func;  // [h]
    const float* marg_cdf;   // [h+1]
    float marg_int;
    int _pad2;
};

// Homogeneous (Mediums/Homogeneous.cs:14-20), coefficients already scaled:
// sa = sigma_a * scale, ss = sigma_s * scale, st = sa + ss.
struct DevMedium {
    float sa[3], ss[3], st[3], g;
};

// Per original primitive (rd_prim_desc order): shading record.
struct DevPrim {
    int kind;      // kTri / kSphere / kInst
    int index;     // triangle, sphere or instance index
    int material;  // -1: no surface
    int light;     // -1: not emissive
    int med_in, med_out;
    float area;
    int flags;     // triangles: bit0 has normal, bit1 has uv
};

// InstancedTransform (Primitive.cs:94-108): 4x3 row-major (rows 1-3 = the
// linear part, row 4 = translation, row vectors), WorldToModel inverted on
// the host; root = node index of the set's BVH root.
struct DevInstance {
    float w2m[12];
    float m2w[12];
    int root;
    int _pad[3];
};

struct DevCamera {
    float clip_to_camera[16];   // row-major M11..M44
    float camera_to_world[16];
    float inv_res_x, inv_res_y, near_z, far_z;
    int width, height;
};

// Integrator parameters (PathTracer ctor, Integrators/PathTracer.cs:34-41)
struct DevIntegrator {
    int type, max_depth, min_depth, method, strategy;
    float rr_threshold;
};

struct DevScene {
    const float4* nodes;   // 2 per node
    const float4* prims;   // 3 per ordered primitive
    int num_nodes;
    int num_prims;
    int stack_needed;      // deepest interior path (entries a DFS stack can hold)
    int num_lights;
    const DevPrim* prim_info;
    const int* tri_idx;
    const float* pos;
    const float* nrm;
    const float* uv;
    const float4* spheres;
    const float* sphere_w2m;   // 9 per sphere (WorldToModel linear part), null = identity
    const DevMaterial* mats;
    const DevTexture* texs;
    const float* texels;
    const DevLight* lights;
    const DevMedium* media;
    const DevEnv* envs;    // Scene.InfiniteLights order
    int num_envs;
    const DevInstance* insts;
    int num_insts;
    int global_medium;
    DevIntegrator integ;   // the scene's integrator (a batch may mix parameters)
    DevCamera cam;
    unsigned mat_mask;     // material kinds the scene holds (bit k: kind k); read on the host
    int octree;            // 1: nodes / prims hold the loose octree of Accelerators/Octree.cs
};


}  // namespace rd
