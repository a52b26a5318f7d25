/*
 * radiantium_hip.h — C ABI of the MI355X (gfx950) wavefront path tracer.
 *
 * This is the drop-in boundary below the reference's plugin surface.  The
 * reference (C#, ksgfk/Radiantium) reaches its hot path through
 *   MonteCarloIntegrator.Li(Ray3F, Scene, Random)      Radiantium.Offline/IIntegrator.cs:13-24
 *   Renderer.{RenderTarget,Start,Wait,Stop,IsSuccess}  Radiantium.Offline/Renderer.cs:5-52
 * and calls Li once per pixel sample from BlockBasedRenderer.RenderTask
 * (Radiantium.Offline/Renderers/BlockBasedRenderer.cs:107-179).  A GPU
 * replacement therefore sits at the Renderer level: a C# WavefrontRenderer
 * calls rd_render() on a worker task and exposes the result as the
 * reference's column-major ColorBuffer (Radiantium.Core/ColorBuffer.cs:25-42).
 *
 * Conventions
 *  - Every call returns an rd_status; on failure rd_last_error() returns a
 *    thread-local message (the reference throws ArgumentException from
 *    RendererBuilder.Build, Config/RendererBuilder.cs:433-470, and turns
 *    render-time exceptions into IsSuccess=false, BlockBasedRenderer.cs:161-166).
 *  - Host buffers passed in are copied; the caller keeps ownership.  No
 *    callbacks into the caller.  One context is used from one host thread
 *    at a time (an internal mutex rejects concurrent use).
 *  - Image output is float[W*H*3] addressed (x*H + y)*3, as ColorBuffer.
 *  - Plain C types only: pointers, sizes, POD structs.
 */
#ifndef RADIANTIUM_HIP_H
#define RADIANTIUM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RD_ABI_VERSION 4

typedef enum rd_status {
    RD_OK = 0,
    RD_ERR_INVALID = 1,     /* bad argument or scene description            */
    RD_ERR_DEVICE = 2,      /* HIP runtime error                            */
    RD_ERR_OOM = 3,         /* device allocation failed                     */
    RD_ERR_CANCELLED = 4,   /* *cancel became non-zero during rd_render     */
    RD_ERR_STACK = 5,       /* traversal stack overflow (Bvh.StackSize=64)  */
    RD_ERR_UNSUPPORTED = 6, /* feature not built for this path              */
    RD_ERR_BUSY = 7,        /* context used from two threads at once        */
    RD_ERR_IO = 8           /* file could not be read or written            */
} rd_status;

/* ---- scene description (mirrors RendererBuilder.Build's products) ------ */

/* Material kinds: the registry names of Config/RendererBuilder.cs:86-213. */
typedef enum rd_material_type {
    RD_MAT_DIFFUSE = 0,        /* "diffuse"        Materials/Diffuse.cs        */
    RD_MAT_ROUGH_METAL = 1,    /* "rough_metal"    Materials/RoughMetal.cs     */
    RD_MAT_ROUGH_GLASS = 2,    /* "rough_glass"    Materials/RoughGlass.cs     */
    RD_MAT_ROUGH_PLASTIC = 3,  /* "rough_plastic"  Materials/RoughPlastic.cs   */
    RD_MAT_PERFECT_GLASS = 4,  /* "perfect_glass"  Materials/PerfectGlass.cs   */
    RD_MAT_PERFECT_MIRROR = 5, /* "perfect_mirror" Materials/PerfectMirror.cs  */
    RD_MAT_DISNEY = 6,         /* "disney"         Materials/Disney.cs         */
    RD_MAT_SUBSURFACE = 7      /* "subsurface"     Materials/Subsurface.cs     */
} rd_material_type;

/* Texture slot meaning per material (index into rd_scene_desc.textures):
 *   DIFFUSE        tex[0]=kd
 *   ROUGH_METAL    tex[0]=roughness tex[1]=anisotropic          (eta,k)
 *   ROUGH_GLASS    tex[0]=r tex[1]=t tex[2]=roughness tex[3]=anisotropic (eta_a,eta_b)
 *   ROUGH_PLASTIC  tex[0]=r tex[1]=roughness tex[2]=anisotropic (kd,ks,eta_a,eta_b)
 *   PERFECT_GLASS  tex[0]=r tex[1]=t                            (eta_a,eta_b)
 *   PERFECT_MIRROR tex[0]=r
 *   DISNEY         tex[0..14] = base_color metallic roughness eta specular
 *                  specular_tint anisotropic sheen sheen_tint clearcoat
 *                  clearcoat_gloss scatting_distance transmission
 *                  transmission_roughness flatness                (thin)
 *   SUBSURFACE     tex[0]=r tex[1]=t tex[2]=a tex[3]=scatting_distance
 *                                                              (eta_a,eta_b)
 * A non-thin Disney material with a non-black scatting_distance and a
 * subsurface material with one sample the BSSRDF of Bssrdf.cs in the path
 * integrator's Mis method (PathTracer.cs:244-259; the reference's branch is
 * marked unfinished and is reproduced as written). */
typedef struct rd_material_desc {
    int32_t type;       /* rd_material_type                                 */
    int32_t two_sided;  /* "is_two_side"                                    */
    int32_t dist;       /* 0 = Beckmann, 1 = GGX (Bxdf/Microfacet.cs:9-13)  */
    int32_t thin;       /* Disney "is_thin"                                 */
    int32_t tex[16];
    float eta_a, eta_b; /* dielectric IORs (defaults 1.000277 / 1.5046)     */
    float kd, ks;       /* plastic lobe weights                             */
    float eta[3];       /* conductor eta (default Cu, Spectrum.cs:64)       */
    float k[3];         /* conductor k                                      */
} rd_material_desc;

/* Texture2D: constant (width==0) or image.  Images are 3-channel float,
 * column-major (x*height + y)*3 like ColorBuffer; sampled by
 * TextureSampler.Sample (TextureSampler.cs:28-72). */
typedef struct rd_texture_desc {
    int32_t width, height; /* 0,0 => constant texture                       */
    int32_t wrap;          /* 0 = Clamp, 1 = Repeat                         */
    int32_t filter;        /* 0 = Nearest, 1 = Linear                       */
    float value[3];        /* constant colour (ConstColorTexture2D)         */
    int32_t _pad;
    const float* data;     /* width*height*3 floats, or NULL for constants  */
} rd_texture_desc;

typedef enum rd_light_type {
    RD_LIGHT_AREA = 0,    /* "diffuse_area"  Lights/DiffuseAreaLight.cs     */
    RD_LIGHT_INFINITE = 1 /* "infinite"      Lights/InfiniteAreaLight.cs    */
} rd_light_type;

typedef struct rd_light_desc {
    int32_t type;       /* rd_light_type                                     */
    int32_t prim;       /* AREA: primitive index of the emitting shape       */
    float le[3];        /* AREA: emitted radiance "le"                       */
    int32_t tex;        /* INFINITE: texture index of the radiance map       */
    float to_world[16]; /* INFINITE: model-to-world, row-major M11..M44      */
} rd_light_desc;

/* Homogeneous medium (Mediums/Homogeneous.cs). */
typedef struct rd_medium_desc {
    float sigma_a[3];
    float sigma_s[3];
    float g;
    float scale;
} rd_medium_desc;

typedef enum rd_prim_kind { RD_PRIM_TRIANGLE = 0, RD_PRIM_SPHERE = 1, RD_PRIM_INSTANCE = 2 } rd_prim_kind;

/* One GeometricPrimitive (Primitive.cs:44-78) or InstancedPrimitive
 * (Primitive.cs:144-189), in the builder's order (entity order, then shapes,
 * then model triangles, then instances; RendererBuilder.cs:940-999).  The
 * position of a primitive in this array is the "primitive index" that
 * rd_trace_closest reports.  For RD_PRIM_INSTANCE, `index` is an
 * rd_scene_desc.instances entry and `light` must be -1 (InstancedPrimitive
 * cannot carry a light). */
typedef struct rd_prim_desc {
    int32_t kind;           /* rd_prim_kind                                   */
    int32_t index;          /* triangle or sphere index                       */
    int32_t material;       /* -1: no surface (a medium boundary)             */
    int32_t light;          /* -1: not emissive                               */
    int32_t medium_inside;  /* -1: vacuum                                     */
    int32_t medium_outside; /* -1: vacuum                                     */
} rd_prim_desc;

/* One instanced mesh set (RendererBuilder's InstancedPrimitivesEntry,
 * Config/RendererBuilder.cs:331-340, built at :895-921): triangles
 * [first_triangle, first_triangle + num_triangles) of rd_scene_desc, already
 * in the set's model space (its local position/rotation/scale applied, as
 * TriangleMesh does), each a ShapeWrapperPrimitive (Primitive.cs:110-142)
 * under the set's own BVH.  These triangles are not scene primitives. */
typedef struct rd_instance_set_desc {
    int32_t first_triangle;
    int32_t num_triangles;
    int32_t bvh_max_prim;   /* the set's accelerator "max_prim"; 0 = 1 (Bvh default) */
    int32_t _pad;
} rd_instance_set_desc;

/* InstancedTransform (Primitive.cs:94-108): the entity's ModelToWorld. */
typedef struct rd_instance_desc {
    int32_t set;            /* rd_scene_desc.instance_sets index            */
    int32_t _pad;
    float to_world[16];     /* row-major M11..M44 (row vectors)             */
} rd_instance_desc;

typedef struct rd_camera_desc {
    float fov, near_z, far_z;     /* degrees; defaults 60, 0.001, 1000        */
    int32_t width, height;        /* "screen"                                 */
    int32_t use_matrix;           /* 0: origin/target/up, 1: to_world         */
    float origin[3], target[3], up[3];
    float to_world[16];           /* row-major M11..M44 (after ReadMat4)      */
} rd_camera_desc;

/* "path" (Integrators/PathTracer.cs), "vol_path" (VolumetricPathTracer.cs),
 * "adjoint_particle" (AdjointParticleIntegrator.cs: light tracing, rendered
 * as AdjointParticleRenderer renders it), "bdpt" (BidirectionalPathTracer.cs
 * under BidirectionalPathTracingRenderer). */
typedef enum rd_integrator_type {
    RD_INTEGRATOR_PATH = 0,
    RD_INTEGRATOR_VOL_PATH = 1,
    RD_INTEGRATOR_PARTICLE = 2,
    RD_INTEGRATOR_BDPT = 3
} rd_integrator_type;
typedef enum rd_path_method { RD_METHOD_BXDF = 0, RD_METHOD_NEE = 1, RD_METHOD_MIS = 2 } rd_path_method;
typedef enum rd_light_strategy { RD_STRATEGY_UNIFORM = 0, RD_STRATEGY_ALL = 1 } rd_light_strategy;

typedef struct rd_integrator_desc {
    int32_t type;         /* rd_integrator_type ("path" / "vol_path")        */
    int32_t max_depth;    /* default -1                                      */
    int32_t min_depth;    /* default 3                                       */
    float rr_threshold;   /* default 0.99                                    */
    int32_t method;       /* rd_path_method, default Mis                     */
    int32_t strategy;     /* rd_light_strategy, default Uniform              */
} rd_integrator_desc;

typedef struct rd_scene_desc {
    int32_t num_vertices;
    const float* positions;      /* [num_vertices][3], world space           */
    const float* normals;        /* [num_vertices][3] or NULL                */
    const float* uvs;            /* [num_vertices][2] or NULL                */
    int32_t num_triangles;
    const int32_t* indices;      /* [num_triangles][3]                       */
    const uint8_t* tri_flags;    /* per triangle: bit0 has normal, bit1 has uv; NULL = all 0 */
    int32_t num_spheres;
    const float* spheres;        /* [num_spheres][4] center.xyz, radius      */
    const float* sphere_w2m;     /* [num_spheres][9] linear part of the sphere's
                                    WorldToModel (Sphere.cs:19-31), used for its
                                    UV (Sphere.cs:40); NULL = identity       */
    int32_t num_prims;
    const rd_prim_desc* prims;
    int32_t num_materials;
    const rd_material_desc* materials;
    int32_t num_textures;
    const rd_texture_desc* textures;
    int32_t num_lights;
    const rd_light_desc* lights; /* Scene.Lights order (RendererBuilder.cs:945-993) */
    int32_t num_media;
    const rd_medium_desc* media;
    int32_t global_medium;       /* -1: none                                 */
    int32_t bvh_max_prim;        /* accel "max_prim", default 1              */
    rd_camera_desc camera;
    rd_integrator_desc integrator;
    int32_t num_instance_sets;
    const rd_instance_set_desc* instance_sets;
    int32_t num_instances;
    const rd_instance_desc* instances;
    /* accelerator of the scene's primitives (the config's "accel",
     * RendererBuilder.cs:86-87): RD_ACCEL_BVH (bvh_max_prim above) or
     * RD_ACCEL_OCTREE (Accelerators/Octree.cs; not with instances) */
    int32_t accel;
    int32_t octree_max_depth;    /* octree "max_depth", default 10          */
    int32_t octree_max_count;    /* octree "max_count", default 10          */
    float octree_out_bound;      /* octree "out_bound", default 2           */
} rd_scene_desc;

#define RD_ACCEL_BVH 0
#define RD_ACCEL_OCTREE 1

/* ---- rays for the traversal oracle hooks -------------------------------- */

/* Ray3F (Radiantium.Core/Ray3F.cs:5-20); InvD is recomputed as 1/d. */
typedef struct rd_ray {
    float o[3];
    float tmin;
    float d[3];
    float tmax;
} rd_ray;

/* Closest hit: t, primitive index (rd_scene_desc.prims order, -1 on miss),
 * barycentric u, v (SurfacePoint, Intersection.cs:6-18). */
typedef struct rd_hit {
    float t;
    int32_t prim;
    float u, v;
} rd_hit;

typedef struct rd_render_params {
    int32_t spp;            /* renderer "spp", default 256                  */
    uint32_t seed;          /* counter-based RNG seed                       */
    int32_t x0, y0, x1, y1; /* pixel crop [x0,x1) x [y0,y1); all 0 = full   */
    int32_t wave_paths;     /* paths in flight per wave; 0 = as many as a
                               quarter of the device memory holds (every
                               path of the render where it fits)            */
    int32_t flags;          /* RD_RENDER_* bits                              */
    int64_t particles;      /* adjoint_particle scenes: the renderer's
                               "particle_count" (AdjointParticleRenderer;
                               0 = its default 16384); spp and the crop are
                               not used by that renderer                    */
} rd_render_params;

#define RD_RENDER_NO_GRAPH 1 /* launch the bounce loop kernel by kernel     */

typedef struct rd_render_stats {
    uint64_t rays_closest;   /* closest-hit BVH queries                      */
    uint64_t rays_any;       /* any-hit BVH queries (shadow rays)            */
    uint64_t paths;          /* camera samples traced                        */
    uint64_t vertices;       /* surface / medium vertices shaded             */
    uint64_t invalid;        /* samples replaced by 0 (Color3F.IsValid)      */
    double render_ms;        /* device time of the last rd_render            */
    int32_t waves;
    int32_t bounces;         /* deepest bounce processed                     */
    /* traversal work per query class: [0] extension (closest), [1] MIS
     * (closest), [2] shadow (any-hit): rays, nodes visited, primitives tested */
    uint64_t class_rays[3];
    uint64_t class_nodes[3];
    uint64_t class_prims[3];
} rd_render_stats;

typedef struct rd_scene_info {
    int32_t num_nodes;
    int32_t num_leaves;
    int32_t max_depth;      /* deepest leaf (root = 0)                        */
    int32_t num_prims;
    double build_ms;
    uint64_t device_bytes;
} rd_scene_info;

typedef struct rd_ctx rd_ctx;
typedef struct rd_scene rd_scene;

/* ---- context ------------------------------------------------------------- */
int rd_abi_version(void);
const char* rd_last_error(void);
/* Creates a context on HIP device `device`.  NULL on failure. */
rd_ctx* rd_create(int device);
void rd_destroy(rd_ctx* ctx);

/* ---- scene: replaces RendererBuilder.Build + new Bvh(...) --------------- */
/* Builds the scene's accelerator (the SAH BVH of Accelerators/Bvh.cs:89-124,
 * or the loose octree of Accelerators/Octree.cs:48-93 when desc->accel says
 * so) and uploads the flattened scene to HBM.  NULL on failure (see
 * rd_last_error). */
rd_scene* rd_scene_create(rd_ctx* ctx, const rd_scene_desc* desc);
void rd_scene_destroy(rd_scene* scene);
int rd_scene_get_info(const rd_scene* scene, rd_scene_info* info);
/* Copies the flattened BVH: nodes as 32-byte LinearBVHNode records
 * (Bvh.cs:18-26) and the BVH-ordered primitive indices (RD_ERR_INVALID for an
 * octree scene: rd_scene_export_octree). */
int rd_scene_export_bvh(const rd_scene* scene, void* nodes32, int32_t max_nodes,
                        int32_t* prim_order, int32_t max_prims);

/* Host-only BVH build over primitive bounds (bmin/bmax: [n][3]), the same
 * builder rd_scene_create uses (Bvh.RecursiveBuild + FlattenBVHTree,
 * Bvh.cs:234-442).  Needs no device.  nodes32 receives num_nodes 32-byte
 * records (capacity max_nodes; 2n-1 always suffices), prim_order n indices. */
int rd_bvh_build_host(const float* bmin, const float* bmax, int32_t n, int32_t max_prim, void* nodes32,
                      int32_t max_nodes, int32_t* prim_order, int32_t* num_nodes, int32_t* max_depth);

/* ---- image files --------------------------------------------------------- */
/* A ColorBuffer (Radiantium.Core/ColorBuffer.cs:8-30): pixel (x, y) at
 * data[(x * height + y) * channels]; y = 0 is the image's bottom row. */
typedef struct rd_image {
    int32_t width, height, channels;
    float* data;   /* owned by the library: rd_image_free */
} rd_image;
/* TextureUtility.LoadImageFromPath(path, needChannel, isFlipY, isCastToLinear)
 * (Radiantium.Core/TextureUtility.cs:118-203): PNG, OpenEXR (scanline;
 * NONE / RLE / ZIPS / ZIP / PIZ), Radiance RGBE (.hdr) and PFM, values as
 * Magick.NET-Q16-HDRI delivers them (OpenEXR through its half RGBA
 * interface: four channels, alpha 1 when absent).  need_channel <= 0 keeps
 * the file's channel count.  Errors: RD_ERR_IO, RD_ERR_INVALID (corrupt),
 * RD_ERR_UNSUPPORTED (tiled / deep / multi-part OpenEXR, PXR24 / B44 / DWA
 * compression, interlaced PNG); rd_image_last_error has the message. */
int rd_image_load(const char* path, int need_channel, int flip_y, int cast_to_linear, rd_image* out);
void rd_image_free(rd_image* img);
const char* rd_image_last_error(void);
/* ColorBuffer.SaveOpenExr (TextureUtility.cs:85-116): half RGBA, rows in
 * ColorBuffer order (its `k = i`: row 0 first, so the file reads upside down
 * next to SavePng's), each channel half(QuantumScale * (color * Quantum.Max))
 * as Magick writes it, alpha 1.  compression: 0 NONE (Magick's default) or
 * 3 ZIP.  data in ColorBuffer layout, channels >= 1 (ColorBuffer.GetRGB). */
int rd_image_save_exr(const char* path, const float* data, int width, int height, int channels, int compression);
/* ColorBuffer.SavePng (TextureUtility.cs:54-83): 8-bit RGBA (Png32), rows
 * flipped, ScaleQuantumToChar(clamp(color * Quantum.Max)).  to_srgb applies
 * Color3F.ToSrgb to every pixel first: what ResultOutput.Save's IsPngToSrgb
 * branch (Radiantium.Offline/ResultOutput.cs:26-38) means to do; its loop, as
 * written, converts only part of column 0 and leaves the rest black. */
int rd_image_save_png(const char* path, const float* data, int width, int height, int channels, int to_srgb);
/* No counterpart in the reference: PFM (float RGB, little-endian, rows from
 * the bottom, so ColorBuffer row 0 first), for lossless float output. */
int rd_image_save_pfm(const char* path, const float* data, int width, int height, int channels);

/* The loose octree of Accelerators/Octree.cs (ctor :48-93, RecursionBuild
 * :197-281, FlattenNode :283-315) over primitive bounds, on the host; needs
 * no device.  sizes3 receives {nodes, leaves, leaf entries}; with a null
 * buffer (or one too small) nothing is copied and RD_OK / RD_ERR_INVALID
 * tells, so a first call with nulls sizes the buffers.  nodes32: 32-byte
 * LinearNode records (bound, NodeCount or LeafIndex, NextHead = -1 for
 * leaves); leaf i holds leaf_prims[leaf_first[i] .. + leaf_count[i]). */
int rd_octree_build_host(const float* bmin, const float* bmax, int32_t n, int32_t max_depth, int32_t max_count,
                         float out_bound, void* nodes32, int32_t max_nodes, int32_t* leaf_first, int32_t* leaf_count,
                         int32_t max_leaves, int32_t* leaf_prims, int32_t max_leaf_prims, int32_t* sizes3);
/* The octree rd_scene_create built for an RD_ACCEL_OCTREE scene, in the
 * same form (RD_ERR_INVALID for a BVH scene). */
int rd_scene_export_octree(const rd_scene* scene, void* nodes32, int32_t max_nodes, int32_t* leaf_first,
                           int32_t* leaf_count, int32_t max_leaves, int32_t* leaf_prims, int32_t max_leaf_prims,
                           int32_t* sizes3);

/* Bvh.RecursiveBuild + FlattenBVHTree (Bvh.cs:234-442) on the device, the
 * tree rd_bvh_build_host builds, node for node: device bounds in (d_bmin,
 * d_bmax: [n][3]), device records out (d_nodes32: capacity max_nodes >= 2n-1
 * records of 32 bytes; d_prim_order: n indices).  Runs on the context's
 * stream and returns when the tree is complete. */
int rd_bvh_build_device(rd_ctx* ctx, const float* d_bmin, const float* d_bmax, int32_t n, int32_t max_prim,
                        void* d_nodes32, int32_t max_nodes, int32_t* d_prim_order, int32_t* num_nodes,
                        int32_t* max_depth);

/* ---- scene configs ------------------------------------------------------- */
/* The reference's scene config (JSON, as Radiantium.Cli/Program.cs:25-140
 * reads it) with its OBJ models and images, built into a scene
 * description the way RendererBuilder.Build builds the scene
 * (Config/RendererBuilder.cs:867-1010).  search_dir: base of relative asset
 * locations (NULL = the config's directory; RendererBuilder.GetPath).
 * params (may be NULL) receives the renderer block: spp (default 256) and
 * particle_count (particles, default 16384).  Images load as rd_image_load
 * loads them.  Integrators "path", "vol_path", "adjoint_particle" and "bdpt"
 * load; others ("ao", "gbuffer") give RD_ERR_UNSUPPORTED. */
typedef struct rd_json_scene rd_json_scene;
int rd_json_load(const char* path, const char* search_dir, rd_json_scene** out, rd_render_params* params);
const rd_scene_desc* rd_json_scene_desc(const rd_json_scene* scene);
void rd_json_scene_free(rd_json_scene* scene);
/* rd_json_load + rd_scene_create: the scene a C# WavefrontRenderer builds
 * from the same config path the CLI receives (Program.cs:27).  NULL on
 * failure (rd_last_error). */
rd_scene* rd_scene_load_json(rd_ctx* ctx, const char* path, const char* search_dir, rd_render_params* params);
/* Spectrum.TryGetMatel (Spectrum.cs:38-41): the RGB eta and k of a named
 * conductor ("Cu", "Au", "Ag", ...), for rd_material_desc.eta / .k. */
int rd_metal_eta_k(const char* name, float* eta, float* k);

/* ---- traversal oracle hooks: Scene.Intersect / IsOccluded ---------------- */
/* Bvh.Intersect(ray, out inct) (Bvh.cs:174-232) for n rays (host buffers). */
int rd_trace_closest(rd_ctx* ctx, rd_scene* scene, const rd_ray* rays, int32_t n, rd_hit* hits);
/* Bvh.Intersect(ray) (Bvh.cs:126-172): occluded[i] = 1 if any hit. */
int rd_trace_any(rd_ctx* ctx, rd_scene* scene, const rd_ray* rays, int32_t n, uint8_t* occluded);
/* rd_trace_closest that also reports, for a hit on an InstancedPrimitive
 * (Primitive.cs:165-187: prim = the instance's primitive index, t = world
 * distance to the transformed hit point, u/v = barycentrics in the set), the
 * hit triangle as an index into its instance set (0 .. num_triangles-1);
 * sub[i] = -1 for other hits and misses. */
int rd_trace_closest_sub(rd_ctx* ctx, rd_scene* scene, const rd_ray* rays, int32_t n, rd_hit* hits, int32_t* sub);
/* Device-pointer variants: rays/hits already in HBM; stream may be NULL.
 * counters (device, may be NULL) receives {nodes visited, prims tested}
 * summed over the batch as two uint64. */
int rd_trace_closest_device(rd_ctx* ctx, rd_scene* scene, const rd_ray* d_rays, int32_t n,
                            rd_hit* d_hits, uint64_t* d_counters, void* stream);
int rd_trace_any_device(rd_ctx* ctx, rd_scene* scene, const rd_ray* d_rays, int32_t n,
                        uint8_t* d_occluded, uint64_t* d_counters, void* stream);

/* ---- rendering: replaces BlockBasedRenderer + PathTracer/VolumetricPathTracer */
/* Renders the scene's camera with its integrator.  out_rgb: host float[W*H*3]
 * column-major, the per-pixel mean (BlockBasedRenderer.cs:167-178).
 * out_sum2 (optional): per-pixel sum of squared samples, same layout.
 * cancel (optional): polled between waves; non-zero => RD_ERR_CANCELLED
 * (Renderer.Stop, BlockBasedRenderer.cs:85-88).  progress (optional):
 * completed waves (Renderer.CompletedTaskCount). */
int rd_render(rd_ctx* ctx, rd_scene* scene, const rd_render_params* params, float* out_rgb,
              float* out_sum2, volatile int32_t* cancel, volatile int32_t* progress);
/* Same, with out_rgb / out_sum2 in device memory (no host copies). */
int rd_render_device(rd_ctx* ctx, rd_scene* scene, const rd_render_params* params,
                     float* d_out_rgb, float* d_out_sum2, volatile int32_t* cancel,
                     volatile int32_t* progress);
int rd_get_render_stats(const rd_ctx* ctx, rd_render_stats* stats);
/* rd_render's per-pixel mean without its final rounding to float: out_mean
 * (host double[W*H*3], same layout) receives sum / spp in float64.  Every
 * render keeps its per-pixel sums in float64 over float32 sample radiance
 * (BlockBasedRenderer.cs:140-178 sums in float32); rd_render rounds the mean
 * to float, this keeps it, e.g. for finite-difference gradient checks. */
int rd_render_mean_f64(rd_ctx* ctx, rd_scene* scene, const rd_render_params* params, double* out_mean);

/* rd_render_device without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The render's working memory (path pool, ray queues,
 * per-sample results, per-pixel sums, segment table) is carved from a fixed
 * work area of RD_NOALLOC_SCRATCH_BYTES that the library holds on every
 * device it runs on until the process ends (looked up once per device by
 * rd_create); the caller passes the scene and the device outputs as for
 * rd_render_device.  Limit: the render's pixels (its crop) must fit with
 * 64 B each (sums and one sample's results) beside a pool of 1024 paths,
 * about 2.09 M pixels (1920x1080 fits); larger renders give RD_ERR_INVALID
 * before any launch.
 * params->wave_paths > 0 caps the pool; the coherence sort and stage timing
 * are off (neither changes a result), so the image equals rd_render_device's
 * bit for bit.  Every context on a device shares that device's work area,
 * which keeps the last call's contents (each call clears what it reads
 * first): calls that overlap on one device, from several host threads or
 * processes' contexts, race. */
#define RD_NOALLOC_SCRATCH_BYTES (128u << 20)
int rd_render_device_noalloc(rd_ctx* ctx, rd_scene* scene, const rd_render_params* params,
                             float* d_out_rgb, float* d_out_sum2, volatile int32_t* cancel,
                             volatile int32_t* progress);

/* ---- tuning of a context's wavefront ------------------------------------
 * No counterpart in the reference (its CPU renderer has none of these
 * choices); for autotuners and benchmarks.  None of them changes a result:
 * every ray's query and every path's random numbers are the same under any
 * setting.  rd_set_tuning leaves fields that are -1 as they are. */
typedef struct rd_tuning {
    int32_t trace_schedule[3]; /* per query class [extension, shadow, MIS]:
                                  0 grid-stride (one ray per lane to its end),
                                  1 idle lanes refilled from a work cursor,
                                  2 64-ray chunks per wavefront from it,
                                  3 256-ray chunks, four rays per lane traced
                                    back to back                            */
    int32_t ray_sort;          /* extension rays traced in coherence order
                                  (direction octant, origin Morton code)    */
    int32_t ray_sort_bits;     /* sort key bits, 3..30                      */
    int32_t ray_sort_min;      /* smallest extension queue that is sorted   */
    int32_t ray_sort_shadow;   /* 1: shadow rays in that sorted order; 0:
                                  in the extension queue's own order        */
    int32_t overlap;           /* a bounce's shadow / MIS traversal and
                                  resolve on a second stream beside the next
                                  bounce's compaction, sort and extension
                                  traversal (once every path has started)   */
    int32_t device_loop;       /* the tail of a pass (every path started, at
                                  most tail_rays rays in the queue) runs
                                  device-driven: each bounce's kernels read
                                  their queue sizes on the device and the
                                  host checks for the end once per 8
                                  bounces; 0 off (one host sync per bounce),
                                  1 kernel launches, 2 two bounces captured
                                  as one hipGraph and replayed (default;
                                  rd_render_params.flags RD_RENDER_NO_GRAPH
                                  makes it 1 for that render)              */
    int32_t tail_rays;         /* queue size below which the tail starts    */
    int32_t chain_lockstep;    /* vol_path chains: 1 one segment per launch
                                  over the chain queues, the host looping
                                  until both are empty; 0 (default) every
                                  chain walked to its end by one lane in
                                  one launch                               */
    int32_t pool_log2;         /* cap of the slot pool when wave_paths is 0:
                                  at most 2^pool_log2 paths in flight
                                  (10..27); 0 = the memory rule alone      */
    int32_t tile;              /* edge of the pixel tiles paths are
                                  numbered in (1..64, default 8)            */
    int32_t stack_min;         /* traversal LDS stack variant at least this
                                  deep (0, or 24..64): a deeper variant
                                  than the tree needs, with 128-thread
                                  blocks from 40 entries on                */
    int32_t bvh_build;         /* where rd_scene_create builds its trees:
                                  0 the host, 1 the device for trees of 2^15
                                  primitives or more (default), 2 the device
                                  always; both give the same tree, node for
                                  node                                      */
} rd_tuning;

int rd_get_tuning(const rd_ctx* ctx, rd_tuning* out);
int rd_set_tuning(rd_ctx* ctx, const rd_tuning* tuning);

/* Autotuner: renders `params` (typically the production render with a few
 * spp) under candidate tunings, one knob at a time from the current setting
 * (extension and shadow schedules, the coherence sort, its key bits and the
 * shadow rays' order, the stream overlap, the slot pool's size (queue
 * capacity), the path tile, the traversal stack variant and with it the
 * block size, the device-driven tail and its threshold), keeps each knob's
 * fastest value by device time (best of two renders, a change must gain
 * 1 %), leaves the result set on the context and copies it to `out`
 * (optional).  The result is also kept in a process-wide cache under the
 * device (architecture, compute units) and the scene class (integrator and
 * method, instancing, log2 of the primitive count and of the paths per
 * render, both in steps of two).  No counterpart in the reference.  Outputs
 * of later renders do not depend on it; rd_get_render_stats afterwards
 * describes the last trial render. */
int rd_autotune(rd_ctx* ctx, rd_scene* scene, const rd_render_params* params, rd_tuning* out);
/* The cached tuning of this device and scene class, set on the context when
 * there is one (*found = 1), else rd_autotune (*found = 0). */
int rd_autotune_cached(rd_ctx* ctx, rd_scene* scene, const rd_render_params* params, rd_tuning* out,
                       int32_t* found);
/* The cache as text, one line per device and scene class, so a server keeps
 * its tunings across restarts; load merges into the cache. */
int rd_tuning_cache_save(const char* path);
int rd_tuning_cache_load(const char* path);
void rd_tuning_cache_clear(void);

/* ---- batched rendering: many scenes / sample counts in shared waves ----- */
/* One request: a scene of this context, its render parameters (spp, seed,
 * crop; wave_paths and flags are ignored) and its output buffers (layout as
 * rd_render; out_sum2 may be NULL).  The reference renders one scene per
 * BlockBasedRenderer (Renderers/BlockBasedRenderer.cs:56-83); a batch is
 * the per-request Start() of many renderers fused into one launch sequence. */
typedef struct rd_request {
    rd_scene* scene;
    rd_render_params params;
    float* out_rgb;
    float* out_sum2;
} rd_request;

/* Renders every request: paths of all requests share waves of wave_paths
 * paths (0 = default), each path traversing and shading its own request's
 * scene.  Each request's image equals what rd_render gives for it alone
 * (same random stream, same per-pixel sample order).  Requests are grouped by
 * integrator kind (path method, vol_path).  Host buffers; rd_render_batch_device
 * takes device buffers.  stats: sums over the batch. */
int rd_render_batch(rd_ctx* ctx, int32_t n, const rd_request* reqs, int32_t wave_paths,
                    volatile int32_t* cancel, volatile int32_t* progress);
int rd_render_batch_device(rd_ctx* ctx, int32_t n, const rd_request* reqs, int32_t wave_paths,
                           volatile int32_t* cancel, volatile int32_t* progress);

/* ---- differentiation: path-replay backpropagation ------------------------ */
/* Gradient of an image loss with respect to the scene's texture parameters
 * (no counterpart in the reference, which renders only).  Given
 * dl_dimage = dLoss/dImage (float[W*H*3], layout as rd_render's output, for
 * the image rd_render produces with the same params), adds dLoss/dtexel to
 * tex_grads[t] (one buffer per scene texture, NULL = no gradient; w*h*3
 * floats laid out as the texture, 3 for a constant).  Differentiable
 * parameters: the albedo texture of "diffuse" (kd) and "rough_plastic" (r),
 * and the roughness texture of "rough_plastic" (its first channel).  Path
 * integrator only.  Each sample's path is replayed twice with the render's
 * random numbers: the primal replay gives the path's radiance, the adjoint
 * replay propagates dLoss/dradiance back through the BSDF factors
 * (Vicini, Speierer, Jakob 2021).  Sampling, MIS weights and Russian
 * roulette are held fixed (detached).  rd_adjoint: host buffers;
 * rd_adjoint_device: device buffers (tex_grads is a host array of device
 * pointers). */
int rd_adjoint(rd_ctx* ctx, rd_scene* scene, const rd_render_params* params, const float* dl_dimage,
               float* const* tex_grads);
int rd_adjoint_device(rd_ctx* ctx, rd_scene* scene, const rd_render_params* params, const float* d_dl_dimage,
                      float* const* d_tex_grads);
/* Replaces texture `tex`'s texels (w*h*3 floats, or 3 for a constant), e.g.
 * after an optimiser step.  Environment-map textures cannot be replaced. */
int rd_scene_set_texture(rd_ctx* ctx, rd_scene* scene, int32_t tex, const float* data);
int rd_scene_set_texture_device(rd_ctx* ctx, rd_scene* scene, int32_t tex, const float* d_data);

/* ---- device memory helpers (so hosts need no GPU framework) ------------- */
void* rd_device_alloc(rd_ctx* ctx, uint64_t bytes);
int rd_device_free(rd_ctx* ctx, void* ptr);
int rd_memcpy_h2d(rd_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int rd_memcpy_d2h(rd_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int rd_synchronize(rd_ctx* ctx);
/* Average device time (ms) of the most recent launches of the named kernel
 * stage ("trace_closest", "trace_any", "shade", ...), measured with HIP events
 * on the stream the kernel ran on; count receives the number of launches. */
int rd_kernel_timing(const rd_ctx* ctx, const char* stage, double* avg_ms, int64_t* count);
int rd_reset_kernel_timing(rd_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* RADIANTIUM_HIP_H */
