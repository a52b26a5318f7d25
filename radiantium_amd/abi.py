"""ctypes mirrors of the C ABI structs declared in include/radiantium_hip.h.

The layouts here must match the header field for field; tests/test_abi.py
checks sizes and offsets against the compiled library.
"""
import ctypes as C

ABI_VERSION = 4   # RD_ABI_VERSION
RD_OK = 0
RD_ERR_INVALID = 1
RD_ERR_DEVICE = 2
RD_ERR_OOM = 3
RD_ERR_CANCELLED = 4
RD_ERR_STACK = 5
RD_ERR_UNSUPPORTED = 6
RD_ERR_BUSY = 7
RD_ERR_IO = 8

STATUS_NAMES = {
    0: "RD_OK", 1: "RD_ERR_INVALID", 2: "RD_ERR_DEVICE", 3: "RD_ERR_OOM",
    4: "RD_ERR_CANCELLED", 5: "RD_ERR_STACK", 6: "RD_ERR_UNSUPPORTED", 7: "RD_ERR_BUSY", 8: "RD_ERR_IO",
}

# material kinds (registry names of Config/RendererBuilder.cs:86-213)
MAT_DIFFUSE = 0
MAT_ROUGH_METAL = 1
MAT_ROUGH_GLASS = 2
MAT_ROUGH_PLASTIC = 3
MAT_PERFECT_GLASS = 4
MAT_PERFECT_MIRROR = 5
MAT_DISNEY = 6
MAT_SUBSURFACE = 7

MATERIAL_NAMES = {
    "diffuse": MAT_DIFFUSE, "rough_metal": MAT_ROUGH_METAL, "rough_glass": MAT_ROUGH_GLASS,
    "rough_plastic": MAT_ROUGH_PLASTIC, "perfect_glass": MAT_PERFECT_GLASS,
    "perfect_mirror": MAT_PERFECT_MIRROR, "disney": MAT_DISNEY, "subsurface": MAT_SUBSURFACE,
}

LIGHT_AREA = 0
LIGHT_INFINITE = 1
PRIM_TRIANGLE = 0
PRIM_SPHERE = 1
PRIM_INSTANCE = 2
ACCEL_BVH = 0
ACCEL_OCTREE = 1
INTEGRATOR_PATH = 0
INTEGRATOR_VOL_PATH = 1
INTEGRATOR_PARTICLE = 2
INTEGRATOR_BDPT = 3
METHOD_BXDF = 0
METHOD_NEE = 1
METHOD_MIS = 2
STRATEGY_UNIFORM = 0
STRATEGY_ALL = 1
RENDER_NO_GRAPH = 1


class rd_material_desc(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("two_sided", C.c_int32), ("dist", C.c_int32), ("thin", C.c_int32),
        ("tex", C.c_int32 * 16),
        ("eta_a", C.c_float), ("eta_b", C.c_float), ("kd", C.c_float), ("ks", C.c_float),
        ("eta", C.c_float * 3), ("k", C.c_float * 3),
    ]


class rd_texture_desc(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("wrap", C.c_int32), ("filter", C.c_int32),
        ("value", C.c_float * 3), ("_pad", C.c_int32), ("data", C.POINTER(C.c_float)),
    ]


class rd_light_desc(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("prim", C.c_int32), ("le", C.c_float * 3), ("tex", C.c_int32),
        ("to_world", C.c_float * 16),
    ]


class rd_medium_desc(C.Structure):
    _fields_ = [("sigma_a", C.c_float * 3), ("sigma_s", C.c_float * 3), ("g", C.c_float), ("scale", C.c_float)]


class rd_prim_desc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("index", C.c_int32), ("material", C.c_int32), ("light", C.c_int32),
        ("medium_inside", C.c_int32), ("medium_outside", C.c_int32),
    ]


class rd_camera_desc(C.Structure):
    _fields_ = [
        ("fov", C.c_float), ("near_z", C.c_float), ("far_z", C.c_float),
        ("width", C.c_int32), ("height", C.c_int32), ("use_matrix", C.c_int32),
        ("origin", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3),
        ("to_world", C.c_float * 16),
    ]


class rd_integrator_desc(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("max_depth", C.c_int32), ("min_depth", C.c_int32),
        ("rr_threshold", C.c_float), ("method", C.c_int32), ("strategy", C.c_int32),
    ]


class rd_instance_set_desc(C.Structure):
    _fields_ = [("first_triangle", C.c_int32), ("num_triangles", C.c_int32), ("bvh_max_prim", C.c_int32),
                ("_pad", C.c_int32)]


class rd_instance_desc(C.Structure):
    _fields_ = [("set", C.c_int32), ("_pad", C.c_int32), ("to_world", C.c_float * 16)]


class rd_scene_desc(C.Structure):
    _fields_ = [
        ("num_vertices", C.c_int32), ("positions", C.POINTER(C.c_float)),
        ("normals", C.POINTER(C.c_float)), ("uvs", C.POINTER(C.c_float)),
        ("num_triangles", C.c_int32), ("indices", C.POINTER(C.c_int32)),
        ("tri_flags", C.POINTER(C.c_uint8)),
        ("num_spheres", C.c_int32), ("spheres", C.POINTER(C.c_float)), ("sphere_w2m", C.POINTER(C.c_float)),
        ("num_prims", C.c_int32), ("prims", C.POINTER(rd_prim_desc)),
        ("num_materials", C.c_int32), ("materials", C.POINTER(rd_material_desc)),
        ("num_textures", C.c_int32), ("textures", C.POINTER(rd_texture_desc)),
        ("num_lights", C.c_int32), ("lights", C.POINTER(rd_light_desc)),
        ("num_media", C.c_int32), ("media", C.POINTER(rd_medium_desc)),
        ("global_medium", C.c_int32), ("bvh_max_prim", C.c_int32),
        ("camera", rd_camera_desc), ("integrator", rd_integrator_desc),
        ("num_instance_sets", C.c_int32), ("instance_sets", C.POINTER(rd_instance_set_desc)),
        ("num_instances", C.c_int32), ("instances", C.POINTER(rd_instance_desc)),
        ("accel", C.c_int32), ("octree_max_depth", C.c_int32), ("octree_max_count", C.c_int32),
        ("octree_out_bound", C.c_float),
    ]


class rd_ray(C.Structure):
    _fields_ = [("o", C.c_float * 3), ("tmin", C.c_float), ("d", C.c_float * 3), ("tmax", C.c_float)]


class rd_hit(C.Structure):
    _fields_ = [("t", C.c_float), ("prim", C.c_int32), ("u", C.c_float), ("v", C.c_float)]


class rd_render_params(C.Structure):
    _fields_ = [
        ("spp", C.c_int32), ("seed", C.c_uint32), ("x0", C.c_int32), ("y0", C.c_int32),
        ("x1", C.c_int32), ("y1", C.c_int32), ("wave_paths", C.c_int32), ("flags", C.c_int32),
        ("particles", C.c_int64),
    ]


class rd_request(C.Structure):
    _fields_ = [
        ("scene", C.c_void_p), ("params", rd_render_params),
        ("out_rgb", C.c_void_p), ("out_sum2", C.c_void_p),
    ]


class rd_image(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("channels", C.c_int32), ("data", C.POINTER(C.c_float))]


class rd_render_stats(C.Structure):
    _fields_ = [
        ("rays_closest", C.c_uint64), ("rays_any", C.c_uint64), ("paths", C.c_uint64),
        ("vertices", C.c_uint64), ("invalid", C.c_uint64), ("render_ms", C.c_double),
        ("waves", C.c_int32), ("bounces", C.c_int32),
        ("class_rays", C.c_uint64 * 3), ("class_nodes", C.c_uint64 * 3), ("class_prims", C.c_uint64 * 3),
    ]


class rd_tuning(C.Structure):
    _fields_ = [
        ("trace_schedule", C.c_int32 * 3), ("ray_sort", C.c_int32), ("ray_sort_bits", C.c_int32),
        ("ray_sort_min", C.c_int32), ("ray_sort_shadow", C.c_int32), ("overlap", C.c_int32),
        ("device_loop", C.c_int32), ("tail_rays", C.c_int32), ("chain_lockstep", C.c_int32),
        ("pool_log2", C.c_int32), ("tile", C.c_int32), ("stack_min", C.c_int32), ("bvh_build", C.c_int32),
    ]


class rd_scene_info(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_int32), ("num_leaves", C.c_int32), ("max_depth", C.c_int32),
        ("num_prims", C.c_int32), ("build_ms", C.c_double), ("device_bytes", C.c_uint64),
    ]


# numpy dtypes with the same layouts, for bulk ray / hit arrays
def ray_dtype():
    import numpy as np
    return np.dtype([("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])


def hit_dtype():
    import numpy as np
    return np.dtype([("t", "<f4"), ("prim", "<i4"), ("u", "<f4"), ("v", "<f4")])
