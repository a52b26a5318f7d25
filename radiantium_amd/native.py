"""ctypes binding of libradiantium_hip.so (the C ABI of include/radiantium_hip.h).

The library is the only compute path: there is no CPU fallback, and every
entry point raises if the library or a HIP device is missing.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libradiantium_hip.so")
_lib = None

EXPORTS = [
    "rd_abi_version", "rd_last_error", "rd_create", "rd_destroy", "rd_scene_create", "rd_scene_destroy",
    "rd_scene_get_info", "rd_scene_export_bvh", "rd_trace_closest", "rd_trace_any",
    "rd_trace_closest_device", "rd_trace_any_device", "rd_render", "rd_render_device",
    "rd_get_render_stats", "rd_bvh_build_host", "rd_device_alloc", "rd_device_free", "rd_memcpy_h2d", "rd_memcpy_d2h",
    "rd_synchronize", "rd_kernel_timing", "rd_reset_kernel_timing", "rd_render_batch", "rd_render_batch_device",
    "rd_adjoint", "rd_adjoint_device", "rd_scene_set_texture", "rd_scene_set_texture_device",
    "rd_trace_closest_sub", "rd_json_load", "rd_json_scene_desc", "rd_json_scene_free", "rd_scene_load_json",
    "rd_metal_eta_k", "rd_get_tuning", "rd_set_tuning", "rd_autotune", "rd_render_device_noalloc",
    "rd_render_mean_f64", "rd_image_load", "rd_image_free", "rd_image_last_error", "rd_image_save_exr",
    "rd_image_save_png", "rd_image_save_pfm", "rd_autotune_cached", "rd_tuning_cache_save", "rd_tuning_cache_load",
    "rd_tuning_cache_clear", "rd_bvh_build_device", "rd_octree_build_host", "rd_scene_export_octree",
]


class RadiantiumError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{abi.STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def lib():
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build it with `python -m radiantium_amd.build`")
        # torch ships its own libamdhip64 under the same soname
        # (libamdhip64.so.7) as /opt/rocm's but opens it by another name, so a
        # process that loads this library first ends up with two HIP runtimes
        # and torch then sees no GPU.  Importing torch first makes this
        # library bind to the runtime torch has already loaded.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp, i32, u64 = C.c_void_p, C.c_int32, C.c_uint64
        L.rd_abi_version.restype = C.c_int
        L.rd_last_error.restype = C.c_char_p
        L.rd_create.restype = vp
        L.rd_create.argtypes = [C.c_int]
        L.rd_destroy.argtypes = [vp]
        L.rd_render_batch.argtypes = [vp, i32, vp, i32, vp, vp]
        L.rd_adjoint.argtypes = [vp, vp, C.POINTER(abi.rd_render_params), vp, vp]
        L.rd_adjoint_device.argtypes = [vp, vp, C.POINTER(abi.rd_render_params), vp, vp]
        L.rd_scene_set_texture.argtypes = [vp, vp, i32, vp]
        L.rd_scene_set_texture_device.argtypes = [vp, vp, i32, vp]
        L.rd_render_batch_device.argtypes = [vp, i32, vp, i32, vp, vp]
        L.rd_scene_create.restype = vp
        L.rd_scene_create.argtypes = [vp, C.POINTER(abi.rd_scene_desc)]
        L.rd_scene_destroy.argtypes = [vp]
        L.rd_scene_get_info.argtypes = [vp, C.POINTER(abi.rd_scene_info)]
        L.rd_scene_export_bvh.argtypes = [vp, vp, i32, vp, i32]
        L.rd_trace_closest.argtypes = [vp, vp, vp, i32, vp]
        L.rd_trace_closest_sub.argtypes = [vp, vp, vp, i32, vp, vp]
        L.rd_trace_any.argtypes = [vp, vp, vp, i32, vp]
        L.rd_trace_closest_device.argtypes = [vp, vp, vp, i32, vp, vp, vp]
        L.rd_trace_any_device.argtypes = [vp, vp, vp, i32, vp, vp, vp]
        L.rd_render.argtypes = [vp, vp, C.POINTER(abi.rd_render_params), vp, vp, vp, vp]
        L.rd_render_device.argtypes = [vp, vp, C.POINTER(abi.rd_render_params), vp, vp, vp, vp]
        L.rd_render_device_noalloc.argtypes = [vp, vp, C.POINTER(abi.rd_render_params), vp, vp, vp, vp]
        L.rd_render_mean_f64.argtypes = [vp, vp, C.POINTER(abi.rd_render_params), vp]
        L.rd_get_render_stats.argtypes = [vp, C.POINTER(abi.rd_render_stats)]
        L.rd_device_alloc.restype = vp
        L.rd_device_alloc.argtypes = [vp, u64]
        L.rd_device_free.argtypes = [vp, vp]
        L.rd_memcpy_h2d.argtypes = [vp, vp, vp, u64]
        L.rd_memcpy_d2h.argtypes = [vp, vp, vp, u64]
        L.rd_synchronize.argtypes = [vp]
        L.rd_kernel_timing.argtypes = [vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.rd_reset_kernel_timing.argtypes = [vp]
        L.rd_bvh_build_host.argtypes = [vp, vp, i32, i32, vp, i32, vp, vp, vp]
        L.rd_octree_build_host.argtypes = [vp, vp, i32, i32, i32, C.c_float, vp, i32, vp, vp, i32, vp, i32, vp]
        L.rd_scene_export_octree.argtypes = [vp, vp, i32, vp, vp, i32, vp, i32, vp]
        L.rd_json_load.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(vp), C.POINTER(abi.rd_render_params)]
        L.rd_json_scene_desc.restype = C.POINTER(abi.rd_scene_desc)
        L.rd_json_scene_desc.argtypes = [vp]
        L.rd_json_scene_free.argtypes = [vp]
        L.rd_scene_load_json.restype = vp
        L.rd_scene_load_json.argtypes = [vp, C.c_char_p, C.c_char_p, C.POINTER(abi.rd_render_params)]
        L.rd_metal_eta_k.argtypes = [C.c_char_p, vp, vp]
        L.rd_get_tuning.argtypes = [vp, C.POINTER(abi.rd_tuning)]
        L.rd_set_tuning.argtypes = [vp, C.POINTER(abi.rd_tuning)]
        L.rd_autotune.argtypes = [vp, vp, C.POINTER(abi.rd_render_params), C.POINTER(abi.rd_tuning)]
        L.rd_bvh_build_device.argtypes = [vp, vp, vp, i32, i32, vp, i32, vp, C.POINTER(i32), C.POINTER(i32)]
        L.rd_autotune_cached.argtypes = [vp, vp, C.POINTER(abi.rd_render_params), C.POINTER(abi.rd_tuning),
                                         C.POINTER(C.c_int32)]
        L.rd_tuning_cache_save.argtypes = [C.c_char_p]
        L.rd_tuning_cache_load.argtypes = [C.c_char_p]
        L.rd_tuning_cache_clear.argtypes = []
        L.rd_image_load.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(abi.rd_image)]
        L.rd_image_free.argtypes = [C.POINTER(abi.rd_image)]
        L.rd_image_last_error.restype = C.c_char_p
        L.rd_image_save_exr.argtypes = [C.c_char_p, vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.rd_image_save_png.argtypes = [C.c_char_p, vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.rd_image_save_pfm.argtypes = [C.c_char_p, vp, C.c_int, C.c_int, C.c_int]
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise RadiantiumError(rc, lib().rd_last_error().decode())
    return rc


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _image_check(rc):
    if rc != 0:
        raise RadiantiumError(rc, lib().rd_image_last_error().decode())


def image_load(path, need_channel=-1, flip_y=True, cast_to_linear=None):
    """TextureUtility.LoadImageFromPath (rd_image_load): a ColorBuffer as
    float32 [width, height, channels], [x, y] with y = 0 the bottom row when
    flip_y.  cast_to_linear None: the reference's default (false for .exr /
    .hdr locations, true otherwise)."""
    if cast_to_linear is None:
        ext = os.path.splitext(str(path))[1]
        cast_to_linear = not ("exr" in ext or "hdr" in ext)
    im = abi.rd_image()
    _image_check(lib().rd_image_load(os.fsencode(path), int(need_channel), int(bool(flip_y)), int(bool(cast_to_linear)),
                                     C.byref(im)))
    try:
        n = im.width * im.height * im.channels
        out = np.ctypeslib.as_array(im.data, shape=(n,)).copy()
    finally:
        lib().rd_image_free(C.byref(im))
    return out.reshape(im.width, im.height, im.channels)


def _colorbuffer(img):
    a = np.ascontiguousarray(img, np.float32)
    if a.ndim == 2:
        a = a[:, :, None]
    return a


def image_save_exr(path, img, compression=0):
    """ColorBuffer.SaveOpenExr (rd_image_save_exr); img float32 [width, height, channels]."""
    a = _colorbuffer(img)
    _image_check(lib().rd_image_save_exr(os.fsencode(path), _ptr(a), a.shape[0], a.shape[1], a.shape[2], compression))


def image_save_png(path, img, to_srgb=False):
    """ColorBuffer.SavePng (rd_image_save_png); to_srgb: Color3F.ToSrgb first."""
    a = _colorbuffer(img)
    _image_check(lib().rd_image_save_png(os.fsencode(path), _ptr(a), a.shape[0], a.shape[1], a.shape[2],
                                         int(bool(to_srgb))))


def image_save_pfm(path, img):
    """rd_image_save_pfm: float RGB, ColorBuffer row 0 first (PFM's bottom row)."""
    a = _colorbuffer(img)
    _image_check(lib().rd_image_save_pfm(os.fsencode(path), _ptr(a), a.shape[0], a.shape[1], a.shape[2]))


def bvh_build_device(ctx, bmin, bmax, max_prim):
    """rd_bvh_build_device from host arrays (uploaded, built, read back):
    (nodes [N,32] uint8, order, max_depth), as bvh_build_host."""
    bmin = np.ascontiguousarray(bmin, np.float32)
    bmax = np.ascontiguousarray(bmax, np.float32)
    n = bmin.shape[0]
    cap = max(2 * n - 1, 1)
    d_lo, d_hi = ctx.alloc(max(bmin.nbytes, 4)), ctx.alloc(max(bmax.nbytes, 4))
    d_nodes, d_order = ctx.alloc(cap * 32), ctx.alloc(max(n, 1) * 4)
    try:
        if n:
            ctx.h2d(d_lo, bmin)
            ctx.h2d(d_hi, bmax)
        nn = C.c_int32(0)
        md = C.c_int32(0)
        check(lib().rd_bvh_build_device(ctx.h, d_lo, d_hi, n, max_prim, d_nodes, cap, d_order, C.byref(nn),
                                        C.byref(md)))
        nodes = np.zeros(max(nn.value, 1) * 32, np.uint8)
        order = np.zeros(max(n, 1), np.int32)
        if nn.value:
            ctx.d2h(nodes[: nn.value * 32], d_nodes)
        if n:
            ctx.d2h(order[:n], d_order)
        return nodes[: nn.value * 32].reshape(-1, 32), order[:n], md.value
    finally:
        for p in (d_lo, d_hi, d_nodes, d_order):
            ctx.free(p)


def tuning_cache_save(path):
    """rd_tuning_cache_save: the process's autotuner results as text."""
    check(lib().rd_tuning_cache_save(os.fsencode(path)))


def tuning_cache_load(path):
    """rd_tuning_cache_load: merge a saved cache into the process's."""
    check(lib().rd_tuning_cache_load(os.fsencode(path)))


def tuning_cache_clear():
    lib().rd_tuning_cache_clear()


def bvh_build_host(bmin, bmax, max_prim):
    """Host SAH build (no device needed): returns (nodes [N,32] uint8, order, max_depth)."""
    bmin = np.ascontiguousarray(bmin, np.float32)
    bmax = np.ascontiguousarray(bmax, np.float32)
    n = bmin.shape[0]
    cap = max(2 * n - 1, 1)
    nodes = np.zeros(cap * 32, np.uint8)
    order = np.zeros(max(n, 1), np.int32)
    nn = C.c_int32(0)
    md = C.c_int32(0)
    check(lib().rd_bvh_build_host(_ptr(bmin), _ptr(bmax), n, max_prim, _ptr(nodes), cap, _ptr(order),
                                  C.byref(nn), C.byref(md)))
    return nodes[: nn.value * 32].reshape(-1, 32), order[:n], md.value


def _octree_arrays(call):
    """Sizes, then fills, an octree export: (nodes [N,32] uint8, leaf_first,
    leaf_count, leaf_prims)."""
    sz = np.zeros(3, np.int32)
    check(call(None, 0, None, None, 0, None, 0, _ptr(sz)))
    nodes = np.zeros(max(sz[0], 1) * 32, np.uint8)
    first = np.zeros(max(sz[1], 1), np.int32)
    count = np.zeros(max(sz[1], 1), np.int32)
    prims = np.zeros(max(sz[2], 1), np.int32)
    check(call(_ptr(nodes), int(sz[0]), _ptr(first), _ptr(count), int(sz[1]), _ptr(prims), int(sz[2]), _ptr(sz)))
    return nodes[: sz[0] * 32].reshape(-1, 32), first[: sz[1]], count[: sz[1]], prims[: sz[2]]


def octree_build_host(bmin, bmax, max_depth=10, max_count=10, out_bound=2.0):
    """Host loose-octree build (Accelerators/Octree.cs; no device needed):
    (nodes [N,32] uint8 LinearNode records, leaf_first, leaf_count, leaf_prims)."""
    bmin = np.ascontiguousarray(bmin, np.float32)
    bmax = np.ascontiguousarray(bmax, np.float32)
    n = bmin.shape[0]
    return _octree_arrays(lambda *a: lib().rd_octree_build_host(_ptr(bmin), _ptr(bmax), n, max_depth, max_count,
                                                                out_bound, *a))


def metal_eta_k(name):
    """rd_metal_eta_k: (eta, k) RGB of a named conductor (Spectrum.NameToEtaAndK)."""
    eta = np.zeros(3, np.float32)
    k = np.zeros(3, np.float32)
    check(lib().rd_metal_eta_k(name.encode(), _ptr(eta), _ptr(k)))
    return tuple(float(v) for v in eta), tuple(float(v) for v in k)


class JsonScene:
    """A scene config of the reference (JSON + OBJ / PNG assets) read by
    rd_json_load into an rd_scene_desc; usable wherever a SceneDesc is (its
    to_c() is the loaded description).  `spp` is the config's renderer spp."""

    def __init__(self, path, search_dir=None):
        self.h = None
        h = C.c_void_p()
        params = abi.rd_render_params()
        check(lib().rd_json_load(os.fsencode(path), os.fsencode(search_dir) if search_dir else None, C.byref(h),
                                 C.byref(params)))
        self.h = h.value
        self.path = path
        self.spp = params.spp
        self._desc = lib().rd_json_scene_desc(self.h).contents
        cam = self._desc.camera
        self.camera = {"width": cam.width, "height": cam.height}
        self.textures = []

    def to_c(self):
        return self._desc

    def close(self):
        if getattr(self, "h", None):
            lib().rd_json_scene_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Context:
    """One HIP device (rd_create)."""

    def __init__(self, device=0):
        self.h = lib().rd_create(device)
        if not self.h:
            raise RadiantiumError(abi.RD_ERR_DEVICE, lib().rd_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            lib().rd_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def render_batch(self, requests, wave_paths=0, with_sum2=True):
        """Batched render of [(scene, spp, seed[, crop])]: returns [(rgb, sum2)]
        (layouts as Scene.render)."""
        reqs = (abi.rd_request * len(requests))()
        outs = []
        for i, r in enumerate(requests):
            scene, spp, seed = r[0], r[1], r[2]
            crop = r[3] if len(r) > 3 else None
            n = scene.width * scene.height * 3
            rgb = np.zeros(n, np.float32)
            s2 = np.zeros(n, np.float32) if with_sum2 else None
            reqs[i].scene = scene.h
            reqs[i].params.spp, reqs[i].params.seed = spp, seed
            if crop is not None:
                reqs[i].params.x0, reqs[i].params.y0, reqs[i].params.x1, reqs[i].params.y1 = crop
            reqs[i].out_rgb = rgb.ctypes.data
            reqs[i].out_sum2 = s2.ctypes.data if s2 is not None else None
            outs.append((rgb, s2))
        check(lib().rd_render_batch(self.h, len(requests), reqs, wave_paths, None, None))
        return outs

    def render_batch_device(self, requests, wave_paths=0):
        """requests: [(scene, spp, seed, d_rgb, d_sum2 or None[, crop])] with device outputs."""
        reqs = (abi.rd_request * len(requests))()
        for i, r in enumerate(requests):
            scene, spp, seed, d_rgb, d_s2 = r[:5]
            reqs[i].scene = scene.h
            reqs[i].params.spp, reqs[i].params.seed = spp, seed
            if len(r) > 5 and r[5] is not None:
                reqs[i].params.x0, reqs[i].params.y0, reqs[i].params.x1, reqs[i].params.y1 = r[5]
            reqs[i].out_rgb = d_rgb
            reqs[i].out_sum2 = d_s2
        check(lib().rd_render_batch_device(self.h, len(requests), reqs, wave_paths, None, None))

    def alloc(self, nbytes):
        p = lib().rd_device_alloc(self.h, nbytes)
        if not p:
            raise RadiantiumError(abi.RD_ERR_OOM, lib().rd_last_error().decode())
        return p

    def free(self, p):
        check(lib().rd_device_free(self.h, p))

    def h2d(self, dst, arr):
        arr = np.ascontiguousarray(arr)
        check(lib().rd_memcpy_h2d(self.h, dst, _ptr(arr), arr.nbytes))

    def d2h(self, arr, src):
        check(lib().rd_memcpy_d2h(self.h, _ptr(arr), src, arr.nbytes))
        return arr

    def synchronize(self):
        check(lib().rd_synchronize(self.h))

    def kernel_timing(self, stage):
        avg = C.c_double(0)
        cnt = C.c_int64(0)
        check(lib().rd_kernel_timing(self.h, stage.encode(), C.byref(avg), C.byref(cnt)))
        return avg.value, cnt.value

    def reset_kernel_timing(self):
        check(lib().rd_reset_kernel_timing(self.h))

    def tuning(self):
        """The context's wavefront tuning (rd_get_tuning) as a dict."""
        t = abi.rd_tuning()
        check(lib().rd_get_tuning(self.h, C.byref(t)))
        return {"trace_schedule": list(t.trace_schedule), "ray_sort": t.ray_sort, "ray_sort_bits": t.ray_sort_bits,
                "ray_sort_min": t.ray_sort_min, "ray_sort_shadow": t.ray_sort_shadow, "overlap": t.overlap,
                "device_loop": t.device_loop, "tail_rays": t.tail_rays, "chain_lockstep": t.chain_lockstep,
                "pool_log2": t.pool_log2, "tile": t.tile, "stack_min": t.stack_min, "bvh_build": t.bvh_build}

    def set_tuning(self, trace_schedule=None, ray_sort=None, ray_sort_bits=None, ray_sort_min=None,
                   ray_sort_shadow=None, overlap=None, device_loop=None, tail_rays=None, chain_lockstep=None,
                   pool_log2=None, tile=None, stack_min=None, bvh_build=None):
        """rd_set_tuning; arguments left None keep their value."""
        t = abi.rd_tuning()
        sched = trace_schedule if trace_schedule is not None else (-1, -1, -1)
        for i in range(3):
            t.trace_schedule[i] = -1 if sched[i] is None else int(sched[i])
        for f, v in (("ray_sort", ray_sort), ("ray_sort_bits", ray_sort_bits), ("ray_sort_min", ray_sort_min),
                     ("ray_sort_shadow", ray_sort_shadow), ("overlap", overlap), ("device_loop", device_loop),
                     ("tail_rays", tail_rays), ("chain_lockstep", chain_lockstep), ("pool_log2", pool_log2),
                     ("tile", tile), ("stack_min", stack_min), ("bvh_build", bvh_build)):
            setattr(t, f, -1 if v is None else int(v))
        check(lib().rd_set_tuning(self.h, C.byref(t)))

    def autotune(self, scene, spp=4, seed=0, wave_paths=0, cached=False):
        """rd_autotune (cached: rd_autotune_cached) on `scene` rendered at
        `spp`; returns the tuning it leaves on the context (as tuning()),
        and with cached=True also whether it came from the cache."""
        p = abi.rd_render_params()
        p.spp, p.seed, p.wave_paths = spp, seed, wave_paths
        t = abi.rd_tuning()
        if cached:
            found = C.c_int32(0)
            check(lib().rd_autotune_cached(self.h, scene.h, C.byref(p), C.byref(t), C.byref(found)))
            return self.tuning(), bool(found.value)
        check(lib().rd_autotune(self.h, scene.h, C.byref(p), C.byref(t)))
        return self.tuning()

    def render_stats(self):
        st = abi.rd_render_stats()
        check(lib().rd_get_render_stats(self.h, C.byref(st)))
        out = {}
        for f, _ in abi.rd_render_stats._fields_:
            v = getattr(st, f)
            out[f] = list(v) if f.startswith("class_") else v
        return out


class Scene:
    """A scene resident in HBM (rd_scene_create: host SAH build + upload)."""

    def __init__(self, ctx, scene_desc):
        self.ctx = ctx
        self.desc = scene_desc
        self._c = scene_desc.to_c()
        self.h = lib().rd_scene_create(ctx.h, C.byref(self._c))
        if not self.h:
            raise RadiantiumError(abi.RD_ERR_INVALID, lib().rd_last_error().decode())
        self.width = scene_desc.camera["width"]
        self.height = scene_desc.camera["height"]

    @classmethod
    def from_json(cls, ctx, path, search_dir=None):
        """rd_scene_load_json: (scene, spp of the config's renderer block)."""
        params = abi.rd_render_params()
        h = lib().rd_scene_load_json(ctx.h, os.fsencode(path), os.fsencode(search_dir) if search_dir else None,
                                     C.byref(params))
        if not h:
            raise RadiantiumError(abi.RD_ERR_INVALID, lib().rd_last_error().decode())
        s = cls.__new__(cls)
        s.ctx, s.desc, s._c, s.h = ctx, None, None, h
        js = JsonScene(path, search_dir)      # the image size (the camera's screen)
        s.width, s.height = js.camera["width"], js.camera["height"]
        js.close()
        return s, params.spp

    def close(self):
        if getattr(self, "h", None):
            lib().rd_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def info(self):
        i = abi.rd_scene_info()
        check(lib().rd_scene_get_info(self.h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in abi.rd_scene_info._fields_}

    def export_octree(self):
        """The loose octree of an octree scene, as octree_build_host returns it."""
        return _octree_arrays(lambda *a: lib().rd_scene_export_octree(self.h, *a))

    def export_bvh(self):
        inf = self.info()
        nodes = np.zeros(inf["num_nodes"] * 32, np.uint8)
        order = np.zeros(inf["num_prims"], np.int32)
        check(lib().rd_scene_export_bvh(self.h, _ptr(nodes), inf["num_nodes"], _ptr(order), inf["num_prims"]))
        return nodes.reshape(-1, 32), order

    def trace_closest(self, rays):
        rays = np.ascontiguousarray(rays, abi.ray_dtype())
        hits = np.zeros(rays.shape[0], abi.hit_dtype())
        check(lib().rd_trace_closest(self.ctx.h, self.h, _ptr(rays), rays.shape[0], _ptr(hits)))
        return hits

    def trace_closest_sub(self, rays):
        """rd_trace_closest_sub: (hits, sub) with sub = the hit triangle's index
        within its instance set, -1 for hits not on an instance."""
        rays = np.ascontiguousarray(rays, abi.ray_dtype())
        hits = np.zeros(rays.shape[0], abi.hit_dtype())
        sub = np.full(rays.shape[0], -1, np.int32)
        check(lib().rd_trace_closest_sub(self.ctx.h, self.h, _ptr(rays), rays.shape[0], _ptr(hits), _ptr(sub)))
        return hits, sub

    def trace_any(self, rays):
        rays = np.ascontiguousarray(rays, abi.ray_dtype())
        occ = np.zeros(rays.shape[0], np.uint8)
        check(lib().rd_trace_any(self.ctx.h, self.h, _ptr(rays), rays.shape[0], _ptr(occ)))
        return occ

    def trace_closest_device(self, d_rays, n, d_hits, d_counters=None):
        check(lib().rd_trace_closest_device(self.ctx.h, self.h, d_rays, n, d_hits, d_counters, None))

    def trace_any_device(self, d_rays, n, d_occ, d_counters=None):
        check(lib().rd_trace_any_device(self.ctx.h, self.h, d_rays, n, d_occ, d_counters, None))

    def render(self, spp, seed=0, crop=None, wave_paths=0, with_sum2=True, cancel=None, progress=None):
        """Returns (mean rgb [W*H*3] column-major, sum2 or None)."""
        p = abi.rd_render_params()
        p.spp, p.seed, p.wave_paths = spp, seed, wave_paths
        if crop is not None:
            p.x0, p.y0, p.x1, p.y1 = crop
        n = self.width * self.height * 3
        rgb = np.zeros(n, np.float32)
        s2 = np.zeros(n, np.float32) if with_sum2 else None
        check(lib().rd_render(self.ctx.h, self.h, C.byref(p), _ptr(rgb), _ptr(s2) if s2 is not None else None,
                              C.cast(cancel, C.c_void_p) if cancel is not None else None,
                              C.cast(progress, C.c_void_p) if progress is not None else None))
        return rgb, s2

    def render_particles(self, particles, seed=0, wave_paths=0, cancel=None, progress=None):
        """Light tracing of an adjoint_particle scene (rd_render with
        params.particles): the image [W*H*3] column-major."""
        p = abi.rd_render_params()
        p.seed, p.wave_paths, p.particles = seed, wave_paths, particles
        rgb = np.zeros(self.width * self.height * 3, np.float32)
        check(lib().rd_render(self.ctx.h, self.h, C.byref(p), _ptr(rgb), None,
                              C.cast(cancel, C.c_void_p) if cancel is not None else None,
                              C.cast(progress, C.c_void_p) if progress is not None else None))
        return rgb

    def render_mean_f64(self, spp, seed=0, crop=None, wave_paths=0):
        """rd_render_mean_f64: the per-pixel mean [W*H*3] column-major in float64."""
        p = abi.rd_render_params()
        p.spp, p.seed, p.wave_paths = spp, seed, wave_paths
        if crop is not None:
            p.x0, p.y0, p.x1, p.y1 = crop
        out = np.zeros(self.width * self.height * 3, np.float64)
        check(lib().rd_render_mean_f64(self.ctx.h, self.h, C.byref(p), _ptr(out)))
        return out

    def adjoint(self, spp, dl_dimage, textures, seed=0, crop=None, wave_paths=0):
        """Path-replay adjoint (rd_adjoint): dLoss/dtexel for the listed texture
        indices, given dLoss/dImage (host float32 [W*H*3]) of the render with
        the same spp / seed / crop.  Returns {tex: float32 [n*3]}."""
        p = abi.rd_render_params()
        p.spp, p.seed, p.wave_paths = spp, seed, wave_paths
        if crop is not None:
            p.x0, p.y0, p.x1, p.y1 = crop
        dl = np.ascontiguousarray(dl_dimage, np.float32)
        ntex = len(self.desc.textures)
        ptrs = (C.c_void_p * max(ntex, 1))()
        out = {}
        for t in textures:
            tx = self.desc.textures[t]
            n = tx.image.shape[0] * tx.image.shape[1] if tx.image is not None else 1
            out[t] = np.zeros(n * 3, np.float32)
            ptrs[t] = out[t].ctypes.data
        check(lib().rd_adjoint(self.ctx.h, self.h, C.byref(p), _ptr(dl), ptrs))
        return out

    def adjoint_device(self, spp, d_dl, d_grads, seed=0, crop=None, wave_paths=0):
        """rd_adjoint_device: d_grads is a list (one per scene texture) of device pointers or None."""
        p = abi.rd_render_params()
        p.spp, p.seed, p.wave_paths = spp, seed, wave_paths
        if crop is not None:
            p.x0, p.y0, p.x1, p.y1 = crop
        ptrs = (C.c_void_p * max(len(d_grads), 1))(*[g or None for g in d_grads])
        check(lib().rd_adjoint_device(self.ctx.h, self.h, C.byref(p), d_dl, ptrs))

    def set_texture(self, tex, data):
        """rd_scene_set_texture: replace texture `tex`'s texels ([w][h][3] or 3 floats)."""
        a = np.ascontiguousarray(data, np.float32)
        check(lib().rd_scene_set_texture(self.ctx.h, self.h, tex, _ptr(a)))
        self.desc.textures[tex] = type(self.desc.textures[tex])(image=a.reshape(self.desc.textures[tex].image.shape),
                                                                wrap=("Clamp", "Repeat")[self.desc.textures[tex].wrap],
                                                                filter=("Nearest", "Linear")[self.desc.textures[tex].filter]) \
            if self.desc.textures[tex].image is not None else self.desc.textures[tex]

    def set_texture_device(self, tex, d_data):
        check(lib().rd_scene_set_texture_device(self.ctx.h, self.h, tex, d_data))

    def render_device(self, spp, d_rgb, d_sum2=None, seed=0, crop=None, wave_paths=0):
        p = abi.rd_render_params()
        p.spp, p.seed, p.wave_paths = spp, seed, wave_paths
        if crop is not None:
            p.x0, p.y0, p.x1, p.y1 = crop
        check(lib().rd_render_device(self.ctx.h, self.h, C.byref(p), d_rgb, d_sum2, None, None))

    def render_particles_device(self, particles, d_rgb, seed=0, wave_paths=0):
        """Light tracing into a device image (rd_render_device with params.particles)."""
        p = abi.rd_render_params()
        p.seed, p.wave_paths, p.particles = seed, wave_paths, particles
        check(lib().rd_render_device(self.ctx.h, self.h, C.byref(p), d_rgb, None, None, None))

    def render_device_noalloc(self, spp, d_rgb, d_sum2=None, seed=0, crop=None, wave_paths=0):
        """rd_render_device_noalloc: as render_device, with no device allocation
        (work buffers from the library's fixed per-device work area)."""
        p = abi.rd_render_params()
        p.spp, p.seed, p.wave_paths = spp, seed, wave_paths
        if crop is not None:
            p.x0, p.y0, p.x1, p.y1 = crop
        check(lib().rd_render_device_noalloc(self.ctx.h, self.h, C.byref(p), d_rgb, d_sum2, None, None))
