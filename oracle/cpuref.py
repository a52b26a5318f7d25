"""TEST INFRASTRUCTURE: ctypes loader for the oracle (oracle/_build/libcpuref.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product package never does.
"""
import ctypes as C
import os

import numpy as np

from radiantium_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libcpuref.so")
_lib = None


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.cr_last_error.restype = C.c_char_p
        L.cr_scene_create.restype = C.c_void_p
        L.cr_scene_create.argtypes = [C.POINTER(abi.rd_scene_desc)]
        L.cr_scene_destroy.argtypes = [C.c_void_p]
        L.cr_scene_num_nodes.argtypes = [C.c_void_p]
        L.cr_scene_max_depth.argtypes = [C.c_void_p]
        L.cr_scene_export_bvh.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.cr_sss_lut.argtypes = [C.c_void_p]
        L.cr_fresnel_moment.restype = C.c_float
        L.cr_fresnel_moment.argtypes = [C.c_float]
        L.cr_sss_profile.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.cr_sss_adapter.argtypes = [C.c_float, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.cr_sss_sample_pi.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_uint64, C.c_int32, C.c_void_p]
        L.cr_scene_export_octree.argtypes = [C.c_void_p] + [C.c_void_p] * 5
        L.cr_octree_build.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_float] + \
            [C.c_void_p] * 6
        for fn in ("cr_trace_closest", "cr_trace_any"):
            getattr(L, fn).argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
        L.cr_trace_brute.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
        L.cr_trace_closest_sub.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.cr_trace_brute_sub.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
        L.cr_camera_rays.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
        L.cr_render.argtypes = [C.c_void_p, C.POINTER(abi.rd_render_params), C.c_int32, C.c_void_p,
                                C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]
        L.cr_sample_li.argtypes = [C.c_void_p, C.c_uint32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]
        L.cr_material_eval.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p]
        L.cr_material_eval_n.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p]
        L.cr_material_sample_n.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_uint64,
                                           C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.cr_material_sample.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_uint64,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.cr_fresnel_dielectric.restype = C.c_float
        L.cr_fresnel_dielectric.argtypes = [C.c_float, C.c_float, C.c_float]
        L.cr_fresnel_conductor.argtypes = [C.c_float, C.c_void_p, C.c_void_p, C.c_void_p]
        L.cr_mf_eval.argtypes = [C.c_int32, C.c_float, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.cr_mf_sample.argtypes = [C.c_int32, C.c_float, C.c_float, C.c_void_p, C.c_uint64, C.c_void_p]
        L.cr_rng_floats.argtypes = [C.c_uint64, C.c_int32, C.c_void_p]
        L.cr_dist1d_sample.argtypes = [C.c_void_p, C.c_int32, C.c_float, C.c_void_p, C.c_void_p]
        L.cr_dist2d_sample.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_void_p,
                                       C.c_void_p, C.c_void_p]
        L.cr_dist2d_pdf.restype = C.c_float
        L.cr_dist2d_pdf.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_float]
        L.cr_env_le.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
        L.cr_env_pdf.restype = C.c_float
        L.cr_env_pdf.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.cr_env_sample.argtypes = [C.c_void_p, C.c_int32, C.c_float, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p]
        L.cr_env_bounds.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
        L.cr_medium_tr.argtypes = [C.POINTER(abi.rd_medium_desc), C.c_float, C.c_float, C.c_void_p]
        L.cr_medium_sample.argtypes = [C.POINTER(abi.rd_medium_desc), C.c_void_p, C.c_void_p, C.c_float, C.c_float,
                                       C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]
        L.cr_hg_p.restype = C.c_float
        L.cr_hg_p.argtypes = [C.c_float, C.c_void_p, C.c_void_p]
        L.cr_hg_sample.restype = C.c_float
        L.cr_hg_sample.argtypes = [C.c_float, C.c_void_p, C.c_uint64, C.c_void_p]
        L.cr_adjoint.argtypes = [C.c_void_p, C.POINTER(abi.rd_render_params), C.c_int32, C.c_void_p, C.c_void_p]
        L.cr_rho_ratio.restype = C.c_float
        L.cr_rho_ratio.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.cr_path_key.restype = C.c_uint64
        L.cr_path_key.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.cr_render_bdpt.argtypes = [C.c_void_p, C.POINTER(abi.rd_render_params), C.c_int32, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p]
        L.cr_render_bdpt_strategy.argtypes = [C.c_void_p, C.POINTER(abi.rd_render_params), C.c_int32, C.c_int32,
                                              C.c_int32, C.c_void_p]
        L.cr_render_particles.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_int32, C.c_void_p, C.c_void_p,
                                          C.c_void_p]
        L.cr_camera_sample_wi.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.cr_camera_importance.restype = C.c_float
        L.cr_camera_importance.argtypes = [C.c_void_p, C.c_void_p]
        L.cr_light_sample_emit.argtypes = [C.c_void_p, C.c_int32, C.c_uint64, C.c_void_p]
        L.cr_material_eval_mode.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p]
        L.cr_material_sample_mode.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_uint64,
                                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def fresnel_dielectric(ci, ei, et):
    return lib().cr_fresnel_dielectric(ci, ei, et)


def fresnel_conductor(ci, eta, k):
    eta = np.asarray(eta, np.float32); k = np.asarray(k, np.float32)
    out = np.zeros(3, np.float32)
    lib().cr_fresnel_conductor(ci, _ptr(eta), _ptr(k), _ptr(out))
    return out


def mf_eval(kind, roughness, aniso, wh, wo, wi):
    wh = np.asarray(wh, np.float32); wo = np.asarray(wo, np.float32); wi = np.asarray(wi, np.float32)
    out = np.zeros(3, np.float32)
    lib().cr_mf_eval(kind, roughness, aniso, _ptr(wh), _ptr(wo), _ptr(wi), _ptr(out))
    return out


def mf_sample(kind, roughness, aniso, wo, key):
    wo = np.asarray(wo, np.float32)
    wh = np.zeros(3, np.float32)
    lib().cr_mf_sample(kind, roughness, aniso, _ptr(wo), C.c_uint64(key), _ptr(wh))
    return wh


def dist1d_sample(f, u):
    """Distribution1D over f at uniform u: (continuous sample, pdf, discrete pdf, offset)."""
    f = np.ascontiguousarray(f, np.float32)
    out = np.zeros(3, np.float32)
    off = C.c_int32(0)
    lib().cr_dist1d_sample(_ptr(f), f.size, C.c_float(u), _ptr(out), C.byref(off))
    return float(out[0]), float(out[1]), float(out[2]), off.value


def dist2d_sample(f2, u0, u1):
    """Distribution2D over f2 [nv][nu]: ((d0, d1), pdf, (vj, vi))."""
    f2 = np.ascontiguousarray(f2, np.float32)
    out = np.zeros(3, np.float32)
    vj, vi = C.c_int32(0), C.c_int32(0)
    lib().cr_dist2d_sample(_ptr(f2), f2.shape[1], f2.shape[0], C.c_float(u0), C.c_float(u1), _ptr(out),
                           C.byref(vj), C.byref(vi))
    return (float(out[0]), float(out[1])), float(out[2]), (vj.value, vi.value)


def dist2d_pdf(f2, u, v):
    f2 = np.ascontiguousarray(f2, np.float32)
    return float(lib().cr_dist2d_pdf(_ptr(f2), f2.shape[1], f2.shape[0], C.c_float(u), C.c_float(v)))


def _medium(sigma_a, sigma_s, g=0.0, scale=1.0):
    m = abi.rd_medium_desc()
    for i in range(3):
        m.sigma_a[i] = sigma_a[i]
        m.sigma_s[i] = sigma_s[i]
    m.g, m.scale = g, scale
    return m


def medium_tr(sigma_a, sigma_s, tmin, tmax, scale=1.0):
    out = np.zeros(3, np.float32)
    lib().cr_medium_tr(C.byref(_medium(sigma_a, sigma_s, 0.0, scale)), C.c_float(tmin), C.c_float(tmax), _ptr(out))
    return out


def medium_sample(sigma_a, sigma_s, o, d, tmin, tmax, key, scale=1.0):
    """Homogeneous.Sample: (is_medium_vertex, p, weight, t)."""
    o = np.asarray(o, np.float32); d = np.asarray(d, np.float32)
    p = np.zeros(3, np.float32); tr = np.zeros(3, np.float32); t = C.c_float(0)
    r = lib().cr_medium_sample(C.byref(_medium(sigma_a, sigma_s, 0.0, scale)), _ptr(o), _ptr(d), C.c_float(tmin),
                               C.c_float(tmax), C.c_uint64(key), _ptr(p), _ptr(tr), C.byref(t))
    return bool(r), p, tr, t.value


def hg_p(g, wo, wi):
    wo = np.asarray(wo, np.float32); wi = np.asarray(wi, np.float32)
    return float(lib().cr_hg_p(C.c_float(g), _ptr(wo), _ptr(wi)))


def hg_sample(g, wo, key):
    wo = np.asarray(wo, np.float32); wi = np.zeros(3, np.float32)
    p = lib().cr_hg_sample(C.c_float(g), _ptr(wo), C.c_uint64(key), _ptr(wi))
    return wi, float(p)


def rng_floats(key, n):
    out = np.zeros(n, np.float32)
    lib().cr_rng_floats(C.c_uint64(key), n, _ptr(out))
    return out


def path_key(seed, pixel, sample):
    return lib().cr_path_key(seed, pixel, sample)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _octree_arrays(call):
    sz = np.zeros(3, np.int32)
    call(None, None, None, None, _ptr(sz))
    nodes = np.zeros(max(sz[0], 1) * 32, np.uint8)
    first = np.zeros(max(sz[1], 1), np.int32)
    count = np.zeros(max(sz[1], 1), np.int32)
    prims = np.zeros(max(sz[2], 1), np.int32)
    call(_ptr(nodes), _ptr(first), _ptr(count), _ptr(prims), _ptr(sz))
    return nodes[: sz[0] * 32].reshape(-1, 32), first[: sz[1]], count[: sz[1]], prims[: sz[2]]


def sss_lut():
    out = np.zeros(2048, np.float32)
    lib().cr_sss_lut(_ptr(out))
    return out


def fresnel_moment(eta):
    return float(lib().cr_fresnel_moment(eta))


def sss_profile(a, d, channel, u):
    """NormalizedDiffusionRadialProfile: rows {SampleSr radius, PdfSr, Sr rgb, D[channel]}."""
    channel = np.ascontiguousarray(channel, np.int32)
    u = np.ascontiguousarray(u, np.float32)
    out = np.zeros((u.size, 6), np.float32)
    lib().cr_sss_profile(_ptr(np.asarray(a, np.float32)), _ptr(np.asarray(d, np.float32)), u.size, _ptr(channel),
                         _ptr(u), _ptr(out))
    return out


def sss_adapter(eta, wo, wi):
    wo = np.ascontiguousarray(wo, np.float32).reshape(-1, 3)
    wi = np.ascontiguousarray(wi, np.float32).reshape(-1, 3)
    out = np.zeros((wo.shape[0], 4), np.float32)
    lib().cr_sss_adapter(eta, wo.shape[0], _ptr(wo), _ptr(wi), _ptr(out))
    return out


def octree_build(bmin, bmax, max_depth=10, max_count=10, out_bound=2.0):
    """The oracle's Octree build over caller bounds: (nodes [N,32] uint8,
    leaf_first, leaf_count, leaf_prims, now_depth)."""
    bmin = np.ascontiguousarray(bmin, np.float32)
    bmax = np.ascontiguousarray(bmax, np.float32)
    depth = np.zeros(1, np.int32)
    def call(a, b, c, d, e):
        if lib().cr_octree_build(_ptr(bmin), _ptr(bmax), bmin.shape[0], max_depth, max_count, out_bound, a, b, c, d, e,
                                 _ptr(depth)) != 0:
            raise RuntimeError("oracle octree: " + lib().cr_last_error().decode())
    out = _octree_arrays(call)
    return out + (int(depth[0]),)


class OracleScene:
    """The oracle's view of a SceneDesc (BVH built by the restated Bvh.cs)."""

    def __init__(self, scene_desc):
        L = lib()
        self.desc = scene_desc
        self._c = scene_desc.to_c()
        self.h = L.cr_scene_create(C.byref(self._c))
        if not self.h:
            raise RuntimeError("oracle scene: " + L.cr_last_error().decode())
        self.width = scene_desc.camera["width"]
        self.height = scene_desc.camera["height"]

    def __del__(self):
        if getattr(self, "h", None):
            lib().cr_scene_destroy(self.h)
            self.h = None

    @property
    def num_nodes(self):
        return lib().cr_scene_num_nodes(self.h)

    @property
    def max_depth(self):
        return lib().cr_scene_max_depth(self.h)

    def export_bvh(self):
        n = self.num_nodes
        nodes = np.zeros(n * 32, np.uint8)
        order = np.zeros(self.desc.num_prims(), np.int32)
        lib().cr_scene_export_bvh(self.h, _ptr(nodes), _ptr(order))
        return nodes.reshape(n, 32), order

    def sss_sample_pi(self, prim, P, N, wr, uv=(0.0, 0.0), key0=1, n=1):
        """SeparableBssrdf.SamplePi at a point of primitive `prim`: rows {ok,
        S rgb, pdf, pi xyz, closest-hit queries}."""
        f = lambda v: np.ascontiguousarray(v, np.float32)
        out = np.zeros((n, 9), np.float32)
        rc = lib().cr_sss_sample_pi(self.h, prim, _ptr(f(P)), _ptr(f(N)), _ptr(f(wr)), _ptr(f(uv)), key0, n, _ptr(out))
        if rc:
            raise RuntimeError("oracle sss_sample_pi: bad argument")
        return out

    def export_octree(self):
        """(nodes [N,32] uint8, leaf_first, leaf_count, leaf_prims) of an octree scene."""
        return _octree_arrays(lambda *a: lib().cr_scene_export_octree(self.h, *a))

    def trace_closest(self, rays, counters=False):
        rays = np.ascontiguousarray(rays, abi.ray_dtype())
        hits = np.zeros(rays.shape[0], abi.hit_dtype())
        cnt = np.zeros(2, np.uint64)
        lib().cr_trace_closest(self.h, _ptr(rays), rays.shape[0], _ptr(hits), _ptr(cnt))
        return (hits, cnt) if counters else hits

    def trace_any(self, rays, counters=False):
        rays = np.ascontiguousarray(rays, abi.ray_dtype())
        occ = np.zeros(rays.shape[0], np.uint8)
        cnt = np.zeros(2, np.uint64)
        lib().cr_trace_any(self.h, _ptr(rays), rays.shape[0], _ptr(occ), _ptr(cnt))
        return (occ, cnt) if counters else occ

    def trace_brute(self, rays):
        rays = np.ascontiguousarray(rays, abi.ray_dtype())
        hits = np.zeros(rays.shape[0], abi.hit_dtype())
        lib().cr_trace_brute(self.h, _ptr(rays), rays.shape[0], _ptr(hits))
        return hits

    def trace_closest_sub(self, rays):
        """(hits, sub, counters): sub = the hit triangle's index within its
        instance set, -1 for other hits."""
        rays = np.ascontiguousarray(rays, abi.ray_dtype())
        hits = np.zeros(rays.shape[0], abi.hit_dtype())
        sub = np.full(rays.shape[0], -1, np.int32)
        cnt = np.zeros(2, np.uint64)
        lib().cr_trace_closest_sub(self.h, _ptr(rays), rays.shape[0], _ptr(hits), _ptr(sub), _ptr(cnt))
        return hits, sub, cnt

    def trace_brute_sub(self, rays):
        rays = np.ascontiguousarray(rays, abi.ray_dtype())
        hits = np.zeros(rays.shape[0], abi.hit_dtype())
        sub = np.full(rays.shape[0], -1, np.int32)
        lib().cr_trace_brute_sub(self.h, _ptr(rays), rays.shape[0], _ptr(hits), _ptr(sub))
        return hits, sub

    def camera_rays(self, xy):
        xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
        rays = np.zeros(xy.shape[0], abi.ray_dtype())
        lib().cr_camera_rays(self.h, _ptr(xy), xy.shape[0], _ptr(rays))
        return rays

    def render(self, spp, seed=0, crop=None, threads=0, with_sum2=True):
        """Returns (mean rgb [W*H*3] column-major, sum2, stats dict, seconds)."""
        p = abi.rd_render_params()
        p.spp, p.seed = spp, seed
        if crop is not None:
            p.x0, p.y0, p.x1, p.y1 = crop
        W, H = self.width, self.height
        rgb = np.zeros(W * H * 3, np.float32)
        s2 = np.zeros(W * H * 3, np.float32) if with_sum2 else None
        st = np.zeros(5, np.uint64)
        secs = C.c_double(0)
        rc = lib().cr_render(self.h, C.byref(p), threads, _ptr(rgb), _ptr(s2) if s2 is not None else None,
                             _ptr(st), C.byref(secs))
        if rc != 0:
            raise RuntimeError("oracle render: " + lib().cr_last_error().decode())
        stats = dict(rays_closest=int(st[0]), rays_any=int(st[1]), paths=int(st[2]), vertices=int(st[3]),
                     invalid=int(st[4]))
        return rgb, s2, stats, secs.value

    def render_bdpt(self, spp, seed=0, crop=None, threads=0, with_sum2=True):
        """Bidirectional path tracing: (mean rgb [W*H*3], sum2, stats dict, seconds)."""
        p = abi.rd_render_params()
        p.spp, p.seed = spp, seed
        if crop is not None:
            p.x0, p.y0, p.x1, p.y1 = crop
        W, H = self.width, self.height
        rgb = np.zeros(W * H * 3, np.float32)
        s2 = np.zeros(W * H * 3, np.float32) if with_sum2 else None
        st = np.zeros(5, np.uint64)
        secs = C.c_double(0)
        rc = lib().cr_render_bdpt(self.h, C.byref(p), threads, _ptr(rgb), _ptr(s2) if s2 is not None else None,
                                  _ptr(st), C.byref(secs))
        if rc != 0:
            raise RuntimeError("oracle bdpt render: " + lib().cr_last_error().decode())
        stats = dict(rays_closest=int(st[0]), rays_any=int(st[1]), paths=int(st[2]), vertices=int(st[3]))
        return rgb, s2, stats, secs.value

    def render_bdpt_strategy(self, spp, s, t, seed=0, threads=0):
        """Strategy (s, t) of the bidirectional estimator alone: mean rgb [W*H*3]."""
        p = abi.rd_render_params()
        p.spp, p.seed = spp, seed
        rgb = np.zeros(self.width * self.height * 3, np.float32)
        rc = lib().cr_render_bdpt_strategy(self.h, C.byref(p), threads, s, t, _ptr(rgb))
        if rc != 0:
            raise RuntimeError("oracle bdpt strategy render failed")
        return rgb

    def render_particles(self, particles, seed=0, threads=0):
        """Light tracing (adjoint_particle): (rgb [W*H*3] column-major, stats dict, seconds)."""
        rgb = np.zeros(self.width * self.height * 3, np.float32)
        st = np.zeros(5, np.uint64)
        secs = C.c_double(0)
        rc = lib().cr_render_particles(self.h, C.c_uint64(particles), seed, threads, _ptr(rgb), _ptr(st),
                                       C.byref(secs))
        if rc != 0:
            raise RuntimeError("oracle particle render: " + lib().cr_last_error().decode())
        stats = dict(rays_closest=int(st[0]), rays_any=int(st[1]), paths=int(st[2]), vertices=int(st[3]))
        return rgb, stats, secs.value

    def camera_sample_wi(self, pos):
        """Camera.SampleWi: None for the pdf-0 result, else (screen x, screen y, We, pdf, distance)."""
        pos = np.asarray(pos, np.float32); out = np.zeros(6, np.float32)
        lib().cr_camera_sample_wi(self.h, _ptr(pos), _ptr(out))
        return None if out[0] == 0 else (float(out[1]), float(out[2]), float(out[3]), float(out[4]), float(out[5]))

    def camera_importance(self, d):
        d = np.asarray(d, np.float32)
        return float(lib().cr_camera_importance(self.h, _ptr(d)))

    def light_sample_emit(self, light, key):
        """Light.SampleEmit: dict(p, dir, n, le, pdf_pos, pdf_dir) or None (pdf 0)."""
        out = np.zeros(14, np.float32)
        if lib().cr_light_sample_emit(self.h, light, C.c_uint64(key), _ptr(out)) != 1:
            return None
        return dict(p=out[0:3].copy(), dir=out[3:6].copy(), n=out[6:9].copy(), le=out[9:12].copy(),
                    pdf_pos=float(out[12]), pdf_dir=float(out[13]))

    def material_eval_mode(self, mat, wo, wi, importance, uv=(0.0, 0.0)):
        wo = np.asarray(wo, np.float32); wi = np.asarray(wi, np.float32); uv = np.asarray(uv, np.float32)
        fr = np.zeros(3, np.float32)
        lib().cr_material_eval_mode(self.h, mat, 1 if importance else 0, _ptr(wo), _ptr(wi), _ptr(uv), _ptr(fr))
        return fr

    def material_sample_mode(self, mat, wo, key, importance, uv=(0.0, 0.0)):
        wo = np.asarray(wo, np.float32); uv = np.asarray(uv, np.float32)
        wi = np.zeros(3, np.float32); fr = np.zeros(3, np.float32); pdf = np.zeros(1, np.float32)
        ty = np.zeros(1, np.int32)
        lib().cr_material_sample_mode(self.h, mat, 1 if importance else 0, _ptr(wo), _ptr(uv), C.c_uint64(key),
                                      _ptr(wi), _ptr(fr), _ptr(pdf), _ptr(ty))
        return wi, fr, float(pdf[0]), int(ty[0])

    def sample_li(self, seed, x, y, sample):
        out = np.zeros(3, np.float32)
        lib().cr_sample_li(self.h, seed, x, y, sample, _ptr(out))
        return out

    def adjoint(self, spp, dl_dimage, textures, seed=0, crop=None, threads=0):
        """dLoss/dtexel for the listed texture indices: {tex: float32 [n*3]}."""
        p = abi.rd_render_params()
        p.spp, p.seed = spp, seed
        if crop is not None:
            p.x0, p.y0, p.x1, p.y1 = crop
        dl = np.ascontiguousarray(dl_dimage, np.float32)
        ntex = len(self.desc.textures)
        out = {}
        ptrs = (C.c_void_p * max(ntex, 1))()
        for t in textures:
            tx = self.desc.textures[t]
            n = tx.image.shape[0] * tx.image.shape[1] if tx.image is not None else 1
            out[t] = np.zeros(n * 3, np.float32)
            ptrs[t] = out[t].ctypes.data
        rc = lib().cr_adjoint(self.h, C.byref(p), threads, _ptr(dl), ptrs)
        if rc != 0:
            raise RuntimeError("oracle adjoint: " + lib().cr_last_error().decode())
        return out

    def rho_ratio(self, mat, wo, wi, uv=(0.0, 0.0)):
        wo = np.asarray(wo, np.float32); wi = np.asarray(wi, np.float32); uv = np.asarray(uv, np.float32)
        return float(lib().cr_rho_ratio(self.h, mat, _ptr(uv), _ptr(wo), _ptr(wi)))

    def env_le(self, d, env=0):
        d = np.asarray(d, np.float32); out = np.zeros(3, np.float32)
        if lib().cr_env_le(self.h, env, _ptr(d), _ptr(out)) != 0:
            raise RuntimeError("env_le")
        return out

    def env_pdf(self, wi, env=0):
        wi = np.asarray(wi, np.float32)
        return float(lib().cr_env_pdf(self.h, env, _ptr(wi)))

    def env_sample(self, u0, u1, env=0):
        """SampleLightDir: (world direction, pdf, corner uv) or None when pdf is 0."""
        w = np.zeros(3, np.float32); uv = np.zeros(2, np.float32); pdf = C.c_float(0)
        ok = lib().cr_env_sample(self.h, env, C.c_float(u0), C.c_float(u1), _ptr(w), C.byref(pdf), _ptr(uv))
        return (w, pdf.value, uv) if ok == 1 else None

    def env_bounds(self, env=0):
        c = np.zeros(3, np.float32); r = C.c_float(0)
        lib().cr_env_bounds(self.h, env, _ptr(c), C.byref(r))
        return c, r.value

    def material_eval(self, mat, wo, wi, uv=(0.0, 0.0)):
        wo = np.asarray(wo, np.float32); wi = np.asarray(wi, np.float32); uv = np.asarray(uv, np.float32)
        fr = np.zeros(3, np.float32); pdf = np.zeros(1, np.float32)
        lib().cr_material_eval(self.h, mat, _ptr(wo), _ptr(wi), _ptr(uv), _ptr(fr), _ptr(pdf))
        return fr, float(pdf[0])

    def material_eval_n(self, mat, wo, wi, uv=(0.0, 0.0)):
        """material_eval over arrays: wo, wi [n,3] (wo may be one direction) -> (fr [n,3], pdf [n])."""
        wi = np.ascontiguousarray(wi, np.float32).reshape(-1, 3)
        wo = np.ascontiguousarray(np.broadcast_to(np.asarray(wo, np.float32).reshape(-1, 3), wi.shape), np.float32)
        uv = np.asarray(uv, np.float32)
        fr = np.zeros((wi.shape[0], 3), np.float32); pdf = np.zeros(wi.shape[0], np.float32)
        lib().cr_material_eval_n(self.h, mat, wi.shape[0], _ptr(wo), _ptr(wi), _ptr(uv), _ptr(fr), _ptr(pdf))
        return fr, pdf

    def material_sample_n(self, mat, wo, n, key0, uv=(0.0, 0.0)):
        """n samples of one wo with keys key0 .. key0+n-1 -> (wi [n,3], fr [n,3], pdf [n], type [n])."""
        wo = np.ascontiguousarray(np.broadcast_to(np.asarray(wo, np.float32).reshape(1, 3), (n, 3)), np.float32)
        uv = np.asarray(uv, np.float32)
        wi = np.zeros((n, 3), np.float32); fr = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32); ty = np.zeros(n, np.int32)
        lib().cr_material_sample_n(self.h, mat, n, _ptr(wo), _ptr(uv), C.c_uint64(key0), _ptr(wi), _ptr(fr),
                                   _ptr(pdf), _ptr(ty))
        return wi, fr, pdf, ty

    def material_sample(self, mat, wo, key, uv=(0.0, 0.0)):
        wo = np.asarray(wo, np.float32); uv = np.asarray(uv, np.float32)
        wi = np.zeros(3, np.float32); fr = np.zeros(3, np.float32); pdf = np.zeros(1, np.float32)
        ty = np.zeros(1, np.int32)
        lib().cr_material_sample(self.h, mat, _ptr(wo), _ptr(uv), C.c_uint64(key), _ptr(wi), _ptr(fr),
                                 _ptr(pdf), _ptr(ty))
        return wi, fr, float(pdf[0]), int(ty[0])
