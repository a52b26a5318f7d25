"""Scene description on the host side of the C ABI.

`SceneDesc` plays the part of the reference's `RendererBuilder.Build`
(Radiantium.Offline/Config/RendererBuilder.cs:870-1010): it collects
entities (shapes, models), materials, area / infinite lights and media, and
produces the flat arrays of `rd_scene_desc` in the builder's primitive order
(entity order, then shapes, then model triangles; :940-999), with one area
light per emissive primitive (:963-993).

Parameter defaults follow the reference registry (:27-233):
  diffuse.kd 0.5; rough_metal roughness 0.3, GGX, Cu eta/k;
  rough_glass r=t=1, roughness 0.3, GGX, etaA 1.000277, etaB 1.5046;
  rough_plastic r 0.5, Beckmann, roughness 0.3, kd=ks=0.5;
  camera fov 60, near 0.001, far 1000; bvh max_prim 1;
  path max_depth -1, min_depth 3, rr_threshold 0.99, method Mis, strategy Uniform.
"""
import ctypes as C

import numpy as np

from . import abi

# Spectrum.NameToEtaAndK["Cu"] (Radiantium.Offline/Spectrum.cs:64)
CU_ETA = (0.1999899, 0.9220823, 1.0998751)
CU_K = (3.9046354, 2.4476316, 2.1376500)

DEFAULT_ETA_A = 1.000277
DEFAULT_ETA_B = 1.5046


class Texture:
    """Texture2D: constant colour or 3-channel float image, column-major
    (x*H + y)*3 as ColorBuffer (Radiantium.Core/ColorBuffer.cs:25-42)."""

    def __init__(self, value=None, image=None, wrap="Clamp", filter="Nearest"):
        self.value = tuple(float(v) for v in (value if value is not None else (0.0, 0.0, 0.0)))
        self.image = None
        if image is not None:
            img = np.ascontiguousarray(image, dtype=np.float32)
            if img.ndim != 3 or img.shape[2] != 3:
                raise ValueError("image textures are [width][height][3] float32 (column-major)")
            self.image = img
        self.wrap = {"Clamp": 0, "Repeat": 1}[wrap]
        self.filter = {"Nearest": 0, "Linear": 1}[filter]


def _as_tex(v, default):
    if v is None:
        v = default
    if isinstance(v, Texture):
        return v
    if np.isscalar(v):
        return Texture((v, v, v))
    return Texture(tuple(v))


class SceneDesc:
    def __init__(self):
        self._positions = []
        self._normals = []
        self._uvs = []
        self._indices = []
        self._tflags = []
        self._nverts = 0
        self._has_normals = False
        self._has_uvs = False
        self.spheres = []
        self.sphere_w2m = []     # 3x3 per sphere
        self.prims = []          # (kind, index, material, light, med_in, med_out)
        self.materials = []      # dicts
        self.textures = []       # Texture
        self.lights = []         # dicts
        self.media = []          # dicts
        self.global_medium = -1
        self.bvh_max_prim = 1
        # accelerator: "bvh" or "octree" (Accelerators/Octree.cs) with its
        # max_depth / max_count / out_bound (RendererBuilder.cs defaults)
        self.accel = dict(type="bvh", max_depth=10, max_count=10, out_bound=2.0)
        self.camera = dict(fov=60.0, near=0.001, far=1000.0, width=1280, height=720,
                           origin=(0.0, 0.0, 3.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0),
                           to_world=None)
        self.integrator = dict(type="path", max_depth=-1, min_depth=3, rr_threshold=0.99,
                               method="Mis", strategy="Uniform")
        self._ntri = 0
        self.instance_sets = []   # (first_triangle, num_triangles, bvh_max_prim)
        self.instances = []       # (set, to_world 4x4)
        self._keep = []

    # ---- textures / materials -------------------------------------------
    def add_texture(self, tex):
        self.textures.append(tex)
        return len(self.textures) - 1

    def _tex(self, v, default):
        return self.add_texture(_as_tex(v, default))

    def add_material(self, kind, **p):
        """Registers a material by its reference registry name and returns its id."""
        t = [0] * 16
        m = dict(type=abi.MATERIAL_NAMES[kind], two_sided=int(bool(p.get("is_two_side", False))),
                 dist=0, thin=0, tex=t, eta_a=DEFAULT_ETA_A, eta_b=DEFAULT_ETA_B, kd=0.5, ks=0.5,
                 eta=CU_ETA, k=CU_K)
        dist = lambda d: {"Beckmann": 0, "GGX": 1}[d]
        if kind == "diffuse":
            t[0] = self._tex(p.get("kd"), 0.5)
        elif kind == "rough_metal":
            m["dist"] = dist(p.get("dist", "GGX"))
            t[0] = self._tex(p.get("roughness"), 0.3)
            t[1] = self._tex(p.get("anisotropic"), 0.0)
            m["eta"] = tuple(p.get("eta", CU_ETA))
            m["k"] = tuple(p.get("k", CU_K))
            if "metal_type" in p:   # a named conductor of Spectrum.NameToEtaAndK
                from .native import metal_eta_k
                m["eta"], m["k"] = metal_eta_k(p["metal_type"])
        elif kind == "rough_glass":
            m["dist"] = dist(p.get("dist", "GGX"))
            t[0] = self._tex(p.get("r"), 1.0)
            t[1] = self._tex(p.get("t"), 1.0)
            t[2] = self._tex(p.get("roughness"), 0.3)
            t[3] = self._tex(p.get("anisotropic"), 0.0)
            m["eta_a"] = float(p.get("etaA", DEFAULT_ETA_A))
            m["eta_b"] = float(p.get("etaB", DEFAULT_ETA_B))
        elif kind == "rough_plastic":
            m["dist"] = dist(p.get("dist", "Beckmann"))
            t[0] = self._tex(p.get("r"), 0.5)
            t[1] = self._tex(p.get("roughness"), 0.3)
            t[2] = self._tex(p.get("anisotropic"), 0.0)
            m["kd"] = float(p.get("kd", 0.5))
            m["ks"] = float(p.get("ks", 0.5))
            m["eta_a"] = float(p.get("etaA", DEFAULT_ETA_A))
            m["eta_b"] = float(p.get("etaB", DEFAULT_ETA_B))
        elif kind == "perfect_glass":
            t[0] = self._tex(p.get("r"), 1.0)
            t[1] = self._tex(p.get("t"), 1.0)
            m["eta_a"] = float(p.get("etaA", DEFAULT_ETA_A))
            m["eta_b"] = float(p.get("etaB", DEFAULT_ETA_B))
        elif kind == "perfect_mirror":
            t[0] = self._tex(p.get("r"), 1.0)
        elif kind == "subsurface":   # RendererBuilder.cs:206-215
            t[0] = self._tex(p.get("r"), 1.0)
            t[1] = self._tex(p.get("t"), 1.0)
            t[2] = self._tex(p.get("a"), 1.0)
            t[3] = self._tex(p.get("scatting_distance"), 0.0)
            m["eta_a"] = float(p.get("etaA", DEFAULT_ETA_A))
            m["eta_b"] = float(p.get("etaB", DEFAULT_ETA_B))
        elif kind == "disney":
            names = [("base_color", 0.5), ("metallic", 0.0), ("roughness", 1.0), ("eta", 1.5),
                     ("specular", 1.0), ("specular_tint", 0.0), ("anisotropic", 0.0), ("sheen", 0.0),
                     ("sheen_tint", 0.0), ("clearcoat", 0.0), ("clearcoat_gloss", 0.0),
                     ("scatting_distance", 0.0), ("transmission", 0.0),
                     ("transmission_roughness", 0.0), ("flatness", 0.0)]
            for i, (n, d) in enumerate(names):
                t[i] = self._tex(p.get(n), d)
            m["thin"] = int(bool(p.get("is_thin", False)))
        else:
            raise ValueError(kind)
        self.materials.append(m)
        return len(self.materials) - 1

    def add_medium(self, sigma_a=(0, 0, 0), sigma_s=(0, 0, 0), g=0.0, scale=1.0):
        self.media.append(dict(sigma_a=tuple(sigma_a), sigma_s=tuple(sigma_s), g=float(g), scale=float(scale)))
        return len(self.media) - 1

    # ---- geometry ---------------------------------------------------------
    def add_mesh(self, positions, indices, material=-1, normals=None, uvs=None, le=None,
                 medium_inside=-1, medium_outside=-1):
        """A scene entity with one model: one primitive per triangle, and with
        `le` one DiffuseAreaLight per triangle (RendererBuilder.cs:976-992).
        Returns the index of its first primitive."""
        first_prim = self.num_prims()
        t0, nt = self._add_triangles(positions, indices, normals, uvs)
        kinds = np.zeros(nt, np.int32)
        tris = np.arange(t0, t0 + nt, dtype=np.int32)
        if le is not None:
            light0 = len(self.lights)
            for i in range(nt):
                self.lights.append(dict(type=abi.LIGHT_AREA, prim=first_prim + i, le=tuple(le), tex=-1, to_world=None))
            lights = np.arange(light0, light0 + nt, dtype=np.int32)
        else:
            lights = np.full(nt, -1, np.int32)
        block = np.stack([kinds, tris, np.full(nt, material, np.int32), lights,
                          np.full(nt, medium_inside, np.int32), np.full(nt, medium_outside, np.int32)], axis=1)
        self.prims.append(block)
        return first_prim

    def _add_triangles(self, positions, indices, normals, uvs):
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(indices, dtype=np.int32).reshape(-1, 3)
        flags = 0
        self._positions.append(pos)
        if normals is not None:
            self._normals.append(np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3))
            self._has_normals = True
            flags |= 1
        else:
            self._normals.append(None)
        if uvs is not None:
            self._uvs.append(np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1, 2))
            self._has_uvs = True
            flags |= 2
        else:
            self._uvs.append(None)
        self._indices.append(idx + self._nverts)
        self._tflags.append(np.full(idx.shape[0], flags, dtype=np.uint8))
        self._nverts += pos.shape[0]
        t0 = self._ntri
        self._ntri += idx.shape[0]
        return t0, idx.shape[0]

    def add_instance_set(self, positions, indices, normals=None, uvs=None, bvh_max_prim=1):
        """An instanced mesh set (RendererBuilder's InstancedPrimitivesEntry,
        Config/RendererBuilder.cs:895-921): triangles already in the set's
        model space (its local transform applied, as TriangleMesh does) under
        the set's own BVH ("accel" max_prim; the reference's default is 1).
        The triangles are not scene primitives.  Returns the set index."""
        t0, nt = self._add_triangles(positions, indices, normals, uvs)
        self.instance_sets.append((t0, nt, int(bvh_max_prim)))
        return len(self.instance_sets) - 1

    def add_instance(self, set_index, to_world, material=-1, medium_inside=-1, medium_outside=-1):
        """An entity instancing a set: one InstancedPrimitive (Primitive.cs:144-189)
        with the entity's ModelToWorld (row-major 4x4, row vectors), material
        and media.  Returns its primitive index."""
        if not 0 <= set_index < len(self.instance_sets):
            raise ValueError("instance set out of range")
        k = len(self.instances)
        self.instances.append((int(set_index), np.asarray(to_world, np.float32).reshape(4, 4)))
        p = self.num_prims()
        self.prims.append(np.array([[abi.PRIM_INSTANCE, k, material, -1, medium_inside, medium_outside]], np.int32))
        return p

    def add_sphere(self, center, radius, material=-1, le=None, medium_inside=-1, medium_outside=-1, to_world=None):
        """A sphere shape (Shapes/Sphere.cs; the builder of RendererBuilder.cs:49-54):
        `center` is moved by the entity's to_world (row-major 4x4, row
        vectors), the radius is kept, and the inverse's linear part rotates
        the hit normal into the sphere's frame for its UV."""
        k = len(self.spheres)
        w2m = np.eye(3, dtype=np.float32)
        if to_world is not None:
            M = np.asarray(to_world, np.float32).reshape(4, 4)
            c = np.asarray(center, np.float32)
            center = tuple(float(v) for v in (c[0] * M[0, :3] + c[1] * M[1, :3] + c[2] * M[2, :3] + M[3, :3]))
            w2m = np.linalg.inv(M.astype(np.float64))[:3, :3].astype(np.float32)
        self.sphere_w2m.append(w2m)
        self.spheres.append(tuple(center) + (float(radius),))
        p = self.num_prims()
        light = -1
        if le is not None:
            light = len(self.lights)
            self.lights.append(dict(type=abi.LIGHT_AREA, prim=p, le=tuple(le), tex=-1, to_world=None))
        self.prims.append(np.array([[abi.PRIM_SPHERE, k, material, light, medium_inside, medium_outside]], np.int32))
        return p

    def add_infinite_light(self, le, to_world=None):
        """An "infinite" light entity (RendererBuilder.cs:221-225, InfiniteAreaLight):
        `le` is a Texture (or colour) holding the equirectangular radiance map,
        to_world the entity's model-to-world matrix (row-major, row vectors).
        It joins Scene.Lights at this point of the entity order.  Returns its
        light index."""
        t = self.add_texture(_as_tex(le, (0.0, 0.0, 0.0)))
        tw = None if to_world is None else np.asarray(to_world, np.float32).reshape(4, 4)
        self.lights.append(dict(type=abi.LIGHT_INFINITE, prim=-1, le=None, tex=t, to_world=tw))
        return len(self.lights) - 1

    def num_prims(self):
        return int(sum(b.shape[0] for b in self.prims))

    def set_camera(self, **kw):
        self.camera.update(kw)

    def set_integrator(self, **kw):
        self.integrator.update(kw)

    # ---- flatten to rd_scene_desc ------------------------------------------
    def arrays(self):
        pos = np.concatenate(self._positions) if self._positions else np.zeros((0, 3), np.float32)
        nv = pos.shape[0]
        nrm = None
        if self._has_normals:
            nrm = np.concatenate([n if n is not None else np.zeros((p.shape[0], 3), np.float32)
                                  for n, p in zip(self._normals, self._positions)])
        uv = None
        if self._has_uvs:
            uv = np.concatenate([u if u is not None else np.zeros((p.shape[0], 2), np.float32)
                                 for u, p in zip(self._uvs, self._positions)])
        idx = np.concatenate(self._indices) if self._indices else np.zeros((0, 3), np.int32)
        tfl = np.concatenate(self._tflags) if self._tflags else np.zeros(0, np.uint8)
        sph = np.array(self.spheres, np.float32).reshape(-1, 4)
        prims = np.concatenate(self.prims) if self.prims else np.zeros((0, 6), np.int32)
        return dict(positions=pos, normals=nrm, uvs=uv, indices=idx, tri_flags=tfl, spheres=sph,
                    prims=np.ascontiguousarray(prims, np.int32), num_vertices=nv)

    def to_c(self):
        """Builds an abi.rd_scene_desc.  The returned object keeps every
        buffer it points to alive (attribute `_keep`)."""
        a = self.arrays()
        keep = []

        def fptr(arr):
            if arr is None:
                return C.POINTER(C.c_float)()
            arr = np.ascontiguousarray(arr, dtype=np.float32)
            keep.append(arr)
            return arr.ctypes.data_as(C.POINTER(C.c_float))

        d = abi.rd_scene_desc()
        d.num_vertices = a["num_vertices"]
        d.positions = fptr(a["positions"])
        d.normals = fptr(a["normals"])
        d.uvs = fptr(a["uvs"])
        idx = np.ascontiguousarray(a["indices"], np.int32)
        keep.append(idx)
        d.num_triangles = idx.shape[0]
        d.indices = idx.ctypes.data_as(C.POINTER(C.c_int32))
        tfl = np.ascontiguousarray(a["tri_flags"], np.uint8)
        keep.append(tfl)
        d.tri_flags = tfl.ctypes.data_as(C.POINTER(C.c_uint8))
        d.num_spheres = a["spheres"].shape[0]
        d.spheres = fptr(a["spheres"])
        w2m = np.array(self.sphere_w2m, np.float32).reshape(-1, 9)
        d.sphere_w2m = fptr(w2m) if (w2m.size and not np.array_equal(w2m, np.tile(np.eye(3, dtype=np.float32).reshape(9), (w2m.shape[0], 1)))) else fptr(None)
        prims = a["prims"]
        n = prims.shape[0]
        parr = (abi.rd_prim_desc * max(n, 1))()
        if n:
            C.memmove(parr, np.ascontiguousarray(prims, np.int32).ctypes.data, prims.nbytes)
        keep.append(parr)
        d.num_prims = n
        d.prims = C.cast(parr, C.POINTER(abi.rd_prim_desc))

        marr = (abi.rd_material_desc * max(len(self.materials), 1))()
        for i, m in enumerate(self.materials):
            md = marr[i]
            md.type, md.two_sided, md.dist, md.thin = m["type"], m["two_sided"], m["dist"], m["thin"]
            for j in range(16):
                md.tex[j] = m["tex"][j]
            md.eta_a, md.eta_b, md.kd, md.ks = m["eta_a"], m["eta_b"], m["kd"], m["ks"]
            for j in range(3):
                md.eta[j] = m["eta"][j]
                md.k[j] = m["k"][j]
        keep.append(marr)
        d.num_materials = len(self.materials)
        d.materials = C.cast(marr, C.POINTER(abi.rd_material_desc))

        tarr = (abi.rd_texture_desc * max(len(self.textures), 1))()
        for i, t in enumerate(self.textures):
            td = tarr[i]
            td.wrap, td.filter = t.wrap, t.filter
            for j in range(3):
                td.value[j] = t.value[j]
            if t.image is not None:
                td.width, td.height = t.image.shape[0], t.image.shape[1]
                td.data = fptr(t.image)
            else:
                td.width = td.height = 0
        keep.append(tarr)
        d.num_textures = len(self.textures)
        d.textures = C.cast(tarr, C.POINTER(abi.rd_texture_desc))

        larr = (abi.rd_light_desc * max(len(self.lights), 1))()
        for i, l in enumerate(self.lights):
            ld = larr[i]
            ld.type, ld.prim, ld.tex = l["type"], l["prim"], l["tex"]
            for j in range(3):
                ld.le[j] = l["le"][j] if l["le"] is not None else 0.0
            tw = l.get("to_world")
            tw = np.eye(4, dtype=np.float32) if tw is None else np.asarray(tw, np.float32)
            for j in range(16):
                ld.to_world[j] = float(tw.reshape(16)[j])
        keep.append(larr)
        d.num_lights = len(self.lights)
        d.lights = C.cast(larr, C.POINTER(abi.rd_light_desc))

        earr = (abi.rd_medium_desc * max(len(self.media), 1))()
        for i, m in enumerate(self.media):
            e = earr[i]
            for j in range(3):
                e.sigma_a[j] = m["sigma_a"][j]
                e.sigma_s[j] = m["sigma_s"][j]
            e.g, e.scale = m["g"], m["scale"]
        keep.append(earr)
        d.num_media = len(self.media)
        d.media = C.cast(earr, C.POINTER(abi.rd_medium_desc))
        d.global_medium = self.global_medium
        d.bvh_max_prim = self.bvh_max_prim

        sarr = (abi.rd_instance_set_desc * max(len(self.instance_sets), 1))()
        for i, (t0, nt, mp) in enumerate(self.instance_sets):
            sarr[i].first_triangle, sarr[i].num_triangles, sarr[i].bvh_max_prim = t0, nt, mp
        keep.append(sarr)
        d.num_instance_sets = len(self.instance_sets)
        d.instance_sets = C.cast(sarr, C.POINTER(abi.rd_instance_set_desc))
        iarr = (abi.rd_instance_desc * max(len(self.instances), 1))()
        for i, (st, tw) in enumerate(self.instances):
            iarr[i].set = st
            for j in range(16):
                iarr[i].to_world[j] = float(tw.reshape(16)[j])
        keep.append(iarr)
        d.num_instances = len(self.instances)
        d.instances = C.cast(iarr, C.POINTER(abi.rd_instance_desc))

        c = self.camera
        cam = d.camera
        cam.fov, cam.near_z, cam.far_z = c["fov"], c["near"], c["far"]
        cam.width, cam.height = c["width"], c["height"]
        if c.get("to_world") is not None:
            cam.use_matrix = 1
            tw = np.asarray(c["to_world"], np.float32).reshape(16)
            for j in range(16):
                cam.to_world[j] = float(tw[j])
        else:
            cam.use_matrix = 0
        for j in range(3):
            cam.origin[j] = c["origin"][j]
            cam.target[j] = c["target"][j]
            cam.up[j] = c["up"][j]

        d.accel = {"bvh": abi.ACCEL_BVH, "octree": abi.ACCEL_OCTREE}[self.accel["type"]]
        d.octree_max_depth = self.accel["max_depth"]
        d.octree_max_count = self.accel["max_count"]
        d.octree_out_bound = self.accel["out_bound"]
        it = self.integrator
        d.integrator.type = {"path": abi.INTEGRATOR_PATH, "vol_path": abi.INTEGRATOR_VOL_PATH,
                             "adjoint_particle": abi.INTEGRATOR_PARTICLE, "bdpt": abi.INTEGRATOR_BDPT}[it["type"]]
        d.integrator.max_depth = it["max_depth"]
        d.integrator.min_depth = it["min_depth"]
        d.integrator.rr_threshold = it["rr_threshold"]
        d.integrator.method = {"Bxdf": abi.METHOD_BXDF, "Nee": abi.METHOD_NEE, "Mis": abi.METHOD_MIS}[it["method"]]
        d.integrator.strategy = {"Uniform": abi.STRATEGY_UNIFORM, "All": abi.STRATEGY_ALL}[it["strategy"]]
        d._keep = keep
        return d
