"""Deterministic procedural scenes for the benchmark workloads.

Every generator is pure numpy with explicit seeds, so the product and the
oracle receive bit-identical float32 vertices.  Geometry is baked to world
space with identity entity transforms, so TriangleMesh keeps positions
verbatim (Radiantium.Offline/Shapes/Triangle.cs:15-29).

Conventions follow the reference's Cornell examples (Example/cbox/*.json):
y up, camera on +z looking toward -z, area lights facing -y, one-sided
diffuse walls whose geometric normal (cross(Pb-Pa, Pc-Pa)) faces the room.
"""
import math

import numpy as np

from .scene import SceneDesc


def _quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3); normal = cross(p1-p0, p2-p0)."""
    pos = np.array([p0, p1, p2, p3], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    return pos, idx


def _box(size, center, angle_deg):
    """Axis-aligned box of extent `size` (sx, sy, sz) resting on y=center[1],
    rotated about y by angle_deg, outward-facing normals, 12 triangles."""
    sx, sy, sz = size
    hx, hz = sx / 2, sz / 2
    c = [(-hx, 0, -hz), (hx, 0, -hz), (hx, 0, hz), (-hx, 0, hz),
         (-hx, sy, -hz), (hx, sy, -hz), (hx, sy, hz), (-hx, sy, hz)]
    a = math.radians(angle_deg)
    ca, sa = math.cos(a), math.sin(a)
    pts = []
    for x, y, z in c:
        rx = ca * x + sa * z
        rz = -sa * x + ca * z
        pts.append((rx + center[0], y + center[1], rz + center[2]))
    pos = np.array(pts, np.float32)
    # faces wound so cross(b-a, c-a) points outward
    quads = [(4, 7, 6, 5),   # top (+y)
             (0, 1, 2, 3),   # bottom (-y)
             (3, 2, 6, 7),   # +z
             (1, 0, 4, 5),   # -z
             (2, 1, 5, 6),   # +x
             (0, 3, 7, 4)]   # -x
    idx = []
    for q in quads:
        idx.append([q[0], q[1], q[2]])
        idx.append([q[0], q[2], q[3]])
    return pos, np.array(idx, np.int32)


def cornell_box(width=256, height=256, max_depth=8, method="Mis", closed=False, light_kd=0.0):
    """Plain path integrator Cornell box: five unit wall quads (grey 0.73,
    red 0.65/0.05/0.05, green 0.12/0.45/0.15), boxes of size 0.3 and 0.6
    (0.3 x 0.6 x 0.3) rotated +-17 degrees about y, a 0.25 x 0.25 ceiling
    quad light of radiance (17,12,4); all Lambertian, one-sided.
    Room: x in [-0.5,0.5], y in [0,1], z in [-0.5,0.5], open at +z.
    closed: a grey front wall at z = 0.5 and the camera inside the room (no
    path leaves it); light_kd: the light's own diffuse albedo."""
    s = SceneDesc()
    grey = s.add_material("diffuse", kd=(0.73, 0.73, 0.73))
    red = s.add_material("diffuse", kd=(0.65, 0.05, 0.05))
    green = s.add_material("diffuse", kd=(0.12, 0.45, 0.15))
    light_mat = s.add_material("diffuse", kd=(light_kd, light_kd, light_kd))
    L, H = -0.5, 0.5
    # floor (normal +y)
    s.add_mesh(*_quad((L, 0, L), (L, 0, H), (H, 0, H), (H, 0, L)), material=grey)
    # ceiling (normal -y)
    s.add_mesh(*_quad((L, 1, L), (H, 1, L), (H, 1, H), (L, 1, H)), material=grey)
    # back wall z=-0.5 (normal +z)
    s.add_mesh(*_quad((L, 0, L), (H, 0, L), (H, 1, L), (L, 1, L)), material=grey)
    # left wall x=-0.5, red (normal +x)
    s.add_mesh(*_quad((L, 0, L), (L, 1, L), (L, 1, H), (L, 0, H)), material=red)
    # right wall x=+0.5, green (normal -x)
    s.add_mesh(*_quad((H, 0, L), (H, 0, H), (H, 1, H), (H, 1, L)), material=green)
    # short box (0.3 cube) rotated +17, tall box (0.3 x 0.6 x 0.3) rotated -17
    s.add_mesh(*_box((0.3, 0.3, 0.3), (0.17, 0.0, 0.12), 17.0), material=grey)
    s.add_mesh(*_box((0.3, 0.6, 0.3), (-0.17, 0.0, -0.15), -17.0), material=grey)
    # ceiling light 0.25 x 0.25 just below the ceiling, facing -y
    q = 0.125
    y = 1.0 - 1e-3
    s.add_mesh(*_quad((-q, y, -q), (q, y, -q), (q, y, q), (-q, y, q)), material=light_mat, le=(17.0, 12.0, 4.0))
    if closed:
        # front wall z=+0.5 (normal -z), the camera just inside it
        s.add_mesh(*_quad((L, 0, H), (L, 1, H), (H, 1, H), (H, 0, H)), material=grey)
        s.set_camera(fov=70.0, width=width, height=height, origin=(0.0, 0.5, 0.45), target=(0.0, 0.45, 0.0),
                     up=(0.0, 1.0, 0.0))
    else:
        s.set_camera(fov=40.0, width=width, height=height, origin=(0.0, 0.5, 1.85), target=(0.0, 0.5, 0.0),
                     up=(0.0, 1.0, 0.0))
    s.set_integrator(type="path", max_depth=max_depth, method=method)
    s.bvh_max_prim = 1
    return s


def checker_texture(n=64, c0=(0.8, 0.8, 0.8), c1=(0.2, 0.3, 0.7), cells=8):
    """n x n checkerboard as a 3-channel column-major image (x*H + y)."""
    x = np.arange(n)[:, None]
    y = np.arange(n)[None, :]
    mask = ((x * cells // n) + (y * cells // n)) % 2 == 0
    img = np.where(mask[..., None], np.array(c0, np.float32), np.array(c1, np.float32))
    return img.astype(np.float32)


def material_box(width=64, height=64, max_depth=8, method="Mis", delta=True):
    """Cornell box exercising every material of the registry and the image
    texture sampler: a textured rough_plastic left wall (Linear filter, Repeat
    wrap, UVs tiled twice), GGX plastic floor, diffuse green right wall,
    a Disney short box (sheen + clearcoat), a perfect_mirror tall box, a
    perfect_glass sphere, an anisotropic rough_metal sphere and a rough_glass
    sphere, lit by the ceiling quad light and a small emissive sphere.
    delta=False makes the mirror box a rough_metal and the glass sphere a
    rough_glass (no delta BSDF left)."""
    from .scene import Texture
    s = SceneDesc()
    grey = s.add_material("diffuse", kd=(0.73, 0.73, 0.73))
    green = s.add_material("diffuse", kd=(0.12, 0.45, 0.15))
    tex = Texture(image=checker_texture(), wrap="Repeat", filter="Linear")
    plastic_tex = s.add_material("rough_plastic", r=tex, roughness=0.35, dist="Beckmann", kd=0.6, ks=0.4)
    plastic_floor = s.add_material("rough_plastic", r=(0.6, 0.5, 0.4), roughness=0.25, dist="GGX")
    disney = s.add_material("disney", base_color=(0.8, 0.3, 0.2), roughness=0.45, sheen=0.6, sheen_tint=0.5,
                            clearcoat=0.7, clearcoat_gloss=0.6, specular_tint=0.3, metallic=0.1)
    if delta:
        mirror = s.add_material("perfect_mirror", r=(0.9, 0.9, 0.9))
        glass = s.add_material("perfect_glass")
    else:
        mirror = s.add_material("rough_metal", roughness=0.3)
        glass = s.add_material("rough_glass", roughness=0.35)
    metal = s.add_material("rough_metal", roughness=0.4, anisotropic=0.5)
    rglass = s.add_material("rough_glass", roughness=0.2)
    light_mat = s.add_material("diffuse", kd=(0.0, 0.0, 0.0))
    L, H = -0.5, 0.5
    uv = np.array([[0, 0], [2, 0], [2, 2], [0, 2]], np.float32)
    s.add_mesh(*_quad((L, 0, L), (L, 0, H), (H, 0, H), (H, 0, L)), material=plastic_floor)
    s.add_mesh(*_quad((L, 1, L), (H, 1, L), (H, 1, H), (L, 1, H)), material=grey)
    s.add_mesh(*_quad((L, 0, L), (H, 0, L), (H, 1, L), (L, 1, L)), material=grey)
    pos, idx = _quad((L, 0, L), (L, 1, L), (L, 1, H), (L, 0, H))
    s.add_mesh(pos, idx, material=plastic_tex, uvs=uv)
    s.add_mesh(*_quad((H, 0, L), (H, 0, H), (H, 1, H), (H, 1, L)), material=green)
    s.add_mesh(*_box((0.3, 0.3, 0.3), (0.17, 0.0, 0.12), 17.0), material=disney)
    s.add_mesh(*_box((0.3, 0.6, 0.3), (-0.17, 0.0, -0.15), -17.0), material=mirror)
    s.add_sphere((0.18, 0.45, 0.15), 0.12, material=glass)
    s.add_sphere((-0.28, 0.1, 0.25), 0.1, material=metal)
    s.add_sphere((0.3, 0.08, -0.3), 0.08, material=rglass)
    s.add_sphere((-0.3, 0.85, 0.3), 0.04, material=light_mat, le=(6.0, 6.0, 6.0))
    q = 0.125
    y = 1.0 - 1e-3
    s.add_mesh(*_quad((-q, y, -q), (q, y, -q), (q, y, q), (-q, y, q)), material=light_mat, le=(17.0, 12.0, 4.0))
    s.set_camera(fov=40.0, width=width, height=height, origin=(0.0, 0.5, 1.85), target=(0.0, 0.5, 0.0),
                 up=(0.0, 1.0, 0.0))
    s.set_integrator(type="path", max_depth=max_depth, method=method)
    s.bvh_max_prim = 1
    return s


def hdr_sky(width, height, seed=0, sigma=1.5, sun_dir=(0.4, 0.75, 0.3), sun_radiance=400.0, sun_deg=2.5):
    """Equirectangular HDR radiance map [width][height][3] (column-major like
    ColorBuffer): per-texel log-normal radiance (mu 0, sigma) with a slight
    colour spread, plus one bright sun disk baked into the map (the reference
    has no sun-light type).  Texel (x, y) covers phi = 2 pi (x+0.5)/W,
    theta = pi (y+0.5)/H; direction (sin phi sin theta, cos theta, -cos phi sin theta)
    as InfiniteAreaLight.Le's inverse mapping."""
    rng = np.random.default_rng(seed)
    lum = np.exp(rng.normal(0.0, sigma, (width, height)))
    tint = rng.uniform(0.8, 1.2, (width, height, 3))
    img = lum[..., None] * tint
    x, y = np.meshgrid((np.arange(width) + 0.5) / width, (np.arange(height) + 0.5) / height, indexing="ij")
    phi, theta = 2 * np.pi * x, np.pi * y
    d = np.stack([np.sin(phi) * np.sin(theta), np.cos(theta), -np.cos(phi) * np.sin(theta)], -1)
    sd = np.asarray(sun_dir, np.float64)
    sd /= np.linalg.norm(sd)
    sun = (d @ sd) > np.cos(np.deg2rad(sun_deg))
    img[sun] = sun_radiance * np.array([1.0, 0.95, 0.85])
    return img.astype(np.float32)


def env_showcase(width=64, height=64, max_depth=8, method="Mis", map_w=64, map_h=32, seed=0,
                 rotate_deg=30.0, area_light=True):
    """Objects under an HDR environment map (rotated about y by `rotate_deg`
    through the light's to_world): Disney, rough plastic and rough metal
    spheres, a perfect-glass sphere, a rough-glass slab on a diffuse ground,
    and (optionally) one small emissive quad, so light selection mixes area
    and infinite lights."""
    from .scene import Texture
    s = SceneDesc()
    ground = s.add_material("diffuse", kd=(0.5, 0.5, 0.5))
    disney = s.add_material("disney", base_color=(0.7, 0.4, 0.2), roughness=0.4, metallic=0.3, clearcoat=0.5)
    plastic = s.add_material("rough_plastic", r=(0.2, 0.4, 0.7), roughness=0.3, dist="GGX")
    metal = s.add_material("rough_metal", roughness=0.3)
    glass = s.add_material("perfect_glass")
    rglass = s.add_material("rough_glass", roughness=0.25)
    black = s.add_material("diffuse", kd=(0.0, 0.0, 0.0))
    W = 4.0
    s.add_mesh(*_quad((-W, 0, -W), (-W, 0, W), (W, 0, W), (W, 0, -W)), material=ground)
    s.add_sphere((-0.9, 0.4, 0.0), 0.4, material=disney)
    s.add_sphere((0.0, 0.4, -0.4), 0.4, material=plastic)
    s.add_sphere((0.9, 0.4, 0.0), 0.4, material=metal)
    s.add_sphere((0.3, 0.25, 0.7), 0.25, material=glass)
    s.add_mesh(*_box((0.5, 0.1, 0.5), (-0.35, 0.0, 0.75), 20.0), material=rglass)
    a = np.deg2rad(rotate_deg)
    R = np.array([[np.cos(a), 0, -np.sin(a), 0], [0, 1, 0, 0], [np.sin(a), 0, np.cos(a), 0], [0, 0, 0, 1]],
                 np.float32)
    s.add_infinite_light(Texture(image=hdr_sky(map_w, map_h, seed)), to_world=R)
    if area_light:
        q = 0.2
        s.add_mesh(*_quad((-q, 2.0, -q), (q, 2.0, -q), (q, 2.0, q), (-q, 2.0, q)), material=black,
                   le=(20.0, 20.0, 20.0))
    s.set_camera(fov=45.0, width=width, height=height, origin=(0.0, 1.2, 3.2), target=(0.0, 0.35, 0.0),
                 up=(0.0, 1.0, 0.0))
    s.set_integrator(type="path", max_depth=max_depth, method=method)
    s.bvh_max_prim = 1
    return s


def sss_cornell(width=64, height=64, max_depth=8, method="Mis", disney=True, subsurface=True,
                scatter=(0.05, 0.03, 0.02)):
    """The Cornell box of cornell_box() with subsurface scattering: the tall
    box a non-thin Disney material with a scattering distance (its BSSRDF
    profile A = base colour * (1 - metallic)), and a sphere of the
    "subsurface" material (FresnelSpecularBsdf at the surface, BSSRDF
    profile A, scattering distance, eta 1.33)."""
    s = cornell_box(width, height, max_depth=max_depth, method=method)
    if disney:
        skin = s.add_material("disney", base_color=(0.85, 0.55, 0.45), roughness=0.6, eta=1.4,
                              scatting_distance=scatter)
        # the tall box (the seventh mesh: twelve triangles) takes the Disney material
        s.prims[6][:, 2] = skin
    if subsurface:
        jade = s.add_material("subsurface", a=(0.4, 0.8, 0.5), scatting_distance=tuple(2 * v for v in scatter),
                              etaA=1.0, etaB=1.33)
        s.add_sphere((0.2, 0.42, 0.15), 0.12, material=jade)
    return s


def vol_cornell(width=64, height=64, max_depth=64, sigma_a=0.05, sigma_s=2.0, g=0.5, radius=0.22):
    """Volumetric Cornell box: the Cornell box of cornell_box() with a
    perfect_glass sphere whose inside holds a homogeneous medium (grey
    sigma_a, sigma_s, Henyey-Greenstein g), rendered by vol_path."""
    s = cornell_box(width, height, max_depth=max_depth)
    glass = s.add_material("perfect_glass")
    m = s.add_medium(sigma_a=(sigma_a,) * 3, sigma_s=(sigma_s,) * 3, g=g)
    s.add_sphere((0.05, 0.55, 0.1), radius, material=glass, medium_inside=m)
    s.set_integrator(type="vol_path", max_depth=max_depth)
    return s


def vol_fog(width=64, height=64, max_depth=32, sigma_a=(0.01, 0.015, 0.02), sigma_s=(0.1, 0.08, 0.06), g=0.3):
    """Spectral fog everywhere (global medium) over material_box-like objects,
    lit by an HDR sky and a quad light, plus a medium slab bounded by a
    surface-less box (a medium boundary the chains must cross)."""
    from .scene import Texture
    s = SceneDesc()
    ground = s.add_material("diffuse", kd=(0.6, 0.6, 0.6))
    plastic = s.add_material("rough_plastic", r=(0.7, 0.3, 0.2), roughness=0.3, dist="GGX")
    metal = s.add_material("rough_metal", roughness=0.25)
    black = s.add_material("diffuse", kd=(0.0, 0.0, 0.0))
    fog = s.add_medium(sigma_a=sigma_a, sigma_s=sigma_s, g=g)
    slab = s.add_medium(sigma_a=(0.4, 0.2, 0.1), sigma_s=(1.0, 1.5, 2.0), g=-0.2)
    W = 4.0
    s.add_mesh(*_quad((-W, 0, -W), (-W, 0, W), (W, 0, W), (W, 0, -W)), material=ground,
               medium_outside=fog, medium_inside=fog)
    s.add_sphere((-0.6, 0.4, 0.0), 0.4, material=plastic, medium_outside=fog)
    s.add_sphere((0.6, 0.4, -0.2), 0.4, material=metal, medium_outside=fog)
    # medium slab: triangles without a surface, slab medium inside, fog outside
    s.add_mesh(*_box((0.6, 0.5, 0.3), (0.0, 0.0, 0.8), 10.0), material=-1, medium_inside=slab, medium_outside=fog)
    q = 0.3
    s.add_mesh(*_quad((-q, 2.2, -q), (q, 2.2, -q), (q, 2.2, q), (-q, 2.2, q)), material=black, le=(40.0, 36.0, 30.0),
               medium_outside=fog, medium_inside=fog)
    s.add_infinite_light(Texture(image=hdr_sky(32, 16, seed=3, sigma=0.8, sun_radiance=50.0)))
    s.global_medium = fog
    s.set_camera(fov=45.0, width=width, height=height, origin=(0.0, 1.2, 3.2), target=(0.0, 0.35, 0.0),
                 up=(0.0, 1.0, 0.0))
    s.set_integrator(type="vol_path", max_depth=max_depth)
    s.bvh_max_prim = 1
    return s


def batch_requests(seed=0, n=16, resolutions=(128, 256, 512, 1024), spps=(16, 64, 256), max_depth=64):
    """The batched workload: `n` vol_path Cornell boxes, each with a
    perfect_glass sphere holding a homogeneous medium whose grey sigma_a ~
    U[0.01, 0.2], sigma_s ~ U[0.5, 5] and g ~ U[-0.3, 0.9]; square resolution
    drawn from `resolutions`, spp from `spps`, max depth 64.  Returns a list of
    (SceneDesc, spp, seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        res = int(rng.choice(resolutions))
        spp = int(rng.choice(spps))
        sa = float(rng.uniform(0.01, 0.2))
        ss = float(rng.uniform(0.5, 5.0))
        g = float(rng.uniform(-0.3, 0.9))
        d = vol_cornell(res, res, max_depth=max_depth, sigma_a=sa, sigma_s=ss, g=g)
        out.append((d, spp, 1000 + i))
    return out


def diff_cornell(width=64, height=64, tex_res=512, seed=0, init=None, max_depth=8, method="Mis", rr=True,
                 filter="Nearest", wall="rough_plastic"):
    """Differentiable-rendering workload: the Cornell box of cornell_box()
    with the left and right walls made rough_plastic (GGX) carrying tex_res^2
    RGB albedo textures `r` and a roughness texture (GGX alpha ~ U[0.05, 0.8]
    stored as roughness = sqrt(alpha), in every channel), drawn from `seed`;
    init = (albedo, roughness) instead fills them uniformly.  Walls get UVs
    spanning [0,1]^2.  rr=False sets min_depth = max_depth (no Russian
    roulette); wall="diffuse" makes the walls Lambertian with the albedo
    textures as kd.  Returns (SceneDesc, {"left_r", "left_rough", "right_r",
    "right_rough"} -> texture index)."""
    from .scene import Texture
    rng = np.random.default_rng(seed)

    def tex_pair():
        if init is None:
            alb = rng.uniform(0.05, 0.95, (tex_res, tex_res, 3))
            rough = np.sqrt(rng.uniform(0.05, 0.8, (tex_res, tex_res)))
        else:
            alb = np.full((tex_res, tex_res, 3), init[0])
            rough = np.full((tex_res, tex_res), init[1])
        rough3 = np.repeat(rough[:, :, None], 3, axis=2)
        return (Texture(image=alb.astype(np.float32), filter=filter),
                Texture(image=rough3.astype(np.float32), filter=filter))

    s = SceneDesc()
    grey = s.add_material("diffuse", kd=(0.73, 0.73, 0.73))
    light_mat = s.add_material("diffuse", kd=(0.0, 0.0, 0.0))
    lr, lro = tex_pair()
    rr_, rro = tex_pair()
    if wall == "diffuse":
        left = s.add_material("diffuse", kd=lr)
        right = s.add_material("diffuse", kd=rr_)
        s.add_texture(lro)
        s.add_texture(rro)
    else:
        left = s.add_material("rough_plastic", r=lr, roughness=lro, dist="GGX")
        right = s.add_material("rough_plastic", r=rr_, roughness=rro, dist="GGX")
    L, H = -0.5, 0.5
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    s.add_mesh(*_quad((L, 0, L), (L, 0, H), (H, 0, H), (H, 0, L)), material=grey)
    s.add_mesh(*_quad((L, 1, L), (H, 1, L), (H, 1, H), (L, 1, H)), material=grey)
    s.add_mesh(*_quad((L, 0, L), (H, 0, L), (H, 1, L), (L, 1, L)), material=grey)
    pos, idx = _quad((L, 0, L), (L, 1, L), (L, 1, H), (L, 0, H))
    s.add_mesh(pos, idx, material=left, uvs=uv)
    pos, idx = _quad((H, 0, L), (H, 0, H), (H, 1, H), (H, 1, L))
    s.add_mesh(pos, idx, material=right, uvs=uv)
    s.add_mesh(*_box((0.3, 0.3, 0.3), (0.17, 0.0, 0.12), 17.0), material=grey)
    s.add_mesh(*_box((0.3, 0.6, 0.3), (-0.17, 0.0, -0.15), -17.0), material=grey)
    q = 0.125
    y = 1.0 - 1e-3
    s.add_mesh(*_quad((-q, y, -q), (q, y, -q), (q, y, q), (-q, y, q)), material=light_mat, le=(17.0, 12.0, 4.0))
    s.set_camera(fov=40.0, width=width, height=height, origin=(0.0, 0.5, 1.85), target=(0.0, 0.5, 0.0),
                 up=(0.0, 1.0, 0.0))
    s.set_integrator(type="path", max_depth=max_depth, method=method, min_depth=3 if rr else max_depth)
    s.bvh_max_prim = 1
    mats = {m: s.materials[i] for i, m in ((left, "left"), (right, "right"))}
    tex = {"left_r": mats["left"]["tex"][0], "right_r": mats["right"]["tex"][0]}
    if wall != "diffuse":
        tex.update(left_rough=mats["left"]["tex"][1], right_rough=mats["right"]["tex"][1])
    return s, tex


# ---------------------------------------------------------------- instancing
def superquadric(e1, e2, nu=64, nv=32, radius=(1.0, 1.0, 1.0)):
    """Superquadric surface (x, y, z) = r * (c(v)^e1 c(u)^e2, s(v)^e1, c(v)^e1 s(u)^e2)
    with signed powers c^e = sign(cos) |cos|^e, u in [0, 2 pi) (nu segments),
    v in [-pi/2, pi/2] (nv rings, poles as single vertices).  Triangles wind
    outward; UVs (u / 2 pi, v / pi + 1/2).  Returns (positions, indices, uvs)."""
    def sp(x, e):
        return np.sign(x) * np.abs(x) ** e
    us = np.arange(nu) * (2 * np.pi / nu)
    vs = -np.pi / 2 + np.arange(1, nv) * (np.pi / nv)
    U, V = np.meshgrid(us, vs, indexing="xy")              # rings x segments
    cv, sv, cu, su = sp(np.cos(V), e1), sp(np.sin(V), e1), sp(np.cos(U), e2), sp(np.sin(U), e2)
    ring = np.stack([radius[0] * cv * cu, radius[1] * sv, radius[2] * cv * su], -1).reshape(-1, 3)
    pos = np.concatenate([[[0.0, -radius[1], 0.0]], ring, [[0.0, radius[1], 0.0]]]).astype(np.float32)
    uv_r = np.stack([U / (2 * np.pi), V / np.pi + 0.5], -1).reshape(-1, 2)
    uvs = np.concatenate([[[0.5, 0.0]], uv_r, [[0.5, 1.0]]]).astype(np.float32)
    nr = nv - 1
    top = 1 + nr * nu
    idx = []
    for j in range(nu):
        jn = (j + 1) % nu
        idx.append((0, 1 + jn, 1 + j))                      # south cap
        idx.append((top, 1 + (nr - 1) * nu + j, 1 + (nr - 1) * nu + jn))   # north cap
    for r in range(nr - 1):
        a0, a1 = 1 + r * nu, 1 + (r + 1) * nu
        for j in range(nu):
            jn = (j + 1) % nu
            idx.append((a0 + j, a0 + jn, a1 + jn))
            idx.append((a0 + j, a1 + jn, a1 + j))
    idx = np.array(idx, np.int32)[:, [0, 2, 1]]
    return pos, np.ascontiguousarray(idx), uvs


def _rand_rotation(rng):
    """Uniform random rotation (unit quaternion) as a row-vector 3x3 matrix."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    return R.T


def instance_matrix(rot, scale, translate):
    """Entity ModelToWorld = Scale * Rotation * Translation in the reference's
    row-vector convention (RendererBuilder.cs:925-938): rows 1-3 = scale *
    rotation rows, row 4 = translation."""
    M = np.eye(4)
    M[:3, :3] = np.asarray(scale, np.float64).reshape(-1, 1) * np.asarray(rot, np.float64) \
        if np.ndim(scale) else scale * np.asarray(rot, np.float64)
    M[3, :3] = translate
    return M.astype(np.float32)


def instanced_scene(width=64, height=64, n_instances=48, n_sets=3, nu=24, nv=12, seed=0, max_depth=16,
                    method="Mis", env=(64, 32), area_light=True, extent=2.5, set_max_prim=4, ground=True,
                    nonuniform=False, integrator="path"):
    """Instanced scene: `n_sets` distinct superquadric meshes (exponents drawn
    from [0.5, 2.5]) as instance sets, `n_instances` InstancedPrimitives with
    seeded random rotations, scales in [0.5, 2] (per axis when `nonuniform`)
    and positions in a cube of half-size `extent`, materials cycling Disney /
    rough plastic / rough metal; optional ground quad and area light at the
    scene level, an HDR sky (log-normal texels plus a sun disk, rotated).  The
    set BVHs use max_prim `set_max_prim`, the scene BVH max_prim 1."""
    from .scene import Texture
    rng = np.random.default_rng(seed)
    s = SceneDesc()
    mats = [s.add_material("disney", base_color=(0.8, 0.5, 0.3), roughness=0.35, metallic=0.2, clearcoat=0.4),
            s.add_material("rough_plastic", r=(0.25, 0.45, 0.75), roughness=0.3, dist="GGX"),
            s.add_material("rough_metal", roughness=0.25)]
    sets = []
    for k in range(n_sets):
        e1, e2 = rng.uniform(0.5, 2.5, 2)
        pos, idx, uvs = superquadric(e1, e2, nu, nv)
        sets.append(s.add_instance_set(pos, idx, uvs=uvs, bvh_max_prim=set_max_prim))
    if ground:
        g = s.add_material("diffuse", kd=(0.45, 0.45, 0.45))
        W = 4 * extent
        s.add_mesh(*_quad((-W, -extent - 0.6, -W), (-W, -extent - 0.6, W), (W, -extent - 0.6, W),
                          (W, -extent - 0.6, -W)), material=g)
    for i in range(n_instances):
        scale = rng.uniform(0.5, 2.0, 3) if nonuniform else rng.uniform(0.5, 2.0)
        M = instance_matrix(_rand_rotation(rng), scale, rng.uniform(-extent, extent, 3))
        s.add_instance(sets[i % n_sets], M, material=mats[i % 3])
    if area_light:
        black = s.add_material("diffuse", kd=(0.0, 0.0, 0.0))
        q, y = 0.6 * extent, 2.5 * extent
        s.add_mesh(*_quad((-q, y, -q), (q, y, -q), (q, y, q), (-q, y, q)), material=black, le=(8.0, 7.5, 7.0))
    if env is not None:
        sky = hdr_sky(env[0], env[1], seed=seed + 1)
        c, sn = math.cos(math.radians(25.0)), math.sin(math.radians(25.0))
        tw = np.array([[c, 0, -sn, 0], [0, 1, 0, 0], [sn, 0, c, 0], [0, 0, 0, 1]], np.float32)
        s.add_infinite_light(Texture(image=sky, filter="Nearest"), to_world=tw)
    d = 3.2 * extent
    s.set_camera(fov=45.0, width=width, height=height, origin=(0.35 * d, 0.3 * d, d), target=(0.0, 0.0, 0.0),
                 up=(0.0, 1.0, 0.0))
    s.set_integrator(type=integrator, max_depth=max_depth, method=method)
    s.bvh_max_prim = 1
    return s


def instanced_workload(width=1024, height=1024, n_instances=1024, n_sets=8, seed=0, max_depth=16,
                       env=(2048, 1024), nu=224, nv=224):
    """The instanced benchmark workload: 1024 instances of 8 superquadrics
    (224 x 224 grid, 99,904 triangles each; exponents in [0.5, 2.5]) with
    seeded random rotations and scales in [0.5, 2] spread through a cube,
    Disney / plastic / rough-metal materials, under a 2048 x 1024 log-normal
    (mu 0, sigma 1.5) HDR sky with a sun disk; 1024 x 1024, max depth 16."""
    return instanced_scene(width, height, n_instances=n_instances, n_sets=n_sets, nu=nu, nv=nv, seed=seed,
                           max_depth=max_depth, env=env, area_light=False, extent=9.0, ground=False)


# ---------------------------------------------------------------- icosphere
def icosphere(level):
    """Unit icosphere: 20*4^level faces, midpoint subdivision with shared
    edge vertices, projected to the sphere after each level."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    for _ in range(level):
        nv = v.shape[0]
        nf = f.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        es = np.sort(e, axis=1)
        key = es[:, 0] * nv + es[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a = uniq // nv
        b = uniq % nv
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        v = np.concatenate([v, mid])
        m01 = inv[0:nf] + nv
        m12 = inv[nf:2 * nf] + nv
        m20 = inv[2 * nf:3 * nf] + nv
        c0 = np.stack([f[:, 0], m01, m20], 1)
        c1 = np.stack([f[:, 1], m12, m01], 1)
        c2 = np.stack([f[:, 2], m20, m12], 1)
        c3 = np.stack([m01, m12, m20], 1)
        f = np.stack([c0, c1, c2, c3], 1).reshape(-1, 3)
    return v, f


def _hash3(ix, iy, iz, seed):
    """Integer lattice hash -> [0,1) (value-noise lattice values)."""
    h = (ix.astype(np.uint64) * np.uint64(0x8DA6B343) ^ iy.astype(np.uint64) * np.uint64(0xD8163841)
         ^ iz.astype(np.uint64) * np.uint64(0xCB1AB31F) ^ np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15))
    h ^= h >> np.uint64(33)
    h *= np.uint64(0xFF51AFD7ED558CCD)
    h ^= h >> np.uint64(33)
    h *= np.uint64(0xC4CEB9FE1A85EC53)
    h ^= h >> np.uint64(33)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def value_noise(p, seed, octaves=3, base_freq=4.0):
    """Seeded multi-octave value noise in [-1,1]: trilinear interpolation of
    lattice values with smoothstep weights, octave k at base_freq*2^k with
    amplitude 2^-k, normalised by the amplitude sum."""
    total = np.zeros(p.shape[0])
    amp_sum = 0.0
    with np.errstate(over="ignore"):
        for k in range(octaves):
            freq = base_freq * (2.0 ** k)
            amp = 0.5 ** k
            q = p * freq + 1000.0
            i0 = np.floor(q).astype(np.int64)
            fr = q - i0
            w = fr * fr * (3.0 - 2.0 * fr)
            acc = np.zeros(p.shape[0])
            for dx in (0, 1):
                for dy in (0, 1):
                    for dz in (0, 1):
                        val = _hash3(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz, seed + k)
                        wx = w[:, 0] if dx else 1 - w[:, 0]
                        wy = w[:, 1] if dy else 1 - w[:, 1]
                        wz = w[:, 2] if dz else 1 - w[:, 2]
                        acc += val * wx * wy * wz
            total += amp * (2.0 * acc - 1.0)
            amp_sum += amp
    return total / amp_sum


def displaced_icosphere_scene(level=8, width=1920, height=1080, max_depth=12, seed=0, groups=64,
                              with_light=True, camera_z=7.0):
    """1M-triangle workload: icosphere subdivided to `level` (level 8 =
    1,310,720 triangles), radially displaced by seeded 3-octave value noise
    of amplitude 0.05, over a ground quad.  The mesh is split into `groups`
    face groups (8 latitude x 8 longitude bins of the face centroid), each a
    GGX rough_metal (default Cu) with alpha ~ U[0.1,0.5] stored as
    roughness = sqrt(alpha) (Microfacet.RoughnessToAlpha squares it).  The
    ground is a rough_glass with etaB = 1.5 (roughness: reference default
    0.3).  The workload names no emitter, so a 2 x 2 quad area light of
    radiance (8,8,8) faces down from y = 4 (one light per triangle, as the
    reference builds emissive meshes).  BVH max_prim 4.  The camera (fov 40)
    looks from (0, 2.2, camera_z) at (0, 0.9, 0): at camera_z = 7.0 about
    22 % of the primary rays meet the mesh (50 % the ground, 28 % nothing),
    at camera_z = 4.6 about half of them."""
    rng = np.random.default_rng(seed)
    v, f = icosphere(level)
    disp = value_noise(v, seed, octaves=3)
    center = np.array([0.0, 1.1, 0.0])
    pos = (v * (1.0 + 0.05 * disp)[:, None] + center).astype(np.float32)
    s = SceneDesc()
    alphas = rng.uniform(0.1, 0.5, size=groups)
    mats = [s.add_material("rough_metal", roughness=float(np.sqrt(np.float32(a))), dist="GGX") for a in alphas]
    cen = v[f].mean(axis=1)
    lat = np.clip(((np.arcsin(np.clip(cen[:, 1] / np.linalg.norm(cen, axis=1), -1, 1)) / np.pi + 0.5) * 8).astype(int), 0, 7)
    lon = np.clip(((np.arctan2(cen[:, 2], cen[:, 0]) / (2 * np.pi) + 0.5) * 8).astype(int), 0, 7)
    gid = (lat * 8 + lon) % groups
    order = np.argsort(gid, kind="stable")
    fs = f[order]
    gs = gid[order]
    bounds = np.searchsorted(gs, np.arange(groups + 1))
    for g in range(groups):
        sub = fs[bounds[g]:bounds[g + 1]]
        if sub.shape[0] == 0:
            continue
        used, remap = np.unique(sub.reshape(-1), return_inverse=True)
        s.add_mesh(pos[used], remap.reshape(-1, 3).astype(np.int32), material=mats[g])
    glass = s.add_material("rough_glass", etaB=1.5)
    G = 10.0
    s.add_mesh(*_quad((-G, 0, -G), (-G, 0, G), (G, 0, G), (G, 0, -G)), material=glass)
    if with_light:
        lm = s.add_material("diffuse", kd=(0.0, 0.0, 0.0))
        y = 4.0
        s.add_mesh(*_quad((-1, y, -1), (1, y, -1), (1, y, 1), (-1, y, 1)), material=lm, le=(8.0, 8.0, 8.0))
    s.set_camera(fov=40.0, width=width, height=height, origin=(0.0, 2.2, float(camera_z)), target=(0.0, 0.9, 0.0),
                 up=(0.0, 1.0, 0.0))
    s.set_integrator(type="path", max_depth=max_depth, method="Mis")
    s.bvh_max_prim = 4
    return s
