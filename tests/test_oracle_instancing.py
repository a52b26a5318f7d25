"""Pins the oracle's instancing (InstancedPrimitive, Primitive.cs:144-189;
Ray3F.Transform, Ray3F.cs:27-34; the per-set Bvh of RendererBuilder.cs:895-921)
against brute force over every instance triangle, against the same meshes
baked into world space, and against the reference's float behaviour for
rays whose MaxT is float.MaxValue.  CPU only."""
import numpy as np
import pytest

from radiantium_amd import abi, scenes
from radiantium_amd.scene import SceneDesc


def _rays(n, seed, extent=3.0, tmax=100.0):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, abi.ray_dtype())
    o = rng.uniform(-2.5 * extent, 2.5 * extent, (n, 3))
    tgt = rng.uniform(-extent, extent, (n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r["o"] = o.astype(np.float32)
    r["d"] = d.astype(np.float32)
    r["tmin"] = 1e-4
    r["tmax"] = tmax
    return r


def test_instance_traversal_matches_brute_force(oracle):
    desc = scenes.instanced_scene(n_instances=40, n_sets=3, nu=20, nv=10, seed=3, env=None, extent=2.5,
                                  nonuniform=True)
    o = oracle.OracleScene(desc)
    rays = _rays(20000, 1)
    h, sub, _ = o.trace_closest_sub(rays)
    b, bsub = o.trace_brute_sub(rays)
    hit = h["prim"] >= 0
    assert np.mean(hit == (b["prim"] >= 0)) > 0.999
    inst = hit & (sub >= 0)
    assert inst.sum() > 5000 and (hit & (sub < 0)).sum() > 500    # instances and the ground both hit
    same = (h["prim"] == b["prim"]) & (sub == bsub)
    assert np.mean(same[hit]) > 0.995
    # the same triangle through the same model ray: identical bits
    m = same & hit
    for f in ("t", "u", "v"):
        assert np.array_equal(h[f][m], b[f][m])


def test_identity_instance_is_the_mesh():
    """An instance with an identity transform finds the triangle the plain
    mesh finds (the model ray is the world ray up to the renormalisation of
    D), at the same distance up to rounding."""
    from oracle import cpuref
    pos, idx, uvs = scenes.superquadric(1.3, 0.8, 24, 12)
    rays = _rays(4000, 2, extent=1.2, tmax=50.0)
    plain = SceneDesc()
    m = plain.add_material("diffuse")
    plain.add_mesh(pos, idx, material=m, uvs=uvs)
    plain.set_camera(width=4, height=4)
    inst = SceneDesc()
    m = inst.add_material("diffuse")
    st = inst.add_instance_set(pos, idx, uvs=uvs, bvh_max_prim=1)
    inst.add_instance(st, np.eye(4), material=m)
    inst.set_camera(width=4, height=4)
    a = cpuref.OracleScene(plain).trace_closest(rays)
    b, sub, _ = cpuref.OracleScene(inst).trace_closest_sub(rays)
    hit = a["prim"] >= 0
    assert hit.sum() > 1000
    assert np.array_equal(hit, b["prim"] >= 0)
    assert np.mean(a["prim"][hit] == sub[hit]) > 0.999
    assert np.allclose(a["t"][hit], b["t"][hit], rtol=1e-5)


def test_instance_distance_is_world_distance(oracle):
    """T of an instance hit is |O - M2W * P_model|: the same point as the
    world-space ray parameter of the transformed mesh's hit (rotation,
    uniform scale 1.7, translation)."""
    from oracle import cpuref
    pos, idx, uvs = scenes.superquadric(0.7, 1.6, 24, 12)
    rng = np.random.default_rng(4)
    M = scenes.instance_matrix(scenes._rand_rotation(rng), 1.7, (0.3, -0.2, 0.5))
    world = (pos.astype(np.float64) @ M[:3, :3].astype(np.float64) + M[3, :3]).astype(np.float32)
    baked = SceneDesc()
    mt = baked.add_material("diffuse")
    baked.add_mesh(world, idx, material=mt)
    baked.set_camera(width=4, height=4)
    inst = SceneDesc()
    mt = inst.add_material("diffuse")
    st = inst.add_instance_set(pos, idx, bvh_max_prim=2)
    inst.add_instance(st, M, material=mt)
    inst.set_camera(width=4, height=4)
    rays = _rays(4000, 5, extent=2.0, tmax=50.0)
    a = cpuref.OracleScene(baked).trace_closest(rays)
    b, sub, _ = cpuref.OracleScene(inst).trace_closest_sub(rays)
    both = (a["prim"] >= 0) & (b["prim"] >= 0)
    assert both.sum() > 1000
    assert np.mean((a["prim"] >= 0) == (b["prim"] >= 0)) > 0.998
    agree = both & (a["prim"] == sub)
    assert np.mean(agree[both]) > 0.995
    assert np.allclose(a["t"][agree], b["t"][agree], rtol=1e-4)


def test_float_max_ray_overflows_into_a_miss():
    """Ray3F.Transform maps ray.At(float.MaxValue) through WorldToModel; with
    a shrinking instance (scale 0.25, rotated 45 degrees: entries +-2.83) the
    x component sums a product that overflows to +inf and one that overflows
    to -inf, MaxT becomes NaN and the set's root box test fails: the reference
    misses the instance.  The same ray with a finite MaxT hits it at 4.75."""
    from oracle import cpuref
    pos, idx, _ = scenes.superquadric(1.0, 1.0, 16, 8)
    c, s_ = np.cos(np.pi / 4), np.sin(np.pi / 4)
    R = np.array([[c, s_, 0], [-s_, c, 0], [0, 0, 1]])
    M = scenes.instance_matrix(R, 0.25, (0.0, 0.0, 0.0))
    inst = SceneDesc()
    mt = inst.add_material("diffuse")
    st = inst.add_instance_set(pos, idx)
    inst.add_instance(st, M, material=mt)
    inst.set_camera(width=4, height=4)
    o = cpuref.OracleScene(inst)
    r = np.zeros(2, abi.ray_dtype())
    d = np.array([1.0, -1.0, 0.0]) / np.sqrt(2.0)
    r["o"] = (-5.0 * d).astype(np.float32)
    r["d"] = d.astype(np.float32)
    r["tmin"] = 1e-4
    r["tmax"] = [np.finfo(np.float32).max, 100.0]
    h, sub, _ = o.trace_closest_sub(r)
    assert h["prim"][0] == -1 and h["prim"][1] == 0
    assert h["t"][1] == pytest.approx(4.75, abs=0.01)


def test_instanced_render_matches_baked_meshes(oracle):
    """The same geometry as instances (rotation, uniform scale, translation)
    and baked into world space renders the same image: both estimators draw
    the same random numbers, so most pixels agree to rounding."""
    from oracle import cpuref
    rng = np.random.default_rng(7)
    sets = [scenes.superquadric(*rng.uniform(0.6, 2.0, 2), 20, 10) for _ in range(2)]
    mats_spec = [("rough_plastic", dict(r=(0.3, 0.5, 0.8), roughness=0.3, dist="GGX")),
                 ("rough_metal", dict(roughness=0.3)), ("disney", dict(base_color=(0.8, 0.4, 0.2), roughness=0.5))]
    Ms = [scenes.instance_matrix(scenes._rand_rotation(rng), rng.uniform(0.5, 1.5), rng.uniform(-1.2, 1.2, 3))
          for _ in range(6)]

    def build(instanced):
        s = SceneDesc()
        mats = [s.add_material(k, **p) for k, p in mats_spec]
        g = s.add_material("diffuse", kd=(0.5, 0.5, 0.5))
        s.add_mesh(*scenes._quad((-6, -2, -6), (-6, -2, 6), (6, -2, 6), (6, -2, -6)), material=g)
        ids = [s.add_instance_set(p, i, uvs=u) for p, i, u in sets] if instanced else None
        for k, M in enumerate(Ms):
            p, i, u = sets[k % 2]
            if instanced:
                s.add_instance(ids[k % 2], M, material=mats[k % 3])
            else:
                w = (p.astype(np.float64) @ M[:3, :3].astype(np.float64) + M[3, :3]).astype(np.float32)
                s.add_mesh(w, i, material=mats[k % 3], uvs=u)
        black = s.add_material("diffuse", kd=(0, 0, 0))
        s.add_mesh(*scenes._quad((-1, 4, -1), (1, 4, -1), (1, 4, 1), (-1, 4, 1)), material=black, le=(12, 11, 10))
        s.set_camera(fov=50.0, width=32, height=32, origin=(2.0, 2.5, 7.0), target=(0, 0, 0), up=(0, 1, 0))
        s.set_integrator(type="path", max_depth=6, method="Mis")
        return s

    a, _, _, _ = cpuref.OracleScene(build(False)).render(16, seed=3, threads=0)
    b, _, _, _ = cpuref.OracleScene(build(True)).render(16, seed=3, threads=0)
    close = np.isclose(a, b, rtol=2e-2, atol=1e-3)
    assert close.mean() > 0.85, close.mean()
    assert abs(a.mean() - b.mean()) < 0.02 * a.mean()


def test_instanced_scene_validation():
    from oracle import cpuref
    s = SceneDesc()
    m = s.add_material("diffuse")
    pos, idx, _ = scenes.superquadric(1, 1, 8, 4)
    st = s.add_instance_set(pos, idx)
    s.add_instance(st, np.zeros((4, 4), np.float32), material=m)   # singular
    s.set_camera(width=4, height=4)
    with pytest.raises(Exception):
        cpuref.OracleScene(s)
