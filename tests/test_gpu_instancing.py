"""GPU parity of instancing (InstancedPrimitive, Primitive.cs:144-189): the
two-level traversal bit for bit against the oracle's nested Bvh.Intersect
(primitive, set triangle, t, u, v and the node / primitive visit counts),
the reference's NaN-MaxT miss for float.MaxValue rays, instanced renders
(twin and independent-seed), and a batch mixing instanced and plain scenes."""
import numpy as np
import pytest

from radiantium_amd import abi, scenes
from radiantium_amd.native import Scene
from radiantium_amd.scene import SceneDesc
from tests.helpers import z_test
from tests.test_gpu_parity import _assert_hits_equal, _render_pair, _trace_both

pytestmark = pytest.mark.gpu

FLT_MAX = np.finfo(np.float32).max


def _rays(n, seed, extent=2.5, tmax=100.0):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, abi.ray_dtype())
    o = rng.uniform(-2.5 * extent, 2.5 * extent, (n, 3))
    d = rng.uniform(-extent, extent, (n, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r["o"] = o.astype(np.float32)
    r["d"] = d.astype(np.float32)
    r["tmin"] = 1e-4
    r["tmax"] = tmax
    return r


@pytest.mark.parametrize("nonuniform", [False, True])
def test_trace_parity_instanced(ctx, oracle, nonuniform):
    desc = scenes.instanced_scene(n_instances=40, n_sets=3, nu=20, nv=10, seed=3, env=None,
                                  nonuniform=nonuniform)
    rays = np.concatenate([_rays(20000, 1), _rays(10000, 2, tmax=FLT_MAX)])
    print(f"[instanced trace] nonuniform={nonuniform} rays={rays.shape[0]}", flush=True)
    gs, os_, gh, oh, go, oo, (gc, oc), (gac, oac), gh2 = _trace_both(ctx, oracle, desc, rays)
    _assert_hits_equal(gh, oh)
    _assert_hits_equal(gh2, oh)
    assert np.array_equal(go, oo)
    assert gc.tolist() == oc.tolist() and gac.tolist() == oac.tolist()
    g, gsub = gs.trace_closest_sub(rays)
    o, osub, _ = os_.trace_closest_sub(rays)
    _assert_hits_equal(g, o)
    assert np.array_equal(gsub, osub)
    assert (osub >= 0).sum() > 5000 and ((o["prim"] >= 0) & (osub < 0)).sum() > 500


def test_float_max_ray_misses_shrunk_instance(ctx, oracle):
    """The reference's Ray3F.Transform of ray.At(float.MaxValue) overflows to
    +inf and -inf in one component for a 0.25-scaled, 45-degree instance:
    MaxT is NaN and the instance is missed; a finite MaxT hits it."""
    pos, idx, _ = scenes.superquadric(1.0, 1.0, 16, 8)
    c, s_ = np.cos(np.pi / 4), np.sin(np.pi / 4)
    M = scenes.instance_matrix(np.array([[c, s_, 0], [-s_, c, 0], [0, 0, 1]]), 0.25, (0, 0, 0))
    desc = SceneDesc()
    mt = desc.add_material("diffuse")
    desc.add_instance(desc.add_instance_set(pos, idx), M, material=mt)
    desc.set_camera(width=4, height=4)
    r = np.zeros(2, abi.ray_dtype())
    d = np.array([1.0, -1.0, 0.0]) / np.sqrt(2.0)
    r["o"] = (-5.0 * d).astype(np.float32)
    r["d"] = d.astype(np.float32)
    r["tmin"] = 1e-4
    r["tmax"] = [FLT_MAX, 100.0]
    g, gsub = Scene(ctx, desc).trace_closest_sub(r)
    o, osub, _ = oracle.OracleScene(desc).trace_closest_sub(r)
    assert g["prim"].tolist() == [-1, 0] and o["prim"].tolist() == [-1, 0]
    _assert_hits_equal(g, o)
    assert np.array_equal(gsub, osub)


@pytest.mark.parametrize("method,integrator", [("Mis", "path"), ("Nee", "path"), ("Mis", "vol_path")])
def test_render_twin_instanced(ctx, oracle, method, integrator):
    """Instances of three superquadrics (Disney / plastic / metal), a ground
    quad, an area light and a rotated HDR sky."""
    desc = scenes.instanced_scene(32, 32, n_instances=24, n_sets=3, nu=16, nv=8, seed=5, method=method,
                                  env=(64, 32), integrator=integrator, max_depth=8)
    print(f"[render twin] instanced 32x32 spp=8 method={method} integrator={integrator}", flush=True)
    g, g2, gst, o, o2, ost = _render_pair(ctx, oracle, desc, 8, seed=17)
    assert np.isfinite(g).all()
    for k in ("rays_closest", "rays_any", "vertices"):
        assert abs(gst[k] - ost[k]) <= max(16, 0.02 * ost[k]), (k, gst[k], ost[k])
    close = np.isclose(g, o, rtol=1e-3, atol=1e-5)
    print(f"[render twin] close {close.mean():.4f}", flush=True)
    assert close.mean() > 0.95, close.mean()
    z = z_test(g, g2, 8, o, o2, 8)
    assert np.mean(np.abs(z) < 4) > 0.99


def test_render_instanced_statistical(ctx, oracle):
    desc = scenes.instanced_scene(24, 24, n_instances=24, n_sets=3, nu=16, nv=8, seed=6, env=(64, 32))
    g, g2 = Scene(ctx, desc).render(64, seed=100)
    o, o2, _, _ = oracle.OracleScene(desc).render(64, seed=200, threads=0)
    z = z_test(g, g2, 64, o, o2, 64)
    assert np.mean(np.abs(z) > 3) < 0.02


def test_batch_mixes_instanced_and_plain_scenes(ctx):
    descs = [(scenes.instanced_scene(20, 16, n_instances=12, nu=12, nv=6, seed=8, env=(32, 16)), 4, 3),
             (scenes.cornell_box(16, 16), 5, 4),
             (scenes.instanced_scene(16, 16, n_instances=8, nu=12, nv=6, seed=9, env=None), 3, 5)]
    sc = [Scene(ctx, d) for d, _, _ in descs]
    singles = [s_.render(spp, seed=seed) for s_, (_, spp, seed) in zip(sc, descs)]
    outs = ctx.render_batch([(s_, spp, seed) for s_, (_, spp, seed) in zip(sc, descs)], wave_paths=700)
    for (g, g2), (o, o2) in zip(outs, singles):
        assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
        assert np.array_equal(g2.view(np.uint32), o2.view(np.uint32))


def test_trace_parity_full_instanced_workload(ctx, oracle):
    """The instanced benchmark geometry at full size (1024 instances of 8
    superquadrics of 99,904 triangles): bit-exact hits and visit counts on
    rays through the instance cloud, finite and float.MaxValue MaxT."""
    desc = scenes.instanced_workload(width=64, height=64)
    rays = np.concatenate([_rays(12000, 11, extent=9.0, tmax=1000.0), _rays(6000, 12, extent=9.0, tmax=FLT_MAX)])
    print(f"[instanced trace] full workload rays={rays.shape[0]}", flush=True)
    gs, os_, gh, oh, go, oo, (gc, oc), (gac, oac), _ = _trace_both(ctx, oracle, desc, rays)
    _assert_hits_equal(gh, oh)
    assert np.array_equal(go, oo)
    assert gc.tolist() == oc.tolist() and gac.tolist() == oac.tolist()
    g, gsub = gs.trace_closest_sub(rays)
    o, osub, _ = os_.trace_closest_sub(rays)
    assert np.array_equal(gsub, osub)
    assert (osub >= 0).sum() > rays.shape[0] // 4
