"""GPU parity of the BSSRDF branch of the path integrator's Mis method
(PathTracer.cs:244-259, Bssrdf.cs, Materials/Subsurface.cs, the subsurface
lobe of DisneyBsdf.Sample): the HIP path through the C ABI against the
oracle with the shared counter-based streams, so most pixels agree to float
rounding; the other methods and integrators, which ignore the subsurface
flag as the reference does, on the same scenes.
"""
import numpy as np
import pytest

from radiantium_amd import abi, scenes
from radiantium_amd.native import RadiantiumError, Scene
from tests.helpers import rel_mse, z_test

pytestmark = pytest.mark.gpu


def _pair(ctx, oracle, desc, spp, seed, wave_paths=0):
    gs = Scene(ctx, desc)
    g, g2 = gs.render(spp, seed=seed, wave_paths=wave_paths)
    gst = ctx.render_stats()
    gs.close()
    o, o2, ost, _ = oracle.OracleScene(desc).render(spp, seed=seed, threads=0)
    return g, g2, gst, o, o2, ost


def _check(g, g2, gst, o, o2, ost, spp, close_frac=0.95, tol=0.02, any_rays=True):
    assert np.isfinite(g).all()
    assert gst["paths"] == ost["paths"]
    for k in ("rays_closest", "rays_any", "vertices") if any_rays else ("rays_closest", "vertices"):
        assert abs(gst[k] - ost[k]) <= max(16, tol * ost[k]), (k, gst[k], ost[k])
    close = np.isclose(g, o, rtol=1e-3, atol=1e-5)
    assert close.mean() > close_frac, close.mean()
    z = z_test(g, g2, spp, o, o2, spp)
    assert np.mean(np.abs(z) < 4) > 0.99
    assert rel_mse(g, o) < 3e-3


CASES = {
    "both": dict(),
    "disney": dict(subsurface=False),
    "subsurface": dict(disney=False),
}


@pytest.mark.parametrize("which", list(CASES))
def test_bssrdf_twin_mis(ctx, oracle, which):
    desc = scenes.sss_cornell(32, 32, **CASES[which])
    print(f"[bssrdf twin] sss_cornell[{which}] 32x32 spp=16 seed=3", flush=True)
    g, g2, gst, o, o2, ost = _pair(ctx, oracle, desc, 16, seed=3)
    _check(g, g2, gst, o, o2, ost, 16)
    # the probes are closest-hit queries of their own: more than the plain box's
    plain = scenes.cornell_box(32, 32)
    _, _, pst, _ = oracle.OracleScene(plain).render(16, seed=3, threads=0, with_sum2=False)
    assert ost["rays_closest"] > pst["rays_closest"]


def test_bssrdf_twin_octree(ctx, oracle):
    """The probe chain over the octree accelerator."""
    desc = scenes.sss_cornell(32, 32)
    desc.accel = dict(type="octree", max_depth=10, max_count=4, out_bound=2.0)
    print("[bssrdf twin] sss_cornell octree 32x32 spp=16 seed=4", flush=True)
    _check(*_pair(ctx, oracle, desc, 16, seed=4), 16)


def test_bssrdf_twin_instanced(ctx, oracle):
    """The probe chain through the two-level traversal: the instances' Disney
    material with a scattering distance (instances of one material share it,
    so a probe counts hits on every instance of that material)."""
    desc = scenes.instanced_scene(32, 32, n_instances=24, n_sets=3, max_depth=6)
    ids = [i for i, m in enumerate(desc.materials) if m["type"] == abi.MAT_DISNEY]
    assert ids
    for i in ids:
        desc.materials[i]["tex"][11] = desc._tex((0.05, 0.04, 0.03), 0.0)
    print("[bssrdf twin] instanced_scene (Disney with scattering) 32x32 spp=8 seed=5", flush=True)
    _check(*_pair(ctx, oracle, desc, 8, seed=5), 8, close_frac=0.93, tol=0.03)


@pytest.mark.parametrize("method", ["Nee", "Bxdf"])
def test_other_methods_ignore_the_flag(ctx, oracle, method):
    """Only Mis has the branch: Nee and Bxdf follow the sampled refraction."""
    desc = scenes.sss_cornell(32, 32, method=method)
    print(f"[bssrdf twin] sss_cornell {method} 32x32 spp=16 seed=6", flush=True)
    _check(*_pair(ctx, oracle, desc, 16, seed=6), 16, any_rays=method != "Bxdf")


def test_vol_path_ignores_the_flag(ctx, oracle):
    desc = scenes.sss_cornell(24, 24, max_depth=12)
    desc.set_integrator(type="vol_path", max_depth=12)
    print("[bssrdf twin] sss_cornell vol_path 24x24 spp=16 seed=7", flush=True)
    _check(*_pair(ctx, oracle, desc, 16, seed=7), 16, any_rays=False)


def test_bssrdf_schedule_and_pass_independence(ctx):
    """A path's BSSRDF stage depends on its own stream only: the host-driven
    loop, the device-driven tail (plain launches or the captured graph) and a
    small slot pool give the same image bit for bit."""
    desc = scenes.sss_cornell(48, 48)
    base = ctx.tuning()
    sc = Scene(ctx, desc)
    try:
        ctx.set_tuning(device_loop=0)
        ref, _ = sc.render(8, seed=11)
        rst = ctx.render_stats()
        print(f"[bssrdf passes] sss_cornell 48x48 spp=8: {rst['bounces']} bounces", flush=True)
        for dl, tail, wave in ((1, 1 << 30, 0), (2, 1 << 30, 0), (2, 1 << 30, 0), (2, None, 4096), (0, None, 4096)):
            ctx.set_tuning(device_loop=dl, tail_rays=tail if tail is not None else base["tail_rays"])
            img, _ = sc.render(8, seed=11, wave_paths=wave)
            assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), (dl, tail, wave)
            assert ctx.render_stats()["rays_closest"] == rst["rays_closest"]
    finally:
        ctx.set_tuning(**base)
        sc.close()


def test_bssrdf_refusals(ctx):
    """Batches, the work-area entry point and the adjoint cover scenes
    without a BSSRDF; such a scene renders alone."""
    gs = Scene(ctx, scenes.sss_cornell(8, 8))
    with pytest.raises(RadiantiumError) as e:
        ctx.render_batch([(gs, 2, 1), (gs, 2, 2)])
    assert e.value.code == abi.RD_ERR_UNSUPPORTED
    d = ctx.alloc(8 * 8 * 12)
    try:
        with pytest.raises(RadiantiumError) as e:
            gs.render_device_noalloc(2, d, seed=1)
        assert e.value.code == abi.RD_ERR_UNSUPPORTED
    finally:
        ctx.free(d)
    gs.close()
