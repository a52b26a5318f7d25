"""GPU parity of bidirectional path tracing (the reference's bdpt integrator
and renderer): the HIP path through the C ABI against the oracle with the
shared counter-based random streams (camera subpath, light subpath and the
s = 1 light samples of each sample), so most pixels agree to float rounding.
"""
import numpy as np
import pytest

from radiantium_amd import scenes
from radiantium_amd.native import Scene
from tests.helpers import rel_mse, z_test

pytestmark = pytest.mark.gpu


def _bd(desc, max_depth=7, min_depth=5):
    desc.integrator.update(type="bdpt", max_depth=max_depth, min_depth=min_depth)
    return desc


def _pair(ctx, oracle, desc, spp, seed, wave_paths=0, crop=None):
    gs = Scene(ctx, desc)
    g, g2 = gs.render(spp, seed=seed, wave_paths=wave_paths, crop=crop)
    gst = ctx.render_stats()
    gs.close()
    o, o2, ost, _ = oracle.OracleScene(desc).render_bdpt(spp, seed=seed, crop=crop, threads=0)
    return g, g2, gst, o, o2, ost


def _check(g, g2, gst, o, o2, ost, spp, close_frac=0.97, count_tol=0.01, mse=1e-3):
    assert np.isfinite(g).all()
    assert gst["paths"] == ost["paths"]
    for k in ("rays_closest", "vertices"):
        assert abs(gst[k] - ost[k]) <= max(16, count_tol * ost[k]), (k, gst[k], ost[k])
    close = np.isclose(g, o, rtol=1e-3, atol=1e-5)
    assert close.mean() > close_frac, close.mean()
    z = z_test(g, g2, spp, o, o2, spp)
    assert np.mean(np.abs(z) < 4) > 0.99
    assert rel_mse(g, o) < mse


def test_bdpt_twin_cornell(ctx, oracle):
    desc = _bd(scenes.cornell_box(32, 32))
    print("[bdpt twin] cornell 32x32 spp=16 seed=3", flush=True)
    _check(*_pair(ctx, oracle, desc, 16, seed=3), 16)


def test_bdpt_twin_closed_room(ctx, oracle):
    """Every strategy active: no path escapes, the emitter reflects."""
    desc = _bd(scenes.cornell_box(24, 24, closed=True, light_kd=0.5))
    print("[bdpt twin] closed cornell 24x24 spp=16 seed=4", flush=True)
    _check(*_pair(ctx, oracle, desc, 16, seed=4), 16)


def test_bdpt_twin_material_box(ctx, oracle):
    """Every material (importance mode on the light subpath), sphere emitters."""
    desc = _bd(scenes.material_box(32, 32), max_depth=6)
    print("[bdpt twin] material_box 32x32 spp=16 seed=5", flush=True)
    _check(*_pair(ctx, oracle, desc, 16, seed=5), 16, close_frac=0.95, count_tol=0.02, mse=2e-3)


def test_bdpt_twin_media(ctx, oracle):
    """Medium vertices (free flight, Henyey-Greenstein) on both subpaths and
    Transmittance chains across the glass sphere's medium boundary."""
    desc = _bd(scenes.vol_cornell(32, 32), max_depth=8)
    print("[bdpt twin] vol_cornell 32x32 spp=16 seed=6", flush=True)
    _check(*_pair(ctx, oracle, desc, 16, seed=6), 16, close_frac=0.95, count_tol=0.02, mse=2e-3)


def test_bdpt_twin_env(ctx, oracle):
    """Light subpaths from the environment map and s = 1 samples of it."""
    desc = _bd(scenes.env_showcase(32, 32, map_w=64, map_h=32), max_depth=6)
    print("[bdpt twin] env_showcase 32x32 spp=16 seed=7", flush=True)
    _check(*_pair(ctx, oracle, desc, 16, seed=7), 16, close_frac=0.95, count_tol=0.02, mse=2e-3)


def test_bdpt_twin_instanced(ctx, oracle):
    desc = _bd(scenes.instanced_scene(32, 32, n_instances=24, n_sets=3), max_depth=6)
    print("[bdpt twin] instanced_scene 32x32 spp=8 seed=8", flush=True)
    _check(*_pair(ctx, oracle, desc, 8, seed=8), 8, close_frac=0.95, count_tol=0.02, mse=2e-3)


def test_bdpt_pass_size_and_crop(ctx):
    """A sample's value depends on its pixel and index only, and a pixel's
    samples are summed in sample order: one sample slice per pass gives the
    same image bit for bit; a crop renders its pixels alone, identically."""
    desc = _bd(scenes.cornell_box(24, 24))
    gs = Scene(ctx, desc)
    print("[bdpt passes] cornell 24x24 spp=8: passes of one sample slice vs one pass; crop", flush=True)
    a, a2 = gs.render(8, seed=9, wave_paths=1)
    b, b2 = gs.render(8, seed=9)
    assert np.array_equal(a, b) and np.array_equal(a2, b2)
    crop = (5, 3, 17, 20)
    c, _ = gs.render(8, seed=9, crop=crop)
    img_a, img_c = a.reshape(24, 24, 3), c.reshape(24, 24, 3)
    assert np.array_equal(img_c[5:17, 3:20], img_a[5:17, 3:20])
    m64 = gs.render_mean_f64(8, seed=9)
    assert np.array_equal(m64.astype(np.float32), a)
    gs.close()
