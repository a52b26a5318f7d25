"""GPU parity of light tracing (the reference's adjoint_particle integrator
and renderer): the HIP path through the C ABI against the oracle with the
shared counter-based random stream (particle i draws from
particle_key(seed, i) on both sides), so most pixels agree to float rounding;
the GPU adds splats with float64 atomics, the oracle per worker in float64.
"""
import numpy as np
import pytest

from radiantium_amd import scenes
from radiantium_amd.native import Scene, RadiantiumError
from tests.helpers import rel_mse

pytestmark = pytest.mark.gpu


def _particle_desc(desc, max_depth=-1, min_depth=3):
    desc.integrator.update(type="adjoint_particle", max_depth=max_depth, min_depth=min_depth)
    return desc


def _pair(ctx, oracle, desc, particles, seed, wave_paths=0):
    gs = Scene(ctx, desc)
    g = gs.render_particles(particles, seed=seed, wave_paths=wave_paths)
    gst = ctx.render_stats()
    gs.close()
    o, ost, _ = oracle.OracleScene(desc).render_particles(particles, seed=seed, threads=0)
    return g, gst, o, ost


def _box3(a, w, h):
    """Sum over each pixel's 3x3 neighbourhood."""
    a = a.reshape(w, h, 3)
    p = np.pad(a, ((1, 1), (1, 1), (0, 0)))
    return sum(p[1 + dx:1 + dx + w, 1 + dy:1 + dy + h] for dx in (-1, 0, 1) for dy in (-1, 0, 1)).reshape(-1)


def _check_twin(g, gst, o, ost, close_frac, count_tol, w, h):
    assert np.isfinite(g).all()
    assert gst["paths"] == ost["paths"]
    for k in ("rays_closest", "rays_any", "vertices"):
        assert abs(gst[k] - ost[k]) <= max(16, count_tol * ost[k]), (k, gst[k], ost[k])
    close = np.isclose(g, o, rtol=1e-3, atol=1e-6)
    assert close.mean() > close_frac, close.mean()
    # a splat lands at floor(screen position): a last-bit difference next to a
    # pixel edge moves a whole (possibly bright) splat one pixel over, so the
    # images are compared through a 3x3 box filter, and by their totals
    assert rel_mse(_box3(g, w, h), _box3(o, w, h)) < 1e-3
    assert abs(float(g.sum()) - float(o.sum())) <= 1e-3 * float(o.sum())


def test_particles_twin_cornell(ctx, oracle):
    desc = _particle_desc(scenes.cornell_box(48, 48))
    print("[light twin] cornell 48x48, 300000 particles, seed 5", flush=True)
    g, gst, o, ost = _pair(ctx, oracle, desc, 300000, seed=5)
    _check_twin(g, gst, o, ost, 0.97, 0.01, 48, 48)


def test_particles_twin_material_box(ctx, oracle):
    """Every material in importance mode (rough and perfect glass without the
    radiance scaling of refraction), sphere emitters, a depth cap."""
    desc = _particle_desc(scenes.material_box(48, 48), max_depth=8)
    print("[light twin] material_box 48x48, 300000 particles, seed 9", flush=True)
    g, gst, o, ost = _pair(ctx, oracle, desc, 300000, seed=9)
    _check_twin(g, gst, o, ost, 0.95, 0.02, 48, 48)


def test_particles_twin_env(ctx, oracle):
    """Emission from the environment map (InfiniteAreaLight.SampleEmit: a disk
    of the scene's bounding sphere) beside an area light."""
    desc = _particle_desc(scenes.env_showcase(48, 48, map_w=64, map_h=32), max_depth=8)
    print("[light twin] env_showcase 48x48, 300000 particles, seed 11", flush=True)
    g, gst, o, ost = _pair(ctx, oracle, desc, 300000, seed=11)
    _check_twin(g, gst, o, ost, 0.95, 0.02, 48, 48)


def test_particles_twin_instanced(ctx, oracle):
    desc = _particle_desc(scenes.instanced_scene(40, 40, n_instances=24, n_sets=3), max_depth=8)
    print("[light twin] instanced_scene 40x40, 200000 particles, seed 2", flush=True)
    g, gst, o, ost = _pair(ctx, oracle, desc, 200000, seed=2)
    _check_twin(g, gst, o, ost, 0.95, 0.02, 40, 40)


def test_particles_wave_size_independent(ctx):
    """Particle i's random numbers depend on i alone: waves of 1000 particles
    give the image of one wave (up to the order of the float64 sums)."""
    desc = _particle_desc(scenes.cornell_box(32, 32))
    gs = Scene(ctx, desc)
    print("[light waves] cornell 32x32, 50000 particles, waves of 1000 vs one wave", flush=True)
    a = gs.render_particles(50000, seed=1, wave_paths=1000)
    b = gs.render_particles(50000, seed=1)
    gs.close()
    assert np.allclose(a, b, rtol=1e-6, atol=1e-9)


def test_particles_rejects_crop_and_sum2(ctx):
    desc = _particle_desc(scenes.cornell_box(16, 16))
    gs = Scene(ctx, desc)
    with pytest.raises(RadiantiumError):
        gs.render(4, crop=(0, 0, 8, 8), with_sum2=False)
    with pytest.raises(RadiantiumError):
        gs.render(4)   # asks for per-pixel sums
    gs.close()
