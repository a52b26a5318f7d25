"""rd_render_device_noalloc, the render entry point that allocates no device
memory (its work buffers come from the library's fixed per-device work area),
gives rd_render_device's image and statistics bit for bit, call after call:
the work area keeps the previous call's contents, so the second call renders
a smaller input over what the first one left there.  Renders too large for
the work area are refused before any launch."""
import numpy as np
import pytest

from radiantium_amd import abi, scenes
from radiantium_amd.native import RadiantiumError, Scene

pytestmark = pytest.mark.gpu

KEYS = ("rays_closest", "rays_any", "paths", "vertices", "invalid", "class_rays", "class_nodes", "class_prims")


def _both(ctx, sc, spp, seed, crop=None, wave_paths=0):
    """(rgb, sum2) of rd_render_device, then of rd_render_device_noalloc, and both stats."""
    n = sc.width * sc.height * 3
    bufs = [ctx.alloc(n * 4) for _ in range(4)]
    try:
        zero = np.zeros(n, np.float32)
        for b in bufs:
            ctx.h2d(b, zero)
        sc.render_device(spp, bufs[0], bufs[1], seed=seed, crop=crop, wave_paths=wave_paths)
        ref_stats = ctx.render_stats()
        sc.render_device_noalloc(spp, bufs[2], bufs[3], seed=seed, crop=crop, wave_paths=wave_paths)
        stats = ctx.render_stats()
        out = [ctx.d2h(np.empty(n, np.float32), b) for b in bufs]
    finally:
        for b in bufs:
            ctx.free(b)
    return out, ref_stats, stats


def _check(case, out, ref_stats, stats):
    rgb, s2, rgb_na, s2_na = out
    assert np.isfinite(rgb).all() and rgb.mean() > 0, case
    assert np.array_equal(rgb.view(np.uint32), rgb_na.view(np.uint32)), case
    assert np.array_equal(s2.view(np.uint32), s2_na.view(np.uint32)), case
    for k in KEYS:
        assert stats[k] == ref_stats[k], (case, k, stats[k], ref_stats[k])


def test_noalloc_equals_render_device_call_after_call(ctx):
    big = Scene(ctx, scenes.cornell_box(96, 96))
    vol = Scene(ctx, scenes.vol_cornell(48, 48, max_depth=24))
    try:
        print("[noalloc] cornell 96x96 spp 8 seed 11 (full frame)", flush=True)
        _check("cornell full", *_both(ctx, big, 8, 11))
        # smaller input: a crop at fewer samples, over a pool of 1024 slots
        # refilled many times, on the work area the first call left behind
        print("[noalloc] cornell crop (10,20)-(50,44) spp 3 seed 12, pool 1024", flush=True)
        _check("cornell crop", *_both(ctx, big, 3, 12, crop=(10, 20, 50, 44), wave_paths=1024))
        print("[noalloc] vol_cornell 48x48 spp 4 seed 5 (vol_path arrays)", flush=True)
        _check("vol_cornell", *_both(ctx, vol, 4, 5))
    finally:
        big.close()
        vol.close()


def test_noalloc_refuses_render_beyond_work_area(ctx):
    sc = Scene(ctx, scenes.cornell_box(2048, 2048))
    n = sc.width * sc.height * 3
    d = ctx.alloc(n * 4)
    try:
        print("[noalloc] cornell 2048x2048: refused before any launch", flush=True)
        with pytest.raises(RadiantiumError) as e:
            sc.render_device_noalloc(1, d, None, seed=0)
        assert e.value.code == abi.RD_ERR_INVALID and "work area" in str(e.value)
        # the context still renders afterwards
        small = Scene(ctx, scenes.cornell_box(32, 32))
        try:
            _check("after refusal", *_both(ctx, small, 2, 1))
        finally:
            small.close()
    finally:
        ctx.free(d)
        sc.close()
