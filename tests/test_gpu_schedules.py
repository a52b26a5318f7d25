"""Traversal schedules, ray ordering and the slot pool size only decide
which lane traces which ray and how many paths are in flight, so renders are
bit-identical under every rd_tuning setting and every wave_paths.  The frames
here keep more rays in flight than the chip has lanes (about 393K), so the
work-cursor schedules run through many chunks per wavefront, and the ranked
shadow / MIS queues hold runs of holes longer than a wavefront."""
import numpy as np
import pytest

from radiantium_amd import scenes
from radiantium_amd.native import RadiantiumError, Scene

pytestmark = pytest.mark.gpu

KEYS = ("rays_closest", "rays_any", "paths", "vertices", "invalid", "class_rays", "class_nodes", "class_prims")


def _render(ctx, sc, spp, **tuning):
    ctx.set_tuning(**tuning)
    img, _ = sc.render(spp, seed=3)
    return img, ctx.render_stats()


@pytest.mark.parametrize("which", ["icosphere_mis", "fog_vol"])
def test_schedules_bit_identical(ctx, which):
    if which == "icosphere_mis":
        desc, spp = scenes.displaced_icosphere_scene(level=5, width=1024, height=576, max_depth=8), 4
    else:
        desc, spp = scenes.vol_fog(512, 512, max_depth=16), 8
    base = ctx.tuning()
    sc = Scene(ctx, desc)
    try:
        ref, rst = _render(ctx, sc, spp, trace_schedule=(0, 0, 0), ray_sort=1, ray_sort_shadow=1, overlap=1)
        print(f"[schedules] {which}: class rays {rst['class_rays']}", flush=True)
        assert rst["paths"] > 2_000_000 and max(rst["class_rays"]) > 1_000_000
        for sched in ((1, 1, 1), (2, 2, 2), (2, 2, 1), (0, 1, 2), (3, 3, 3), (3, 2, 0)):
            img, st = _render(ctx, sc, spp, trace_schedule=sched)
            print(f"[schedules] {which}: {sched}", flush=True)
            assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), sched
            for k in KEYS:
                assert st[k] == rst[k], (sched, k, st[k], rst[k])
        # no overlap of a bounce's shadow / MIS stage with the next bounce
        img, st = _render(ctx, sc, spp, trace_schedule=(2, 2, 2), overlap=0)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
        for k in KEYS:
            assert st[k] == rst[k], ("no overlap", k)
        ctx.set_tuning(overlap=1)
        # queue order instead of the coherence order: same rays, same image
        img, st = _render(ctx, sc, spp, trace_schedule=(2, 2, 2), ray_sort=0)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
        for k in KEYS:
            assert st[k] == rst[k], ("unsorted", k)
        # sorted extension rays, shadow rays in the queue's own order
        img, st = _render(ctx, sc, spp, trace_schedule=(3, 2, 2), ray_sort=1, ray_sort_shadow=0)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
        for k in KEYS:
            assert st[k] == rst[k], ("shadow in queue order", k)
    finally:
        ctx.set_tuning(**base)


def test_tuning_roundtrip_and_validation(ctx):
    base = ctx.tuning()
    try:
        ctx.set_tuning(trace_schedule=(1, None, 0), ray_sort_bits=12)
        t = ctx.tuning()
        assert t["trace_schedule"] == [1, base["trace_schedule"][1], 0] and t["ray_sort_bits"] == 12
        for bad in (dict(trace_schedule=(4, 0, 0)), dict(ray_sort_bits=31), dict(ray_sort_bits=2),
                    dict(ray_sort_min=-2)):
            with pytest.raises(RadiantiumError):
                ctx.set_tuning(**bad)
        assert ctx.tuning() == t   # a rejected call changes nothing
    finally:
        ctx.set_tuning(**base)
    assert ctx.tuning() == base


@pytest.mark.parametrize("which", ["icosphere_mis", "fog_vol"])
def test_pool_size_bit_identical(ctx, which):
    """Small pools (many waves of paths, new paths started in freed slots
    every bounce) and the default pool (every path of the render at once)
    give the same image and the same query counts."""
    if which == "icosphere_mis":
        desc, spp = scenes.displaced_icosphere_scene(level=5, width=512, height=288, max_depth=8), 4
    else:
        desc, spp = scenes.vol_fog(256, 256, max_depth=16), 8
    sc = Scene(ctx, desc)
    ref, _ = sc.render(spp, seed=9)
    rst = ctx.render_stats()
    for wp in (1 << 16, 100_000, 1 << 19):
        img, _ = sc.render(spp, seed=9, wave_paths=wp)
        st = ctx.render_stats()
        print(f"[pool] {which}: wave_paths {wp}: {st['waves']} waves", flush=True)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), wp
        for k in KEYS:
            assert st[k] == rst[k], (wp, k, st[k], rst[k])


def test_autotune(ctx):
    """rd_autotune leaves a valid tuning on the context; renders after it are
    the same as before it."""
    base = ctx.tuning()
    desc = scenes.displaced_icosphere_scene(level=5, width=512, height=288, max_depth=8)
    sc = Scene(ctx, desc)
    try:
        ref, _ = sc.render(4, seed=11)
        t = ctx.autotune(sc, spp=2)
        print(f"[autotune] {t}", flush=True)
        assert all(v in (0, 1, 2, 3) for v in t["trace_schedule"])
        assert t["ray_sort"] in (0, 1) and 3 <= t["ray_sort_bits"] <= 30
        img, _ = sc.render(4, seed=11)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
    finally:
        ctx.set_tuning(**base)


def test_instanced_schedules_bit_identical(ctx):
    """The two-level traversal under every schedule (sequences, chunks,
    grid-stride) gives the same image and query counts; more rays than the
    chip has lanes."""
    base = ctx.tuning()
    desc = scenes.instanced_scene(448, 448, n_instances=96, n_sets=3, nu=64, nv=32, seed=2, max_depth=8)
    sc = Scene(ctx, desc)
    try:
        ref, rst = _render(ctx, sc, 4, trace_schedule=(0, 0, 0))
        print(f"[schedules] instanced: class rays {rst['class_rays']}", flush=True)
        assert rst["paths"] > 700_000
        for sched in ((3, 2, 2), (2, 2, 2), (3, 3, 3), (1, 1, 1), (2, 3, 3)):
            img, st = _render(ctx, sc, 4, trace_schedule=sched)
            print(f"[schedules] instanced: {sched}", flush=True)
            assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), sched
            for k in KEYS:
                assert st[k] == rst[k], (sched, k, st[k], rst[k])
    finally:
        ctx.set_tuning(**base)
