"""Tuning knobs that only schedule work (slot pool cap, path tile, traversal
stack variant and block size) leave images and query counts bit-identical,
and the autotuner's results are kept per device and scene class."""
import numpy as np
import pytest

from radiantium_amd import native, scenes
from radiantium_amd.native import Scene

pytestmark = pytest.mark.gpu

KEYS = ("rays_closest", "rays_any", "vertices", "invalid", "class_rays", "class_nodes", "class_prims")


@pytest.mark.parametrize("which", ["cornell", "vol", "instanced"])
def test_schedule_knobs_bit_identical(ctx, which):
    desc, spp = {"cornell": (scenes.cornell_box(70, 50, method="Mis"), 6),
                 "vol": (scenes.vol_cornell(48, 40), 5),
                 "instanced": (scenes.instanced_scene(64, 48, seed=2), 3)}[which]
    base = ctx.tuning()
    sc = Scene(ctx, desc)
    try:
        ref, ref2 = sc.render(spp, seed=3)
        rst = ctx.render_stats()
        print(f"[tuning] {which}: {rst['rays_closest'] + rst['rays_any']} rays", flush=True)
        for kw in ({"tile": 4}, {"tile": 16}, {"tile": 3}, {"tile": 1}, {"tile": 64}, {"pool_log2": 10},
                   {"pool_log2": 12, "tile": 5}, {"stack_min": 40}, {"stack_min": 48}, {"stack_min": 64},
                   {"stack_min": 32, "tile": 16}):
            ctx.set_tuning(**base)
            ctx.set_tuning(**kw)
            img, s2 = sc.render(spp, seed=3)
            st = ctx.render_stats()
            assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), kw
            assert np.array_equal(s2.view(np.uint32), ref2.view(np.uint32)), kw
            for k in KEYS:
                assert st[k] == rst[k], (kw, k)
        # a crop and a batch under a non-default tile
        ctx.set_tuning(**base)
        crop = (5, 7, 41, 33)
        c0, _ = sc.render(spp, seed=3, crop=crop)
        ctx.set_tuning(tile=4)
        c1, _ = sc.render(spp, seed=3, crop=crop)
        assert np.array_equal(c0.view(np.uint32), c1.view(np.uint32))
        outs = ctx.render_batch([(sc, spp, 3), (sc, spp + 1, 4, crop)])
        ctx.set_tuning(**base)
        outs0 = ctx.render_batch([(sc, spp, 3), (sc, spp + 1, 4, crop)])
        for (a, _), (b, _) in zip(outs, outs0):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    finally:
        ctx.set_tuning(**base)


def test_autotune_cache(ctx, tmp_path):
    base = ctx.tuning()
    sc = Scene(ctx, scenes.cornell_box(64, 64, method="Mis"))
    try:
        native.tuning_cache_clear()
        t1, found = ctx.autotune(sc, spp=2, cached=True)
        assert not found
        ctx.set_tuning(**base)
        t2, found = ctx.autotune(sc, spp=2, cached=True)
        assert found and t2 == t1
        p = tmp_path / "tune.txt"
        native.tuning_cache_save(p)
        lines = p.read_text().strip().splitlines()
        assert len(lines) == 1 and "|path/2|inst0|" in lines[0]
        native.tuning_cache_clear()
        native.tuning_cache_load(p)
        ctx.set_tuning(**base)
        t3, found = ctx.autotune(sc, spp=2, cached=True)
        assert found and t3 == t1
        # the tuned context renders the same image
        img, _ = sc.render(4, seed=1)
        ctx.set_tuning(**base)
        ref, _ = sc.render(4, seed=1)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
    finally:
        ctx.set_tuning(**base)
        native.tuning_cache_clear()
