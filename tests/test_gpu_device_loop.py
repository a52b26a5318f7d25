"""The device-driven tail of a pass (rd_tuning.device_loop: the bounces once
every path has started and the queue is small run with their queue sizes read
on the device, as plain launches or as a replayed hipGraph of two bounces) and
the per-lane vol_path chains (every Transmittance / IntersectTr chain walked
to its end by one lane, instead of one segment per launch with the host
looping) give the host-driven loop's images and query counts bit for bit.
Renders follow one another, so a tail graph is captured, kept on the context
and replayed many times with the host's own synchronizations and other
renders between replays, and the renders check every one of them."""
import numpy as np
import pytest

from radiantium_amd import scenes
from radiantium_amd.native import Scene

pytestmark = pytest.mark.gpu

KEYS = ("rays_closest", "rays_any", "paths", "vertices", "invalid", "class_rays", "class_nodes", "class_prims",
        "bounces")

CASES = {
    "cornell_mis": lambda: (scenes.cornell_box(96, 96, method="Mis"), 8),
    "cornell_nee": lambda: (scenes.cornell_box(80, 80, method="Nee"), 8),
    "cornell_bxdf": lambda: (scenes.cornell_box(80, 80, method="Bxdf"), 8),
    "cornell_deep": lambda: (scenes.cornell_box(64, 64, max_depth=40, method="Mis"), 8),
    "cornell_deep_nee": lambda: (scenes.cornell_box(64, 64, max_depth=40, method="Nee"), 8),
    "material_box": lambda: (scenes.material_box(64, 64), 8),
    "vol_cornell": lambda: (scenes.vol_cornell(64, 64, max_depth=64, sigma_s=4.0, g=0.8), 8),
    "vol_fog": lambda: (scenes.vol_fog(64, 64, max_depth=24), 8),
    "instanced": lambda: (scenes.instanced_scene(96, 96, seed=3), 4),
}


@pytest.mark.parametrize("which", list(CASES))
def test_device_loop_bit_identical(ctx, which):
    desc, spp = CASES[which]()
    base = ctx.tuning()
    sc = Scene(ctx, desc)
    try:
        # the host-driven loop with lockstep chains: one host sync per bounce and
        # chain step.  "bounces" counts the bounce loop's iterations of the
        # longest pass, which depend on the slot pool (wave_paths), so each
        # pool size has its own reference.
        refs = {}
        for wave in (0, 4096):
            ctx.set_tuning(device_loop=0, chain_lockstep=1)
            refs[wave] = (sc.render(spp, seed=13, wave_paths=wave)[0], ctx.render_stats())
        ref = refs[0][0]
        assert np.array_equal(refs[4096][0].view(np.uint32), ref.view(np.uint32))
        rst = refs[0][1]
        print(f"[device loop] {which}: {rst['bounces']} bounces, {rst['rays_closest'] + rst['rays_any']} rays",
              flush=True)
        for dl, lock, tail, wave in ((0, 0, None, 0), (1, 0, None, 0), (2, 0, None, 0), (2, 0, None, 0),
                                     (2, 0, 1 << 30, 0), (2, 0, 1 << 30, 4096), (2, 0, 1 << 30, 4096),
                                     (2, 1, None, 0), (1, 0, 1 << 30, 0), (2, 0, None, 0)):
            ctx.set_tuning(device_loop=dl, chain_lockstep=lock, tail_rays=tail if tail is not None else base["tail_rays"])
            img, _ = sc.render(spp, seed=13, wave_paths=wave)
            st = ctx.render_stats()
            case = (which, dl, lock, tail, wave)
            assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), case
            for k in KEYS:
                assert st[k] == refs[wave][1][k], (case, k, st[k], refs[wave][1][k])
    finally:
        ctx.set_tuning(**base)


def test_device_loop_graph_flag_and_batch(ctx):
    """RD_RENDER_NO_GRAPH launches the tail kernel by kernel, and a batch of
    requests (per-path scenes) through the tail equals the host-driven one."""
    from radiantium_amd import abi
    import ctypes as C
    from radiantium_amd.native import lib, check
    base = ctx.tuning()
    descs = [(scenes.cornell_box(48, 40, method="Mis"), 6, 3), (scenes.vol_cornell(40, 40), 5, 4),
             (scenes.cornell_box(32, 32, method="Nee"), 7, 5), (scenes.vol_fog(36, 36, max_depth=16), 4, 6)]
    sc = [Scene(ctx, d) for d, _, _ in descs]
    try:
        ctx.set_tuning(device_loop=0, chain_lockstep=1)
        ref = ctx.render_batch([(s_, spp, seed) for s_, (d, spp, seed) in zip(sc, descs)])
        ctx.set_tuning(device_loop=2, chain_lockstep=0)
        for _ in range(2):
            outs = ctx.render_batch([(s_, spp, seed) for s_, (d, spp, seed) in zip(sc, descs)])
            for (g, g2), (o, o2) in zip(outs, ref):
                assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
                assert np.array_equal(g2.view(np.uint32), o2.view(np.uint32))
        # single render with the no-graph flag
        s0 = sc[0]
        p = abi.rd_render_params()
        p.spp, p.seed, p.flags = 6, 3, abi.RENDER_NO_GRAPH
        n = s0.width * s0.height * 3
        rgb = np.zeros(n, np.float32)
        check(lib().rd_render(ctx.h, s0.h, C.byref(p), rgb.ctypes.data_as(C.c_void_p), None, None, None))
        assert np.array_equal(rgb.view(np.uint32), ref[0][0].view(np.uint32))
    finally:
        ctx.set_tuning(**base)
