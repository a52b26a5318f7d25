"""The device BVH build (rd_bvh_build_device, rd_bvh_gpu.hip) against the
host builder (rd_bvh_build_host, itself checked node for node against the
oracle's restatement of Bvh.RecursiveBuild + FlattenBVHTree): the same
records byte for byte, the same primitive order and depth, on random bounds
with ties (equal centroids, +0 / -0 coordinates), the 1.31 M-triangle
displaced icosphere and an instanced set; scenes built either way render the
same image."""
import time

import numpy as np
import pytest

from radiantium_amd import native, scenes
from radiantium_amd.native import Scene

pytestmark = pytest.mark.gpu


def tri_bounds(pos, idx):
    p = pos[idx]
    return p.min(1).astype(np.float32), p.max(1).astype(np.float32)


def compare(ctx, lo, hi, max_prim, label):
    t0 = time.perf_counter()
    hn, ho, hd = native.bvh_build_host(lo, hi, max_prim)
    t1 = time.perf_counter()
    dn, do, dd = native.bvh_build_device(ctx, lo, hi, max_prim)
    t2 = time.perf_counter()
    print(f"[bvh] {label}: n={len(lo)} max_prim={max_prim} nodes={len(hn)} depth={hd} "
          f"host {1e3 * (t1 - t0):.1f} ms, device {1e3 * (t2 - t1):.1f} ms (with transfers)", flush=True)
    assert dn.shape == hn.shape, label
    assert np.array_equal(dn, hn), (label, int(np.argmax(np.any(dn != hn, axis=1))))
    assert np.array_equal(do, ho), label
    assert dd == hd, label


@pytest.mark.parametrize("n,max_prim,seed", [(1, 1, 0), (2, 1, 1), (3, 4, 2), (7, 1, 3), (100, 4, 4), (5000, 1, 5),
                                             (5000, 8, 6), (70000, 4, 7)])
def test_random_bounds(ctx, n, max_prim, seed):
    rng = np.random.default_rng(seed)
    c = rng.normal(0, 1, (n, 3)).astype(np.float32)
    c[rng.random(n) < 0.1] = 0.0                         # centroids at the origin
    c[rng.random(n) < 0.05] *= -0.0                      # and signed zeros
    k = max(1, n // 10)
    c[:k] = c[0]                                         # a cluster of equal centroids
    e = np.abs(rng.normal(0, 0.05, (n, 3))).astype(np.float32)
    e[rng.random(n) < 0.2] = 0.0                         # flat boxes
    lo, hi = c - e, c + e
    lo[rng.random((n, 3)) < 0.05] = -0.0
    hi[rng.random((n, 3)) < 0.05] = 0.0
    compare(ctx, lo, hi, max_prim, f"random {n}")


def test_all_equal_and_grid(ctx):
    n = 4096
    lo = np.zeros((n, 3), np.float32)
    hi = np.ones((n, 3), np.float32)
    compare(ctx, lo, hi, 1, "all equal")                 # one leaf of 4096 (equal centroids)
    g = np.stack(np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij"), -1).reshape(-1, 3)
    compare(ctx, g.astype(np.float32), (g + 1).astype(np.float32), 2, "grid")   # many equal SAH costs


def test_icosphere_level8(ctx):
    desc = scenes.displaced_icosphere_scene(level=8, width=64, height=36)
    a = desc.arrays()
    tris = [index for kind, index, *_ in a["prims"] if kind == 0]
    lo, hi = tri_bounds(a["positions"], a["indices"][np.array(tris)])
    compare(ctx, lo, hi, 4, "icosphere level 8")


def test_instanced_set(ctx):
    pos, idx, _ = scenes.superquadric(0.7, 1.8, nu=160, nv=80)
    lo, hi = tri_bounds(np.asarray(pos, np.float32), np.asarray(idx).reshape(-1, 3))
    compare(ctx, lo, hi, 4, "superquadric set")


@pytest.mark.parametrize("which", ["cornell", "instanced"])
def test_scene_builds_render_the_same(ctx, which):
    desc = scenes.cornell_box(48, 48) if which == "cornell" else scenes.instanced_scene(48, 40, seed=4)
    base = ctx.tuning()
    try:
        ctx.set_tuning(bvh_build=0)
        s0 = Scene(ctx, desc)
        ctx.set_tuning(bvh_build=2)
        s1 = Scene(ctx, desc)
        assert s0.info()["num_nodes"] == s1.info()["num_nodes"]
        assert np.array_equal(s0.export_bvh()[0], s1.export_bvh()[0])
        a, _ = s0.render(4, seed=2)
        b, _ = s1.render(4, seed=2)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    finally:
        ctx.set_tuning(**base)
