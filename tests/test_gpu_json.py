"""GPU renders of scene configs read by the native loader: a config in the
reference's schema (OBJ walls, a rotated PNG-textured sphere whose UV goes
through its WorldToModel, media, an instanced entry, area and infinite
lights) rendered through rd_scene_load_json and through rd_json_load +
rd_scene_create (bit-identical), and twin-checked against the oracle fed the
same loaded description."""
import json

import numpy as np
import pytest

from radiantium_amd.native import JsonScene, Scene
from tests.helpers import z_test
from tests.test_json_loader import _png, _write

pytestmark = pytest.mark.gpu

WALLS = "\n".join(["o box",
                   "v -1 0 -1", "v 1 0 -1", "v 1 2 -1", "v -1 2 -1", "v -1 0 1", "v 1 0 1", "v 1 2 1", "v -1 2 1",
                   "vt 0 0", "vt 1 0", "vt 1 1", "vt 0 1",
                   "f 1/1 5/2 6/3", "f 1/1 6/3 2/4",      # floor
                   "f 1/1 2/2 3/3", "f 1/1 3/3 4/4",      # back
                   "f 4/1 3/2 7/3", "f 4/1 7/3 8/4",      # ceiling
                   "f 1/1 4/2 8/3", "f 1/1 8/3 5/4", ""])  # left
RIGHT = "\n".join(["v 1 0 -1", "v 1 0 1", "v 1 2 1", "v 1 2 -1", "f 1 4 3", "f 1 3 2", ""])
LIGHT = "\n".join(["v -0.3 1.98 -0.3", "v 0.3 1.98 -0.3", "v 0.3 1.98 0.3", "v -0.3 1.98 0.3",
                   "f 1 2 3", "f 1 3 4", ""])
TET = "\n".join(["v 0 0 0", "v 0.3 0 0", "v 0 0.3 0", "v 0 0 0.3", "f 1 3 2", "f 1 2 4", "f 1 4 3", "f 2 3 4", ""])


def _config(tmp_path, integrator):
    for name, text in (("walls.obj", WALLS), ("right.obj", RIGHT), ("light.obj", LIGHT), ("tet.obj", TET)):
        _write(tmp_path / name, text)
    rng = np.random.default_rng(3)
    _png(tmp_path / "tex.png", rng.integers(0, 256, (16, 32, 3)), 2, 8)
    hom = {"type": "homogeneous", "sigma_a": [0.05, 0.05, 0.05], "sigma_s": [1.5, 1.0, 0.5], "g": 0.4}
    cfg = {
        "renderer": {"spp": 8},
        "integrator": integrator,
        "camera": {"type": "perspective", "screen": [40, 32], "fov": 45, "origin": [0, 1, 3.6], "target": [0, 1, 0],
                   "up": [0, 1, 0]},
        "accel": {"type": "bvh", "max_prim": 2},
        "assets": [{"type": "model", "name": "walls", "location": "walls.obj"},
                   {"type": "model", "name": "right", "location": "right.obj"},
                   {"type": "model", "name": "light", "location": "light.obj"},
                   {"type": "model", "name": "tet", "location": "tet.obj"},
                   {"type": "image", "name": "tex", "location": "tex.png"}],
        "instanced": [{"shape": {"type": "model", "name": "tet"}, "accel": {"type": "bvh", "max_prim": 1}}],
        "scene": [
            {"material": {"type": "diffuse", "kd": [0.7, 0.7, 0.7]}, "shape": {"type": "model", "name": "walls"}},
            {"material": {"type": "rough_plastic", "r": [0.2, 0.6, 0.3], "dist": "GGX"},
             "shape": {"type": "model", "name": "right"}},
            {"material": {"type": "diffuse", "kd": 0}, "shape": {"type": "model", "name": "light"},
             "light": {"type": "diffuse_area", "le": [12, 11, 9]}},
            {"material": {"type": "diffuse", "kd": {"type": "image", "name": "tex", "sampler": {"filter": "Linear"}}},
             "shape": {"type": "sphere", "radius": 0.3, "center": [0.3, 0, 0]},
             "transform": {"position": [0, 0.35, 0], "rotation": {"axis": [0, 1, 0], "angle": 70}}},
            {"material": {"type": "perfect_glass"}, "shape": {"type": "sphere", "radius": 0.25, "center": [-0.45, 0.3, 0.3]},
             "medium": {"inside": hom}},
            {"material": {"type": "rough_metal", "metal_type": "Au", "roughness": 0.2},
             "transform": {"position": [0.2, 1.2, -0.3], "rotation": {"axis": [1, 0, 0], "angle": 30}},
             "children": [{"material": {"type": "disney", "base_color": [0.8, 0.3, 0.2], "roughness": 0.4},
                           "instanced": {"id": 0}, "transform": {"position": [-0.8, 0, 0]}}],
             "instanced": {"id": 0}},
            {"light": {"type": "infinite", "le": [0.2, 0.25, 0.3]}},
        ],
    }
    p = tmp_path / "scene.json"
    _write(p, json.dumps(cfg))
    return str(p)


@pytest.mark.parametrize("integrator", [{"type": "path", "max_depth": 6, "method": "Mis"},
                                        {"type": "path", "max_depth": 6, "method": "Nee"},
                                        {"type": "vol_path", "max_depth": 8}])
def test_json_scene_twin(ctx, oracle, tmp_path, integrator):
    path = _config(tmp_path, integrator)
    print(f"[json scene] {integrator}", flush=True)
    js = JsonScene(path)
    assert js.spp == 8
    gs = Scene(ctx, js)
    g, g2 = gs.render(js.spp, seed=5)
    gst = ctx.render_stats()
    sl, spp = Scene.from_json(ctx, path)
    h, h2 = sl.render(spp, seed=5)
    assert np.array_equal(g.view(np.uint32), h.view(np.uint32))
    o, o2, ost, _ = oracle.OracleScene(js).render(js.spp, seed=5, threads=0)
    assert np.isfinite(g).all() and g.mean() > 0.01
    for k in ("rays_closest", "vertices"):
        assert abs(gst[k] - ost[k]) <= max(16, 0.02 * ost[k]), (k, gst[k], ost[k])
    close = np.isclose(g, o, rtol=1e-3, atol=1e-5)
    print(f"[json scene] close {close.mean():.4f}", flush=True)
    assert close.mean() > 0.95, close.mean()
    z = z_test(g, g2, 8, o, o2, 8)
    assert np.mean(np.abs(z) < 4) > 0.99
