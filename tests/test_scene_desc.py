"""Scene description invariants (RendererBuilder.Build's primitive and light
lists, Config/RendererBuilder.cs:940-999) and the procedural workloads.  CPU only."""
import numpy as np

from radiantium_amd import abi, scenes
from radiantium_amd.scene import SceneDesc


def test_lights_point_at_their_primitives():
    """Every emissive triangle is its own DiffuseAreaLight and the light's
    shape is that primitive (RendererBuilder.cs:976-992)."""
    for desc in (scenes.cornell_box(8, 8), scenes.displaced_icosphere_scene(level=2, width=8, height=8)):
        a = desc.arrays()
        for li, light in enumerate(desc.lights):
            assert a["prims"][light["prim"]][3] == li
        assert sum(1 for p in a["prims"] if p[3] >= 0) == len(desc.lights)


def test_primitive_order_is_entity_order():
    s = SceneDesc()
    m = s.add_material("diffuse")
    p0 = s.add_mesh(np.eye(3, dtype=np.float32), [[0, 1, 2]], material=m)
    p1 = s.add_sphere((0, 0, 0), 1.0, material=m, le=(1, 1, 1))
    p2 = s.add_mesh(np.eye(3, dtype=np.float32), [[0, 1, 2], [0, 2, 1]], material=m, le=(2, 2, 2))
    assert (p0, p1, p2) == (0, 1, 2)
    a = s.arrays()
    assert a["prims"][:, 0].tolist() == [abi.PRIM_TRIANGLE, abi.PRIM_SPHERE, abi.PRIM_TRIANGLE, abi.PRIM_TRIANGLE]
    assert [l["prim"] for l in s.lights] == [1, 2, 3]
    assert a["indices"][1:].min() == 3   # second mesh indexes its own vertices


def test_cornell_walls_face_the_room():
    desc = scenes.cornell_box(8, 8)
    a = desc.arrays()
    pos, idx = a["positions"], a["indices"]
    for t in idx[:10]:   # the five wall quads
        p = pos[t]
        n = np.cross(p[1] - p[0], p[2] - p[0])
        c = p.mean(0)
        to_center = np.array([0.0, 0.5, 0.0]) - c
        assert np.dot(n, to_center) > 0


def test_icosphere_level8_size_and_orientation():
    v, f = scenes.icosphere(3)
    n = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    assert (np.sum(n * v[f].mean(1), 1) > 0).all()
    assert f.shape[0] == 20 * 4 ** 3
    assert len(np.unique(f)) == 10 * 4 ** 3 + 2
    desc = scenes.displaced_icosphere_scene(level=8, width=8, height=8)
    # 1,310,720 sphere triangles + 2 ground + 2 light
    assert desc.arrays()["indices"].shape[0] == 20 * 4 ** 8 + 4
    assert desc.bvh_max_prim == 4


def test_value_noise_range_and_determinism():
    p = np.random.default_rng(0).normal(size=(1000, 3))
    a = scenes.value_noise(p, 0)
    b = scenes.value_noise(p, 0)
    c = scenes.value_noise(p, 1)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= -1 and a.max() <= 1
