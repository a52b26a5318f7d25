"""Pins the oracle's homogeneous medium, Henyey-Greenstein phase function
(Mediums/Homogeneous.cs:14-71, IPhaseFunction.cs:39-81) and the volumetric
path integrator (Integrators/VolumetricPathTracer.cs:28-314) against closed
forms: Beer-Lambert, mean free path, HG normalisation and mean cosine, and a
furnace filled with a non-absorbing medium.  CPU only."""
import numpy as np
import pytest

from radiantium_amd import scenes
from radiantium_amd.scene import SceneDesc
from radiantium_amd.scenes import _box, _quad


def cpuref():
    from oracle import cpuref as c
    return c


def test_tr_is_beer_lambert():
    tr = cpuref().medium_tr((0.1, 0.0, 0.3), (0.2, 0.0, 0.7), 0.5, 2.5, scale=2.0)
    st = 2.0 * np.array([0.3, 0.0, 1.0])
    assert np.allclose(tr, np.exp(-st * 2.0), rtol=1e-6)
    assert tr[1] == 1.0                     # sigma_t == 0 -> exactly 1


def test_free_flight_grey_weights_and_mean_free_path():
    """Grey medium: a scattering vertex carries exactly sigma_s/sigma_t, a
    pass carries 1; distances are exponential with mean 1/sigma_t."""
    sa, ss = 0.3, 1.2
    st = sa + ss
    c = cpuref()
    ts, n_pass = [], 0
    for k in range(20000):
        med, p, w, t = c.medium_sample((sa,) * 3, (ss,) * 3, (0, 0, 0), (0, 0, 1), 0.0, 1e30, 1000 + k)
        assert med
        assert np.allclose(w, ss / st, rtol=1e-5)
        ts.append(t)
        assert p[2] == pytest.approx(t, rel=1e-6)
    ts = np.array(ts)
    assert ts.mean() == pytest.approx(1 / st, rel=0.03)
    # finite segment: P(pass) = exp(-st d), weight of a pass = tr/pdf = 1 for grey
    d = 0.8
    for k in range(4000):
        med, _, w, _ = c.medium_sample((sa,) * 3, (ss,) * 3, (0, 0, 0), (0, 0, 1), 0.0, d, 50000 + k)
        if not med:
            n_pass += 1
            assert np.allclose(w, 1.0, rtol=1e-6)
    assert n_pass / 4000 == pytest.approx(np.exp(-st * d), abs=0.03)


def test_free_flight_spectral_is_unbiased():
    """Coloured sigma_t: the passing weight tr/pdf averages to exp(-sigma_c d)
    per channel (the spectral MIS of Homogeneous.Sample)."""
    sa, ss, d = np.array([0.1, 0.6, 1.5]), np.array([0.4, 0.2, 0.1]), 0.7
    c = cpuref()
    acc = np.zeros(3)
    N = 40000
    for k in range(N):
        med, _, w, _ = c.medium_sample(sa, ss, (0, 0, 0), (1, 0, 0), 0.0, d, 7 * k + 3)
        if not med:
            acc += w
    assert np.allclose(acc / N, np.exp(-(sa + ss) * d), rtol=0.03)


@pytest.mark.parametrize("g", [-0.3, 0.0, 0.5, 0.9])
def test_hg_normalised_and_mean_cosine(g):
    c = cpuref()
    wo = np.array([0.0, 0.0, 1.0])
    # normalisation by quadrature in cos(theta)
    n = 4000
    mu = (np.arange(n) + 0.5) / n * 2 - 1
    vals = np.array([c.hg_p(g, wo, (np.sqrt(1 - m * m), 0.0, m)) for m in mu])
    assert np.sum(vals) * 2 * np.pi * (2 / n) == pytest.approx(1.0, rel=2e-3)
    # sampling: <cos(wi, -wo)> = g and the returned p equals P(wo, wi)
    cs = []
    for k in range(8000):
        wi, p = c.hg_sample(g, wo, 99 + k)
        assert np.linalg.norm(wi) == pytest.approx(1.0, abs=1e-5)
        assert p == pytest.approx(c.hg_p(g, wo, wi), rel=1e-3)
        cs.append(-np.dot(wi, wo))
    assert np.mean(cs) == pytest.approx(g, abs=0.02)


def _slab_scene(sigma_a, le=(4.0, 4.0, 4.0), res=12):
    """An absorbing box (no surface, medium inside) between the camera and an
    emissive quad: every camera ray through the box carries Le exp(-sigma d)."""
    s = SceneDesc()
    black = s.add_material("diffuse", kd=(0.0, 0.0, 0.0))
    m = s.add_medium(sigma_a=sigma_a, sigma_s=(0.0, 0.0, 0.0))
    s.add_mesh(*_box((1.0, 1.0, 1.0), (0.0, -0.5, 0.0), 0.0), material=-1, medium_inside=m)
    # emitter at z = -1.5 facing +z, larger than the view
    W = 3.0
    s.add_mesh(*_quad((-W, -W, -1.5), (W, -W, -1.5), (W, W, -1.5), (-W, W, -1.5)), material=black, le=le)
    s.set_camera(fov=20.0, width=res, height=res, origin=(0.0, 0.0, 4.0), target=(0.0, 0.0, 0.0), up=(0, 1, 0))
    s.set_integrator(type="vol_path", max_depth=16)
    s.bvh_max_prim = 1
    return s


def test_beer_lambert_through_absorbing_box(oracle):
    sig = np.array([0.2, 0.7, 1.6])
    res, le = 12, 4.0
    desc = _slab_scene(tuple(sig), (le, le, le), res)
    o = oracle.OracleScene(desc)
    spp = 256
    rgb, s2, _, _ = o.render(spp, seed=9, threads=0)
    img = rgb.reshape(res, res, 3)
    # expected per pixel: average over 4x4 sub-pixel rays of le * exp(-sigma * chord in the box)
    sub = (np.arange(4) + 0.5) / 4
    exp_img = np.zeros((res, res, 3))
    lo, hi = np.array([-0.5, -0.5, -0.5]), np.array([0.5, 0.5, 0.5])
    for sx in sub:
        for sy in sub:
            xy = np.array([(x + sx, y + sy) for x in range(res) for y in range(res)], np.float32)
            r = o.camera_rays(xy)
            org, d = r["o"].astype(np.float64), r["d"].astype(np.float64)
            with np.errstate(divide="ignore"):
                t0, t1 = (lo - org) / d, (hi - org) / d
            tn = np.max(np.minimum(t0, t1), axis=1)
            tf = np.min(np.maximum(t0, t1), axis=1)
            chord = np.clip(tf - tn, 0, None)
            exp_img += (le * np.exp(-np.outer(chord, sig))).reshape(res, res, 3) / 16
    var = np.maximum(s2.reshape(res, res, 3) / spp - img.astype(np.float64) ** 2, 0) / spp
    z = (img - exp_img) / np.sqrt(var + 1e-6)
    assert np.mean(np.abs(z) > 4) < 0.01
    assert np.abs(img.mean(axis=(0, 1)) - exp_img.mean(axis=(0, 1))).max() < 0.02 * le


def _medium_furnace(sigma_s, g, rho=0.5, le=1.0, level=2, res=8, max_depth=100000):
    """Closed emissive diffuse sphere (albedo rho, emission le, normals
    inward) filled with a non-absorbing HG medium.  The medium sits on the
    walls' normal side, so GetMedium keeps every interior ray in it, and the
    camera starts in it (global medium)."""
    v, f = scenes.icosphere(level)
    n = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    outward = np.einsum("ij,ij->i", n, v[f].mean(axis=1)) > 0
    f = np.where(outward[:, None], f[:, ::-1], f)
    s = SceneDesc()
    mat = s.add_material("diffuse", kd=(rho, rho, rho))
    m = s.add_medium(sigma_a=(0, 0, 0), sigma_s=(sigma_s,) * 3, g=g)
    s.add_mesh(v.astype(np.float32), f, material=mat, le=(le, le, le), medium_outside=m)
    s.global_medium = m
    s.set_camera(fov=60.0, width=res, height=res, origin=(0, 0, 0), target=(0, 0, 1), up=(0, 1, 0))
    s.set_integrator(type="vol_path", max_depth=max_depth)
    s.bvh_max_prim = 1
    return s


def test_furnace_with_scattering_medium(oracle):
    """Radiance in the medium-filled furnace stays le / (1 - rho) (the
    medium neither absorbs nor emits).  Triangles are small, so the light-pdf
    side of every MIS pair carries almost all the weight and the phase /
    BSDF sides (whose IntersectTr uses |ray.At(t)| as the medium length,
    Scene.cs:121) contribute ~1e-4 of it."""
    desc = _medium_furnace(1.5, 0.4)
    o = oracle.OracleScene(desc)
    spp = 32
    rgb, s2, st, _ = o.render(spp, seed=5, threads=0)
    n = rgb.size // 3 * spp
    mean = float(rgb.mean())
    var = float(s2.sum() / 3 / n - mean ** 2)
    se = (max(var, 1e-12) / n) ** 0.5
    assert abs(mean - 2.0) < 4 * se + 1e-3, (mean, se)
    assert st["invalid"] == 0


def test_vol_path_without_media_matches_path(oracle):
    """With no media, vol_path is the Mis path integrator (its rays keep the
    camera's MinT instead of 1e-4): the two images agree statistically."""
    d_path = scenes.cornell_box(16, 16, method="Mis")
    d_vol = scenes.cornell_box(16, 16, method="Mis")
    d_vol.set_integrator(type="vol_path", max_depth=8)
    a, a2, _, _ = oracle.OracleScene(d_path).render(256, seed=1, threads=0)
    b, b2, _, _ = oracle.OracleScene(d_vol).render(256, seed=2, threads=0)
    va = np.maximum(a2 / 256 - a.astype(np.float64) ** 2, 0) / 256
    vb = np.maximum(b2 / 256 - b.astype(np.float64) ** 2, 0) / 256
    z = (a - b) / np.sqrt(va + vb + 1e-8)
    assert np.mean(np.abs(z) > 4) < 0.01
