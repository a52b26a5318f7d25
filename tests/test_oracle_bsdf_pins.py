"""Pins the oracle's rough-plastic, rough-glass and Disney BSDFs, whose
sample functions fill their value and pdf by calling the same code as
Fr / Pdf, so a self-comparison cannot catch a slip.  CPU only.

- Values: Fr and Pdf at random direction pairs (both hemispheres) against
  an independent float64 numpy statement of the reference's formulas
  (tests/bsdf_numpy.py, written from RoughPlastic.cs, RoughGlass.cs,
  DisneyBsdf.cs, MicrofacetTransmissionBtdf.cs).
- Sampling: a chi-square test of each material's full Sample against its
  Pdf, integrated by quadrature over bins of the sphere (a wrong lobe
  weight, Jacobian or a wrongly normalised D moves the sampled density
  away from Pdf).
- Mathematics: reciprocity of reflection, the (eta_a/eta_b)^2 radiance
  ratio of transmission, the smooth limits (rough glass -> F + (1-F)(eta_a/
  eta_b)^2, plastic with kd = 0 -> R F, Disney metallic -> Schlick(C)), the
  lobe weights with the normalised GTR1 / GTR2 distributions.
"""
import math
import zlib

import numpy as np
import pytest
from scipy import stats

from radiantium_amd.scene import SceneDesc

import bsdf_numpy as B

ETA_A, ETA_B = 1.000277, 1.5046


def _mat(oracle, kind, **p):
    s = SceneDesc()
    m = s.add_material(kind, **p)
    s.set_camera(width=4, height=4)
    return oracle.OracleScene(s), m


def _dirs(rng, n, side=None):
    w = rng.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    if side is not None:
        w[:, 2] = side * np.abs(w[:, 2])
    # the oracle works in float32: compare on the directions it sees
    return w.astype(np.float32).astype(np.float64)


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


# ------------------------------------------------------------------ values vs the numpy statement
PLASTIC = [
    dict(r=(0.7, 0.5, 0.3), roughness=0.4, kd=0.6, ks=0.4),
    dict(r=(0.2, 0.8, 0.5), roughness=0.25, kd=0.3, ks=0.7, dist="GGX", anisotropic=0.3),
    dict(r=(0.9, 0.9, 0.9), roughness=0.6, kd=0.5, ks=0.5, dist="Beckmann", anisotropic=-0.4, etaB=1.33),
    dict(r=(0.5, 0.4, 0.3), roughness=0.3, kd=0.5, ks=0.5, is_two_side=True),
]


def _plastic_np(p):
    return dict(R=p["r"], kd=p["kd"], ks=p["ks"], roughness=p["roughness"], aniso=p.get("anisotropic", 0.0),
                kind=1 if p.get("dist") == "GGX" else 0, eta_b=p.get("etaB", ETA_B),
                two_sided=p.get("is_two_side", False))


GLASS = [
    dict(roughness=0.4, etaB=1.5),
    dict(roughness=0.25, etaB=1.7, anisotropic=0.5),
    dict(roughness=0.5, dist="Beckmann", r=(0.9, 0.8, 0.7), t=(0.6, 0.7, 0.8)),
]


def _glass_np(p):
    return dict(R=p.get("r", 1.0), T=p.get("t", 1.0), roughness=p["roughness"], aniso=p.get("anisotropic", 0.0),
                kind=0 if p.get("dist") == "Beckmann" else 1, eta_b=p.get("etaB", ETA_B))


DISNEY = [
    dict(base_color=(0.7, 0.4, 0.2), roughness=0.5, sheen=0.5, clearcoat=0.6, clearcoat_gloss=0.5),
    dict(base_color=(0.5, 0.6, 0.3), roughness=0.4, metallic=0.3, transmission=0.5, transmission_roughness=0.3,
         specular_tint=0.4, anisotropic=0.5),
    dict(base_color=(0.5, 0.6, 0.3), roughness=0.6, is_thin=True, flatness=0.4, sheen=0.3, sheen_tint=0.5),
    dict(base_color=(0.9, 0.6, 0.2), roughness=0.3, metallic=1.0, specular=0.7),
    dict(base_color=(0.2, 0.3, 0.9), roughness=0.8, eta=1.3, clearcoat=1.0, clearcoat_gloss=1.0, specular_tint=1.0,
         sheen=1.0, sheen_tint=1.0),
]


@pytest.mark.parametrize("kind,params", [("rough_plastic", p) for p in PLASTIC] + [("rough_glass", p) for p in GLASS]
                         + [("disney", p) for p in DISNEY])
def test_values_match_independent_statement(oracle, kind, params):
    o, m = _mat(oracle, kind, **params)
    rng = np.random.default_rng(zlib.crc32(repr(sorted(params.items())).encode()))
    n = 8000
    # most pairs with wo outside (one-sided materials vanish below), the rest anywhere
    wo = np.concatenate([_dirs(rng, 3 * n // 4, 1), _dirs(rng, n - 3 * n // 4)])
    wi = _dirs(rng, n)
    fr, pdf = o.material_eval_n(m, wo, wi)
    if kind == "rough_plastic":
        fr2, pdf2 = B.plastic(wo, wi, **_plastic_np(params))
    elif kind == "rough_glass":
        fr2, pdf2 = B.glass(wo, wi, **_glass_np(params))
    else:
        fr2, pdf2 = B.disney(wo, wi, **params)
    # float32 against float64: relative 1e-3, and an absolute floor far below
    # the values' scale (a term dropped or mis-signed moves values by O(1))
    assert (fr2.max(1) > 0).sum() > n // 3 and (pdf2 > 0).sum() > n // 3
    np.testing.assert_allclose(fr, fr2, rtol=1e-3, atol=1e-6 * fr2.max())
    np.testing.assert_allclose(pdf, pdf2, rtol=1e-3, atol=1e-6 * pdf2.max())


# ------------------------------------------------------------------ chi-square: Sample against Pdf
def _chi2(o, m, wo, n, key0, nth=36, nph=72, sub=6, restrict=None):
    """Chi-square of n samples of wo against n * (integral of Pdf over each
    (theta, phi) bin); returns (p-value, chi2, dof, fraction of samples kept,
    fraction of valid samples whose own pdf is not Pdf's).  restrict(wi) ->
    bool mask limits both sides to a region.  Samples whose pdf disagrees
    with Pdf at their direction are not counted: where the reference's own
    sampler and Pdf disagree (see CHI2), the sample does not come from the
    density Pdf describes."""
    wi, fr, pdf, ty = o.material_sample_n(m, wo, n, key0)
    ok = pdf > 0
    _, pe = o.material_eval_n(m, wo, wi)
    agree = np.abs(pe.astype(np.float64) - pdf) <= 1e-3 * np.maximum(pe, pdf)
    unsupported = float((ok & ~agree).mean())
    ok &= agree
    if restrict is not None:
        ok &= restrict(wi.astype(np.float64))
    wi = wi[ok].astype(np.float64)
    th = np.arccos(np.clip(wi[:, 2], -1, 1))
    ph = np.arctan2(wi[:, 1], wi[:, 0])
    obs, _, _ = np.histogram2d(th, ph, [nth, nph], [[0, np.pi], [-np.pi, np.pi]])
    # expected: midpoint quadrature with sub x sub points per bin
    mt, mp = nth * sub, nph * sub
    tc = (np.arange(mt) + 0.5) * np.pi / mt
    pc = -np.pi + (np.arange(mp) + 0.5) * 2 * np.pi / mp
    T, P = np.meshgrid(tc, pc, indexing="ij")
    g = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    _, pg = o.material_eval_n(m, wo, g)
    pg = pg.astype(np.float64)
    if restrict is not None:
        pg = np.where(restrict(g), pg, 0.0)
    dens = (pg * np.sin(T).reshape(-1)).reshape(mt, mp) * (np.pi / mt) * (2 * np.pi / mp)
    exp = dens.reshape(nth, sub, nph, sub).sum(axis=(1, 3)) * n
    # pool cells expected to hold fewer than 5 samples into one
    big = exp >= 5
    o_c = np.concatenate([obs[big], [obs[~big].sum()]])
    e_c = np.concatenate([exp[big], [exp[~big].sum()]])
    keep = e_c > 0
    chi2 = float(np.sum((o_c[keep] - e_c[keep]) ** 2 / e_c[keep]))
    dof = int(keep.sum()) - 1
    return float(stats.chi2.sf(chi2, dof)), chi2, dof, ok.mean(), unsupported


CHI2 = [
    ("rough_plastic", dict(r=(0.7, 0.5, 0.3), roughness=0.45, kd=0.6, ks=0.4), (0.3, 0.1, 0.9), None),
    ("rough_plastic", dict(r=(0.7, 0.5, 0.3), roughness=0.35, kd=0.3, ks=0.7, dist="GGX", anisotropic=0.4),
     (0.6, -0.2, 0.5), None),
    ("disney", dict(base_color=(0.7, 0.4, 0.2), roughness=0.55, sheen=0.5, clearcoat=0.6, clearcoat_gloss=0.3),
     (0.2, 0.3, 0.9), None),
    ("disney", dict(base_color=(0.9, 0.6, 0.2), roughness=0.4, metallic=1.0), (0.5, 0.0, 0.7), None),
    ("disney", dict(base_color=(0.5, 0.6, 0.3), roughness=0.6, is_thin=True, flatness=0.4, anisotropic=0.5),
     (-0.3, 0.4, 0.8), None),
    ("disney", dict(base_color=(0.5, 0.6, 0.3), roughness=0.5, transmission=0.6, transmission_roughness=0.45),
     (0.3, 0.0, 0.8), None),
    ("rough_glass", dict(roughness=0.45, etaB=1.5), (0.4, 0.1, 0.8), None),
    ("rough_glass", dict(roughness=0.4, etaB=1.5, dist="Beckmann"), (0.0, 0.3, 0.9), None),
    # from inside, only the transmitted side: RoughGlass's total-internal-
    # reflection fallback (RoughGlass.cs:160-164) samples reflections with
    # the weight 1 - F that its Pdf scores as F, so the reflected side of the
    # reference's own sampler does not follow its Pdf there; and a reflection
    # off a steep microfacet can cross to the far side, typed as a
    # transmission, where Pdf (whose half-vector checks reject it) is 0 or
    # another value: with the fallback, about 7 % of the samples disagree with
    # Pdf there
    ("rough_glass", dict(roughness=0.45, etaB=1.5), (0.3, 0.0, -0.9), "up"),
]


@pytest.mark.parametrize("kind,params,wo,region", CHI2)
def test_sampling_follows_pdf(oracle, kind, params, wo, region):
    o, m = _mat(oracle, kind, **params)
    wo = _unit(wo).astype(np.float32)
    restrict = (lambda w: w[..., 2] > 0) if region == "up" else None
    p, chi2, dof, kept, unsupported = _chi2(o, m, wo, 300_000, 0x51A7 + len(kind), restrict=restrict)
    assert kept > 0.3
    assert unsupported < (0.1 if region == "up" else 1e-4), unsupported
    assert p > 1e-4, (chi2, dof, p)


def test_chi2_detects_a_wrong_pdf(oracle):
    """The test has power: the same samples against a pdf whose glossy lobe
    weight is off by 10 % fail it."""
    o, m = _mat(oracle, "rough_plastic", r=(0.7, 0.5, 0.3), roughness=0.45, kd=0.6, ks=0.4)
    o2, m2 = _mat(oracle, "rough_plastic", r=(0.7, 0.5, 0.3), roughness=0.45, kd=0.6, ks=0.44)

    class Mixed:   # samples of the first material, Pdf of the second
        def material_sample_n(self, _m, wo, n, key0):
            return o.material_sample_n(m, wo, n, key0)

        def material_eval_n(self, _m, wo, wi):
            return o2.material_eval_n(m2, wo, wi)

    p, chi2, dof, _, _ = _chi2(Mixed(), 0, _unit((0.3, 0.1, 0.9)).astype(np.float32), 300_000, 7)
    assert p < 1e-6, (chi2, dof)


# ------------------------------------------------------------------ reciprocity
@pytest.mark.parametrize("kind,params", [("rough_plastic", PLASTIC[0]), ("rough_plastic", PLASTIC[1]),
                                         ("rough_glass", GLASS[0]), ("disney", DISNEY[0]), ("disney", DISNEY[2]),
                                         ("disney", DISNEY[4])])
def test_reflection_is_reciprocal(oracle, kind, params):
    o, m = _mat(oracle, kind, **params)
    rng = np.random.default_rng(11)
    wo, wi = _dirs(rng, 3000, 1), _dirs(rng, 3000, 1)
    f1, _ = o.material_eval_n(m, wo, wi)
    f2, _ = o.material_eval_n(m, wi, wo)
    assert f1.max() > 0
    np.testing.assert_allclose(f1, f2, rtol=2e-4, atol=1e-6 * f1.max())


@pytest.mark.parametrize("eta_b", [1.5, 1.33, 2.0])
def test_glass_transmission_radiance_ratio(oracle, eta_b):
    """Radiance transport through a rough dielectric, wo outside, wi inside:
    f(wo, wi) (1 - F_i) / (f(wi, wo) (1 - F_o)) = (eta_a / eta_b)^2, with
    the (1/eta)^2 factor of RoughGlass.cs:62-66 (without it the ratio is
    (eta_b / eta_a)^2).  F_o, F_i: the Fresnel terms each direction's Fr
    uses.  The reference orients the etas by the side of wo and
    DielectricFunc orients them again by the sign of wo . wh (RoughGlass.cs:
    49-58, Fresnel.cs:53-60), so from inside both land on eta_a -> eta_b at
    the inside angle: F_i = F(|wi . wh|, eta_a, eta_b), not the physical
    F(|wo . wh|); the Fresnel formula itself is pinned in test_oracle_shading.
    Pairs that see the back of their microfacet (wo . wh < 0 with wh
    facing up), which the reference does not reject, are left out."""
    o, m = _mat(oracle, "rough_glass", roughness=0.4, etaB=eta_b)
    rng = np.random.default_rng(3)
    wo, wi = _dirs(rng, 4000, 1), _dirs(rng, 4000, -1)
    f1, _ = o.material_eval_n(m, wo, wi)
    f2, _ = o.material_eval_n(m, wi, wo)
    eta = eta_b / ETA_A
    wh = wo + wi * eta
    wh /= np.linalg.norm(wh, axis=1, keepdims=True)
    wh *= np.sign(wh[:, 2:3])
    front = (np.sum(wo * wh, 1) > 0) & (np.sum(wi * wh, 1) < 0)
    use = (f1[:, 0] > 1e-4) & (f2[:, 0] > 1e-4) & front
    assert use.sum() > 500
    f_o = B.fresnel_dielectric(np.abs(np.sum(wo * wh, 1)), ETA_A, eta_b)
    f_i = B.fresnel_dielectric(np.abs(np.sum(wi * wh, 1)), ETA_A, eta_b)
    ratio = f1[use, 0] * (1 - f_i[use]) / (f2[use, 0] * (1 - f_o[use]))
    np.testing.assert_allclose(ratio, (ETA_A / eta_b) ** 2, rtol=2e-3)


# ------------------------------------------------------------------ smooth limits
def _albedo(o, m, wo, n=20000, key0=0x9000):
    wi, fr, pdf, _ = o.material_sample_n(m, wo, n, key0)
    ok = pdf > 0
    w = np.zeros((n, 3))
    w[ok] = fr[ok] * np.abs(wi[ok, 2:3]) / pdf[ok, None]
    return w.mean(0), w.std(0) / math.sqrt(n)


@pytest.mark.parametrize("cos_o", [0.9, 0.5, -0.95])
def test_smooth_glass_limit(oracle, cos_o):
    """roughness -> 0: E[f |cos| / pdf] -> F + (1 - F)(eta_i / eta_t)^2."""
    o, m = _mat(oracle, "rough_glass", roughness=0.02, etaB=1.5)
    wo = np.array([math.sqrt(1 - cos_o ** 2), 0, cos_o], np.float32)
    ei, et = (ETA_A, 1.5) if cos_o > 0 else (1.5, ETA_A)
    F = float(B.fresnel_dielectric(abs(cos_o), ei, et))
    mean, se = _albedo(o, m, wo)
    assert mean[0] == pytest.approx(F + (1 - F) * (ei / et) ** 2, rel=5e-3, abs=4 * se[0])


@pytest.mark.parametrize("dist", ["GGX", "Beckmann"])
@pytest.mark.parametrize("cos_o", [0.95, 0.6, 0.3])
def test_smooth_plastic_specular_limit(oracle, dist, cos_o):
    """kd = 0, roughness -> 0: E[f cos / pdf] -> R F_dielectric(cos_o)."""
    R = np.array([0.9, 0.6, 0.3])
    o, m = _mat(oracle, "rough_plastic", r=tuple(R), kd=0.0, ks=1.0, roughness=0.03, dist=dist)
    wo = np.array([math.sqrt(1 - cos_o ** 2), 0, cos_o], np.float32)
    mean, se = _albedo(o, m, wo)
    F = float(B.fresnel_dielectric(cos_o, ETA_A, ETA_B))
    np.testing.assert_allclose(mean, R * F, rtol=1e-2, atol=4 * se.max())


@pytest.mark.parametrize("cos_o", [0.95, 0.5, 0.25])
def test_disney_metallic_is_a_schlick_conductor(oracle, cos_o):
    """metallic = 1, roughness -> 0: the Disney BSDF is a near-mirror GGX lobe
    with Schlick's Fresnel of the base colour: albedo -> C + (1-C)(1-cos)^5."""
    C = np.array([0.9, 0.6, 0.2])
    o, m = _mat(oracle, "disney", base_color=tuple(C), metallic=1.0, roughness=0.05)
    wo = np.array([math.sqrt(1 - cos_o ** 2), 0, cos_o], np.float32)
    mean, se = _albedo(o, m, wo, n=40000)
    np.testing.assert_allclose(mean, C + (1 - C) * (1 - cos_o) ** 5, rtol=1.5e-2, atol=4 * se.max())


def _lobe_mass_45(d):
    """Mass of a normalised isotropic D(cos) cos over the half vectors whose
    mirror of the normal stays above the horizon (theta_h < 45 deg)."""
    th = np.linspace(0, np.pi / 4, 200001)
    f = d(np.cos(th)) * np.cos(th) * np.sin(th) * 2 * np.pi
    return np.trapezoid(f, th)


@pytest.mark.parametrize("gloss_param,roughness,cc", [(0.0, 0.5, 1.0), (0.7, 0.3, 0.5), (1.0, 0.6, 0.25)])
def test_disney_pdf_mass_from_lobe_weights(oracle, gloss_param, roughness, cc):
    """At normal incidence the Pdf integrates to w_d + w_s M_s + w_c M_c, where
    w are GetSampleWeight's lobe weights (DisneyBsdf.cs:503-512) and M the
    mass of the textbook GGX (GTR2) and GTR1 (Burley 2012) distributions that
    reflects above the horizon.  A misnormalised GTR1 clearcoat or a wrong
    weight changes the integral."""
    C = (0.6, 0.5, 0.4)
    o, m = _mat(oracle, "disney", base_color=C, roughness=roughness, clearcoat=cc, clearcoat_gloss=gloss_param)
    wo = np.array([0, 0, 1.0], np.float32)
    # peaked quadrature over the upper hemisphere (theta = pi/2 u^2)
    n = 1200
    u = (np.arange(n) + 0.5) / n
    th = 0.5 * np.pi * u ** 2
    dth = np.pi * u / n
    ph = 2 * np.pi * (np.arange(64) + 0.5) / 64
    T, P = np.meshgrid(th, ph, indexing="ij")
    g = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    _, pdf = o.material_eval_n(m, wo, g)
    tot = float(np.sum(pdf.reshape(n, 64) * (np.sin(T) * dth[:, None]) * (2 * np.pi / 64)))
    a2 = roughness ** 4                                      # alpha = roughness^2
    gtr2 = lambda c: a2 / (np.pi * (1 + (a2 - 1) * c * c) ** 2)
    ga = (1 - gloss_param) * 0.1 + gloss_param * 0.001
    gtr1 = lambda c: (ga * ga - 1) / (np.pi * np.log(ga * ga) * (1 + (ga * ga - 1) * c * c))
    lum = 0.212671 * C[0] + 0.715160 * C[1] + 0.072169 * C[2]
    a = min(max(lum, 0.3), 0.7)
    w_d, w_s, w_c = a, (1 - a) * 2 / (2 + cc), (1 - a) * cc / (2 + cc)
    expect = w_d + w_s * _lobe_mass_45(gtr2) + w_c * _lobe_mass_45(gtr1)
    assert tot == pytest.approx(expect, rel=3e-3)
