"""An independent numpy (float64) statement of the reference's rough-plastic,
rough-glass and Disney BSDF values and pdfs, written from the C# sources:

- Radiantium.Offline/Materials/RoughPlastic.cs:54-128 (FrImpl, PdfImpl)
- Radiantium.Offline/Materials/RoughGlass.cs:37-118 (FrImpl, PdfImpl)
- Radiantium.Offline/Bxdf/DisneyBsdf.cs:13-512 (the lobes, FrImpl, PdfImpl,
  GetSampleWeight) with Bxdf/MicrofacetTransmissionBtdf.cs:25-63
- Radiantium.Offline/Bxdf/Microfacet.cs:93-196 (GGX / Beckmann D, Lambda,
  RoughnessToAlpha), Bxdf/MicrofacetReflectionBrdf.cs:26-47,
  Bxdf/Fresnel.cs:50-68, Radiantium.Offline/Coordinate.cs:66-104.

It shares no code with the oracle (which is C++) and is used only by the
tests, to catch a slip in the oracle's restatement of these formulas: the
oracle's float32 values must agree with these float64 ones.  Vectorised over
direction pairs wo, wi of shape [n, 3] in the local shading frame; radiance
transport.
"""
import numpy as np

def _dot(a, b):
    return np.sum(a * b, -1)


def _normalize(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _same(a, b):
    return a[..., 2] * b[..., 2] > 0


def _phi2(w):
    """Cos2Phi, Sin2Phi (Coordinate.cs:94-104).  The reference's threshold is
    4 * float.Epsilon (a denormal): only an exactly vertical w takes it."""
    s2 = w[..., 0] ** 2 + w[..., 1] ** 2
    tiny = s2 <= 0.0
    with np.errstate(divide="ignore", invalid="ignore"):
        c2 = np.where(tiny, 1.0, np.clip(w[..., 0] ** 2 / np.where(tiny, 1, s2), -1, 1))
        sn2 = np.where(tiny, 0.0, np.clip(w[..., 1] ** 2 / np.where(tiny, 1, s2), -1, 1))
    return c2, sn2


def _tan2(w):
    with np.errstate(divide="ignore"):
        return np.maximum(1 - w[..., 2] ** 2, 0) / w[..., 2] ** 2


def roughness_to_alpha(r, k):
    """Microfacet.RoughnessToAlpha (Microfacet.cs:187-193)."""
    return max(r * r * (1 + k), 0.001), max(r * r * (1 - k), 0.001)


def d_ggx(wh, ax, ay):
    t2 = _tan2(wh)
    c2, s2 = _phi2(wh)
    cos4 = (wh[..., 2] ** 2) ** 2
    e = (c2 / (ax * ax) + s2 / (ay * ay)) * t2
    with np.errstate(divide="ignore", invalid="ignore"):
        d = 1 / (np.pi * ax * ay * cos4 * (1 + e) ** 2)
    return np.where(np.isinf(t2), 0.0, d)


def d_beckmann(wh, ax, ay):
    t2 = _tan2(wh)
    c2, s2 = _phi2(wh)
    cos4 = (wh[..., 2] ** 2) ** 2
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        d = np.exp(-t2 * (c2 / (ax * ax) + s2 / (ay * ay))) / (np.pi * ax * ay * cos4)
    return np.where(np.isinf(t2), 0.0, d)


def _abs_tan(w):
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.abs(np.sqrt(np.maximum(1 - w[..., 2] ** 2, 0)) / w[..., 2])


def lambda_ggx(w, ax, ay):
    at = _abs_tan(w)
    c2, s2 = _phi2(w)
    alpha = np.sqrt(c2 * ax * ax + s2 * ay * ay)
    with np.errstate(invalid="ignore"):
        lam = (-1 + np.sqrt(1 + (alpha * at) ** 2)) / 2
    return np.where(np.isinf(at), 0.0, lam)


def lambda_beckmann(w, ax, ay):
    at = _abs_tan(w)
    c2, s2 = _phi2(w)
    alpha = np.sqrt(c2 * ax * ax + s2 * ay * ay)
    with np.errstate(divide="ignore", invalid="ignore"):
        a = 1 / (alpha * at)
        lam = (1 - 1.259 * a + 0.396 * a * a) / (3.535 * a + 2.181 * a * a)
    return np.where(np.isinf(at) | (a >= 1.6), 0.0, lam)


class Dist:
    """IMicrofacetDistribution: kind 0 Beckmann, 1 GGX; `separable` is
    DisneyDistributionGtr2's G = G1(wo) G1(wi) instead of 1/(1+Lo+Li)."""

    def __init__(self, kind, ax, ay, separable=False):
        self.kind, self.ax, self.ay, self.separable = kind, ax, ay, separable

    def D(self, wh):
        return (d_ggx if self.kind == 1 else d_beckmann)(wh, self.ax, self.ay)

    def lam(self, w):
        return (lambda_ggx if self.kind == 1 else lambda_beckmann)(w, self.ax, self.ay)

    def G(self, wo, wi):
        if self.separable:
            return 1 / (1 + self.lam(wo)) * (1 / (1 + self.lam(wi)))
        return 1 / (1 + self.lam(wo) + self.lam(wi))

    def pdf(self, wh):
        return self.D(wh) * np.abs(wh[..., 2])


def fresnel_dielectric(ci, ei, et):
    """Fresnel.DielectricFunc (Fresnel.cs:50-68), vectorised over ci."""
    ci = np.clip(np.asarray(ci, np.float64), -1, 1)
    ent = ci > 0
    e_i = np.where(ent, ei, et)
    e_t = np.where(ent, et, ei)
    ci = np.abs(ci)
    si = np.sqrt(np.maximum(0, 1 - ci * ci))
    st = e_i / e_t * si
    ct = np.sqrt(np.maximum(0, 1 - st * st))
    rpar = (e_t * ci - e_i * ct) / (e_t * ci + e_i * ct)
    rper = (e_i * ci - e_t * ct) / (e_i * ci + e_t * ct)
    return np.where(st >= 1, 1.0, (rpar ** 2 + rper ** 2) / 2)


def _faceforward_z(wh):
    return np.where((wh[..., 2] < 0)[..., None], -wh, wh)


def mf_reflection_fr(wo, wi, dist, fresnel):
    """MicrofacetReflectionBrdf.Fr with R = 1 (MicrofacetReflectionBrdf.cs:26-40);
    fresnel(cos) -> [n] or [n, 3]."""
    co, ci = np.abs(wo[:, 2]), np.abs(wi[:, 2])
    wh = wo + wi
    ok = _same(wo, wi) & (co != 0) & (ci != 0) & np.any(wh != 0, -1)
    wh = _normalize(np.where(ok[:, None], wh, [[0, 0, 1.0]]))
    f = fresnel(_dot(wi, _faceforward_z(wh)))
    if f.ndim == 1:
        f = f[:, None]
    val = np.abs(dist.D(wh) * dist.G(wo, wi))[:, None] * np.abs(f) / (4 * ci * co)[:, None]
    return np.where(ok[:, None], val, 0.0)


def mf_reflection_pdf(wo, wi, dist):
    """MicrofacetReflectionBrdf.Pdf (MicrofacetReflectionBrdf.cs:42-47)."""
    wh = _normalize(wo + wi)
    with np.errstate(divide="ignore", invalid="ignore"):
        p = dist.pdf(wh) / (4 * _dot(wo, wh))
    return np.where(_same(wo, wi), p, 0.0)


def lambert_pdf(wo, wi):
    return np.where(_same(wo, wi), np.abs(wi[:, 2]) / np.pi, 0.0)


# ------------------------------------------------------------------ RoughPlastic
def plastic(wo, wi, R, kd, ks, roughness, aniso=0.0, kind=0, eta_a=1.000277, eta_b=1.5046, two_sided=False):
    """RoughPlastic Fr [n,3] and Pdf [n] (RoughPlastic.cs:54-128)."""
    R = np.asarray(R, np.float64)
    dist = Dist(kind, *roughness_to_alpha(roughness, aniso))
    off = (wo[:, 2] <= 0) | (wi[:, 2] <= 0)
    diffuse = np.where(_same(wo, wi)[:, None], R / np.pi, 0.0)
    glossy = R * mf_reflection_fr(wo, wi, dist, lambda c: fresnel_dielectric(c, eta_a, eta_b))
    fr = kd * diffuse + ks * glossy
    p = ks / (kd + ks)
    pdf = (1 - p) * lambert_pdf(wo, wi) + p * mf_reflection_pdf(wo, wi, dist)
    if not two_sided:
        fr = np.where(off[:, None], 0.0, fr)
        pdf = np.where(off, 0.0, pdf)
    return fr, pdf


# ------------------------------------------------------------------ RoughGlass
def glass(wo, wi, R=1.0, T=1.0, roughness=0.3, aniso=0.0, kind=1, eta_a=1.000277, eta_b=1.5046):
    """RoughGlass Fr [n,3] and Pdf [n] in radiance transport (RoughGlass.cs:37-118)."""
    R = np.broadcast_to(np.asarray(R, np.float64), (3,))
    T = np.broadcast_to(np.asarray(T, np.float64), (3,))
    dist = Dist(kind, *roughness_to_alpha(roughness, aniso))
    refl = _same(wo, wi)
    ent = wo[:, 2] > 0
    ei = np.where(ent, eta_a, eta_b)
    et = np.where(ent, eta_b, eta_a)
    # Fr
    eta = np.where(refl, 1.0, et / ei)
    wh = _normalize(wo + wi * eta[:, None])
    wh = _faceforward_z(wh)
    odh, idh = _dot(wo, wh), _dot(wi, wh)
    sd = odh + eta * idh
    d = dist.D(wh)
    g = dist.G(wo, wi)
    f = fresnel_dielectric(odh, ei, et)
    with np.errstate(divide="ignore", invalid="ignore"):
        brdf = (d * g * f / (4.0 * wi[:, 2] * wo[:, 2]))[:, None] * R
        factor = 1 / eta
        btdf = ((1 - f)[:, None] * T) * np.abs(d * g * eta * eta * np.abs(idh) * np.abs(odh) * factor * factor /
                                               (wo[:, 2] * wi[:, 2] * sd * sd))[:, None]
    btdf = np.where((odh * idh > 0)[:, None], 0.0, btdf)
    fr = np.where(refl[:, None], brdf, btdf)
    fr = np.where(np.all(np.isfinite(fr) & (fr >= 0), -1)[:, None], fr, 0.0)   # Color3F.IsValid
    # Pdf
    whr = _faceforward_z(_normalize(wo + wi))
    with np.errstate(divide="ignore", invalid="ignore"):
        ok_r = ~((_dot(wo, whr) * wo[:, 2] <= 0) | (_dot(wi, whr) * wi[:, 2] <= 0))
        pdf_r = dist.pdf(whr) * (1.0 / (4.0 * np.abs(_dot(whr, wi)))) * fresnel_dielectric(_dot(wo, whr), ei, et)
        eta_t = et / ei
        wht = _faceforward_z(_normalize(wo + wi * eta_t[:, None]))
        o_t, i_t = _dot(wo, wht), _dot(wi, wht)
        ok_t = ~((o_t * i_t > 0) | (o_t * wo[:, 2] <= 0) | (i_t * wi[:, 2] <= 0))
        sdt = o_t + eta_t * i_t
        pdf_t = dist.pdf(wht) * np.abs(eta_t * eta_t * i_t / (sdt * sdt)) * (1 - fresnel_dielectric(o_t, ei, et))
    pdf = np.where(refl, np.where(ok_r, pdf_r, 0.0), np.where(ok_t, pdf_t, 0.0))
    pdf = np.where(np.isfinite(pdf), pdf, 0.0)
    return fr, pdf


# ------------------------------------------------------------------ Disney
def _schlick_w(c):
    m = np.clip(1 - c, 0, 1)
    return (m * m) * (m * m) * m


def _lum(c):
    return c[0] * 0.212671 + c[1] * 0.715160 + c[2] * 0.072169


def disney_weights(color, metallic, transmission, clearcoat):
    """DisneyBsdf.GetSampleWeight (DisneyBsdf.cs:503-512)."""
    a = np.clip(_lum(np.asarray(color, np.float64)) * (1 - metallic), 0.3, 0.7)
    b = 1 - a
    return a * (1 - transmission), a * transmission, b * 2 / (2 + clearcoat), b * clearcoat / (2 + clearcoat)


def _gtr1(c, alpha):
    a2 = alpha * alpha
    return (a2 - 1) / (np.pi * np.log(a2) * (1 + (a2 - 1) * c * c))


def _g1_ggx(c, alpha):
    c2 = c * c
    return 2 / (1 + np.sqrt(1 + alpha * alpha * (1 - c2) / c2))


def disney(wo, wi, base_color=(0.5, 0.5, 0.5), metallic=0.0, eta=1.5, roughness=1.0, specular=1.0,
           specular_tint=0.0, anisotropic=0.0, sheen=0.0, sheen_tint=0.0, clearcoat=0.0, clearcoat_gloss=0.0,
           transmission=0.0, transmission_roughness=0.0, flatness=0.0, is_thin=False):
    """DisneyBsdf Fr [n,3] and Pdf [n] (DisneyBsdf.cs:230-348), scatting distance 0."""
    C = np.asarray(base_color, np.float64)
    lum = _lum(C)
    tint = C / lum if lum > 0 else np.ones(3)
    dw = 1 - metallic

    def anis_alpha(r):
        aspect = np.sqrt(1 - anisotropic * 0.9)
        return max(0.001, r * r / aspect), max(0.001, r * r * aspect)

    spec_dist = Dist(1, *anis_alpha(roughness), separable=True)
    tr_dist = Dist(1, *anis_alpha(transmission_roughness), separable=True)
    r0 = tint * ((eta - 1) ** 2 / (eta + 1) ** 2)

    def schlick(r0c, c):
        w = _schlick_w(c)[:, None]
        return (1 - w) * r0c + w * 1.0

    def disney_fresnel(c):
        fd = (1 - specular_tint) * fresnel_dielectric(c, 1.0, eta)[:, None] + specular_tint * schlick(r0, c)
        return (1 - metallic) * (fd * specular) + metallic * schlick(C, c)

    wh = wo + wi
    nz = np.any(wh != 0, -1)
    whn = _normalize(np.where(nz[:, None], wh, [[0, 0, 1.0]]))
    cd = _dot(wi, whn)
    fo, fi = _schlick_w(np.abs(wo[:, 2])), _schlick_w(np.abs(wi[:, 2]))
    # lobes of the upper hemisphere
    diff = (dw * C / np.pi) * ((1 - fo / 2) * (1 - fi / 2))[:, None]
    rr = 2 * roughness * cd * cd
    retro = (dw * C / np.pi) * (rr * (fo + fi + fo * fi * (rr - 1)))[:, None]
    fss90 = cd * cd * roughness
    fss = ((1 - fo) + fo * fss90) * ((1 - fi) + fi * fss90)
    ss = 1.25 * (fss * (1 / (np.abs(wo[:, 2]) + np.abs(wi[:, 2])) - 0.5) + 0.5)
    fake_ss = (dw * C / np.pi) * ss[:, None]
    sheen_col = (1 - sheen_tint) * 1.0 + sheen_tint * tint
    sheen_l = (dw * sheen * sheen_col) * _schlick_w(cd)[:, None] if sheen > 0 else 0.0 * diff
    for lobe in (retro, fake_ss, sheen_l):
        lobe[~nz] = 0.0
    dis_diff = diff * (1 - flatness) + fake_ss * flatness if is_thin else diff
    diffuse = (retro + sheen_l + dis_diff) * (1 - transmission) if dw > 0 else 0.0 * diff
    spec = mf_reflection_fr(wo, wi, spec_dist, disney_fresnel)
    gloss = (1 - clearcoat_gloss) * 0.1 + clearcoat_gloss * 0.001
    cc_d = _gtr1(np.abs(whn[:, 2]), gloss)
    cc_f = 0.04 + (1 - 0.04) * _schlick_w(_dot(wo, whn))
    cc_g = _g1_ggx(np.abs(wo[:, 2]), 0.25) * _g1_ggx(np.abs(wi[:, 2]), 0.25)
    cc = np.where(nz, clearcoat * cc_d * cc_f * cc_g / 4, 0.0)[:, None] * np.ones(3) if clearcoat > 0 else 0.0 * diff
    fr_up = diffuse + spec + cc
    # transmission (MicrofacetTransmissionBtdf, etaA 1, etaB eta) and the inner reflection
    T = (1 - metallic) * transmission * np.sqrt(C)

    def btdf(wo, wi):
        ent = wo[:, 2] > 0
        ei = np.where(ent, 1.0, eta)
        et = np.where(ent, eta, 1.0)
        e = et / ei
        h = _normalize(wo + wi * e[:, None])
        h = _faceforward_z(h)
        o, i = _dot(wo, h), _dot(wi, h)
        f = fresnel_dielectric(o, 1.0, eta)   # Fresnel.Dielectric(1, eta).Eval: DielectricFunc orients by sign
        sd = o + e * i
        with np.errstate(divide="ignore", invalid="ignore"):
            val = ((1 - f)[:, None] * T) * np.abs(tr_dist.D(h) * tr_dist.G(wo, wi) * e * e * np.abs(i) * np.abs(o) /
                                                 (e * e) / (wi[:, 2] * wo[:, 2] * sd * sd))[:, None]
            h2 = _normalize(wo + wi * e[:, None])
            o2, i2 = _dot(wo, h2), _dot(wi, h2)
            sd2 = o2 + e * i2
            pdf = tr_dist.pdf(h2) * np.abs(e * e * i2 / (sd2 * sd2))
        bad = (o * i > 0) | (wo[:, 2] == 0) | (wi[:, 2] == 0) | _same(wo, wi)
        val = np.where(bad[:, None], 0.0, val)
        pdf = np.where((o2 * i2 > 0) | _same(wo, wi), 0.0, pdf)
        return val, pdf

    inner_R = transmission * C
    inner_fr = inner_R * mf_reflection_fr(wo, wi, tr_dist, lambda c: fresnel_dielectric(c, 1.0, eta))
    inner_pdf = mf_reflection_pdf(wo, wi, tr_dist)
    bt_fr, bt_pdf = btdf(wo, wi)
    same = _same(wo, wi)
    both_down = (wo[:, 2] < 0) & (wi[:, 2] < 0)
    fr = np.where(~same[:, None], bt_fr if transmission != 0 else 0.0,
                  np.where(both_down[:, None], inner_fr, fr_up))
    # Pdf
    w_d, w_t, w_s, w_c = disney_weights(C, metallic, transmission, clearcoat)
    with np.errstate(divide="ignore", invalid="ignore"):
        cc_pdf = np.where(same & nz, _gtr1(np.abs(whn[:, 2]), gloss) * np.abs(whn[:, 2]) / (4 * _dot(wo, whn)), 0.0)
    pdf_up = lambert_pdf(wo, wi) * w_d + mf_reflection_pdf(wo, wi, spec_dist) * w_s + cc_pdf * w_c
    fres_o = np.clip(fresnel_dielectric(wo[:, 2], 1.0, eta), 0.1, 0.9)
    pdf_down = np.where(wi[:, 2] > 0, bt_pdf * (1 - fres_o), inner_pdf * fres_o) if transmission != 0 else 0.0 * fres_o
    pdf = np.where(wo[:, 2] < 0, pdf_down, np.where(wi[:, 2] < 0, bt_pdf * w_t, pdf_up))
    return fr, pdf
