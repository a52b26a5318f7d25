"""GPU parity of rd_adjoint (path-replay backpropagation): the gradient
against the oracle's adjoint on the same random numbers, and against central
finite differences of the GPU render (h = 1e-3, loss in float64) on 256
random albedo texels."""
import numpy as np
import pytest

from radiantium_amd import scenes
from radiantium_amd.native import Scene

pytestmark = pytest.mark.gpu


def _loss(img, target):
    d = img.astype(np.float64) - target
    return float(np.sum(d * d))


@pytest.mark.parametrize("wall,filt,rr", [("rough_plastic", "Nearest", True), ("rough_plastic", "Linear", False),
                                          ("diffuse", "Nearest", True)])
def test_adjoint_twin_with_oracle(ctx, oracle, wall, filt, rr):
    spp, seed = 8, 4
    desc, tex = scenes.diff_cornell(32, 32, tex_res=16, seed=1, rr=rr, filter=filt, wall=wall)
    tdesc, _ = scenes.diff_cornell(32, 32, tex_res=16, seed=9, rr=rr, filter=filt, wall=wall)
    print(f"[adjoint twin] {wall} {filt} rr={rr}", flush=True)
    gs = Scene(ctx, desc)
    target, _ = Scene(ctx, tdesc).render(64, seed=77)
    img, _ = gs.render(spp, seed=seed)
    dl = (2 * (img.astype(np.float64) - target)).astype(np.float32)
    ids = list(tex.values())
    g = gs.adjoint(spp, dl, ids, seed=seed)
    o = oracle.OracleScene(desc).adjoint(spp, dl, ids, seed=seed, threads=8)
    for t in ids:
        scale = max(np.abs(o[t]).max(), 1e-12)
        close = np.isclose(g[t], o[t], rtol=1e-3, atol=1e-4 * scale)
        assert close.mean() > 0.97, (t, close.mean())
        assert abs(g[t].sum() - o[t].sum()) <= 1e-2 * np.abs(o[t]).sum()


def test_adjoint_matches_central_differences_256_texels(ctx):
    """Without Russian roulette the sample estimator of a diffuse wall is a
    polynomial in every albedo texel (degree = how often one path meets the
    texel, almost always 1 or 2), so central differences of the GPU render
    (common random numbers) are exact up to third-order terms.  h = 1e-3,
    1e-2 and 5e-2 are taken: with h = 1e-3 a texel seen only through indirect
    bounces moves a pixel's float32 sample sum by a few of its ulps, so that
    difference carries the image's rounding (its error falls as 1/h); the
    tight check is made at h = 5e-2, where albedos stay inside [0, 1]."""
    spp, seed, res, tr = 16, 3, 64, 32
    desc, tex = scenes.diff_cornell(res, res, tex_res=tr, seed=1, rr=False)
    tdesc, _ = scenes.diff_cornell(res, res, tex_res=tr, seed=9, rr=False)
    gs = Scene(ctx, desc)
    target, _ = Scene(ctx, tdesc).render(64, seed=77)
    target = target.astype(np.float64)
    img, _ = gs.render(spp, seed=seed, with_sum2=False)
    base = _loss(img, target)
    dl = (2 * (img.astype(np.float64) - target)).astype(np.float32)
    ids = [tex["left_r"], tex["right_r"]]
    g = gs.adjoint(spp, dl, ids, seed=seed)
    rng = np.random.default_rng(2)
    hs = (1e-3, 1e-2, 5e-2)
    fds, gvs = {h: [] for h in hs}, []
    print(f"[adjoint fd] {res}x{res} spp={spp} texture {tr}^2, loss {base:.4f}", flush=True)
    for k in range(256):
        t = ids[k % 2]
        x, y, c = rng.integers(0, tr), rng.integers(0, tr), rng.integers(0, 3)
        orig = desc.textures[t].image.copy()
        for h in hs:
            lo = []
            for d in (h, -h):
                im = orig.copy()
                im[x, y, c] += d
                gs.set_texture(t, im)
                r, _ = gs.render(spp, seed=seed, with_sum2=False)
                lo.append(_loss(r, target))
            fds[h].append((lo[0] - lo[1]) / (2 * h))
        gs.set_texture(t, orig)
        gvs.append(float(g[t][(x * tr + y) * 3 + c]))
    gvs = np.array(gvs)
    stats = {}
    for h in hs:
        fd = np.array(fds[h])
        big = np.abs(fd).max()
        sig = np.abs(fd) >= 1e-2 * big
        rel = np.abs(gvs[sig] - fd[sig]) / np.abs(fd[sig])
        stats[h] = (fd, big, sig, rel)
        worst = np.argsort(-np.abs(gvs - fd))[:4]
        print(f"[adjoint fd] h={h:g}: {sig.sum()} of 256 texels significant; rel err median {np.median(rel):.2e}, "
              f"99% {np.quantile(rel, 0.99):.2e}, max {rel.max():.2e}; max abs err elsewhere "
              f"{np.abs(gvs[~sig] - fd[~sig]).max() / big:.2e} of the largest gradient; worst (g, fd): "
              + ", ".join(f"({gvs[i]:.4g}, {fd[i]:.4g})" for i in worst), flush=True)
    fd, big, sig, rel = stats[5e-2]
    assert sig.sum() > 100
    assert rel.max() < 1e-2 and np.median(rel) < 1e-3
    assert np.all(np.abs(gvs - fd) <= 1e-2 * np.abs(fd) + 1e-3 * big)
    # h = 1e-3: the same agreement for the bulk, the tail bounded by rounding
    fd3, big3, sig3, rel3 = stats[1e-3]
    assert np.median(rel3) < 1e-3 and np.quantile(rel3, 0.9) < 1e-2
