"""GPU parity of rd_adjoint (path-replay backpropagation): the gradient
against the oracle's adjoint on the same random numbers, against central
finite differences of the GPU render (h = 1e-3, float64 means and loss) on
256 random albedo texels, and statistically against finite differences on
roughness texels."""
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
    """Without Russian roulette the sample estimator of a rough-plastic wall
    is a polynomial in every albedo texel (degree = how often one path meets
    the texel, almost always 1 or 2), so central differences of the GPU
    render (common random numbers) are exact up to third-order terms.  The
    loss and the renders are taken in float64 (rd_render_mean_f64: float32
    sample radiance, float64 per-pixel sums and mean); h = 1e-3 is checked
    on every significant texel, with 5e-2 beside it."""
    spp, seed, res, tr = 16, 3, 64, 32
    desc, tex = scenes.diff_cornell(res, res, tex_res=tr, seed=1, rr=False)
    tdesc, _ = scenes.diff_cornell(res, res, tex_res=tr, seed=9, rr=False)
    gs = Scene(ctx, desc)
    target = Scene(ctx, tdesc).render_mean_f64(64, seed=77)
    img = gs.render_mean_f64(spp, seed=seed)
    base = _loss(img, target)
    dl = (2 * (img - target)).astype(np.float32)
    ids = [tex["left_r"], tex["right_r"]]
    g = gs.adjoint(spp, dl, ids, seed=seed)
    rng = np.random.default_rng(2)
    hs = (1e-3, 5e-2)
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
                lo.append(_loss(gs.render_mean_f64(spp, seed=seed), target))
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
    for h in hs:
        fd, big, sig, rel = stats[h]
        assert sig.sum() > 100, h
        # every significant texel within 1e-2 relative; the rest within 1e-3 of the largest gradient
        assert rel.max() < 1e-2 and np.median(rel) < 1e-3, (h, rel.max(), np.median(rel))
        assert np.all(np.abs(gvs - fd) <= 1e-2 * np.abs(fd) + 1e-3 * big), h


def test_roughness_gradient_matches_finite_differences_statistically(ctx):
    """Roughness moves the sampled directions, so a finite difference with
    common random numbers and the detached adjoint are different estimators
    of the same derivative; their means over independent seeds agree.  The
    functional is linear in the image (fixed weights dl from an unrelated
    render), so both estimate sum(dl * dE[image]/drho) without a
    noise-squared term.  Twelve roughness texels of the left and right walls
    (those with the largest gradients), six seeds of 2048 spp each, h = 1e-2,
    float64 image means."""
    res, tr, spp, h = 48, 8, 2048, 1e-2
    desc, tex = scenes.diff_cornell(res, res, tex_res=tr, seed=1, rr=False)
    tdesc, _ = scenes.diff_cornell(res, res, tex_res=tr, seed=9, rr=False)
    gs = Scene(ctx, desc)
    target = Scene(ctx, tdesc).render_mean_f64(256, seed=77)
    ref = gs.render_mean_f64(256, seed=78)
    dl = (2 * (ref - target)).astype(np.float32)
    ids = [tex["left_rough"], tex["right_rough"]]
    pick = gs.adjoint(spp, dl, ids, seed=99)
    cand = [(t, i) for t in ids for i in range(tr * tr)]
    cand.sort(key=lambda ti: -abs(pick[ti[0]][ti[1] * 3]))
    cand = cand[:12]
    print(f"[rough fd] {res}x{res} spp={spp}, h={h}, texels {cand}", flush=True)
    G, F = [], []
    for seed in range(6):
        g = gs.adjoint(spp, dl, ids, seed=seed)
        G.append([g[t][i * 3] for t, i in cand])
        row = []
        for t, i in cand:
            orig = desc.textures[t].image.copy()
            val = []
            for d in (h, -h):
                im = orig.copy()
                im[i // tr, i % tr, 0] += d
                gs.set_texture(t, im)
                val.append(float(np.dot(dl.astype(np.float64), gs.render_mean_f64(spp, seed=seed))))
            gs.set_texture(t, orig)
            row.append((val[0] - val[1]) / (2 * h))
        F.append(row)
        print(f"[rough fd] seed {seed}: adjoint {np.round(G[-1], 4).tolist()} fd {np.round(row, 4).tolist()}",
              flush=True)
    G, F = np.array(G, np.float64), np.array(F)
    gm, fm = G.mean(0), F.mean(0)
    se = np.sqrt(G.var(0, ddof=1) / len(G) + F.var(0, ddof=1) / len(F))
    z = (gm - fm) / se
    print(f"[rough fd] adjoint mean {np.round(gm, 4).tolist()}\n[rough fd] fd mean {np.round(fm, 4).tolist()}\n"
          f"[rough fd] z {np.round(z, 2).tolist()}", flush=True)
    assert np.all(np.abs(fm) > 3 * se), "texels too noisy to compare"
    assert np.all(np.abs(gm - fm) <= 4.5 * se + 2e-2 * np.abs(fm)), z
    # and over the twelve texels together, within a few percent
    assert abs(gm.sum() - fm.sum()) <= 0.03 * np.abs(fm).sum() + 4 * np.sqrt(np.sum(se ** 2))


def test_render_function_gradient_is_rd_adjoint(ctx):
    """The autograd Function's backward equals a direct rd_adjoint of the same
    dLoss/dImage (torch's L2 gradient), and its forward equals rd_render."""
    import torch
    from radiantium_amd.autograd import render
    spp, seed = 8, 5
    desc, tex = scenes.diff_cornell(32, 32, tex_res=16, seed=1)
    tdesc, _ = scenes.diff_cornell(32, 32, tex_res=16, seed=9)
    target, _ = Scene(ctx, tdesc).render(64, seed=77)
    gs = Scene(ctx, desc)
    ids = sorted(tex.values())
    params = {t: torch.tensor(desc.textures[t].image.reshape(-1), device="cuda", requires_grad=True) for t in ids}
    tgt = torch.tensor(target.reshape(32, 32, 3), device="cuda")
    img = render(gs, params, spp, seed=seed)
    loss = ((img - tgt) ** 2).sum()
    loss.backward()
    ref_img, _ = gs.render(spp, seed=seed, with_sum2=False)
    assert np.array_equal(img.detach().cpu().numpy().reshape(-1), ref_img.reshape(-1))
    dl = (2 * (ref_img.astype(np.float32) - target.astype(np.float32))).astype(np.float32)
    g = gs.adjoint(spp, dl, ids, seed=seed)
    for t in ids:
        got = params[t].grad.cpu().numpy()
        scale = np.abs(g[t]).max()
        assert scale > 0
        assert np.allclose(got, g[t], rtol=1e-5, atol=1e-6 * scale), t


def test_render_function_optimisation_recovers_albedo(ctx):
    """Adam through the autograd Function, from uniform 0.5 textures toward a
    render of random 8x8 textures: the albedo texels move toward the target's
    (the per-step loss at 64 spp is dominated by the render's own noise, so the
    texels, not the loss, are checked)."""
    import torch
    from radiantium_amd.autograd import render
    desc, tex = scenes.diff_cornell(48, 48, tex_res=8, init=(0.5, 0.5))
    tdesc, ttex = scenes.diff_cornell(48, 48, tex_res=8, seed=9)
    target, _ = Scene(ctx, tdesc).render(256, seed=77)
    tgt = torch.tensor(target.reshape(48, 48, 3), device="cuda")
    gs = Scene(ctx, desc)
    params = {t: torch.full((8 * 8 * 3,), 0.5, device="cuda", requires_grad=True) for t in tex.values()}
    alb = {tex[k]: tdesc.textures[ttex[k]].image.reshape(-1) for k in ("left_r", "right_r")}
    err0 = np.mean([np.abs(0.5 - a).mean() for a in alb.values()])
    opt = torch.optim.Adam(list(params.values()), lr=0.03)
    losses = []
    for k in range(16):
        opt.zero_grad()
        loss = ((render(gs, params, 64, seed=k) - tgt) ** 2).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            for x in params.values():
                x.clamp_(0.05, 1.0)
        losses.append(float(loss.detach()))
    err = np.mean([np.abs(params[t].detach().cpu().numpy() - a).mean() for t, a in alb.items()])
    print(f"[adjoint opt] albedo error {err0:.4f} -> {err:.4f}; losses {['%.2f' % v for v in losses]}", flush=True)
    assert err < 0.8 * err0
