"""Differentiable rendering over the C-ABI: a torch autograd Function whose
forward is rd_render_device and whose backward is rd_adjoint_device
(path-replay backpropagation of dLoss/dImage into texture texels).

The integrator is the one of Integrators/PathTracer.cs:26-413 that
IIntegrator.Li (IIntegrator.cs:13-24) drives; forward renders with the
textures given as device tensors, backward replays the same paths (same spp,
seed and crop) and scatters each vertex's contribution into the texels it
read.  Sampling, MIS weights and Russian roulette are treated as constants
(detached), so the gradient is that of the sample estimator itself.

Textures are float32 CUDA tensors of the scene texture's size (w*h*3, any
shape with that many elements, [x][y][c] order).  A roughness texture
carries its gradient in channel 0 (the only channel RoughPlastic reads,
Materials/RoughPlastic.cs:42-45).  torch is plumbing here: memory and the
optimiser; every pixel and every gradient is computed by the HIP kernels.
"""
import torch

from .native import Scene


def _check_texture(scene, t, x):
    img = scene.desc.textures[t].image
    if img is None:
        raise ValueError(f"texture {t} is a constant, not a texel image")
    if not (x.is_cuda and x.dtype == torch.float32 and x.is_contiguous()):
        raise ValueError(f"texture {t}: a contiguous float32 CUDA tensor is required")
    if x.numel() != img.size:
        raise ValueError(f"texture {t}: {x.numel()} elements, the scene texture has {img.size}")


def _upload(scene, tex_ids, textures):
    torch.cuda.current_stream().synchronize()
    for t, x in zip(tex_ids, textures):
        _check_texture(scene, t, x)
        scene.set_texture_device(t, x.data_ptr())


class RenderFunction(torch.autograd.Function):
    """image = RenderFunction.apply(scene, spp, seed, tex_ids, *textures)

    `image` is [W, H, 3] float32 (the ColorBuffer's column-major layout,
    ColorBuffer.cs:35), the mean of spp samples per pixel.  `tex_ids` lists
    the scene texture index of each tensor in `textures`."""

    @staticmethod
    def forward(fctx, scene, spp, seed, tex_ids, *textures):
        if not isinstance(scene, Scene):
            raise TypeError("scene must be a radiantium_amd.native.Scene")
        tex_ids = tuple(int(t) for t in tex_ids)
        if len(tex_ids) != len(textures):
            raise ValueError("one texture index per texture tensor")
        _upload(scene, tex_ids, textures)
        img = torch.empty(scene.width * scene.height * 3, dtype=torch.float32, device=textures[0].device
                          if textures else "cuda")
        torch.cuda.current_stream().synchronize()
        scene.render_device(spp, img.data_ptr(), seed=seed)
        scene.ctx.synchronize()
        fctx.scene, fctx.spp, fctx.seed, fctx.tex_ids = scene, spp, seed, tex_ids
        fctx.save_for_backward(*textures)
        return img.view(scene.width, scene.height, 3)

    @staticmethod
    def backward(fctx, grad_img):
        scene, tex_ids = fctx.scene, fctx.tex_ids
        textures = fctx.saved_tensors
        # the replay must see the textures the forward rendered with
        _upload(scene, tex_ids, textures)
        dl = grad_img.detach().to(torch.float32).contiguous()
        grads = [torch.zeros_like(x) for x in textures]
        ptrs = [None] * len(scene.desc.textures)
        for t, g in zip(tex_ids, grads):
            ptrs[t] = g.data_ptr()
        torch.cuda.current_stream().synchronize()
        scene.adjoint_device(fctx.spp, dl.data_ptr(), ptrs, seed=fctx.seed)
        scene.ctx.synchronize()
        return (None, None, None, None, *grads)


def render(scene, textures, spp, seed=0):
    """Differentiable render: textures = {scene texture index: tensor}."""
    ids = list(textures.keys())
    return RenderFunction.apply(scene, spp, seed, ids, *[textures[t] for t in ids])
