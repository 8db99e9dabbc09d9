"""A thin torch.autograd.Function over the tracer's differentiable render (rpx_adjoint).

The forward renders the mean image of samples [first, first + count) at the given material
parameters (materials x 5: GGX roughness, base colour r, g, b, IOR) and base-colour texture; the
backward hands d loss / d image to the path-replay adjoint, which returns d loss / d parameter
and d loss / d texel. Tensors cross the boundary through host memory (the image and the gradients
are small next to the render), and the GPU work runs on the scene's own stream.
"""
import numpy as np
import torch

from . import native


class TexturedScene:
    """A loaded scene plus where its differentiable texture goes (material, planar mapping)."""

    def __init__(self, scene: native.Scene, tex_material=None, origin=(0.0, 0.0), scale=(1.0, 1.0)):
        self.scene = scene
        self.tex_material = tex_material
        self.origin = origin
        self.scale = scale


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, texture, target, first, count):
        sc = target.scene
        sc.set_material_params(params.detach().to("cpu", torch.float32).numpy())
        if texture is not None:
            sc.set_texture(target.tex_material, texture.detach().to("cpu", torch.float32).numpy(), target.origin,
                           target.scale)
        sc.reset_film()
        sc.render_samples(first, count)
        image = sc.resolve()
        ctx.target, ctx.first, ctx.count = target, first, count
        ctx.has_texture = texture is not None
        ctx.devices = (params.device, texture.device if texture is not None else None)
        return torch.from_numpy(image).to(params.device)

    @staticmethod
    def backward(ctx, grad_image):
        sc = ctx.target.scene
        gm, gt = sc.adjoint(ctx.first, ctx.count, grad_image.detach().to("cpu", torch.float32).numpy())
        g_params = torch.from_numpy(gm).to(ctx.devices[0])
        g_tex = torch.from_numpy(gt).to(ctx.devices[1]) if ctx.has_texture and gt is not None else None
        return g_params, g_tex, None, None, None


def render(target: TexturedScene, params, texture=None, first=0, count=1):
    """The differentiable image (H, W, 3) of `target` at `params` (and `texture`)."""
    return RenderFunction.apply(params, texture, target, first, count)


def linear_gradient(scene: native.Scene, weights, first, count):
    """d (sum weights * image) / d parameters: the adjoint of a fixed linear loss."""
    return scene.adjoint(first, count, np.ascontiguousarray(weights, np.float32))
