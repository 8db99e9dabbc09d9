"""Inverse-material fitting over the path-replay adjoint (rpx_adjoint): Adam on an L2 image loss.

Each step renders the mean image of one sample range at the current parameters, forms
d loss / d image = 2 (image - target) / N, and replays a second, disjoint sample range for the
gradient (the same samples in both would bias the gradient of an L2 loss by the image's variance).
Parameters are the materials' (roughness, base colour r, g, b, IOR) and, optionally, one planar
base-colour texture; every step keeps them inside their bounds.
"""
import numpy as np

from . import native

# Bounds of (roughness, base r, g, b, IOR) and of a texel.
PARAM_LO = np.array([0.05, 0.01, 0.01, 0.01, 1.0], np.float32)
PARAM_HI = np.array([1.0, 0.99, 0.99, 0.99, 2.5], np.float32)
TEXEL_LO, TEXEL_HI = 0.01, 0.99


class Adam:
    """Adam (Kingma and Ba 2015) on one float32 array, in place."""

    def __init__(self, x, b1=0.9, b2=0.999, eps=1e-8):
        self.x, self.b1, self.b2, self.eps = x, b1, b2, eps
        self.m = np.zeros_like(x)
        self.v = np.zeros_like(x)
        self.t = 0

    def step(self, g, lr):
        self.t += 1
        self.m *= self.b1
        self.m += (1.0 - self.b1) * g
        self.v *= self.b2
        self.v += (1.0 - self.b2) * g * g
        mh = self.m / (1.0 - self.b1 ** self.t)
        vh = self.v / (1.0 - self.b2 ** self.t)
        self.x -= (lr * mh / (np.sqrt(vh) + self.eps)).astype(np.float32)


def fit(scene: native.Scene, target, params, texture=None, tex_material=None, origin=(0.0, 0.0), scale=(1.0, 1.0),
        steps=300, spp=16, lr=0.02, lr_texture=0.004, decay=0.1, frozen=(), first=0, callback=None):
    """Fits `params` ((materials, 5), float32) and `texture` ((h, w, 3) or None) so that the scene's
    image approaches `target` ((H, W, 3)). The learning rates fall geometrically to `decay` times
    their start over the steps; materials listed in `frozen` keep their parameters. Step k renders
    samples first + 2k·spp .. and replays first + (2k + 1)·spp ..; returns the fitted arrays and the
    loss of every step."""
    p = np.array(params, np.float32, copy=True)
    tex = None if texture is None else np.array(texture, np.float32, copy=True)
    target = np.asarray(target, np.float64)
    n = target.size
    opt_p = Adam(p)
    opt_t = None if tex is None else Adam(tex)
    losses = []
    for k in range(steps):
        scene.set_material_params(p)
        if tex is not None:
            scene.set_texture(tex_material, tex, origin, scale)
        scene.reset_film()
        base = first + 2 * k * spp
        scene.render_samples(base, spp)
        img = scene.resolve().astype(np.float64)
        diff = img - target
        losses.append(float(np.mean(diff * diff)))
        gm, gt = scene.adjoint(base + spp, spp, (2.0 * diff / n).astype(np.float32))
        for m in frozen:
            gm[m] = 0.0
        f = decay ** (k / max(steps, 1))
        opt_p.step(gm, lr * f)
        np.clip(p, PARAM_LO, PARAM_HI, out=p)
        if tex is not None and gt is not None:
            opt_t.step(gt, lr_texture * f)
            np.clip(tex, TEXEL_LO, TEXEL_HI, out=tex)
        if callback is not None:
            callback(k, p, tex, losses[-1])
    scene.set_material_params(p)
    if tex is not None:
        scene.set_texture(tex_material, tex, origin, scale)
    return p, tex, losses


def identifiable(params, metals=()):
    """The (material, parameter) pairs an image depends on, given the true `params`: a metal's IOR
    (its Fresnel term is its base colour's) and the roughness of a surface of IOR 1 without metal
    (no specular lobe) do not change what it reflects."""
    mask = np.ones(params.shape, bool)
    for m in range(params.shape[0]):
        if m in metals:
            mask[m, 4] = False
        elif params[m, 4] <= 1.0:
            mask[m, 0] = False
    return mask
