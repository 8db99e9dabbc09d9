"""The differentiable render on the GPU, through the C ABI: textured and re-parameterised films
against the oracle (bit-exact), the path-replay adjoint against the oracle's CPU adjoint (the same
per-path terms; sums differ only in accumulation order), and against central finite differences
of the forward renderer on the inverse-material fit's scene (statistical: independent sample
ranges give each estimate an error bar), plus the torch.autograd wrapper."""
import numpy as np
import pytest

from oracle import oracle as orc
from repackshot_amd.rsb import SceneWriter, uv_sphere

import bench
import test_oracle_adjoint as lam

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))


def glossy_scene(depth=5):
    """A rough conductor, rough glass, a rough opaque dielectric and a textured Lambert floor, at a
    depth Russian roulette reaches."""
    w = SceneWriter()
    w.camera((250, -700, 330), (0, 0, 90), focal_mm=35.0)
    w.material(base_color=(0.9, 0.7, 0.5), metallic=1.0, roughness=0.3)
    w.material(base_color=(0.9, 0.92, 0.95), transmission=1.0, roughness=0.15, ior=1.5)
    w.material(base_color=(0.5, 0.5, 0.5), ior=1.0, roughness=1.0)
    w.material(base_color=(0.4, 0.6, 0.3), roughness=0.4, ior=1.45)
    p, n, i = uv_sphere((0, 0, 100), 100.0, 12, 24)
    w.mesh(p, i, n, 0)
    quad = np.array([[-900, -900, 0], [900, -900, 0], [900, 900, 0], [-900, 900, 0]], np.float32)
    w.mesh(quad, [0, 1, 2, 0, 2, 3], np.tile([0, 0, 1], (4, 1)), 2)
    w.instance(0, material=0)
    w.instance(0, rows=((0.5, 0, 0, -230), (0, 0.5, 0, 90), (0, 0, 0.5, 0)), material=1)
    w.instance(0, rows=((0.5, 0, 0, 230), (0, 0.5, 0, 120), (0, 0, 0.5, 0)), material=3)
    w.instance(1, material=2)
    w.light(position=(200, -300, 600), direction=(-0.2, 0.3, -1), size_x=300, size_y=300, power=6.0)
    w.environment(kind=0, floor=0.05, horizon=0.2, zenith=0.4)
    w.render(lam.W, lam.H, lam.SPP, 7, depth)
    return w.bytes()


@pytest.mark.parametrize("name", ["lambert", "glossy"])
def test_textured_films_and_adjoint_match_the_oracle(gpu, name):
    rsb = lam.lambert_scene() if name == "lambert" else glossy_scene()
    g, o = gpu.Scene(rsb), orc.OracleScene(rsb)
    tex = lam.texture()
    g.set_texture(2, tex, lam.TEX_ORIGIN, lam.TEX_SCALE)
    o.set_texture(2, tex, lam.TEX_ORIGIN, lam.TEX_SCALE)
    # Parameters moved away from the file's: both sides rebuild their BSDFs from them.
    params = o.material_params()
    params[:, 0] = np.clip(params[:, 0] * 0.9, 0.05, 1.0)
    params[:, 1:4] *= 0.95
    params[:, 4] = np.where(params[:, 4] > 1.01, params[:, 4] + 0.05, params[:, 4])
    g.set_material_params(params)
    o.set_material_params(params)
    assert np.array_equal(g.material_params(), o.material_params())
    g.render_samples(0, lam.SPP)
    of, _ = o.render_film(0, lam.SPP)
    assert all(bits_equal(a, b) for a, b in zip(g.film(), of))
    weights = np.random.default_rng(11).uniform(-1, 1, size=(lam.H, lam.W, 3)).astype(np.float32)
    gm, gt = g.adjoint(0, lam.SPP, weights)
    om, ot = o.adjoint(0, lam.SPP, weights, threads=8)
    scale_m, scale_t = np.max(np.abs(om)), np.max(np.abs(ot))
    assert scale_m > 0 and scale_t > 0
    assert np.max(np.abs(gm - om)) <= 1e-5 * scale_m, (gm, om)
    assert np.max(np.abs(gt - ot)) <= 1e-5 * scale_t
    # Parameters the image cannot depend on have exactly zero gradient.
    assert np.all(gm[2, 1:4] == 0.0)  # the textured floor's own base colour


FIT_PARAMS = [(m, 0) for m in range(6)] + [(m, 1) for m in range(8)] + [(4, 4), (5, 4)]


def fit_setup(gpu):
    s = gpu.Scene(gpu.write_scene(4, 0))
    rng = np.random.default_rng(3)
    tex = rng.uniform(0.2, 0.8, size=(256, 256, 3)).astype(np.float32)
    s.set_texture(8, tex, (-4000.0, -4000.0), (1.0 / 8000.0, 1.0 / 8000.0))
    return s, tex


def test_adjoint_agrees_with_finite_differences_on_the_fit_scene(gpu):
    """A fixed linear loss (seeded weights per pixel and channel), K independent sample ranges:
    the adjoint and the central difference (step 1e-2) of the forward render must agree within
    their combined error bars, and within 2% where the difference is resolved to 1%."""
    s, tex = fit_setup(gpu)
    spp = s.info()["max_samples"]
    K, h = 6, 1.0e-2
    # The image is piecewise smooth in roughness (energy-table nodes every 1/31, the compensation
    # from roughness 0.0447): the comparison is made in the middle of each roughness's cell.
    params = bench.cell_centred(s.material_params(), h)
    s.set_material_params(params)

    def image(first):
        s.reset_film()
        s.render_samples(first, spp)
        rgb, wt, _, _ = s.film()
        return rgb.astype(np.float64) * (1.0 / wt.astype(np.float64))[..., None]

    weights = np.random.default_rng(5).uniform(0.5, 1.5, size=(s.height, s.width, 3))
    A, F = [], []
    for k in range(K):
        first = 1000 * k
        gm, _ = s.adjoint(first, spp, weights.astype(np.float32))
        A.append([gm[m, j] for m, j in FIT_PARAMS])
        row = []
        for m, j in FIT_PARAMS:
            up, dn = params.copy(), params.copy()
            up[m, j] += h
            dn[m, j] -= h
            s.set_material_params(up)
            lu = float(np.sum(weights * image(first)))
            s.set_material_params(dn)
            ld = float(np.sum(weights * image(first)))
            row.append((lu - ld) / (2 * h))
        s.set_material_params(params)
        F.append(row)
    A, F = np.array(A), np.array(F)
    am, fm = A.mean(0), F.mean(0)
    sigma = np.hypot(A.std(0, ddof=1), F.std(0, ddof=1)) / np.sqrt(K)
    assert np.all(np.abs(am - fm) <= 4.0 * sigma + 0.02 * np.abs(fm)), (am, fm, sigma)
    resolved = sigma <= 0.01 * np.abs(fm)
    assert resolved.sum() >= 4
    assert np.all(np.abs(am - fm)[resolved] <= 0.02 * np.abs(fm)[resolved] + 2.0 * sigma[resolved])


def test_torch_function_gradients_and_a_fit_step(gpu):
    torch = pytest.importorskip("torch")
    from repackshot_amd import diff

    s, tex = fit_setup(gpu)
    spp = 4
    truth = torch.tensor(s.material_params())
    target_scene = diff.TexturedScene(s, 8, (-4000.0, -4000.0), (1.0 / 8000.0, 1.0 / 8000.0))
    with torch.no_grad():
        target = diff.render(target_scene, truth, torch.tensor(tex), first=5000, count=16)
    start = truth.clone()
    start[:8, 1:4] = (start[:8, 1:4] * 0.75).clamp(0.05, 0.95)
    params = start.clone().requires_grad_(True)
    texture = torch.tensor(np.clip(tex * 0.8, 0.05, 0.95).astype(np.float32), requires_grad=True)
    image = diff.render(target_scene, params, texture, first=0, count=spp)
    loss = ((image - target) ** 2).sum()
    loss.backward()
    # The wrapper hands back what the C ABI computes for d loss / d image = 2 (image - target) (the
    # sums' float atomics may add in another order).
    gm, gt = s.adjoint(0, spp, (2.0 * (image - target)).detach().numpy())
    assert np.allclose(params.grad.numpy(), gm, rtol=1e-4, atol=1e-4 * np.abs(gm).max())
    assert np.allclose(texture.grad.numpy(), gt, rtol=1e-4, atol=1e-4 * np.abs(gt).max())
    # Gradient descent on the base colours lowers the loss.
    opt = torch.optim.Adam([params, texture], lr=0.02)
    losses = []
    for it in range(6):
        opt.zero_grad()
        image = diff.render(target_scene, params, texture, first=100 * it, count=spp)
        loss = ((image - target) ** 2).sum()
        loss.backward()
        params.grad[:, 0] = 0.0  # roughness and IOR held; base colours and texels fitted
        params.grad[:, 4] = 0.0
        opt.step()
        with torch.no_grad():
            params[:, 1:4].clamp_(0.02, 0.98)
            texture.clamp_(0.02, 0.98)
        losses.append(float(loss.detach()))
    assert losses[-1] < 0.8 * losses[0], losses
