"""Convergence: the GPU render against a 4096-spp CPU image accumulated in float64 (the oracle's
convergence reference). At the same sample indices the two differ only by the film's f32 sums; at
independent sample indices the GPU image's error must fall like the Monte Carlo noise of both
images, sqrt(a / n + a / 4096), i.e. the GPU converges to the oracle's image, not near it."""
import numpy as np
import pytest

from oracle import oracle as orc
from repackshot_amd.rsb import SceneWriter, uv_sphere

pytestmark = pytest.mark.gpu

W, H, REF_SPP = 64, 48, 4096


def scene():
    w = SceneWriter()
    w.camera((250, -700, 330), (0, 0, 90), focal_mm=35.0)
    w.material(base_color=(0.9, 0.7, 0.5), metallic=1.0, roughness=0.3)
    w.material(base_color=(0.9, 0.92, 0.95), transmission=1.0, roughness=0.15, ior=1.5)
    w.material(base_color=(0.6, 0.6, 0.6), ior=1.0, roughness=1.0)
    w.material(base_color=(0.4, 0.6, 0.3), roughness=0.4, ior=1.45)
    p, n, i = uv_sphere((0, 0, 100), 100.0, 16, 32)
    w.mesh(p, i, n, 0)
    quad = np.array([[-900, -900, 0], [900, -900, 0], [900, 900, 0], [-900, 900, 0]], np.float32)
    w.mesh(quad, [0, 1, 2, 0, 2, 3], np.tile([0, 0, 1], (4, 1)), 2)
    w.instance(0, material=0)
    w.instance(0, rows=((0.5, 0, 0, -230), (0, 0.5, 0, 90), (0, 0, 0.5, 0)), material=1)
    w.instance(0, rows=((0.5, 0, 0, 230), (0, 0.5, 0, 120), (0, 0, 0.5, 0)), material=3)
    w.instance(1, material=2)
    w.light(position=(200, -300, 600), direction=(-0.2, 0.3, -1), size_x=300, size_y=300, power=6.0)
    w.environment(kind=0, floor=0.05, horizon=0.2, zenith=0.4)
    w.render(W, H, REF_SPP, 3, 8)
    return w.bytes()


def rrmse(a, ref):
    return float(np.sqrt(np.mean((a - ref) ** 2)) / np.sqrt(np.mean(ref ** 2)))


@pytest.fixture(scope="module")
def reference():
    rsb = scene()
    return rsb, orc.OracleScene(rsb).render_f64(0, REF_SPP)


def gpu_image(gpu, rsb, first, count):
    s = gpu.Scene(rsb)
    s.render_samples(first, count)
    return s.resolve().astype(np.float64)


def test_same_samples_match_the_float64_reference(gpu, reference):
    rsb, ref = reference
    img = gpu_image(gpu, rsb, 0, REF_SPP)
    assert rrmse(img, ref) <= 1e-4


def test_independent_samples_converge_to_the_reference(gpu, reference):
    rsb, ref = reference
    errs = {n: rrmse(gpu_image(gpu, rsb, 1_000_000, n), ref) for n in (16, 64, 256, 1024)}
    # a from the noisiest image; every other one must sit on sqrt(a / n + a / 4096).
    a = errs[16] ** 2 / (1.0 / 16 + 1.0 / REF_SPP)
    for n, e in errs.items():
        model = np.sqrt(a / n + a / REF_SPP)
        assert 0.7 * model <= e <= 1.3 * model, (n, e, model, errs)
    assert errs[1024] < errs[16] / 3.0
