"""The oracle's float64 convergence film against its own f32 film at the same samples (they
differ only by the f32 sums), and the sample mean's error falling as 1 / sqrt(n)."""
import numpy as np

from oracle import oracle as orc

import test_gpu_convergence as conv


def test_float64_film_agrees_with_the_f32_film():
    rsb = conv.scene()
    o = orc.OracleScene(rsb)
    ref = o.render_f64(0, 64, threads=8)
    film, _ = o.render_film(0, 64, threads=8)
    img = film[0] * (np.float32(1.0) / film[1])[..., None]
    assert conv.rrmse(img.astype(np.float64), ref) <= 1e-6


def test_oracle_noise_falls_as_one_over_sqrt_n():
    o = orc.OracleScene(conv.scene())
    ref = o.render_f64(100_000, 256, threads=8)
    e4 = conv.rrmse(o.render_f64(0, 4, threads=8), ref)
    e64 = conv.rrmse(o.render_f64(0, 64, threads=8), ref)
    model = lambda n: np.sqrt(1.0 / n + 1.0 / 256)
    ratio = (e4 / e64) / (model(4) / model(64))
    assert 0.7 <= ratio <= 1.3, (e4, e64)
