"""The oracle's BSDF and lights against textbook closed forms and the invariants the reference's
own BSDF tests check (bsdf/tests.rs:60-113, 196, 277, 361; integrator/tests.rs:183-282): the
exact dielectric Fresnel term, reciprocity, the sample weight being f·cos/pdf, the pdf being a
density, no energy gained, a white metal white in a furnace, and a rect light's analytic
irradiance."""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from repackshot_amd.rsb import SceneWriter, icosphere

GLASS = 1


def textbook_fresnel(cos_i, eta):
    """Unpolarised reflectance of a smooth interface, relative index eta = n_t / n_i (Hecht)."""
    sin_t2 = (1.0 - cos_i * cos_i) / (eta * eta)
    if sin_t2 >= 1.0:
        return 1.0
    cos_t = math.sqrt(1.0 - sin_t2)
    rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t)
    rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t)
    return 0.5 * (rs * rs + rp * rp)


def test_fresnel_is_the_textbook_dielectric_term():
    # bsdf/tests.rs:277: 4 % at normal incidence for glass, total internal reflection from inside.
    assert abs(orc.fresnel_dielectric(1.0, 1.5) - 0.04) < 1e-6
    for eta in (1.33, 1.5, 2.4, 1 / 1.5, 1 / 1.33):
        for cos_i in (1.0, 0.9, 0.7, 0.5, 0.3, 0.1, 0.02):
            want = textbook_fresnel(cos_i, eta)
            assert abs(orc.fresnel_dielectric(cos_i, eta) - want) < 2e-6 + 1e-5 * want, (eta, cos_i)
    # Beyond the critical angle of glass seen from inside (41.8 degrees): all reflected.
    assert orc.fresnel_dielectric(math.cos(math.radians(45.0)), 1 / 1.5) == 1.0


def unit(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v)).astype(np.float32)


def hemisphere_dirs(n, rng, lower=False):
    out = []
    while len(out) < n:
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        if (v[2] > 0.05) != lower and abs(v[2]) > 0.05:
            out.append(unit(v))
    return out


REFLECTORS = {
    "lambert": dict(base=(0.7, 0.5, 0.3), ior=1.0),
    "rough metal": dict(base=(0.95, 0.64, 0.54), metallic=1.0, roughness=0.3),
    "very rough metal": dict(base=(0.9, 0.9, 0.9), metallic=1.0, roughness=0.9),
    "plastic": dict(base=(0.2, 0.6, 0.8), roughness=0.45, ior=1.5),
    "coated plastic": dict(base=(0.8, 0.2, 0.2), roughness=0.6, ior=1.5, coat=1.0, coat_roughness=0.25),
}


@pytest.mark.parametrize("name", sorted(REFLECTORS))
def test_reflection_is_reciprocal(name):
    # bsdf/tests.rs:60-113: f(wo, wi) = f(wi, wo) for the reflecting lobes, the energy terms included.
    b = orc.Bsdf(**REFLECTORS[name])
    rng = np.random.default_rng(3)
    for wo, wi in zip(hemisphere_dirs(60, rng), hemisphere_dirs(60, rng)):
        a, c = b.eval(wo, wi), b.eval(wi, wo)
        assert np.allclose(a, c, rtol=2e-5, atol=1e-7), (name, wo, wi, a, c)


SAMPLED = dict(REFLECTORS, **{
    "rough glass from outside": dict(kind=GLASS, base=(1, 1, 1), roughness=0.35, ior=1.5, transmission=1.0),
    "rough glass from inside": dict(kind=GLASS, base=(1, 1, 1), roughness=0.35, ior=1.5, transmission=1.0,
                                    entering=False),
})


@pytest.mark.parametrize("name", sorted(SAMPLED))
def test_sample_weight_is_f_cos_over_pdf(name):
    # bsdf/tests.rs:196: a sampled direction's weight is eval * |cos| / pdf at that direction.
    b = orc.Bsdf(**SAMPLED[name])
    rng = np.random.default_rng(11)
    checked = 0
    for wo in hemisphere_dirs(12, rng):
        for u0, u1 in rng.random((40, 2)):
            s = b.sample(wo, float(u0), float(u1))
            if s is None:
                continue
            wi, weight, pdf = s
            f = b.eval(wo, wi)
            p = b.pdf(wo, wi)
            assert abs(p / pdf - 1.0) < 2e-3, (name, p, pdf)
            want = f * abs(float(wi[2])) / p
            assert np.allclose(weight, want, rtol=3e-3, atol=1e-6), (name, weight, want)
            checked += 1
    assert checked > 300


@pytest.mark.parametrize("name", ["lambert", "rough metal", "plastic", "rough glass from outside",
                                  "rough glass from inside"])
def test_pdf_integrates_to_one_over_the_sphere(name):
    # bsdf/tests.rs:361: the pdf is a density over all directions (reflection and transmission),
    # midpoint rule in (cos theta, phi) on a 300 x 400 grid.
    b = orc.Bsdf(**SAMPLED[name])
    wo = unit([0.35, -0.2, 0.9])
    nz, nphi = 300, 400
    total = 0.0
    for i in range(nz):
        z = -1.0 + (i + 0.5) * 2.0 / nz
        r = math.sqrt(max(1.0 - z * z, 0.0))
        for j in range(nphi):
            ph = (j + 0.5) * 2.0 * math.pi / nphi
            total += b.pdf(wo, np.array([r * math.cos(ph), r * math.sin(ph), z], np.float32))
    total *= (2.0 / nz) * (2.0 * math.pi / nphi)
    assert abs(total - 1.0) < 0.02, (name, total)


@pytest.mark.parametrize("name,white", [("lambert", 1.0), ("very rough metal", None), ("rough metal", None)])
def test_no_energy_is_gained_and_white_stays_white(name, white):
    # The mean sample weight is the directional albedo (stratified over a 80 x 80 grid): exactly
    # one for a white Lambert; for a white metal one within 2 % at every angle (the Kulla-Conty
    # lobe restores what single scattering loses; its 16-bit tables are interpolated), and never
    # above that.
    params = dict(REFLECTORS[name], base=(1.0, 1.0, 1.0))
    b = orc.Bsdf(**params)
    rng = np.random.default_rng(5)
    n = 80
    grid = [((i + 0.5) / n, (j + 0.5) / n) for i in range(n) for j in range(n)]
    for wo in hemisphere_dirs(6, rng):
        ws = []
        for u0, u1 in grid:
            s = b.sample(wo, u0, u1)
            ws.append(0.0 if s is None else float(s[1][1]))
        mean = float(np.mean(ws))
        if white is not None:
            assert abs(mean - white) < 1e-5, (name, mean)
        else:
            assert abs(mean - 1.0) < 0.02, (name, wo, mean)


def uniform_world(w):
    w.environment(kind=1, horizon=1.0, strength=1.0, background_mode=0)


def test_rough_metal_never_gains_energy_in_a_furnace():
    # integrator/tests.rs:183-224: a white metal sphere in a uniform world of radiance 1 is white
    # (0.99..1.005) at every roughness, and no pixel exceeds 1.03.
    p, i = icosphere((0, 0, 0), 100.0, 3)
    for roughness in (0.05, 0.3, 1.0):
        w = SceneWriter()
        w.camera((0, -1000.0, 0), (0, 0, 0), up=(0, 0, 1), ortho=True, ortho_width=220.0)
        w.material(base_color=(1, 1, 1), metallic=1.0, roughness=roughness)
        w.mesh(p, i, None, 0)
        uniform_world(w)
        w.render(16, 16, 256, 0, 8)
        img = orc.OracleScene(w.bytes()).render_f64(0, 256)
        yy, xx = np.indices((16, 16))
        inside = np.hypot((xx + 0.5) / 16 - 0.5, (yy + 0.5) / 16 - 0.5) * 220.0 < 90.0
        mean = float(img[..., 1][inside].mean())
        assert 0.99 < mean <= 1.005, (roughness, mean)
        assert float(img.max()) <= 1.03, roughness


def rect_irradiance_factor(a, b, h):
    """E / L below the centre of an a x b Lambertian rectangle at height h, parallel: pi times
    four corner view factors of a differential element parallel to a rectangle at a corner (the
    textbook closed form, e.g. Howell's catalogue of view factors)."""
    x, y = 0.5 * a / h, 0.5 * b / h
    corner = (x / math.sqrt(1 + x * x) * math.atan(y / math.sqrt(1 + x * x))
              + y / math.sqrt(1 + y * y) * math.atan(x / math.sqrt(1 + y * y))) / (2 * math.pi)
    return 4 * math.pi * corner


def test_a_rect_light_gives_the_analytic_irradiance():
    # integrator/tests.rs:258-282: a 400 x 200 mm panel of 8 W, 300 mm above a white floor, seen
    # at a 2 mm spot straight below: radiance L F / pi within 2 %.
    a, b, h, power = 400.0, 200.0, 300.0, 8.0
    w = SceneWriter()
    w.camera((0, 0, 2000.0), (0, 0, 0), up=(0, 1, 0), ortho=True, ortho_width=2.0)
    w.material(base_color=(1, 1, 1), ior=1.0)
    q = np.array([[-1000, -1000, 0], [1000, -1000, 0], [1000, 1000, 0], [-1000, 1000, 0]], np.float32)
    w.mesh(q, [0, 1, 2, 0, 2, 3], None, 0)
    w.light(kind=0, position=(0, 0, h), direction=(0, 0, -1), up=(0, 1, 0), size_x=a, size_y=b, power=power,
            flags=1)
    w.environment(kind=0, strength=0.0, background_color=(0, 0, 0), background_mode=1)
    w.render(8, 8, 4096, 0, 8)
    got = float(orc.OracleScene(w.bytes()).render_f64(0, 4096)[..., 1].mean())
    radiance = power / (math.pi * a * b * 1e-6)
    want = radiance * rect_irradiance_factor(a, b, h) / math.pi
    assert abs(got / want - 1.0) < 0.02, (got, want)


# ---------------------------------------------------------------- coloured conductors
# A metal's specular lobe is GGX with height-correlated Smith masking and Schlick's Fresnel per
# channel, F0 the base colour and F90 = min(1, 50 F0) (bsdf.rs:195-200, 416-420, 603-612); the
# Kulla-Conty lobe adds back what single scattering loses, in the colour F_avg^2 E_avg /
# (1 - F_avg (1 - E_avg)) with F_avg the hemispherical average of Schlick's term
# (energy.rs:500-514, 545-554). Both are restated here in float64 from the published formulas
# (Walter et al. 2007, Heitz 2014, Schlick 1994, Kulla and Conty 2017), with E(mu) integrated
# numerically, independent of the oracle's tables.
COLOURED_F0 = [(0.95, 0.64, 0.54), (0.01, 0.2, 0.005), (0.04, 0.9, 0.3)]


def ggx_terms(wo, wi, alpha):
    """D G2 / (4 mu_o mu_i) and the microfacet cosine wo . h, float64 (wo, wi: (..., 3))."""
    wo = np.asarray(wo, np.float64)
    wi = np.asarray(wi, np.float64)
    h = wo + wi
    h = h / np.linalg.norm(h, axis=-1, keepdims=True)
    a2 = alpha * alpha
    c2 = h[..., 2] ** 2
    d = a2 / (math.pi * ((a2 - 1.0) * c2 + 1.0) ** 2)

    def lam(w):
        z2 = w[..., 2] ** 2
        return (np.sqrt(1.0 + a2 * (1.0 - z2) / z2) - 1.0) * 0.5

    g2 = 1.0 / (1.0 + lam(wo) + lam(wi))
    cos_h = np.maximum(np.sum(wo * h, axis=-1), 0.0)
    return d * g2 / (4.0 * wo[..., 2] * wi[..., 2]), cos_h


def schlick(f0, cos):
    f90 = min(1.0, 50.0 * f0)
    return f0 + (f90 - f0) * (1.0 - cos) ** 5


def hemisphere_grid(n_mu, n_phi):
    mu = (np.arange(n_mu) + 0.5) / n_mu
    phi = (np.arange(n_phi) + 0.5) * 2.0 * math.pi / n_phi
    m, p = np.meshgrid(mu, phi, indexing="ij")
    s = np.sqrt(1.0 - m * m)
    dirs = np.stack([s * np.cos(p), s * np.sin(p), m], axis=-1).reshape(-1, 3)
    weight = (1.0 / n_mu) * (2.0 * math.pi / n_phi)  # d omega = d mu d phi
    return dirs, weight


def textbook_albedo(mu_o, alpha, f0=None, grid=(400, 800)):
    """Directional albedo of the single-scattering lobe towards mu_o: with Fresnel one (f0 None),
    or with Schlick's term of f0 (midpoint rule on a mu x phi grid)."""
    dirs, dw = hemisphere_grid(*grid)
    wo = np.array([math.sqrt(1.0 - mu_o * mu_o), 0.0, mu_o])
    f, cos_h = ggx_terms(np.broadcast_to(wo, dirs.shape), dirs, alpha)
    if f0 is not None:
        f = f * schlick(f0, cos_h)
    return float(np.sum(f * dirs[:, 2]) * dw)


def test_coloured_conductor_single_scattering_is_textbook_ggx():
    rng = np.random.default_rng(21)
    worst = 0.0
    for f0 in COLOURED_F0:
        for roughness in (0.1, 0.3, 0.6, 0.9):
            b = orc.Bsdf(base=f0, metallic=1.0, roughness=roughness)
            alpha = max(roughness * roughness, 1e-3)
            for wo, wi in zip(hemisphere_dirs(40, rng), hemisphere_dirs(40, rng)):
                got = b.eval(wo, wi, single=True).astype(np.float64)
                ft, cos_h = ggx_terms(wo[None], wi[None], alpha)
                want = np.array([schlick(c, float(cos_h[0])) for c in f0]) * float(ft[0])
                err = np.abs(got - want) / np.maximum(want, 1e-30)
                worst = max(worst, float(err.max()))
                assert np.allclose(got, want, rtol=2e-4, atol=1e-9), (f0, roughness, wo, wi, got, want)
    assert worst > 0.0


def test_coloured_conductor_multiple_scattering_is_kulla_conty():
    rng = np.random.default_rng(22)
    for roughness in (0.3, 0.6):
        alpha = roughness * roughness
        mus = (np.arange(100) + 0.5) / 100
        e = np.array([textbook_albedo(float(m), alpha) for m in mus])
        e_avg = float(2.0 * np.sum(e * mus) / len(mus))
        for f0 in COLOURED_F0:
            b = orc.Bsdf(base=f0, metallic=1.0, roughness=roughness)
            f_avg = np.array([c + (min(1.0, 50.0 * c) - c) / 21.0 for c in f0])
            f_ms = f_avg * f_avg * e_avg / (1.0 - f_avg * (1.0 - e_avg))
            for wo, wi in zip(hemisphere_dirs(25, rng), hemisphere_dirs(25, rng)):
                ms = (b.eval(wo, wi) - b.eval(wo, wi, single=True)).astype(np.float64)
                # The lobe's colour: the channels' ratios are the Kulla-Conty colour's.
                assert np.allclose(ms / ms.max(), f_ms / f_ms.max(), rtol=2e-3, atol=1e-6), (f0, ms, f_ms)
                # Its shape: (1 - E(mu_o)) (1 - E(mu_i)) / (pi (1 - E_avg)), E integrated here.
                eo = np.interp(float(wo[2]), mus, e)
                ei = np.interp(float(wi[2]), mus, e)
                shape = (1.0 - eo) * (1.0 - ei) / (math.pi * (1.0 - e_avg))
                assert np.allclose(ms, f_ms * shape, rtol=0.06, atol=2e-4), (f0, roughness, ms, f_ms * shape)


def test_coloured_conductor_albedo_per_channel():
    # The mean sample weight per channel (stratified 80 x 80) is the compensated directional
    # albedo: single scattering with Schlick's Fresnel plus F_ms (1 - E(mu_o)), within 2 %, and
    # lies between F0 and 1 (bsdf/tests.rs:60-113 for white; here each channel its own colour).
    rng = np.random.default_rng(23)
    n = 80
    grid = [((i + 0.5) / n, (j + 0.5) / n) for i in range(n) for j in range(n)]
    roughness = 0.3
    alpha = roughness * roughness
    mus = (np.arange(100) + 0.5) / 100
    e_avg = float(2.0 * np.sum([textbook_albedo(float(m), alpha) * m for m in mus]) / len(mus))
    for f0 in COLOURED_F0[:2]:
        b = orc.Bsdf(base=f0, metallic=1.0, roughness=roughness)
        f_avg = np.array([c + (min(1.0, 50.0 * c) - c) / 21.0 for c in f0])
        f_ms = f_avg * f_avg * e_avg / (1.0 - f_avg * (1.0 - e_avg))
        for wo in hemisphere_dirs(3, rng):
            ws = np.zeros(3)
            for u0, u1 in grid:
                s = b.sample(wo, u0, u1)
                if s is not None:
                    ws += s[1]
            mean = ws / len(grid)
            mu = float(wo[2])
            want = np.array([textbook_albedo(mu, alpha, c) for c in f0]) + f_ms * (1.0 - textbook_albedo(mu, alpha))
            assert np.allclose(mean, want, rtol=0.02, atol=1e-3), (f0, mu, mean, want)
            assert np.all(mean >= np.array(f0) * 0.98) and np.all(mean <= 1.0 + 1e-3), (f0, mean)


def test_coloured_metal_sphere_in_a_furnace_has_the_textbook_albedo():
    # integrator/tests.rs:183-224 with the configs' coloured conductor: in a uniform world of
    # radiance 1 a convex metal body shows its compensated directional albedo. Over the disc the
    # camera sees (radius 0.9 of the sphere's, uniform in area) each channel's mean is the
    # textbook albedo's mean within 3 %, in F0's order. (Below F0 at roughness 1: light that
    # scatters between microfacets meets the Fresnel term more than once.)
    p, i = icosphere((0, 0, 0), 100.0, 3)
    f0 = (0.95, 0.64, 0.54)
    rs = (np.arange(24) + 0.5) / 24 * 0.9
    mu_of_r = np.sqrt(1.0 - rs * rs)
    area = 2.0 * rs / 0.81 * (0.9 / 24)
    for roughness in (0.05, 0.3, 1.0):
        w = SceneWriter()
        w.camera((0, -1000.0, 0), (0, 0, 0), up=(0, 0, 1), ortho=True, ortho_width=220.0)
        w.material(base_color=f0, metallic=1.0, roughness=roughness)
        w.mesh(p, i, None, 0)
        uniform_world(w)
        w.render(16, 16, 256, 0, 8)
        img = orc.OracleScene(w.bytes()).render_f64(0, 256)
        yy, xx = np.indices((16, 16))
        inside = np.hypot((xx + 0.5) / 16 - 0.5, (yy + 0.5) / 16 - 0.5) * 220.0 < 90.0
        means = [float(img[..., c][inside].mean()) for c in range(3)]
        alpha = roughness * roughness
        if roughness < 0.1:
            # A near mirror: Schlick's term at the view angle.
            want = [float(np.sum([schlick(c, m) for m in mu_of_r] * area)) for c in f0]
        else:
            mus = (np.arange(60) + 0.5) / 60
            e = np.array([textbook_albedo(float(m), alpha, grid=(200, 400)) for m in mus])
            e_avg = float(2.0 * np.sum(e * mus) / len(mus))
            want = []
            for c in f0:
                f_avg = c + (min(1.0, 50.0 * c) - c) / 21.0
                f_ms = f_avg * f_avg * e_avg / (1.0 - f_avg * (1.0 - e_avg))
                alb = [textbook_albedo(float(m), alpha, c, grid=(200, 400)) + f_ms * (1.0 - np.interp(m, mus, e))
                       for m in mu_of_r]
                want.append(float(np.sum(np.array(alb) * area)))
        for c in range(3):
            assert abs(means[c] / want[c] - 1.0) < 0.03, (roughness, means, want)
        assert means[0] > means[1] > means[2], (roughness, means)
