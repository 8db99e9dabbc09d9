"""The oracle's scattering media and sphere/spot lights against closed forms.

The reference pins its media (integrator/medium.rs) and sphere lights (light.rs) by physics, not by
hashes: a white medium in a white furnace is invisible, a grey half-space reflects Chandrasekhar's
plane albedo, an absorbing medium is Beer-Lambert per channel, a sparse slab passes the
unscattered light straight, a point or sphere light gives I / d2 and a spot fades by its
smoothstep. These are the reference's own tests (integrator/tests.rs:310-348, 1663-1999;
light.rs:1813-1898), restated on the oracle at the same tolerances; a dropped Fresnel term, a wrong
phase-function sign or a lost MIS weight fails one of them."""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from repackshot_amd.rsb import SceneWriter, box_mesh, icosphere

GLASS, EMISSIVE = 1, 2
POINT, SPOT = 3, 4
GLOSSY, CAMERA = 1, 2


def fresnel_dielectric(eta):
    """Unpolarised Fresnel reflectance at normal incidence, (eta - 1)^2 / (eta + 1)^2."""
    return ((eta - 1.0) / (eta + 1.0)) ** 2


def mean_of(rsb, spp, channel=1, where=None):
    img = orc.OracleScene(rsb).render_f64(0, spp)[..., channel]
    return float(img.mean() if where is None else img[where(*np.indices(img.shape)[::-1])].mean())


def ortho_down(w, x, y, width):
    w.camera((x, y, 2000.0), (x, y, 0.0), up=(0, 1, 0), ortho=True, ortho_width=width)


def ortho_front(w, width):
    w.camera((0, -1000.0, 0), (0, 0, 0), up=(0, 0, 1), ortho=True, ortho_width=width)


def uniform_world(w, value=1.0):
    w.environment(kind=1, horizon=value, strength=1.0, background_mode=0)


def black_world(w):
    w.environment(kind=0, strength=0.0, background_color=(0, 0, 0), background_mode=1)


def milky(w, scatter_mm, albedo, g, roughness, ior=1.5, **extra):
    w.material(kind=GLASS, base_color=(1, 1, 1), roughness=roughness, ior=ior, transmission=1.0,
               scatter_mm=scatter_mm, scatter_color=albedo, scatter_g=g, **extra)


def furnace(scatter_mm, g, roughness, albedo=(1, 1, 1), size=16, depth=64):
    # integrator/tests.rs:1582-1602: a faceted sphere of radius 100 seen by a 260 mm orthographic
    # camera in a uniform world of radiance 1.
    w = SceneWriter()
    ortho_front(w, 260.0)
    milky(w, scatter_mm, albedo, g, roughness)
    p, i = icosphere((0, 0, 0), 100.0, 3)
    w.mesh(p, i, None, 0)
    uniform_world(w)
    w.render(size, size, 1024, 0, depth)
    return w.bytes()


def inside_disk(size):
    def where(x, y):
        u = (x + 0.5) / size - 0.5
        v = (y + 0.5) / size - 0.5
        return np.sqrt(u * u + v * v) * 260.0 < 85.0
    return where


@pytest.mark.parametrize("mfp,g,roughness,spp", [(10.0, 0.0, 0.0, 512), (10.0, 0.4, 0.0, 256),
                                                 (3.0, -0.3, 0.3, 256), (0.1, 0.0, 0.0, 64),
                                                 (60.0, 0.0, 0.0, 256)])
def test_a_white_medium_in_a_furnace_neither_creates_nor_loses_energy(mfp, g, roughness, spp):
    # integrator/tests.rs:1672-1700: invisible within 2 %.
    got = mean_of(furnace(mfp, g, roughness), spp, where=inside_disk(16))
    assert abs(got - 1.0) < 0.02, (mfp, g, roughness, got)


def test_a_grey_medium_in_a_furnace_is_dark():
    # The exit lobe carries the walk's throughput: a medium of albedo 0.9 and mean free path 1 mm
    # is well below white (integrator/tests.rs:1726: under 0.9 by the exact boundary).
    got = mean_of(furnace(1.0, 0.0, 0.0, albedo=(0.9, 0.9, 0.9)), 64, where=inside_disk(16))
    assert 0.2 < got < 0.9, got


def plane_albedo(a):
    """Chandrasekhar's plane albedo for light along the normal, 1 - H(1) sqrt(1 - a), with H
    solved by iteration on a midpoint grid (Radiative Transfer, 1950)."""
    n = 400
    mu = (np.arange(n) + 0.5) / n
    kernel = 1.0 / (mu[:, None] + mu[None, :]) / n
    h = np.ones(n)
    for _ in range(500):
        h = 1.0 / (1.0 - a / 2.0 * mu * (kernel @ h))
    h1 = 1.0 / (1.0 - a / 2.0 * float(np.sum(h / (1.0 + mu))) / n)
    return 1.0 - h1 * math.sqrt(1.0 - a)


def test_chandrasekhar_plane_albedo_closed_form():
    assert abs(plane_albedo(0.0)) < 1e-12 and abs(plane_albedo(1.0) - 1.0) < 1e-9
    # The published values for isotropic scattering: 0.415 at albedo 0.9, 0.753 at 0.99.
    assert abs(plane_albedo(0.9) - 0.415) < 1e-3 and abs(plane_albedo(0.99) - 0.753) < 1e-3


@pytest.mark.parametrize("a", [0.9, 0.99])
def test_a_grey_half_space_reflects_chandrasekhars_plane_albedo(a):
    # integrator/tests.rs:1749-1780: 400 mean free paths thick behind an index of 1.001, seen
    # straight down: within 2 % of the plane albedo. Walks run far past the roulette's start.
    w = SceneWriter()
    ortho_down(w, 0, 0, 20.0)
    milky(w, 1.0, (a, a, a), 0.0, 0.0, ior=1.001)
    p, i = box_mesh((-2000, -2000, -400), (2000, 2000, 0))
    w.mesh(p, i, None, 0)
    uniform_world(w)
    w.render(8, 8, 512, 0, 64)
    got = mean_of(w.bytes(), 512)
    want = plane_albedo(a)
    assert abs(got / want - 1.0) < 0.02, (a, got, want)


def slab_scene(material, emitter=True):
    # integrator/tests.rs:1039-1073: a 10 mm slab over a 6 m emissive panel 50 mm below, seen
    # straight down by a 40 mm orthographic camera in a black world.
    w = SceneWriter()
    ortho_down(w, 0, 0, 40.0)
    material(w)
    w.material(kind=EMISSIVE, emission=(1, 1, 1), emission_power=1.0)
    p, i = box_mesh((-1000, -1000, 0), (1000, 1000, 10))
    w.mesh(p, i, None, 0)
    if emitter:
        q = np.array([[-3000, -3000, -50], [3000, -3000, -50], [3000, 3000, -50], [-3000, 3000, -50]], np.float32)
        w.mesh(q, [0, 1, 2, 0, 2, 3], None, 1)
    black_world(w)
    w.render(8, 8, 1024, 0, 64)
    return w


def test_a_medium_that_only_absorbs_is_beer_lambert_in_every_channel():
    # integrator/tests.rs:1832-1856: albedo 0, sigma_t = 1/20 + ln(1/c)/10 per mm through 10 mm,
    # each channel its own transmittance by the walk's spectral MIS; within 2 %.
    r = fresnel_dielectric(1.5)
    w = slab_scene(lambda w: w.material(kind=GLASS, base_color=(0.5, 0.5, 0.25), roughness=0.0, ior=1.5,
                                        transmission=1.0, absorption_mm=10.0, scatter_mm=20.0,
                                        scatter_color=(0, 0, 0)))
    img = orc.OracleScene(w.bytes()).render_f64(0, 1024)
    for channel, tau in ((1, 0.5 * math.exp(-0.5)), (2, 0.25 * math.exp(-0.5))):
        want = (1 - r) ** 2 * tau / (1 - r * r * tau * tau)
        got = float(img[..., channel].mean())
        assert abs(got / want - 1.0) < 0.02, (channel, got, want)


def test_a_sparse_medium_lets_what_does_not_scatter_through_straight():
    # integrator/tests.rs:1786-1830: a 20 mm emitter seen through a slab of 0.01 mean free paths
    # stays sharp and keeps (1 - R)^2 tau / (1 - R^2 tau^2), tau = e^-0.01, within 3 %.
    w = SceneWriter()
    ortho_down(w, 0, 0, 40.0)
    milky(w, 1000.0, (1, 1, 1), 0.0, 0.0)
    w.material(kind=EMISSIVE, emission=(1, 1, 1), emission_power=1.0)
    p, i = box_mesh((-1000, -1000, 0), (1000, 1000, 10))
    w.mesh(p, i, None, 0)
    q = np.array([[-10, -10, -50], [10, -10, -50], [10, 10, -50], [-10, 10, -50]], np.float32)
    w.mesh(q, [0, 1, 2, 0, 2, 3], None, 1)
    black_world(w)
    w.render(8, 8, 64, 0, 64)
    img = orc.OracleScene(w.bytes()).render_f64(0, 64)[..., 1]
    r, tau = fresnel_dielectric(1.5), math.exp(-0.01)
    want = (1 - r) ** 2 * tau / (1 - r * r * tau * tau)
    middle = float(img[3:5, 3:5].mean())
    outside = float(img[[0, 0, 7, 7], [0, 7, 0, 7]].mean())
    assert abs(middle / want - 1.0) < 0.03, (middle, want)
    assert outside < 0.05 * want, outside


def test_an_emitter_that_is_not_a_light_is_found_after_a_medium_event():
    # integrator/tests.rs:1985-1999: one mean free path of white medium over the emissive panel
    # passes more than 1.5 times the unscattered light and less than all of it.
    w = slab_scene(lambda w: milky(w, 10.0, (1, 1, 1), 0.0, 0.0))
    got = mean_of(w.bytes(), 512)
    r, tau = fresnel_dielectric(1.5), math.exp(-1.0)
    unscattered = (1 - r) ** 2 * tau / (1 - r * r * tau * tau)
    assert 1.5 * unscattered < got < 1.0, (got, unscattered)


def medium_slab(g):
    # integrator/tests.rs:1910-1934: the slab lit from below by a 1 m glossy panel of radiance 1.
    w = slab_scene(lambda w: milky(w, 10.0, (1, 1, 1), g, 0.0), emitter=False)
    w.light(kind=0, position=(0, 0, -50.0), direction=(0, 0, 1), up=(0, 1, 0), size_x=1000.0, size_y=1000.0,
            power=math.pi, flags=GLOSSY)
    return w.bytes()


def test_forward_scattering_sends_the_light_onwards():
    # integrator/tests.rs:1936-1948.
    fwd, iso, back = (mean_of(medium_slab(g), 256) for g in (0.8, 0.0, -0.8))
    assert fwd > 1.05 * iso and iso > 1.2 * back, (fwd, iso, back)


def test_the_phase_function_is_a_density_with_mean_cosine_g():
    # Henyey-Greenstein (integrator/medium.rs:143-169): integrates to 1 over the sphere, its mean
    # cosine is g, and a sampled direction's density is the function at that direction.
    c = (np.arange(20000) + 0.5) / 10000 - 1.0
    rng = np.random.default_rng(5)
    for g in (-0.9, -0.3, 0.0, 0.4, 0.9):
        f = np.array([orc.phase(float(x), g) for x in c])
        total = float(np.sum(f) * (2 / 20000) * 2 * math.pi)
        mean_cos = float(np.sum(f * c) * (2 / 20000) * 2 * math.pi)
        assert abs(total - 1.0) < 2e-3, (g, total)
        assert abs(mean_cos - g) < 2e-3, (g, mean_cos)
        d = np.array([0.3, -0.5, 0.8], np.float32)
        d /= np.linalg.norm(d)
        cs = []
        for u0, u1 in rng.random((2000, 2)):
            out, pdf = orc.sample_phase(d, g, float(u0), float(u1))
            assert abs(np.linalg.norm(out) - 1.0) < 1e-5
            cos = float(np.dot(out, d))
            assert abs(pdf / orc.phase(cos, g) - 1.0) < 1e-3, (g, cos, pdf)
            cs.append(cos)
        assert abs(np.mean(cs) - g) < 0.03, (g, np.mean(cs))


def point_scene(kind, diameter, h, power, flags, x=0.0, view=2.0, size=8):
    # integrator/tests.rs:284-306: a light `h` above a white Lambertian floor seen straight down.
    w = SceneWriter()
    ortho_down(w, x, 0.0, view)
    w.material(base_color=(1, 1, 1), ior=1.0)
    q = np.array([[-2000, -2000, 0], [2000, -2000, 0], [2000, 2000, 0], [-2000, 2000, 0]], np.float32)
    w.mesh(q, [0, 1, 2, 0, 2, 3], None, 0)
    w.light(kind=kind, position=(0, 0, h), direction=(0, 0, -1), up=(0, 1, 0), size_x=diameter, size_y=diameter,
            power=power, spread_deg=50.0, angle_deg=5.0, flags=flags)
    black_world(w)
    w.render(size, size, 4096, 0, 8)
    return w.bytes()


@pytest.mark.parametrize("kind,diameter,spp", [(POINT, 0.0, 4), (POINT, 80.0, 1024), (SPOT, 0.0, 4),
                                               (SPOT, 80.0, 1024)])
def test_point_sphere_and_spot_lights_give_the_inverse_square_irradiance(kind, diameter, spp):
    # integrator/tests.rs:310-333: 6 W at 300 mm, E = 6 / (4 pi 0.3^2) W/m^2, floor radiance E / pi,
    # within 1 %.
    want = 6.0 / (4 * math.pi * 0.09) / math.pi
    got = mean_of(point_scene(kind, diameter, 300.0, 6.0, GLOSSY), spp)
    assert abs(got / want - 1.0) < 0.01, (kind, diameter, got, want)


def test_sphere_lights_are_counted_once_whichever_way_they_are_found():
    # integrator/tests.rs:336-348: next-event estimation with BSDF-sampled hits under MIS equals
    # next-event estimation alone, within 2 %.
    both = mean_of(point_scene(POINT, 300.0, 250.0, 6.0, GLOSSY, view=1200.0, size=24), 256)
    alone = mean_of(point_scene(POINT, 300.0, 250.0, 6.0, 0, view=1200.0, size=24), 256)
    assert both > 0.01
    assert abs(both / alone - 1.0) < 0.02, (both, alone)


def test_a_spot_fades_by_its_smoothstep_and_leaves_dark_what_is_outside_its_cone():
    # integrator/tests.rs:351-364 and light.rs:1852-1898: a 50 degree spot with a 5 degree soft
    # edge 300 mm up. At 15 degrees off its axis the point light's E = I cos / d^2; at 22.5
    # degrees (inside the edge, 20 to 25 degrees) times the smoothstep in the cosine; at 34
    # degrees nothing.
    h, power = 300.0, 6.0
    cos = lambda deg: math.cos(math.radians(deg))  # noqa: E731
    t = (cos(22.5) - cos(25.0)) / (cos(20.0) - cos(25.0))
    for deg, mask in ((15.0, 1.0), (22.5, t * t * (3 - 2 * t)), (34.0, 0.0)):
        x = h * math.tan(math.radians(deg))
        d2 = (x * x + h * h) * 1e-6
        want = power / (4 * math.pi) * math.cos(math.radians(deg)) / d2 / math.pi * mask
        got = mean_of(point_scene(SPOT, 0.0, h, power, GLOSSY, x=x, view=0.05), 4)
        if mask == 0.0:
            assert got == 0.0, deg
        else:
            assert abs(got / want - 1.0) < 2e-3, (deg, got, want)
    assert mean_of(point_scene(POINT, 0.0, h, power, GLOSSY, x=200.0), 4) > 0.01


def test_a_sphere_light_shows_its_radiance_to_the_camera():
    # light.rs:1813-1838: a 100 mm sphere of 12 W has the radiance I / (pi r^2) on its surface; a
    # camera flag off hides it.
    def view(flags):
        w = SceneWriter()
        ortho_down(w, 0, 0, 10.0)
        w.material(base_color=(0, 0, 0), ior=1.0)
        w.light(kind=POINT, position=(0, 0, 400.0), size_x=100.0, size_y=100.0, power=12.0, flags=flags)
        black_world(w)
        w.render(4, 4, 4, 0, 8)
        return mean_of(w.bytes(), 4)

    radiance = 12.0 / (4 * math.pi) / (math.pi * 0.05 ** 2)
    assert abs(view(GLOSSY | CAMERA) / radiance - 1.0) < 1e-4
    assert view(GLOSSY) == 0.0
