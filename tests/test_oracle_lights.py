"""The oracle's light selection and multiple importance sampling against the reference's own tests
and closed forms (paths inside engine/crates/rt-core/src):

- light.rs:1707-1750 (selection_follows_power): the static cdf, the per-point pmf, pick's bucket
  boundaries and the rescaled remainder;
- integrator/tests.rs:2636-2700 (light_selection_is_unbiased_with_either_technique_and_with_both):
  next-event estimation alone, BSDF sampling alone and both under MIS converge to the same image,
  in the white furnace and over a floor lit by a panel, a spot and a dim world; here also with
  three panels and the studio dome, the light configuration of the large render;
- a floor under three panels gets the sum of their analytic irradiances (view factors of offset
  rectangles by superposition of corner rectangles).
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from repackshot_amd.rsb import SceneWriter, icosphere

MIS, LIGHTS_ONLY, BSDF_ONLY = 0, 1, 2
GLOSSY = 1


def floor_quad(w, half):
    q = np.array([[-half, -half, 0], [half, -half, 0], [half, half, 0], [-half, half, 0]], np.float32)
    w.mesh(q, [0, 1, 2, 0, 2, 3], None, 0)


def ortho_down(w, x, y, width):
    w.camera((x, y, 2000.0), (x, y, 0.0), up=(0, 1, 0), ortho=True, ortho_width=width)


def unit(v):
    v = np.asarray(v, np.float32)
    return tuple(float(c) for c in v / np.float32(math.sqrt(float(np.dot(v, v)))))


def test_selection_follows_power():
    # light.rs:1707-1750: two rects in the same place, 30 W and 10 W (a third of no power is
    # dropped), a black world: by flux 0.75 / 0.25, per point the same, pick's boundaries.
    w = SceneWriter()
    w.camera((0, -1000.0, 0), (0, 0, 0))
    w.material()
    floor_quad(w, 50.0)
    for power, z in ((30.0, 1000.0), (10.0, 1000.0), (0.0, 0.0)):
        w.light(kind=0, position=(0, 0, z), direction=(0, 0, -1), up=(0, 1, 0), size_x=100.0, size_y=100.0,
                power=power)
    w.environment(strength=0.0)
    w.render(8, 8, 1, 0, 4)
    o = orc.OracleScene(w.bytes())
    per_point, cdf, pmf = o.lights()
    assert len(cdf) == 2
    assert abs(cdf[0] - 0.75) < 1e-6 and abs(cdf[1] - cdf[0] - 0.25) < 1e-6
    assert abs(pmf[0] - 0.75) < 1e-6 and abs(pmf[1] - 0.25) < 1e-6
    assert per_point
    assert abs(o.light_pmf((0, 0, 0), 0) - 0.75) < 1e-6
    assert abs(o.light_pmf((0, 0, 0), 1) - 0.25) < 1e-6
    assert o.light_pick(0.0)[0] == 0
    assert o.light_pick(0.74)[0] == 0
    assert o.light_pick(0.76)[0] == 1
    assert o.light_pick(0.9999999)[0] == 1
    assert abs(o.light_pick(0.375)[1] - 0.5) < 1e-6
    # The per-point pick gives the same buckets here.
    for u, e in ((0.0, 0), (0.74, 0), (0.76, 1), (0.9999999, 1)):
        assert o.light_pick_at((0, 0, 0), u)[0] == e
    assert abs(o.light_pick_at((0, 0, 0), 0.375)[1] - 0.5) < 1e-6


def three_panels(w):
    """Three square panels of different size, height, offset and power, facing down."""
    panels = [((-300.0, 100.0, 600.0), 400.0, 12.0), ((250.0, -150.0, 900.0), 700.0, 30.0),
              ((50.0, 400.0, 450.0), 200.0, 3.0)]
    for pos, size, power in panels:
        w.light(kind=0, position=pos, direction=(0, 0, -1), up=(0, 1, 0), size_x=size, size_y=size, power=power,
                flags=GLOSSY)
    return panels


def test_selection_among_more_than_two_entries_is_a_partition_of_the_unit_interval():
    # Per point, entry k is picked for u in [sum of the earlier pmfs, that plus its own), the
    # remainder rescaled inside the bucket, the probability its pmf (light.rs:1118-1141).
    w = SceneWriter()
    w.camera((0, -1000.0, 0), (0, 0, 0))
    w.material()
    floor_quad(w, 2000.0)
    three_panels(w)
    w.environment(kind=0, strength=1.0)
    w.render(8, 8, 1, 0, 4)
    o = orc.OracleScene(w.bytes())
    per_point, cdf, _ = o.lights()
    assert per_point and len(cdf) == 4
    rng = np.random.default_rng(7)
    for point in rng.uniform(-800, 800, size=(20, 3)) * np.array([1, 1, 0]) + np.array([0, 0, 1.0]):
        pm = [o.light_pmf(point, e) for e in range(4)]
        assert abs(sum(pm) - 1.0) < 1e-5 and min(pm) > 0
        lo = 0.0
        for e in range(4):
            hi = 1.0 if e == 3 else lo + pm[e]
            for u in (lo + 1e-4, 0.5 * (lo + hi), hi - 1e-4):
                k, rest, p = o.light_pick_at(point, u)
                assert k == e, (point, u, k, e)
                assert abs(p - pm[e]) < 1e-6
                assert abs(rest - (u - lo) / (hi - lo)) < 1e-3
            lo = hi
    # The weights differ from point to point.
    spread = [o.light_pmf((x, 0, 1), 0) for x in (-800.0, 0.0, 800.0)]
    assert max(spread) - min(spread) > 0.05


def offset_rect_factor(x1, x2, y1, y2, h):
    """View factor from a floor element at the origin (facing up) to the parallel rectangle
    [x1, x2] x [y1, y2] at height h facing down: signed corner rectangles superposed (the corner
    form is the textbook one, e.g. Howell's catalogue of view factors)."""
    def corner(a, b):
        A, B = abs(a) / h, abs(b) / h
        c = (A / math.sqrt(1 + A * A) * math.atan(B / math.sqrt(1 + A * A))
             + B / math.sqrt(1 + B * B) * math.atan(A / math.sqrt(1 + B * B))) / (2 * math.pi)
        return math.copysign(1.0, a) * math.copysign(1.0, b) * c
    return corner(x2, y2) - corner(x1, y2) - corner(x2, y1) + corner(x1, y1)


def test_three_panels_give_the_sum_of_their_analytic_irradiances():
    # The floor point at (x0, y0) sees the three panels; a white Lambert floor's radiance is
    # sum_k L_k F_k (F_k the view factor), within 2 %.
    x0, y0 = 40.0, -30.0
    w = SceneWriter()
    ortho_down(w, x0, y0, 2.0)
    w.material(base_color=(1, 1, 1), ior=1.0)
    floor_quad(w, 3000.0)
    panels = three_panels(w)
    w.environment(kind=0, strength=0.0, background_color=(0, 0, 0), background_mode=1)
    w.render(8, 8, 4096, 0, 8)
    got = float(orc.OracleScene(w.bytes()).render_f64(0, 4096)[..., 1].mean())
    want = 0.0
    for (px, py, pz), size, power in panels:
        radiance = power / (math.pi * size * size * 1e-6)
        h = size / 2
        want += radiance * offset_rect_factor(px - h - x0, px + h - x0, py - h - y0, py + h - y0, pz)
    assert abs(got / want - 1.0) < 0.02, (got, want)


def furnace_scene():
    p, i = icosphere((0, 0, 0), 100.0, 3)
    w = SceneWriter()
    w.camera((0, -1000.0, 0), (0, 0, 0), up=(0, 0, 1), ortho=True, ortho_width=260.0)
    w.material(base_color=(1, 1, 1), ior=1.0)
    w.mesh(p, i, None, 0)
    w.environment(kind=1, horizon=1.0, strength=1.0, background_mode=0)
    w.render(16, 16, 4096, 0, 8)
    return w.bytes(), (slice(4, 12), slice(4, 12))


def panel_and_spot_scene():
    # integrator/tests.rs:2599-2632: a white Lambert floor under a dim uniform world (0.2), an 800 mm
    # panel 900 mm up tilted towards the spot, a 160 mm spot of 60/20 degrees 400 mm up, both found
    # by both techniques (glossy), seen straight down over the edge of the spot's pool.
    w = SceneWriter()
    ortho_down(w, 150.0, 0.0, 120.0)
    w.material(base_color=(1, 1, 1), ior=1.0)
    floor_quad(w, 2000.0)
    w.light(kind=0, position=(-500.0, 0.0, 900.0), direction=unit((0.5, 0.0, -1.0)), up=(0, 1, 0), size_x=800.0,
            size_y=800.0, power=20.0, flags=GLOSSY)
    w.light(kind=4, position=(0.0, 0.0, 400.0), direction=(0, 0, -1), up=(0, 1, 0), size_x=160.0, size_y=160.0,
            power=2.0, spread_deg=60.0, angle_deg=20.0, flags=GLOSSY)
    w.environment(kind=1, horizon=0.2, strength=1.0, background_mode=0)
    w.render(8, 8, 4096, 0, 8)
    return w.bytes(), (slice(0, 8), slice(0, 8))


def panels_and_dome_scene():
    # The large render's light configuration: three rect panels and the studio dome (four entries,
    # chosen per point), over a white floor with a rough coloured metal sphere on it.
    p, i = icosphere((0, 0, 300.0), 300.0, 3)
    w = SceneWriter()
    ortho_down(w, 0.0, 0.0, 1600.0)
    w.material(base_color=(0.9, 0.9, 0.9), ior=1.0)
    w.material(base_color=(0.95, 0.64, 0.54), metallic=1.0, roughness=0.6)
    floor_quad(w, 3000.0)
    w.mesh(p, i, None, 1)
    three_panels(w)
    w.environment(kind=0, strength=1.0)
    w.render(8, 8, 4096, 0, 8)
    return w.bytes(), (slice(0, 8), slice(0, 8))


@pytest.mark.parametrize("name,make", [("furnace", furnace_scene), ("panel and spot", panel_and_spot_scene),
                                       ("three panels and the dome", panels_and_dome_scene)])
def test_light_selection_is_unbiased_with_either_technique_and_with_both(name, make):
    # integrator/tests.rs:2636-2700: means of 4096 samples of 8 x 8 pixels agree to 1 %, the
    # selection per point.
    rsb, crop = make()
    o = orc.OracleScene(rsb)
    means = {}
    for tech in (MIS, LIGHTS_ONLY, BSDF_ONLY):
        o.set_techniques(tech, per_point=True)
        img = o.render_f64(0, 4096)
        means[tech] = float(img[crop[0], crop[1], 1].mean())
    both = means[MIS]
    if name == "furnace":
        assert abs(both - 1.0) < 0.01, both
    assert both > 0.05, means
    assert abs(means[LIGHTS_ONLY] / both - 1.0) < 0.01, (name, means)
    assert abs(means[BSDF_ONLY] / both - 1.0) < 0.01, (name, means)


# ---------------------------------------------------------------- suns, emissive meshes, many lights
SUN, POINT, SPOT, DISK = 2, 3, 4, 1
EMISSIVE = 2
MATERIAL_FLAG_LIGHT = 1


def lights_scene(lights, env_strength=0.0, meshes=(), materials=None):
    """A scene of the given lights (SceneWriter.light keyword dicts), a small floor far below (or
    the given meshes with their materials) and an environment of the given strength."""
    w = SceneWriter()
    w.camera((0, -1000.0, 0), (0, 0, 0))
    for m in (materials or [dict()]):
        w.material(**m)
    if meshes:
        for p, i, mat in meshes:
            w.mesh(p, i, None, mat)
    else:
        q = np.array([[-10, -10, -5000], [10, -10, -5000], [10, 10, -5000], [-10, 10, -5000]], np.float32)
        w.mesh(q, [0, 1, 2, 0, 2, 3], None, 0)
    for l in lights:
        w.light(**l)
    w.environment(kind=0, strength=env_strength)
    w.render(8, 8, 1, 0, 4)
    return orc.OracleScene(w.bytes())


def test_suns_give_their_irradiance_and_cover_their_disc():
    # light.rs:1663-1705: a 0.53 degree sun of 3 W/m^2 straight up: cone samples give E = 3 on a
    # horizontal surface; a camera sees its disc straight up and nothing beside it; a delta sun's
    # sample is the irradiance with the selection probability as density, and no disc to see.
    sun = dict(kind=SUN, direction=(0, 0, -1), power=3.0, angle_deg=0.53)
    o = lights_scene([sun])
    e = 0.0
    for i in range(64):
        for j in range(64):
            s = o.light_sample((0, 0, 0), 0.3, (i + 0.5) / 64, (j + 0.5) / 64)
            assert s is not None and s["kind"] == 1
            cos_max = math.cos(math.radians(0.265))
            assert float(s["wi"][2]) >= cos_max - 1e-6
            e += s["radiance"][1] * s["wi"][2] / s["pdf"]
    assert abs(e / 4096 - 3.0) < 1e-3, e / 4096
    assert o.light_escaped((0, 0, 1))[1] > 1e3
    assert np.all(o.light_escaped(unit((0, 0.1, 1))) == 0)
    o = lights_scene([dict(sun, angle_deg=0.0)])
    s = o.light_sample((0, 0, 0), 0.9, 0.1, 0.2)
    assert s["delta"] and s["pdf"] == 1.0 and s["radiance"][1] == np.float32(3.0)
    assert np.all(o.light_escaped((0, 0, 1)) == 0)


def quad_mesh(corners, order=(0, 2, 1, 0, 3, 2)):
    return np.array(corners, np.float32), list(order)


def test_emissive_meshes_are_sampled_by_area():
    # light.rs:2031-2091: an emissive 100 x 100 mm quad of radiance 2 at z = 100 facing down, two
    # triangles of different size: stratified samples give the corner-form-factor irradiance
    # below its centre within 1 %; a BSDF ray that finds a non-emitter gets weight 1, one that finds
    # the emitter from 100 mm less.
    quad = quad_mesh([[-50, -50, 100], [50, -50, 100], [50, 50, 100], [-50, 50, 100]])
    o = lights_scene([], meshes=[(*quad, 0)],
                     materials=[dict(kind=EMISSIVE, emission=(1, 1, 1), emission_power=2.0, flags=MATERIAL_FLAG_LIGHT)])
    per_point, cdf, _ = o.lights()
    assert len(cdf) == 1
    e = 0.0
    for i in range(64):
        for j in range(64):
            sel = ((i * 64 + j) + 0.5) / 4096.0
            s = o.light_sample((0, 0, 0), sel, (i + 0.5) / 64, (j + 0.5) / 64)
            assert s is not None and s["kind"] == 3
            e += float(s["radiance"][1]) * float(s["wi"][2]) / s["pdf"]
            # pdf = d^2 / (A cos): the quad's whole area, 10,000 mm^2.
            area = s["distance"] ** 2 / (s["pdf"] * abs(float(s["wi"][2])))
            assert abs(area - 10000.0) < 1.0, area
    x = 0.5
    f = (x / math.sqrt(1 + x * x) * math.atan(x / math.sqrt(1 + x * x))) * 2 / (2 * math.pi)
    want = 2.0 * 4.0 * math.pi * f
    assert abs(e / 4096 / want - 1.0) < 0.01, (e / 4096, want)
    assert o.emitter_weight((0, 0, 0), 5, 1.0, 1.0, 1.0) == 1.0
    assert o.emitter_weight((0, 0, 0), 0, 1.0, 100.0, 1.0) < 1.0


def test_beyond_sixty_four_lights_every_point_chooses_by_flux():
    # light.rs:2482-2508: up to 64 entries the choice is per point, from 65 the static one, for
    # next-event estimation and the MIS weights alike.
    for n, per_point in ((64, True), (65, False)):
        row = [dict(kind=POINT, position=(100.0 * i, 0.0, 200.0), size_x=0.0, size_y=0.0, power=1.0, flags=GLOSSY)
               for i in range(n)]
        o = lights_scene(row)
        pp, cdf, _ = o.lights()
        assert (len(cdf), pp) == (n, per_point)
        s = o.light_sample((0, 0, 0), 0.0, 0.5, 0.5)
        assert s is not None and s["entry"] == 0
        assert s["pdf"] == o.light_pmf((0, 0, 0), 0)
        if per_point:
            assert s["pdf"] > 0.2
        else:
            assert s["pdf"] == cdf[0] and s["pdf"] < 0.1


def every_kind_scene():
    """light.rs:2329-2393 without the IES light: two panels, a spot of 80 mm, a sun disc of 8 degrees,
    an emissive quad sampled as a light, the studio dome."""
    quad = quad_mesh([[-400, 300, 100], [-200, 300, 100], [-200, 300, 300], [-400, 300, 300]], (0, 1, 2, 0, 2, 3))
    lights = [dict(kind=0, position=(0, 0, 800.0), direction=(0, 0, -1), up=(0, 1, 0), size_x=600.0, size_y=600.0,
                   power=20.0),
              dict(kind=DISK, position=(700.0, 0, 300.0), direction=(-1.0, 0.0, -0.2), up=(0, 1, 0), size_x=300.0,
                   size_y=300.0, power=8.0),
              dict(kind=SPOT, position=(-300.0, -300.0, 250.0), direction=(0, 0, -1), size_x=80.0, size_y=80.0,
                   power=3.0, spread_deg=120.0, angle_deg=20.0, flags=GLOSSY),
              dict(kind=SUN, direction=(0.3, 0.2, -1.0), power=2.0, angle_deg=8.0)]
    return lights_scene(lights, env_strength=1.0, meshes=[(*quad, 0)],
                        materials=[dict(kind=EMISSIVE, emission=(1, 1, 1), emission_power=4.0,
                                        flags=MATERIAL_FLAG_LIGHT)])


def test_next_event_estimation_and_bsdf_rays_agree_on_every_probability():
    # light.rs:2394-2480: a BSDF ray of the very density of a light sample, sent from the same vertex
    # along the sample's direction, gets the power heuristic's 1/2 for what it finds, for every kind.
    o = every_kind_scene()
    per_point, cdf, _ = o.lights()
    assert len(cdf) == 6 and per_point
    rng = np.random.default_rng(6)
    n = 4000
    pts = np.stack([600.0 * (rng.random(n) - 0.5), 400.0 * (rng.random(n) - 0.5), 10.0 + 50.0 * rng.random(n)], 1)
    tuples = np.concatenate([pts, rng.random((n, 3))], 1)
    kinds, weights = o.light_mis_probe(tuples)
    done = kinds >= 0
    assert np.all(np.abs(weights[done] - 0.5) < 2e-3), (kinds[done][np.abs(weights[done] - 0.5) >= 2e-3],
                                                       weights[done][np.abs(weights[done] - 0.5) >= 2e-3])
    counts = [int(np.sum(kinds == k)) for k in range(5)]
    assert all(c > 10 for c in counts), counts


def test_a_delta_sun_lights_a_floor_by_its_irradiance_times_the_cosine():
    # A white Lambert floor under a delta sun of irradiance 2 at 30 degrees from the zenith shows
    # radiance E cos / pi (black world), exactly up to f32 rounding.
    w = SceneWriter()
    ortho_down(w, 0.0, 0.0, 100.0)
    w.material(base_color=(1, 1, 1), ior=1.0)
    floor_quad(w, 3000.0)
    th = math.radians(30.0)
    w.light(kind=SUN, direction=(-math.sin(th), 0.0, -math.cos(th)), power=2.0, angle_deg=0.0)
    w.environment(kind=0, strength=0.0, background_color=(0, 0, 0), background_mode=1)
    w.render(4, 4, 16, 0, 2)
    img = orc.OracleScene(w.bytes()).render_f64(0, 16)
    want = 2.0 * math.cos(th) / math.pi
    assert np.allclose(img[..., 1], want, rtol=1e-5), (img[..., 1], want)


@pytest.mark.parametrize("name", ["sun disc", "emissive quad"])
def test_suns_and_emissive_meshes_are_unbiased_with_either_technique(name):
    # integrator/tests.rs:2636-2700 for the two new kinds: NEE alone, BSDF sampling alone and MIS
    # agree to 1 % over a white floor with a rough metal sphere on it (the emissive quad is seen
    # by the floor and the sphere, never by the camera).
    p, i = icosphere((0, 0, 150.0), 150.0, 3)
    w = SceneWriter()
    ortho_down(w, 0.0, 0.0, 700.0)
    w.material(base_color=(0.9, 0.9, 0.9), ior=1.0)
    w.material(base_color=(0.95, 0.64, 0.54), metallic=1.0, roughness=0.6)
    w.material(kind=EMISSIVE, emission=(1, 0.9, 0.8), emission_power=30.0, flags=MATERIAL_FLAG_LIGHT)
    floor_quad(w, 3000.0)
    w.mesh(p, i, None, 1)
    if name == "sun disc":
        w.light(kind=SUN, direction=(0.4, 0.3, -1.0), power=2.0, angle_deg=12.0, flags=GLOSSY)
        w.environment(kind=1, horizon=0.05, strength=1.0, background_mode=0)
    else:
        quad = np.array([[-400, 500, 50], [400, 500, 50], [400, 500, 450], [-400, 500, 450]], np.float32)
        w.mesh(quad, [0, 2, 1, 0, 3, 2], None, 2)
        w.environment(kind=0, strength=0.0, background_color=(0, 0, 0), background_mode=1)
    w.render(8, 8, 4096, 0, 8)
    o = orc.OracleScene(w.bytes())
    means = {}
    for tech in (MIS, LIGHTS_ONLY, BSDF_ONLY):
        o.set_techniques(tech, per_point=True)
        means[tech] = float(o.render_f64(0, 4096)[..., 1].mean())
    both = means[MIS]
    assert both > 0.05, means
    assert abs(means[LIGHTS_ONLY] / both - 1.0) < 0.01, (name, means)
    assert abs(means[BSDF_ONLY] / both - 1.0) < 0.01, (name, means)
