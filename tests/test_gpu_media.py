"""GPU scattering media and sphere/point/spot lights, through the C ABI, against the oracle.

Bar: bit-exact films (sample-order f32 sums) and equal ray counts on every scene of the oracle's
closed-form media and light tests (integrator/tests.rs:284-364, 1576-1999; light.rs:1813-1898) and on
the bottle workload at a reduced size. The oracle's own closed-form checks pin what these films
mean; here the GPU must reproduce the oracle bit for bit."""
import numpy as np
import pytest

from oracle import oracle as orc
from repackshot_amd.rsb import SceneWriter, with_size
from test_oracle_media import (CAMERA, GLOSSY, POINT, SPOT, black_world, furnace, medium_slab, milky,
                               ortho_down, point_scene, slab_scene)

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def check_film(gpu, rsb, spp, split=None):
    """Renders samples [0, spp) on the GPU (in two calls at `split`) and compares film and rays
    with the oracle's: once with the drain (small images: every path walks and shades in the drain
    kernel) and once with the iteration loop only (the walk kernel between extend and shade)."""
    s = gpu.Scene(rsb)
    (orgb, ow, oa, ol2), stats = orc.OracleScene(rsb).render_film(0, spp)
    for drain in (True, False):
        s.set_options(drain=drain)
        s.reset_film()
        before = s.stats()["rays"]
        if split:
            s.render_samples(0, split)
            s.render_samples(split, spp - split)
        else:
            s.render_samples(0, spp)
        rgb, w, a, l2 = s.film()
        assert same_bits(rgb, orgb) and same_bits(w, ow) and same_bits(a, oa) and same_bits(l2, ol2), drain
        assert s.stats()["rays"] - before == stats["rays"], drain
    return s, rgb


FURNACES = [(10.0, 0.0, 0.0, (1, 1, 1)), (3.0, -0.3, 0.3, (1, 1, 1)), (0.1, 0.0, 0.0, (1, 1, 1)),
            (10.0, 0.4, 0.0, (1, 1, 1)), (1.0, 0.0, 0.0, (0.9, 0.9, 0.9)), (4.0, 0.3, 0.0, (0.97, 0.8, 0.6))]


@pytest.mark.parametrize("mfp,g,roughness,albedo", FURNACES)
def test_medium_spheres_in_a_furnace_are_bit_exact(gpu, mfp, g, roughness, albedo):
    # Dense media reach the 1024-step cap and the roulette after 32 events; rough boundaries
    # reflect back in through the Fresnel test.
    check_film(gpu, furnace(mfp, g, roughness, albedo=albedo), 8, split=3)


def test_a_grey_half_space_is_bit_exact(gpu):
    from repackshot_amd.rsb import box_mesh
    w = SceneWriter()
    ortho_down(w, 0, 0, 20.0)
    milky(w, 1.0, (0.99, 0.99, 0.99), 0.0, 0.0, ior=1.001)
    p, i = box_mesh((-2000, -2000, -400), (2000, 2000, 0))
    w.mesh(p, i, None, 0)
    w.environment(kind=1, horizon=1.0, strength=1.0, background_mode=0)
    w.render(8, 8, 64, 0, 64)
    check_film(gpu, w.bytes(), 16)


@pytest.mark.parametrize("which", ["absorbing", "sparse", "emitter", "forward", "backward"])
def test_slab_scenes_are_bit_exact(gpu, which):
    if which == "absorbing":
        rsb = slab_scene(lambda w: w.material(kind=1, base_color=(0.5, 0.5, 0.25), roughness=0.0, ior=1.5,
                                              transmission=1.0, absorption_mm=10.0, scatter_mm=20.0,
                                              scatter_color=(0, 0, 0))).bytes()
    elif which == "sparse":
        rsb = slab_scene(lambda w: milky(w, 1000.0, (1, 1, 1), 0.0, 0.0)).bytes()
    elif which == "emitter":
        rsb = slab_scene(lambda w: milky(w, 10.0, (1, 1, 1), 0.0, 0.0)).bytes()
    else:
        rsb = medium_slab(0.8 if which == "forward" else -0.8)
    check_film(gpu, rsb, 32)


@pytest.mark.parametrize("kind,diameter,flags", [(POINT, 0.0, GLOSSY), (POINT, 80.0, GLOSSY), (SPOT, 0.0, GLOSSY),
                                                 (SPOT, 80.0, GLOSSY), (POINT, 300.0, 0),
                                                 (POINT, 300.0, GLOSSY | CAMERA)])
def test_point_sphere_and_spot_lights_are_bit_exact(gpu, kind, diameter, flags):
    view = 1200.0 if diameter == 300.0 else 300.0
    check_film(gpu, point_scene(kind, diameter, 250.0, 6.0, flags, view=view, size=24), 16)


def test_spot_edge_and_sphere_radiance_views_are_bit_exact(gpu):
    for x in (0.0, 123.0, 170.0):
        check_film(gpu, point_scene(SPOT, 0.0, 300.0, 6.0, GLOSSY, x=x, view=40.0), 4)
    w = SceneWriter()
    ortho_down(w, 0, 0, 160.0)
    w.material(base_color=(0.7, 0.7, 0.7), ior=1.0)
    q = np.array([[-500, -500, 0], [500, -500, 0], [500, 500, 0], [-500, 500, 0]], np.float32)
    w.mesh(q, [0, 1, 2, 0, 2, 3], None, 0)
    w.light(kind=POINT, position=(0, 0, 400.0), size_x=100.0, size_y=100.0, power=12.0, flags=GLOSSY | CAMERA)
    w.light(kind=SPOT, position=(60, 0, 300.0), direction=(0, 0, -1), size_x=20.0, power=3.0, spread_deg=40.0,
            angle_deg=8.0, flags=GLOSSY)
    w.light(kind=0, position=(-200, 0, 500.0), direction=(0, 0, -1), size_x=100.0, size_y=50.0, power=4.0)
    black_world(w)
    w.render(32, 32, 16, 3, 8)
    s, _ = check_film(gpu, w.bytes(), 8)
    # The passes see a camera-visible sphere as background (light.rs:1479-1488).
    s.render_passes(0, 2)
    got = s.passes()
    want = orc.OracleScene(w.bytes()).render_passes(0, 2)
    for k in want:
        assert same_bits(got[k], want[k]), k


def bottle(width=128, height=72):
    from repackshot_amd import native
    return with_size(native.write_scene(5, 0), width, height)


def test_the_bottle_workload_is_bit_exact(gpu):
    # The media workload (200,502 triangles, sphere light, scattering glass) at 128 x 72.
    s, rgb = check_film(gpu, bottle(), 4, split=1)
    assert float(rgb.mean()) > 0


def test_options_and_queue_keep_media_films(gpu, example_rsb):
    rsb = furnace(3.0, -0.3, 0.3)
    (orgb, ow, oa, ol2), _ = orc.OracleScene(rsb).render_film(0, 6)
    s = gpu.Scene(rsb)
    for opts in (dict(pool=1024), dict(pool=1024, drain=False), dict(per_pass=True), dict(graph=True, group=4),
                 dict(shape=5, shadow_shape=6), dict(tail=200)):
        s.set_options(**opts)
        s.reset_film()
        s.render_samples(0, 6)
        rgb, _, _, _ = s.film()
        assert same_bits(rgb, orgb), opts
    # A queue mixing a media scene with a plain one: each film is its own render's.
    q = gpu.Queue()
    a = gpu.Scene(rsb)
    b = gpu.Scene(example_rsb)
    ja = q.add(a, 0, 6)
    jb = q.add(b, 0, 2)
    q.run()
    assert same_bits(q.film(ja)[0], orgb)
    (ergb, _, _, _), _ = orc.OracleScene(example_rsb).render_film(0, 2)
    assert same_bits(q.film(jb)[0], ergb)


def test_media_scenes_refuse_the_adjoint_and_new_media(gpu):
    s = gpu.Scene(furnace(10.0, 0.0, 0.0))
    with pytest.raises(gpu.RenderError, match="status 4"):
        s.adjoint(0, 1, np.zeros((16, 16, 3), np.float32))
    # A scattering material of index 1 has no medium (not a clear solid); moving its index into
    # range would give a scene without the walk one.
    w = SceneWriter()
    ortho_down(w, 0, 0, 40.0)
    milky(w, 10.0, (1, 1, 1), 0.0, 0.0, ior=1.0)
    from repackshot_amd.rsb import box_mesh
    p, i = box_mesh((-100, -100, 0), (100, 100, 10))
    w.mesh(p, i, None, 0)
    black_world(w)
    w.render(8, 8, 4, 0, 8)
    plain = gpu.Scene(w.bytes())
    params = plain.material_params()
    params[0, 4] = 1.5
    with pytest.raises(gpu.RenderError, match="scattering medium"):
        plain.set_material_params(params)
    params[0, 4] = 1.0
    plain.set_material_params(params)
