"""GPU tracer parity, through the C ABI, against the oracle and the reference's pinned hashes.

Bar: bit-exact. Hit (instance, triangle, t) and any-hit of fixed ray sets, per-sample radiance of
tiles, and whole films are compared as raw f32 bits; the example scene must reproduce the
reference's own film hashes (render.rs:991-1065)."""
import numpy as np
import pytest

from oracle import oracle as orc
from repackshot_amd.rsb import SceneWriter, uv_sphere

pytestmark = pytest.mark.gpu

PINNED_CENTRE = 0xE2515C9A
PINNED_SMALL_SPHERE = 0x7B62574D
PINNED_EXAMPLE_4SPP = 0x75BA4CB4


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def mismatch(a, b):
    a = np.ascontiguousarray(a, np.float32).view(np.uint32)
    b = np.ascontiguousarray(b, np.float32).view(np.uint32)
    return int(np.sum(a != b))


def test_energy_tables_are_the_oracles(gpu):
    g = gpu.energy_tables()
    o = orc.energy_tables()
    for k in g:
        assert bits_equal(g[k], o[k]), k


def test_example_tiles_have_the_pinned_hashes(gpu, example_rsb):
    s = gpu.Scene(example_rsb)
    o = orc.OracleScene(example_rsb)
    for (x, y), pin in [((240, 176), PINNED_CENTRE), ((48, 144), PINNED_SMALL_SPHERE)]:
        got = s.render_tile(x, y, 32, 32, 3)
        want, alpha = o.render_tile(x, y, 32, 32, 3)
        assert mismatch(got, want) == 0
        assert gpu.fnv1a(got) == pin
        rgba = s.render_tile_rgba(x, y, 32, 32, 3)
        assert bits_equal(rgba[..., :3], want) and bits_equal(rgba[..., 3], alpha)


def test_example_film_at_four_samples_has_the_pinned_hash(gpu, example_rsb):
    s = gpu.Scene(example_rsb)
    s.render_samples(0, 4)
    rgb, w, a, l2 = s.film()
    assert np.all(w == 4.0)
    assert gpu.fnv1a(gpu.resolve(rgb, w)) == PINNED_EXAMPLE_4SPP
    # Any split into consecutive calls gives the same film.
    s.reset_film()
    s.render_samples(0, 1)
    s.render_samples(1, 3)
    rgb2, w2, a2, l22 = s.film()
    assert bits_equal(rgb, rgb2) and bits_equal(a, a2) and bits_equal(l2, l22)
    (orgb, ow, oa, ol2), stats = orc.OracleScene(example_rsb).render_film(0, 4)
    assert bits_equal(rgb, orgb) and bits_equal(a, oa) and bits_equal(l2, ol2)
    # Rays: camera + extension + shadow, as the reference counts them.
    assert s.stats()["rays"] >= stats["rays"]


def test_rays_counted_as_the_reference_counts_them(gpu, example_rsb):
    s = gpu.Scene(example_rsb)
    s.render_samples(0, 1)
    _, stats = orc.OracleScene(example_rsb).render_film(0, 1)
    assert s.stats()["rays"] == stats["rays"]


def _rays(n, seed, centre, spread, targets):
    rng = np.random.default_rng(seed)
    o = (np.asarray(centre, np.float64) + rng.uniform(-spread, spread, size=(n, 3))).astype(np.float32)
    d = targets[rng.integers(0, len(targets), n)] - o + rng.normal(scale=spread * 0.05, size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = o
    rays[:, 3:6] = d
    rays[:, 6] = 1e-3
    rays[:, 7] = np.inf
    rays[1::4, 7] = spread
    return rays


@pytest.mark.parametrize("kind", [0, 1])
def test_hits_are_bit_exact(gpu, kind):
    rsb = gpu.write_scene(kind, 0)
    s = gpu.Scene(rsb)
    o = orc.OracleScene(rsb)
    if kind == 0:
        rays = _rays(20000, 3, (0, 0, 200), 600, np.array([[0, 0, 100], [-330, 120, 50], [300, 200, 0]], np.float32))
    else:
        rays = _rays(20000, 4, (0, 0, 1500), 4000, np.array([[0, 0, 1000], [0, 0, 0], [2000, 0, 0]], np.float32))
    g = s.trace_rays(rays)
    r = o.trace_rays(rays)
    for name, x, y in zip(["instance", "triangle", "t", "any"], g, r):
        assert bits_equal(x, y), (name, int(np.sum(x != y)))
    assert np.mean(g[0] != 0xFFFFFFFF) > 0.2


def test_plain_scene_films_are_bit_exact(gpu):
    rsb = gpu.write_scene(1, 0)
    s = gpu.Scene(rsb)
    s.render_samples(0, 2)
    rgb, w, a, l2 = s.film()
    (orgb, ow, oa, ol2), _ = orc.OracleScene(rsb).render_film(0, 2)
    assert mismatch(rgb, orgb) == 0 and bits_equal(w, ow) and bits_equal(a, oa) and bits_equal(l2, ol2)


def feature_scene(background_mode=0, floor=0, ortho=False, fstop=0.0, disk=False):
    """Glass with absorption, a coated dielectric, a rough metal, a transformed instance, a visible
    floor or not, a disk or rect light, a clamp of indirect light: the branches the integrator has."""
    w = SceneWriter()
    w.camera((300, -900, 350), (0, 0, 80), focal_mm=40.0, fstop=fstop, focus_dist=900.0, ortho=ortho,
             ortho_width=900.0)
    w.material(base_color=(0.6, 0.8, 0.9), transmission=1.0, roughness=0.0, ior=1.5, absorption_mm=80.0)
    w.material(base_color=(0.9, 0.2, 0.1), roughness=0.4, coat=0.8, coat_roughness=0.1)
    w.material(base_color=(0.9, 0.9, 0.9), metallic=1.0, roughness=0.2)
    w.material(base_color=(0.7, 0.7, 0.7), ior=1.0, roughness=1.0)
    w.material(kind=1, base_color=(1.0, 1.0, 1.0), transmission=1.0, roughness=0.3, ior=1.33)
    p, n, i = uv_sphere((0, 0, 100), 100.0, 20, 40)
    w.mesh(p, i, n, 0)
    quad = np.array([[-2000, -2000, 0], [2000, -2000, 0], [2000, 2000, 0], [-2000, 2000, 0]], np.float32)
    w.mesh(quad, [0, 1, 2, 0, 2, 3], np.tile([0, 0, 1], (4, 1)), 3)
    w.instance(0, material=0)
    w.instance(0, rows=((0.6, 0, 0, -260), (0, 0.6, 0, 80), (0, 0, 0.6, 0)), material=1)
    w.instance(0, rows=((0.5, 0, 0, 230), (0, 0.5, 0, 150), (0, 0, 0.5, 0)), material=2)
    w.instance(0, rows=((0.4, 0, 0, 120), (0, 0.4, 0, -260), (0, 0, 0.4, 0)), material=4, flags=1)
    if floor:
        w.floor(mode=1, z=-1.0, color=(0.5, 0.5, 0.5))
    else:
        w.instance(1, material=3)
    w.light(kind=1 if disk else 0, position=(300, -300, 700), direction=(-0.3, 0.3, -1), size_x=300, size_y=200,
            power=8.0, spread_deg=120.0)
    w.light(position=(-600, 0, 500), direction=(1, 0, -0.4), size_x=150, size_y=150, power=3.0, flags=1)
    w.environment(kind=0, floor=0.05, horizon=0.2, zenith=0.5, background_mode=background_mode,
                  background_color=(0.1, 0.2, 0.3))
    w.render(48, 40, 4, 5, 10, clamp_indirect=4.0)
    return w.bytes()


@pytest.mark.parametrize("variant", [dict(), dict(background_mode=1, floor=1, disk=True),
                                     dict(background_mode=2, fstop=4.0), dict(ortho=True)])
def test_feature_scenes_are_bit_exact(gpu, variant):
    rsb = feature_scene(**variant)
    s = gpu.Scene(rsb)
    s.render_samples(0, 3)
    rgb, w, a, l2 = s.film()
    (orgb, ow, oa, ol2), _ = orc.OracleScene(rsb).render_film(0, 3)
    assert mismatch(rgb, orgb) == 0, variant
    assert bits_equal(a, oa) and bits_equal(l2, ol2)
    assert float(rgb.max()) > 0


def test_tiles_outside_the_image_are_zero_and_bad_tiles_refused(gpu, example_rsb):
    s = gpu.Scene(example_rsb)
    t = s.render_tile(500, 370, 32, 32, 0)
    o, _ = orc.OracleScene(example_rsb).render_tile(500, 370, 12, 14, 0)
    assert bits_equal(t[:14, :12], o)
    assert np.all(t[14:] == 0) and np.all(t[:, 12:] == 0)
    lib = gpu.library()
    out = np.zeros(16, np.float32)
    assert lib.rt_render_tile(s.handle, 0, 0, 0, 4, 0, out.ctypes.data) == 2
    assert lib.rt_render_tile(s.handle, 0, 0, 4096, 4096, 0, out.ctypes.data) == 2


def test_bad_scene_bytes_are_refused_with_the_reason(gpu):
    with pytest.raises(gpu.RsbError, match="magic"):
        gpu.Scene(b"NOPE" + bytes(8))
    w = SceneWriter().material(anisotropy=0.5).render(4, 4)
    with pytest.raises(gpu.RsbError, match="anisotropic"):
        gpu.Scene(w.bytes())


def test_scene_update_keeps_the_meshes(gpu, example_rsb):
    s = gpu.Scene(example_rsb)
    tris = s.info()["triangles"]
    # Only a new camera and render settings: the meshes stay.
    w = SceneWriter()
    w.camera((560, -900, 400), (0, 0, 90), focal_mm=50.0)
    for _ in range(3):
        w.material()
    w.instance(0, material=0).instance(1, material=1)
    w.render(64, 48, 2, 1)
    s.update(w.bytes())
    assert s.info()["triangles"] == tris and (s.width, s.height) == (64, 48)
