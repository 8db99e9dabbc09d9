"""GPU caustics (the photon map of every sample pass, traced and built on the device, and the split
of the path tracer around it) against the oracle through the C ABI. Bar: bit-exact films.

The scenes are the reference's caustic test scenes (caustics/tests.rs:26-63, 389-449, 522-620) at
small photon counts: a glass sphere under a disc light, under a sun straight above (delta) and
under a sun disc, two overlapping casters, polished metal lit from below, and an emissive mesh as
the photons' light beside an absorbing glass block."""
import numpy as np
import pytest

from oracle import oracle as orc
from repackshot_amd import native
from repackshot_amd.rsb import SceneWriter, icosphere

from test_gpu_parity import bits_equal
from test_oracle_caustics import disc_light, sphere_scene, sun_above

pytestmark = pytest.mark.gpu

SUN, GLOSSY, EMISSIVE, LIGHT_FLAG = 2, 1, 2, 1


def sun_disc_scene():
    light = dict(kind=SUN, direction=(0.2, -0.3, -1.0), power=800.0, angle_deg=4.0, flags=GLOSSY)
    return sphere_scene(30.0, light, 3000, (40, 30))


def emitter_scene():
    """An emissive quad sampled as a light (its photons by triangle area) over an absorbing glass
    block and a glass ball on the floor, under a dim sky."""
    w = SceneWriter()
    w.camera((0.0, 260.0, 160.0), (0.0, 15.0, 10.0), focal_mm=50.0, sensor_mm=36.0)
    w.material(kind=1, base_color=(1, 1, 1), roughness=0.0, ior=1.5, transmission=1.0)             # 0 glass
    w.material(base_color=(0.6, 0.9, 0.7), roughness=0.05, ior=1.45, transmission=1.0, absorption_mm=60.0)  # 1 tinted
    w.material(kind=EMISSIVE, emission=(1.0, 0.9, 0.8), emission_power=40.0, flags=LIGHT_FLAG)      # 2 light
    p, i = icosphere((-25.0, 0.0, 20.0), 20.0, 4)
    w.mesh(p, i, None, 0)
    box = np.array([[10, -15, 0], [50, -15, 0], [50, 15, 0], [10, 15, 0],
                    [10, -15, 30], [50, -15, 30], [50, 15, 30], [10, 15, 30]], np.float32)
    faces = [0, 2, 1, 0, 3, 2, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4, 1, 2, 6, 1, 6, 5, 2, 3, 7, 2, 7, 6, 3, 0, 4, 3, 4, 7]
    w.mesh(box, faces, None, 1)
    quad = np.array([[-40, -60, 160], [40, -60, 160], [40, 20, 160], [-40, 20, 160]], np.float32)
    w.mesh(quad, [0, 2, 1, 0, 3, 2], None, 2)
    w.environment(kind=0, strength=0.05)
    w.floor(mode=1, z=0.0, color=(0.8, 0.8, 0.8))
    w.render(40, 30, 2, 5, 10, caustic_photons=3000)
    return w.bytes()


SCENES = {
    "glass under a disc": lambda: sphere_scene(30.0, disc_light(), 3000, (40, 30)),
    "glass under the delta sun": lambda: sphere_scene(30.0, sun_above(1000.0), 3000, (40, 30)),
    "glass under a sun disc": sun_disc_scene,
    "overlapping casters": lambda: sphere_scene(30.0, disc_light(), 3000, (40, 30),
                                                spheres=[((0.0, 0.0, 30.0), 30.0, 4), ((0.0, -40.0, 72.5), 12.0, 3)]),
    "polished metal from below": lambda: sphere_scene(
        30.0, disc_light(position=(0.0, 0.0, 100.0), target_dir=(0.0, 0.0, 1.0), size=40.0, up=(0, 1, 0)), 3000,
        (40, 30), spheres=[((0.0, 0.0, 190.0), 40.0, 4)],
        material=dict(base_color=(0.9, 0.9, 0.9), metallic=1.0, roughness=0.2)),
    "emissive mesh and absorbing glass": emitter_scene,
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_caustic_films_equal_the_oracle(gpu, name):
    rsb = SCENES[name]()
    print("caustic scene", name, flush=True)
    o = orc.OracleScene(rsb)
    assert o.caustics()["on"]
    want, stats = o.render_film(0, 3)
    s = gpu.Scene(rsb)
    s.render_samples(0, 3)
    got = s.film()
    for plane, a, b in zip(["rgb", "weight", "alpha", "lum2"], got, want):
        assert bits_equal(a, b), (name, plane, int(np.sum(a.view(np.uint32) != b.view(np.uint32))))
    assert float(np.mean(got[0])) > 0.0
    assert s.stats()["rays"] == stats["rays"]
    # Another pool and launch shape give the same bits (the photon maps do not depend on them).
    s.set_options(pool=777, shape=5, shadow_shape=6)
    s.reset_film()
    s.render_samples(0, 3)
    assert all(bits_equal(a, b) for a, b in zip(s.film(), want)), name
    s.close()


def test_a_quarter_million_photons_per_pass_equal_the_oracle(gpu):
    # 2^18 photons a pass: a table of 2^19 buckets, sorted and gathered at scale.
    rsb = sphere_scene(30.0, disc_light(size=60.0), 1 << 18, (48, 36))
    want, _ = orc.OracleScene(rsb).render_film(3, 2)
    s = gpu.Scene(rsb)
    s.render_samples(3, 2)
    got = s.film()
    assert all(bits_equal(a, b) for a, b in zip(got, want))
    s.close()


def test_caustic_tiles_blocks_and_the_allocation_free_tile_equal_the_oracle(gpu):
    rsb = SCENES["glass under a disc"]()
    o = orc.OracleScene(rsb)
    s = gpu.Scene(rsb)
    for k in (0, 4):
        want_rgb, want_alpha = o.render_tile(8, 6, 24, 18, k)[:2]
        got = s.render_tile(8, 6, 24, 18, k)
        assert bits_equal(got, want_rgb), k
        rgba = s.render_tile_rgba(8, 6, 24, 18, k)
        assert bits_equal(np.ascontiguousarray(rgba[..., :3]), want_rgb), k
        assert bits_equal(np.ascontiguousarray(rgba[..., 3]), want_alpha), k
        # A block of pass k with the map of pass k is the tile; the map of another pass differs.
        assert bits_equal(s.render_block_rgba(8, 6, 24, 18, k, iteration=k), rgba), k
        assert not bits_equal(s.render_block_rgba(8, 6, 24, 18, k, iteration=k + 7), rgba), k
        assert bits_equal(s.render_tile_noalloc(8, 6, 24, 18, k), want_rgb), k
    s.close()


def test_tiles_of_passes_out_of_order_reuse_their_kept_maps(gpu):
    # rt-wasm/src/lib.rs:1311-1351: tiles of samples 0, 1, 0, 2 each equal the engine's tile with
    # that sample's map; an update to no photons leaves a scene without caustics.
    rsb = SCENES["overlapping casters"]()
    o = orc.OracleScene(rsb)
    s = gpu.Scene(rsb)
    assert s.caustics() == 3000
    assert native.library().rt_scene_caustics(999) == 0xFFFFFFFF
    for k in (0, 1, 0, 2, 1):
        assert bits_equal(s.render_tile(12, 4, 20, 16, k), o.render_tile(12, 4, 20, 16, k)[0]), k
    plain = sphere_scene(30.0, disc_light(), 0, (40, 30),
                         spheres=[((0.0, 0.0, 30.0), 30.0, 4), ((0.0, -40.0, 72.5), 12.0, 3)])
    s.update(plain)
    assert s.caustics() == 0
    assert bits_equal(s.render_tile(12, 4, 20, 16, 0), orc.OracleScene(plain).render_tile(12, 4, 20, 16, 0)[0])
    s.close()


def test_material_changes_rebuild_the_caster_set(gpu):
    # Roughening the glass past 0.25 leaves nothing that focuses light: the films follow the oracle
    # of the changed scene either way.
    rsb = SCENES["glass under a disc"]()
    s = gpu.Scene(rsb)
    params = s.material_params()
    s.render_tile(0, 0, 40, 30, 1)
    rough = params.copy()
    rough[0, 0] = 0.4
    s.set_material_params(rough)
    glass = dict(kind=1, base_color=(1, 1, 1), roughness=0.4, ior=1.5, transmission=1.0)
    changed = sphere_scene(30.0, disc_light(), 3000, (40, 30), material=glass)
    assert not orc.OracleScene(changed).caustics()["on"]
    assert bits_equal(s.render_tile(0, 0, 40, 30, 1), orc.OracleScene(changed).render_tile(0, 0, 40, 30, 1)[0])
    s.set_material_params(params)
    assert bits_equal(s.render_tile(0, 0, 40, 30, 1), orc.OracleScene(rsb).render_tile(0, 0, 40, 30, 1)[0])
    s.close()


def test_queue_and_adaptive_refuse_a_scene_with_caustics(gpu):
    s = gpu.Scene(SCENES["glass under a disc"]())
    q = gpu.Queue()
    with pytest.raises(native.RenderError):
        q.add(s, 0, 1)
    with pytest.raises(native.RenderError):
        s.render_adaptive(4)
    q.close()
    s.close()
