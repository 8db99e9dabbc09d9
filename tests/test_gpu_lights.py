"""GPU films of scenes lit by suns (disc and delta), emissive meshes sampled as lights (one- and
two-sided, transformed instances), more than 64 lights (the static selection by flux) and one light
of every kind, against the oracle through the C ABI. Bar: bit-exact film sums."""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from repackshot_amd.rsb import SceneWriter, icosphere

from test_gpu_parity import bits_equal

pytestmark = pytest.mark.gpu

SUN, POINT, SPOT, DISK, RECT = 2, 3, 4, 1, 0
EMISSIVE, LIGHT_FLAG = 2, 1
GLOSSY, CAMERA = 1, 2


def base_scene(w, camera="side"):
    if camera == "side":
        w.camera((0.0, -1600.0, 500.0), (0.0, 0.0, 250.0), focal_mm=30.0)
    else:
        w.camera((0.0, 0.0, 2000.0), (0.0, 0.0, 0.0), up=(0, 1, 0), ortho=True, ortho_width=1400.0)
    w.material(base_color=(0.8, 0.8, 0.8), ior=1.0)                              # 0 floor
    w.material(base_color=(0.95, 0.64, 0.54), metallic=1.0, roughness=0.3)       # 1 metal
    w.material(kind=1, base_color=(1, 1, 1), transmission=1.0, roughness=0.0, ior=1.5)  # 2 glass
    w.material(base_color=(0.2, 0.5, 0.8), roughness=0.4, ior=1.5)               # 3 plastic
    q = np.array([[-3000, -3000, 0], [3000, -3000, 0], [3000, 3000, 0], [-3000, 3000, 0]], np.float32)
    w.mesh(q, [0, 1, 2, 0, 2, 3], None, 0)
    p, i = icosphere((-250.0, 0.0, 200.0), 200.0, 3)
    w.mesh(p, i, None, 1)
    p, i = icosphere((250.0, 100.0, 150.0), 150.0, 3)
    w.mesh(p, i, None, 2)
    p, i = icosphere((0.0, 400.0, 120.0), 120.0, 2)
    w.mesh(p, i, None, 3)


def sun_scene(angle):
    w = SceneWriter()
    base_scene(w)
    w.light(kind=SUN, direction=(0.3, 0.9, -0.35), power=3.0, angle_deg=angle, flags=GLOSSY | CAMERA)
    w.environment(kind=0, strength=0.3)
    w.render(64, 48, 4, 0, 8)
    return w.bytes()


def emitter_scene():
    w = SceneWriter()
    base_scene(w, camera="top")
    w.material(kind=EMISSIVE, emission=(1.0, 0.8, 0.6), emission_power=20.0, flags=LIGHT_FLAG)   # 4
    w.material(kind=EMISSIVE, emission=(0.5, 0.7, 1.0), emission_power=6.0, flags=LIGHT_FLAG)    # 5
    w.material(kind=EMISSIVE, emission=(1.0, 1.0, 1.0), emission_power=3.0)                       # 6: not a light
    quad = np.array([[-300, -500, 600], [300, -500, 600], [300, -300, 600], [-300, -300, 600]], np.float32)
    w.mesh(quad, [0, 2, 1, 0, 3, 2], None, 4)          # mesh 4: facing down
    p, i = icosphere((0.0, 0.0, 0.0), 60.0, 2)
    w.mesh(p, i, None, 5)                               # mesh 5: a glowing ball (instanced below)
    small = np.array([[500, 300, 10], [600, 300, 10], [600, 400, 10]], np.float32)
    w.mesh(small, [0, 1, 2], None, 6)                   # mesh 6: emissive, not sampled
    for m in range(4):
        w.instance(m, material=[0, 1, 2, 3][m])
    w.instance(4, material=4, flags=1)                   # one-sided
    w.instance(5, rows=((1.0, 0, 0, -450.0), (0, 0.5, 0, 350.0), (0, 0, 1.5, 350.0)), material=5)
    w.instance(6, material=6)
    w.environment(kind=0, strength=0.05)
    w.render(64, 48, 4, 0, 8)
    return w.bytes()


def many_lights_scene(n=70):
    w = SceneWriter()
    base_scene(w, camera="top")
    for k in range(n):
        x, y = -1000.0 + 250.0 * (k % 9), -1000.0 + 250.0 * (k // 9)
        if k % 3 == 0:
            w.light(kind=POINT, position=(x, y, 500.0), size_x=40.0, size_y=40.0, power=2.0 + 0.1 * k, flags=GLOSSY)
        else:
            w.light(kind=RECT, position=(x, y, 700.0), direction=(0, 0, -1), size_x=120.0, size_y=80.0,
                    power=1.0 + 0.05 * k, flags=GLOSSY)
    w.environment(kind=0, strength=0.1)
    w.render(64, 48, 4, 0, 8)
    return w.bytes()


def every_kind_scene():
    w = SceneWriter()
    base_scene(w)
    w.material(kind=EMISSIVE, emission=(1, 1, 1), emission_power=4.0, flags=LIGHT_FLAG)  # 4
    quad = np.array([[-400, 300, 100], [-200, 300, 100], [-200, 300, 300], [-400, 300, 300]], np.float32)
    w.mesh(quad, [0, 1, 2, 0, 2, 3], None, 4)
    w.light(kind=RECT, position=(0, 0, 800.0), direction=(0, 0, -1), up=(0, 1, 0), size_x=600.0, size_y=600.0,
            power=20.0)
    w.light(kind=DISK, position=(700.0, 0, 300.0), direction=(-1.0, 0.0, -0.2), size_x=300.0, size_y=300.0, power=8.0)
    w.light(kind=SPOT, position=(-300.0, -300.0, 250.0), direction=(0, 0, -1), size_x=80.0, size_y=80.0, power=3.0,
            spread_deg=120.0, angle_deg=20.0, flags=GLOSSY)
    w.light(kind=SUN, direction=(0.3, 0.2, -1.0), power=2.0, angle_deg=8.0)
    w.environment(kind=0, strength=1.0)
    w.render(64, 48, 4, 0, 8)
    return w.bytes()


SCENES = {"sun disc": lambda: sun_scene(6.0), "delta sun": lambda: sun_scene(0.0), "emissive meshes": emitter_scene,
          "70 lights by flux": many_lights_scene, "every kind": every_kind_scene}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_lit_films_equal_the_oracle(gpu, name):
    rsb = SCENES[name]()
    print("lights scene", name, flush=True)
    want, _ = orc.OracleScene(rsb).render_film(0, 4)
    s = gpu.Scene(rsb)
    s.render_samples(0, 4)
    got = s.film()
    for plane, a, b in zip(["rgb", "weight", "alpha", "lum2"], got, want):
        assert bits_equal(a, b), (name, plane, int(np.sum(a.view(np.uint32) != b.view(np.uint32))))
    assert float(np.mean(got[0])) > 0.0
    # The same under another pool and launch shape, and in a queue beside the plain scene.
    s.set_options(pool=999, shape=5, shadow_shape=6)
    s.reset_film()
    s.render_samples(0, 4)
    assert all(bits_equal(a, b) for a, b in zip(s.film(), want)), name
    s.close()
