"""GPU closest and any hit against the oracle's brute force on the reference's own BVH test cases
(geometry/tests.rs:145-224): 10,000 random rays per geometry over the level-4 icosphere, the
example scene with transformed instances and the 5,000-triangle soup of degenerate triangles.
Bar: bit-exact in instance, triangle, t and any."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.geometry_cases import CASES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,make,seed", CASES)
def test_gpu_hits_equal_brute_force_on_the_references_cases(gpu, name, make, seed, example_rsb):
    rsb = example_rsb if make is None else make()
    o = orc.OracleScene(rsb)
    rays = o.random_rays(10000, seed)
    brute = o.trace_rays(rays, brute=True)
    print("geometry case", name, "rays", rays.shape[0], flush=True)
    s = gpu.Scene(rsb)
    got = s.trace_rays(rays)
    s.close()
    assert int(np.sum(brute[0] != 0xFFFFFFFF)) > 2000
    for what, a, b in zip(["instance", "triangle", "t", "any"], got, brute):
        a = np.ascontiguousarray(a)
        b = np.ascontiguousarray(b)
        assert a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8)), \
            (name, what, int(np.sum(a != b)))


@pytest.mark.parametrize("stack", [1024, 8])
@pytest.mark.parametrize("name,make,seed", CASES)
def test_one_wave_per_ray_equals_brute_force(gpu, name, make, seed, example_rsb, stack):
    # The drain's traversal of its last paths (a whole wave per ray over a shared stack): the same
    # bits; with 8 stack entries most rays overflow it and finish on one lane.
    rsb = example_rsb if make is None else make()
    o = orc.OracleScene(rsb)
    rays = o.random_rays(10000, seed)
    brute = o.trace_rays(rays, brute=True)
    print("geometry case", name, "one wave per ray, stack", stack, flush=True)
    s = gpu.Scene(rsb)
    got = s.trace_rays(rays, wave_stack=stack)
    s.close()
    for what, a, b in zip(["instance", "triangle", "t", "any"], got, brute):
        a = np.ascontiguousarray(a)
        b = np.ascontiguousarray(b)
        assert a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8)), \
            (name, stack, what, int(np.sum(a != b)))


def test_one_wave_per_ray_inside_the_media_bottle(gpu):
    # Rays from inside the 200,000-triangle bottle, where the longest walks run: the wave's hits are
    # the per-lane traversal's.
    rsb = gpu.write_scene(5, 3)
    o = orc.OracleScene(rsb)
    rays = o.random_rays(20000, 7)
    rng = np.random.default_rng(5)
    inside = rays.copy()
    # The bottle (mm): radius 19 mm or more from z = 1 to 301 about the z axis.
    inside[:, 0] = rng.uniform(-12.0, 12.0, len(rays))
    inside[:, 1] = rng.uniform(-12.0, 12.0, len(rays))
    inside[:, 2] = rng.uniform(20.0, 280.0, len(rays))
    inside[:, 6] = 1e-3
    inside[:, 7] = np.inf
    print("media bottle: 20,000 rays from inside, one wave per ray", flush=True)
    s = gpu.Scene(rsb)
    lane = s.trace_rays(inside)
    wave = s.trace_rays(inside, wave_stack=1024)
    s.close()
    ref = o.trace_rays(inside)
    assert int(np.sum(ref[0] != 0xFFFFFFFF)) > 10000
    for what, a, b, c in zip(["instance", "triangle", "t", "any"], wave, lane, ref):
        for x, y in ((a, b), (a, c)):
            x = np.ascontiguousarray(x)
            y = np.ascontiguousarray(y)
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (what, int(np.sum(x != y)))
