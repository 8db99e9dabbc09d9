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
