"""The oracle's tree traversal against its brute force on the reference's own BVH test cases
(geometry/tests.rs:145-224): 10,000 random rays per geometry (aimed at random triangles of random
instances, axis-aligned with signed zeros, random intervals) over a level-4 icosphere, the example
scene with transformed instances and a 5,000-triangle soup of points, repeated corners,
collinear corners, slivers and duplicates. Nearest (instance, triangle, t) and any hit must agree
exactly."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.geometry_cases import CASES


def test_the_soup_is_the_references():
    # geometry/tests.rs:48-88: 5,000 triangles, among them degenerate ones of every kind.
    p, i = orc.soup()
    tris = p[i.reshape(-1, 3)].astype(np.float64)
    assert tris.shape == (5000, 3, 3)
    area2 = np.linalg.norm(np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]), axis=1)
    points = np.all(tris[:, 0] == tris[:, 1], axis=1) & np.all(tris[:, 0] == tris[:, 2], axis=1)
    repeated = np.all(tris[:, 0] == tris[:, 2], axis=1) & ~points
    assert points.sum() > 100 and repeated.sum() > 100
    edge2 = np.max(np.sum((tris - np.roll(tris, 1, axis=1)) ** 2, axis=2), axis=1)
    flat = (area2 <= 1e-6 * edge2) & (edge2 > 0) & ~repeated
    assert flat.sum() > 100  # collinear corners
    keys = [tuple(t) for t in i.reshape(-1, 3)]
    assert len(set(keys)) < len(keys) - 100  # exact duplicates


@pytest.mark.parametrize("name,make,seed", CASES)
def test_tree_equals_brute_force_on_the_references_cases(name, make, seed, example_rsb):
    rsb = example_rsb if make is None else make()
    o = orc.OracleScene(rsb)
    rays = o.random_rays(10000, seed)
    axis = np.sum(rays[:, 3:6] != 0, axis=1) == 1
    assert axis.sum() > 200 and np.any(np.signbit(rays[axis, 3:6]) & (rays[axis, 3:6] == 0))
    tree = o.trace_rays(rays)
    brute = o.trace_rays(rays, brute=True)
    hits = int(np.sum(brute[0] != 0xFFFFFFFF))
    assert hits > 2000, (name, hits)
    assert np.any((brute[0] == 0xFFFFFFFF) & (brute[3] == 0))
    for what, a, b in zip(["instance", "triangle", "t", "any"], tree, brute):
        assert np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8)), \
            (name, what, int(np.sum(a != b)))
