"""Scenes of the reference's BVH-against-brute-force tests (geometry/tests.rs:36-88), written as RSB
bytes for the oracle and the GPU: the level-4 icosphere, the example scene with its instances,
and the 5,000-triangle soup of random and degenerate triangles. The rays come from the oracle's
restatement of the same tests' random_rays (geometry/tests.rs:92-140)."""
from oracle import oracle as orc
from repackshot_amd.rsb import SceneWriter, icosphere


def _scene(positions, indices):
    w = SceneWriter()
    w.camera((0, -1000.0, 0), (0, 0, 0))
    w.material()
    w.mesh(positions, indices, None, 0)
    w.render(64, 64, 1, 0, 4)
    return w.bytes()


def icosphere_rsb():
    p, i = icosphere((10.0, -20.0, 100.0), 100.0, 4)
    return _scene(p, i)


def soup_rsb():
    p, i = orc.soup()
    return _scene(p, i)


# (name, scene bytes or "example", seed of random_rays) as in geometry/tests.rs:206-224.
CASES = [("icosphere", icosphere_rsb, 1), ("example", None, 2), ("soup", soup_rsb, 3)]
