"""The path pool and the render queue, through the C ABI: films must not depend on the pool size,
the traversal launch shape, the chunking of the sample range, per-pass or persistent scheduling,
graph launches or the drain of the last paths, and a queue's films must equal each scene rendered
on its own and the oracle.

Bar: bit-exact (raw f32 bits of every film sum)."""
import numpy as np
import pytest

from oracle import oracle as orc

from test_gpu_parity import bits_equal, feature_scene

pytestmark = pytest.mark.gpu


def films_equal(a, b):
    return all(bits_equal(x, y) for x, y in zip(a, b))


@pytest.fixture(scope="module")
def plain(gpu):
    rsb = gpu.write_scene(1, 0)
    film, _ = orc.OracleScene(rsb).render_film(0, 3)
    return rsb, film


@pytest.mark.parametrize("opts", [
    dict(pool=64),                      # many regenerations per sample
    dict(pool=5000, group=2),
    dict(pool=1 << 16, shape=1),
    dict(pool=1 << 16, shape=2),
    dict(pool=1 << 16, shape=3),
    dict(pool=1 << 16, shape=4),
    dict(pool=1 << 16, shape=5, shadow_shape=6),        # resident lanes
    dict(pool=3000, shape=6, shadow_shape=7),
    dict(pool=1 << 16, shape=7, shadow_shape=5),
    dict(pool=1 << 16, shape=1, shadow_shape=4),
    dict(per_pass=True),
    dict(graph=True, group=4),
    dict(pool=3000, graph=True),
    dict(chunk=2 * 256 * 256 + 17),     # chunks of two sample passes
    dict(drain=False),                  # the iteration loop to the end
    dict(pool=5000, tail=3000),         # drain after the last regeneration
    dict(pool=1 << 16, tail=1 << 30),   # the whole chunk drained at once
    dict(tail=64, shape=5),
    dict(pool=3000, graph=True, group=2, tail=2000),
])
def test_pool_options_do_not_change_the_film(gpu, plain, opts):
    rsb, want = plain
    s = gpu.Scene(rsb)
    s.set_options(**opts)
    s.render_samples(0, 3)
    assert films_equal(s.film(), want), opts


def test_graph_replays_match_eager_renders(gpu):
    rsb = feature_scene()
    s = gpu.Scene(rsb)
    (want, _) = orc.OracleScene(rsb).render_film(0, 2)
    # Several graph renders in a row, with eager renders and synchronisations between them.
    for k in range(3):
        s.set_options(pool=700, graph=True, group=2)
        s.reset_film()
        s.render_samples(0, 2)
        assert films_equal(s.film(), want), k
        s.set_options(pool=700)
        s.reset_film()
        s.render_samples(0, 2)
        assert films_equal(s.film(), want), k


def test_profiled_render_counts_its_launches(gpu, plain):
    rsb, want = plain
    s = gpu.Scene(rsb)
    s.set_profiling(True)
    s.render_samples(0, 3)
    t = s.last_timing()
    assert t["extend_launches"] == t["shade_launches"] == t["shadow_launches"] == t["iterations"] > 0
    assert t["extend_ms"] > 0 and t["total_ms"] >= t["extend_ms"]
    assert films_equal(s.film(), want)


def test_queue_films_equal_single_scene_renders(gpu, example_rsb):
    scenes = [gpu.Scene(example_rsb), gpu.Scene(gpu.write_scene(1, 0)), gpu.Scene(feature_scene()),
              gpu.Scene(feature_scene(background_mode=1, floor=1, disk=True))]
    jobs = [(0, 0, 2), (1, 1, 3), (2, 0, 4), (3, 2, 2), (2, 4, 1)]
    q = gpu.Queue()
    q.set_options(pool=20000)
    ids = [q.add(scenes[si], first, count) for si, first, count in jobs]
    q.run()
    for jid, (si, first, count) in zip(ids, jobs):
        s = scenes[si]
        s.reset_film()
        s.render_samples(first, count)
        assert films_equal(q.film(jid), s.film()), (si, first, count)
    # The oracle agrees for the small scene, samples 2..3.
    (want, _) = orc.OracleScene(feature_scene(background_mode=1, floor=1, disk=True)).render_film(2, 2)
    assert films_equal(q.film(ids[3]), want)
    # A second run adds the same samples again: every weight doubles.
    q.run()
    for jid, (si, first, count) in zip(ids, jobs):
        assert np.all(q.film(jid)[1] == 2.0 * count)
    st = q.stats()
    assert st["rays"] > 0 and st["samples"] == sum(c for _, _, c in jobs)


def test_queue_graph_and_per_pass_agree(gpu):
    a = gpu.Scene(gpu.write_scene(1, 0))
    b = gpu.Scene(feature_scene(ortho=True))
    films = []
    for opts in (dict(pool=4000), dict(pool=4000, graph=True), dict(per_pass=True), dict(chunk=100000)):
        q = gpu.Queue()
        q.set_options(**opts)
        ja = q.add(a, 0, 2)
        jb = q.add(b, 1, 3)
        q.run()
        films.append((q.film(ja), q.film(jb)))
        q.close()
    for f in films[1:]:
        assert films_equal(f[0], films[0][0]) and films_equal(f[1], films[0][1])


@pytest.mark.parametrize("variant,splits", [(dict(), (16, 54)), (dict(fstop=4.0), (70,)),
                                            (dict(background_mode=1, floor=1, disk=True), (5, 20, 45))])
def test_many_samples_per_call_are_bit_exact(gpu, variant, splits):
    # From 16 samples per call the film takes each pixel's samples with one wave (a contiguous read
    # of its results, added in sample order), in rounds of 64 (70 is ragged); calls add onto the
    # film. Every work order, chunking and option gives the oracle's sums.
    rsb = feature_scene(**variant)
    total = sum(splits)
    want, _ = orc.OracleScene(rsb).render_film(0, total)
    s = gpu.Scene(rsb)
    for opts in (dict(), dict(pool=700), dict(chunk=3000), dict(per_pass=True)):
        s.set_options(**opts)
        s.reset_film()
        first = 0
        for n in splits:
            s.render_samples(first, n)
            first += n
        assert films_equal(s.film(), want), (variant, splits, opts)
