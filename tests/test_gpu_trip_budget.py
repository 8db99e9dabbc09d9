"""The closest-hit trip budget: a walk that has made its budget of loop trips is set aside, with its
stack, for a second launch that takes it to its end. The closest hit is the least (t, instance,
triangle), which no order of visits changes, so films must not depend on the budget, the launch
shape, the pool, graph launches or the continuation store filling up (walks that do not fit go on
in the first launch); the resident-lane shapes run unsplit.

Bar: bit-exact (raw f32 bits of every film sum) against the oracle."""
import numpy as np
import pytest

from oracle import oracle as orc

from test_gpu_parity import bits_equal, feature_scene

pytestmark = pytest.mark.gpu


def films_equal(a, b):
    return all(bits_equal(x, y) for x, y in zip(a, b))


@pytest.fixture(scope="module")
def featured(gpu):
    rsb = feature_scene()  # a transformed instance: walks set aside in its object space
    film, _ = orc.OracleScene(rsb).render_film(0, 2)
    return rsb, film


@pytest.fixture(scope="module")
def plain(gpu):
    rsb = gpu.write_scene(1, 0)
    film, _ = orc.OracleScene(rsb).render_film(0, 2)
    return rsb, film


CASES = [
    (1, dict(pool=1 << 16, shape=2)),   # nearly every walk set aside, more than the store holds
    (3, dict(pool=5000, shape=1)),      # a store of 1,250 walks
    (16, dict(pool=1 << 16, shape=3)),
    (48, dict(pool=1 << 16, shape=4)),
    (7, dict(pool=1 << 16, shape=0)),
    (48, dict(pool=700, graph=True, group=2)),
    (2, dict(pool=3000, tail=2000)),    # with the drain
    (5, dict(pool=1 << 16, shape=5, shadow_shape=6)),  # resident lanes: one launch
]


@pytest.mark.parametrize("budget,opts", CASES)
def test_trip_budget_does_not_change_the_film(gpu, featured, plain, budget, opts):
    for rsb, want in (featured, plain):
        s = gpu.Scene(rsb)
        s.set_options(**opts)
        s.set_trip_budget(budget)
        assert s.trip_budget() == budget
        s.render_samples(0, 2)
        assert films_equal(s.film(), want), (budget, opts)
        s.close()


def test_trip_budget_in_a_queue_and_the_adjoint(gpu, featured):
    """The queue and the adjoint's forward pass trace with the engine's budget (default 48)."""
    rsb, want = featured
    s = gpu.Scene(rsb)
    assert s.trip_budget() == 48
    q = gpu.Queue()
    q.set_options(pool=3000)
    j = q.add(s, 0, 2)
    q.run()
    assert films_equal(q.film(j), want)
    q.close()
    d = np.ones((s.height, s.width, 3), np.float32)
    g48 = s.adjoint(0, 2, d)
    s.set_trip_budget(0)
    g0 = s.adjoint(0, 2, d)
    # Gradient sums are float atomics: equal up to the rounding of their order.
    assert np.abs(g48[0] - g0[0]).max() <= 1e-4 * np.abs(g0[0]).max()
    assert np.abs(g0[0]).max() > 0


def test_tuned_trip_budget_is_a_candidate_and_keeps_the_film(gpu, plain):
    rsb, want = plain
    s = gpu.Scene(rsb)
    t = s.tune_trip_budget(2)
    assert t["trip_budget"] in (0, 48) and s.trip_budget() == t["trip_budget"]
    assert all(v > 0 for v in t["ms"].values())
    assert t["ms"][t["trip_budget"]] == min(t["ms"].values())
    s.reset_film()
    s.render_samples(0, 2)
    assert films_equal(s.film(), want)
    a = s.autotune(2)
    assert a["trip_budget"] in (0, 48) and s.trip_budget() == a["trip_budget"]
    s.reset_film()
    s.render_samples(0, 2)
    assert films_equal(s.film(), want)
