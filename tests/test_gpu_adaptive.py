"""Adaptive sampling on the GPU (adaptive.rs, the reference's E10) against the oracle and the
reference's pinned decisions: the GPU's decisions over the synthetic image hash to
PINNED_E10_DECISIONS (adaptive/tests.rs:455-534); decisions over random planes with ragged edges,
unmeasured pixels, non-finite sums, capped blocks and a zero target equal the oracle's bit for bit;
and an adaptive render (rounds of blocks in HBM) gives the oracle's film and block counts bit for
bit."""
import numpy as np
import pytest

from oracle import oracle as orc
from test_oracle_adaptive import PINNED_E10_DECISIONS, adaptive_scene, synthetic_decisions

pytestmark = pytest.mark.gpu


def test_gpu_decisions_have_the_pinned_hash(gpu):
    assert synthetic_decisions(6, gpu.Adaptive) == PINNED_E10_DECISIONS


@pytest.mark.parametrize("w,h,target,ev,cap,seed", [
    (45, 37, 0.02, 0.5, 200, 1),   # ragged edges, capped
    (64, 64, 0.0, 0.0, 0, 2),      # no target, no cap: every measured block is a candidate
    (130, 17, 0.1, -2.0, 40, 3),   # wide and short, low exposure
    (8, 8, 0.05, 0.0, 100, 4),     # one block
])
def test_gpu_decisions_equal_the_oracle(gpu, w, h, target, ev, cap, seed):
    rng = np.random.default_rng(seed)
    o = orc.Adaptive(w, h, target, ev, cap)
    g = gpu.Adaptive(w, h, target, ev, cap)
    for rnd in range(5):
        o.advance(o.left())
        g.advance(g.left())
        n = np.array([[o.samples_of_pixel(x, y) for x in range(w)] for y in range(h)], np.float32)
        gn = np.array([[g.samples_of_pixel(x, y) for x in range(w)] for y in range(h)], np.float32)
        assert np.array_equal(n, gn)
        c = rng.random((h, w, 3), np.float32) * rng.choice([0.0, 0.01, 1.0, 5.0], (h, w, 1)).astype(np.float32)
        spread = rng.random((h, w), np.float32) * np.float32(0.4)
        sums = np.zeros((h, w, 4), np.float32)
        sums[..., :3] = c * n[..., None]
        sums[..., 3] = n
        l = np.float32(0.2126) * c[..., 0] + np.float32(0.7152) * c[..., 1] + np.float32(0.0722) * c[..., 2]
        lum2 = (l * l + spread * spread) * n
        if rnd == 1:
            sums[0, 0, 0] = np.inf       # a non-finite pixel is unmeasured
            lum2[h - 1, w - 1] = np.nan
        out_o, act_o = o.decide(sums, 4, lum2)
        out_g, act_g = g.decide(sums, 4, lum2)
        assert np.array_equal(out_o.view(np.uint32), out_g.view(np.uint32)), (rnd, out_o, out_g)
        assert np.array_equal(act_o, act_g), rnd
        if out_o[0] == 0:
            break


def test_decide_rejects_an_empty_image(gpu):
    z = np.zeros(4, np.uint8)
    st = gpu.library().rt_adaptive_decide(0, 8, 0.0, 0.1, 16, gpu._ptr(z), gpu._ptr(z), gpu._ptr(z), gpu._ptr(z),
                                          gpu._ptr(z))
    assert st == 2
    st = gpu.library().rt_adaptive_decide(8, 8, 0.0, 0.1, 16, None, gpu._ptr(z), gpu._ptr(z), gpu._ptr(z),
                                          gpu._ptr(z))
    assert st == 3


def test_block_render_is_the_tile_render(gpu):
    s = gpu.Scene(adaptive_scene())
    a = s.render_block_rgba(8, 16, 8, 8, 5, iteration=3)
    b = s.render_tile_rgba(8, 16, 8, 8, 5)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("cap", [64, 24])
def test_adaptive_render_equals_the_oracle(gpu, cap):
    rsb = adaptive_scene()
    o = orc.OracleScene(rsb)
    film_o, counts_o, decisions_o = o.render_adaptive(cap=cap, threads=8)
    s = gpu.Scene(rsb)
    stats = s.render_adaptive(cap)
    counts = s.adaptive_counts().reshape(-1)
    assert np.array_equal(counts, counts_o)
    assert stats["decisions"] == decisions_o
    assert stats["min_samples"] == counts_o.min() and stats["max_samples"] == counts_o.max()
    cols = (s.width + 7) // 8
    bw = np.array([min(8, s.width - (b % cols) * 8) for b in range(counts.size)])
    bh = np.array([min(8, s.height - (b // cols) * 8) for b in range(counts.size)])
    assert stats["paths"] == int((counts.astype(np.int64) * bw * bh).sum())
    film_g = s.film()
    for a, b in zip(film_g, film_o):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    if cap == 64:
        assert counts.max() > counts.min()


def test_adaptive_render_twice_is_the_same(gpu):
    """The film is cleared by each adaptive render, so a second one repeats the first."""
    s = gpu.Scene(adaptive_scene())
    s.render_samples(0, 3)
    a = s.render_adaptive(32)
    fa = s.film()
    b = s.render_adaptive(32)
    fb = s.film()
    assert {k: a[k] for k in a if k != "gpu_ms"} == {k: b[k] for k in b if k != "gpu_ms"}
    for x, y in zip(fa, fb):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


@pytest.mark.parametrize("cap,chunk", [(64, 0), (24, 0), (64, 40000)])
def test_samples_rendered_ahead_give_the_rounds_film(gpu, cap, chunk):
    """A candidate block's samples up to the cap rendered in one run and kept for its rounds (the
    default) against every round rendering only its own samples: the same decisions, block counts
    and film bit for bit, also when the kept samples need several runs (a small chunk)."""
    rsb = adaptive_scene()
    s = gpu.Scene(rsb)
    s.set_options(chunk=chunk, adaptive_ahead=False)
    a = s.render_adaptive(cap)
    wa = s.adaptive_work()
    fa, ca = s.film(), s.adaptive_counts()
    s.set_options(chunk=chunk)
    b = s.render_adaptive(cap)
    wb = s.adaptive_work()
    fb, cb = s.film(), s.adaptive_counts()
    assert {k: a[k] for k in a if k != "gpu_ms"} == {k: b[k] for k in b if k != "gpu_ms"}
    assert np.array_equal(ca, cb)
    for x, y in zip(fa, fb):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    # Without rendering ahead every rendered sample is used; with it, the used ones are the same.
    assert wa["rendered"] == a["paths"] and wa["rendered_ahead"] == 0 and wa["unused"] == 0
    assert wb["rendered"] - wb["unused"] == b["paths"]
    assert wb["rendered_ahead"] > 0 and wb["runs_ahead"] >= 1
    if chunk:
        assert wb["runs_ahead"] >= 2
