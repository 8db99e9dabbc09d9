"""The oracle's passes and denoiser against the reference's own pins and tests (CPU only).

Pins: the seven pass planes of the example at 4 samples of albedo
(rt-core/src/render/passes_tests.rs:155-181, 367-381, 841-866) and the denoised example at 4 samples
with the full filter and at 2 and 4 samples with the viewport's preset
(rt-post/src/denoise/tests.rs:369-392, 489-528). Properties: the reference tests' synthetic film
(rt-post/src/denoise/tests.rs:22-298, 531-620): the filter lowers the error, never mixes two parts,
leaves the background and the coverage alone, a band is exactly those rows of the whole image, the
strength blends bit for bit, per-pixel counts weigh the variance, and a clean image stays clean."""
import numpy as np
import pytest

from oracle import oracle as orc

PINNED_PASSES = {"albedo": 0x43571DA9, "normal": 0x0AABFD24, "depth": 0x25F8DD11, "partId": 0xCCD3D6C7,
                 "materialId": 0xCCD3D6C7, "coc": 0xDF87494C, "alpha": 0xD4EAAC78}
PINNED_DENOISED_EXAMPLE_4SPP = 0x2A430F92
PINNED_DENOISED_VIEWPORT_EXAMPLE_2SPP = 0x7000CEB2
PINNED_DENOISED_VIEWPORT_EXAMPLE_4SPP = 0x5318D4C5
PINNED_EXAMPLE_4SPP = 0x75BA4CB4
HALO, VIEWPORT_HALO = 72, 24


@pytest.fixture(scope="module")
def example(example_rsb):
    return orc.OracleScene(example_rsb)


def example_planes(o, spp):
    film, _ = o.render_film(0, spp)
    rgb, _, alpha, lum2 = film
    ps = o.render_passes(0, spp)
    return {"sums": np.concatenate([rgb, alpha[..., None]], axis=2), "lum2": lum2, "albedo": ps["albedo"],
            "normal": ps["normal"], "depth": ps["depth"], "partId": ps["partId"], "materialId": ps["materialId"]}


def means(out, spp):
    inv = np.float32(1.0) / np.float32(spp)
    return (out[..., :3] * inv).astype(np.float32)


def band(planes, y0, y1):
    return {k: v[y0:y1] for k, v in planes.items()}


def test_example_passes_have_the_pinned_hashes(example):
    ps = example.render_passes(0, 4)
    for k, pin in PINNED_PASSES.items():
        assert orc.fnv1a(ps[k]) == pin, k
    # Background, object and edge pixels in the alpha plane; the floor's ids.
    assert (ps["alpha"] == 0).any() and (ps["alpha"] == 1).any() and ((ps["alpha"] > 0) & (ps["alpha"] < 1)).any()
    assert np.isinf(ps["depth"]).any() and (ps["partId"] == 0).any()


@pytest.mark.parametrize("spp,preset,pin", [(4, orc.FULL, PINNED_DENOISED_EXAMPLE_4SPP),
                                            (2, orc.VIEWPORT, PINNED_DENOISED_VIEWPORT_EXAMPLE_2SPP),
                                            (4, orc.VIEWPORT, PINNED_DENOISED_VIEWPORT_EXAMPLE_4SPP)])
def test_denoised_example_has_the_pinned_hash(example, spp, preset, pin):
    W, H = example.width, example.height
    planes = example_planes(example, spp)
    st, out = orc.denoise(W, H, 0, H, spp, 0, H, planes, preset)
    assert st == 0
    assert orc.fnv1a(means(out, spp)) == pin
    # Splitting the rows into bands, tall or short, changes no pixel.
    halo = HALO if preset == orc.FULL else VIEWPORT_HALO
    parts = []
    for k in range(5):
        y0, y1 = H * k // 5, H * (k + 1) // 5
        lo, hi = max(y0 - halo, 0), min(y1 + halo, H)
        st, rows = orc.denoise(W, H, lo, hi - lo, spp, y0, y1, band(planes, lo, hi), preset)
        assert st == 0
        parts.append(rows)
    assert np.array_equal(np.concatenate(parts).view(np.uint32), out.view(np.uint32))
    if spp == 4 and preset == orc.FULL:
        film, _ = example.render_film(0, 4)
        inv = np.float32(1.0) / np.float32(4.0)
        assert orc.fnv1a((film[0] * inv).astype(np.float32)) == PINNED_EXAMPLE_4SPP


def synthetic(w, h, spp, left, right, spread):
    return orc.denoise_synthetic(w, h, spp, left, right, spread)


def run(planes, spp, preset=orc.FULL, strength=1.0, counts=None, origin=0, rows=None, y0=None, y1=None, height=None):
    h, w = planes["depth"].shape
    rows = h if rows is None else rows
    height = h if height is None else height
    y0 = origin if y0 is None else y0
    y1 = origin + rows if y1 is None else y1
    st, out = orc.denoise(w, height, origin, rows, spp, y0, y1, planes, preset, strength, counts)
    return st, out


def mse(planes, out, id_, truth, spp):
    m = planes["partId"] == id_
    inv = np.float32(1.0) / np.float32(spp)
    d = (out[..., 1] * inv - np.float32(truth)).astype(np.float64)[m]
    return float(np.mean(d * d))


@pytest.mark.parametrize("spp", [1, 2, 4, 16])
def test_it_lowers_the_error_of_a_noisy_flat_surface(spp):
    s = synthetic(48, 40, spp, 0.5, 2.0, 1.0)
    for preset, bound in ((orc.FULL, 0.25), (orc.VIEWPORT, 0.35)):
        st, out = run(s, spp, preset)
        assert st == 0
        for id_, truth in ((1, 0.5), (2, 2.0)):
            before, after = mse(s, s["sums"], id_, truth, spp), mse(s, out, id_, truth, spp)
            assert after < bound * before, (spp, preset, id_, before, after)


def test_it_never_mixes_two_parts_and_leaves_the_background_alone():
    s = synthetic(40, 32, 8, 0.0, 3.0, 1.5)
    st, out = run(s, 8)
    assert st == 0
    bg = s["partId"] == 0
    assert np.array_equal(out[bg].view(np.uint32), s["sums"][bg].view(np.uint32))
    assert np.all(out[s["partId"] == 1][:, :3] == 0.0)
    assert np.array_equal(out[..., 3], s["sums"][..., 3])
    brighter = {k: v.copy() for k, v in s.items()}
    left = brighter["partId"] == 1
    brighter["sums"][left, 0] = 5.0
    brighter["lum2"][left] = 40.0
    _, other = run(brighter, 8)
    right = s["partId"] == 2
    assert np.array_equal(other[right].view(np.uint32), out[right].view(np.uint32))
    # Two materials sharing one part id are still told apart.
    two = synthetic(40, 32, 8, 3.0, 3.0, 1.5)
    fg = two["partId"] != 0
    two["partId"][fg] = 1
    cols = np.arange(40)[None, :].repeat(32, 0)
    dark = fg & (cols < 20)
    two["sums"][dark, :3] = 0.0
    two["lum2"][dark] = 0.0
    _, mixed = run(two, 8)
    assert np.all(mixed[(two["partId"] == 1) & (two["materialId"] == 1)][:, :3] == 0.0)


@pytest.mark.parametrize("preset,halo", [(orc.FULL, HALO), (orc.VIEWPORT, VIEWPORT_HALO)])
def test_a_band_of_rows_is_exactly_those_rows_of_the_whole_image(preset, halo):
    for spp in (2, 6):
        s = synthetic(30, 200, spp, 0.7, 1.4, 1.0)
        st, whole = run(s, spp, preset)
        assert st == 0
        for y0, y1 in ((0, 40), (40, 41), (100, 160), (150, 200), (0, 200)):
            lo, hi = max(y0 - halo, 0), min(y1 + halo, 200)
            st, rows = run(band(s, lo, hi), spp, preset, origin=lo, rows=hi - lo, y0=y0, y1=y1, height=200)
            assert st == 0
            assert np.array_equal(rows.view(np.uint32), whole[y0:y1].view(np.uint32)), (y0, y1)
        # Missing halo rows make the call fail instead of filtering with too little context.
        lo = 100 - halo + 1
        st, _ = run(band(s, lo, 200), spp, preset, origin=lo, rows=200 - lo, y0=100, y1=110, height=200)
        assert st == 2
        st, _ = run(band(s, 0, 120), spp, preset, origin=0, rows=120, y0=100, y1=110, height=200)
        assert st == 2


def test_bad_shapes_are_refused():
    s = synthetic(8, 8, 2, 1.0, 1.0, 0.5)
    st, _ = orc.denoise(8, 8, 0, 9, 2, 0, 8, s)
    assert st == 1
    st, _ = orc.denoise(0, 8, 0, 8, 2, 0, 8, s)
    assert st == 1


def test_half_strength_is_the_mean_of_the_two_images_none_is_the_noisy_one_and_all_the_filter():
    s = synthetic(30, 200, 6, 0.7, 1.4, 1.0)
    for preset in (orc.FULL, orc.VIEWPORT):
        _, clean = run(s, 6, preset)
        noisy = s["sums"]
        _, half = run(s, 6, preset, 0.5)
        want = (noisy + clean) * np.float32(0.5)
        want[..., 3] = noisy[..., 3]
        assert np.array_equal(half.view(np.uint32), want.astype(np.float32).view(np.uint32))
        for strength, ref in ((0.0, noisy), (1.0, clean), (7.0, clean), (-1.0, noisy), (float("nan"), clean)):
            _, got = run(s, 6, preset, strength)
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), strength


def test_each_pixels_own_count_weighs_its_variance_and_equal_counts_change_nothing():
    s = synthetic(40, 120, 16, 0.5, 2.0, 1.0)
    _, plain = run(s, 16)
    n = 40 * 120
    _, same = run(s, 16, counts=np.full(n, 16.0, np.float32))
    assert np.array_equal(same.view(np.uint32), plain.view(np.uint32))
    counts = np.where(s["partId"].reshape(-1) == 2, 64.0, 16.0).astype(np.float32)
    _, weighed = run(s, 16, counts=counts)
    right, left = s["partId"] == 2, s["partId"] == 1

    def moved(out):
        return float(np.abs(out[..., 1] - s["sums"][..., 1])[right].astype(np.float64).sum())
    assert moved(weighed) < 0.9 * moved(plain)
    assert np.array_equal(weighed[left].view(np.uint32), plain[left].view(np.uint32))
    lo, hi = 80 - HALO, min(100 + HALO, 120)
    st, rows = run(band(s, lo, hi), 16, counts=counts[lo * 40:hi * 40], origin=lo, rows=hi - lo, y0=80, y1=100,
                   height=120)
    assert st == 0
    assert np.array_equal(rows.view(np.uint32), weighed[80:100].view(np.uint32))


def test_a_clean_image_stays_clean():
    s = synthetic(32, 24, 16, 1.0, 1.0, 0.0)
    for y in range(1, 24):
        for x in range(32):
            v = np.float32(0.1) + np.float32(0.05) * np.float32(x)
            s["sums"][y, x, :3] = v * np.float32(16.0)
            l = np.float32(0.2126) * v + np.float32(0.7152) * v + np.float32(0.0722) * v
            s["lum2"][y, x] = l * l * np.float32(16.0)
    _, out = run(s, 16)
    a, b = out[1:, :, 1] / 16.0, s["sums"][1:, :, 1] / 16.0
    assert np.all(np.abs(a - b) <= 1.0e-3 * b)


def test_a_firefly_is_brought_down_to_its_neighbours():
    # A pixel 50 x brighter than its flat surroundings loses its excess; the same pixel at 2 x is kept
    # (it is no firefly) and smoothed only by the a-trous passes.
    s = synthetic(24, 24, 16, 1.0, 1.0, 0.0)
    for bright, clamped in ((50.0, True), (2.0, False)):
        t = {k: v.copy() for k, v in s.items()}
        t["sums"][12, 6, :3] *= np.float32(bright)
        t["lum2"][12, 6] *= np.float32(bright * bright)
        _, out = run(t, 16)
        peak = out[12, 6, 1] / 16.0
        if clamped:
            assert peak < 5.0 * 1.0 + 0.1
        else:
            assert peak > 1.0
