"""GPU passes and denoiser, through the C ABI, against the oracle and the reference's pins.

Bar: bit-exact. The seven pass planes of the example (rt-core/src/render/passes_tests.rs:367-381,
866) and the denoised example (rt-post/src/denoise/tests.rs:392, 527-528) must reproduce the
reference's hashes; every other case compares raw f32 bits with the oracle on the same inputs."""
import numpy as np
import pytest

from oracle import oracle as orc
from repackshot_amd import denoise as dn
from test_gpu_parity import feature_scene

pytestmark = pytest.mark.gpu

PINNED_PASSES = {"albedo": 0x43571DA9, "normal": 0x0AABFD24, "depth": 0x25F8DD11, "partId": 0xCCD3D6C7,
                 "materialId": 0xCCD3D6C7, "coc": 0xDF87494C, "alpha": 0xD4EAAC78}
PINNED_DENOISED = {(4, 0): 0x2A430F92, (2, 1): 0x7000CEB2, (4, 1): 0x5318D4C5}


def same_bits(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_example_passes_are_bit_exact_and_pinned(gpu, example_rsb):
    s = gpu.Scene(example_rsb)
    s.render_passes(0, 1)
    s.render_passes(1, 3)  # sums add across calls in sample order
    got = s.passes()
    want = orc.OracleScene(example_rsb).render_passes(0, 4)
    for k, pin in PINNED_PASSES.items():
        assert same_bits(got[k], want[k]), k
        assert orc.fnv1a(got[k]) == pin, k
    s.reset_passes()
    s.render_passes(0, 2)
    again = s.passes()
    assert same_bits(again["albedo"], orc.OracleScene(example_rsb).render_passes(0, 2)["albedo"])
    assert same_bits(again["depth"], got["depth"])


def test_tile_passes_albedo_and_alpha(gpu, example_rsb):
    s = gpu.Scene(example_rsb)
    o = orc.OracleScene(example_rsb)
    whole = o.render_passes(3, 1)  # one sample: the resolved albedo is the sample's albedo
    for x, y, w, h in ((240, 176, 32, 32), (0, 0, 17, 9), (500, 370, 32, 32)):
        t = s.render_tile_passes(x, y, w, h)
        cw, ch = max(min(w, o.width - x), 0), max(min(h, o.height - y), 0)
        for k in ("normal", "depth", "partId", "materialId", "coc"):
            assert same_bits(t[k][:ch, :cw], whole[k][y:y + ch, x:x + cw]), (x, y, k)
        # Outside the image: the background.
        assert np.all(np.isinf(t["depth"][ch:])) and np.all(t["partId"][:, cw:] == 0)
        a = s.render_tile_albedo(x, y, w, h, 3)
        assert same_bits(a[:ch, :cw], whole["albedo"][y:y + ch, x:x + cw])
        assert np.all(a[ch:] == 0) and np.all(a[:, cw:] == 0)
        al = s.render_tile_alpha(x, y, w, h, 3)
        assert same_bits(al[:ch, :cw], whole["alpha"][y:y + ch, x:x + cw])
    only = s.render_tile_passes(240, 176, 8, 8, gpu.PASS_DEPTH | gpu.PASS_COC | gpu.PASS_ALBEDO)
    assert sorted(only) == ["coc", "depth"]
    lib = gpu.library()
    buf = np.zeros(64, np.uint32)
    assert lib.rt_render_tile_passes(s.handle, 0, 0, 0, 4, gpu.PASS_DEPTH, buf.ctypes.data) == 2
    assert lib.rt_render_tile_passes(s.handle, 0, 0, 4, 4, gpu.PASS_DEPTH, None) == 3
    assert lib.rt_render_tile_albedo(98765, 0, 0, 4, 4, 0, buf.ctypes.data) == 1


@pytest.mark.parametrize("variant", [dict(), dict(background_mode=1, floor=1, disk=True),
                                     dict(background_mode=2, fstop=4.0), dict(ortho=True)])
def test_feature_scene_passes_are_bit_exact(gpu, variant):
    rsb = feature_scene(**variant)
    s = gpu.Scene(rsb)
    s.render_passes(0, 3)
    got = s.passes()
    want = orc.OracleScene(rsb).render_passes(0, 3)
    for k in want:
        assert same_bits(got[k], want[k]), (variant, k)
    if variant.get("fstop"):
        assert float(got["coc"].max()) > 0
    if variant.get("floor"):
        assert (got["partId"] == gpu.FLOOR_ID).any()


@pytest.mark.parametrize("variant,first,count,more", [(dict(), 0, 16, 20), (dict(fstop=4.0), 5, 70, 64),
                                                       (dict(background_mode=1, floor=1, disk=True), 3, 130, 16)])
def test_many_sample_passes_are_bit_exact(gpu, variant, first, count, more):
    # From 16 samples a pixel's samples go one per lane of a wave, added in sample order after the
    # wave (in rounds of 64: 70 and 130 are ragged); further calls add onto the sums.
    rsb = feature_scene(**variant)
    s = gpu.Scene(rsb)
    o = orc.OracleScene(rsb)
    s.render_passes(first, count)
    got, want = s.passes(), o.render_passes(first, count)
    for k in want:
        assert same_bits(got[k], want[k]), (variant, k)
    s.render_passes(first + count, more)
    got, want = s.passes(), o.render_passes(first, count + more)
    for k in ("albedo", "alpha"):
        assert same_bits(got[k], want[k]), (variant, k)


def oracle_planes(o, spp):
    film, _ = o.render_film(0, spp)
    rgb, _, alpha, lum2 = film
    ps = o.render_passes(0, spp)
    return {"sums": np.concatenate([rgb, alpha[..., None]], axis=2), "lum2": lum2, "albedo": ps["albedo"],
            "normal": ps["normal"], "depth": ps["depth"], "partId": ps["partId"], "materialId": ps["materialId"]}


@pytest.mark.parametrize("spp,preset", [(4, 0), (2, 1), (4, 1)])
def test_resident_denoise_of_the_example_is_pinned(gpu, example_rsb, spp, preset):
    s = gpu.Scene(example_rsb)
    s.render_samples(0, spp)
    s.render_passes(0, spp)
    out = s.denoise(preset)
    o = orc.OracleScene(example_rsb)
    W, H = o.width, o.height
    st, want = orc.denoise(W, H, 0, H, spp, 0, H, oracle_planes(o, spp), orc.FULL if preset == 0 else orc.VIEWPORT)
    assert st == 0
    assert same_bits(out, want)
    assert orc.fnv1a(dn.means(out, spp)) == PINNED_DENOISED[(spp, preset)]
    t = s.denoise_timing()
    assert t["total_ms"] > 0 and t["levels"] == (5 if preset == 0 else 3)


def band_input(planes, spp, counts=None):
    h, w = planes["depth"].shape
    inp = dn.DenoiseInput(w, h, 0, h, spp, *(np.ascontiguousarray(planes[k]).reshape(-1) for k in
                                               ("sums", "lum2", "albedo", "normal", "depth", "partId",
                                                "materialId")))
    return inp if counts is None else inp.with_counts(counts)


@pytest.mark.parametrize("spp", [1, 2, 3, 6, 16])
def test_c_abi_bands_presets_and_strengths_are_the_oracles(gpu, spp):
    planes = orc.denoise_synthetic(30, 200, spp, 0.7, 1.4, 1.0)
    inp = band_input(planes, spp)
    for preset, op in ((dn.FULL, orc.FULL), (dn.VIEWPORT, orc.VIEWPORT)):
        for strength in (1.0, 0.5, 0.3, 0.0, float("nan")):
            got = dn.denoise_rows_at(inp, 0, 200, preset, strength)
            st, want = orc.denoise(30, 200, 0, 200, spp, 0, 200, planes, op, strength)
            assert st == 0 and same_bits(got, want.reshape(-1)), (preset, strength)
        whole = dn.denoise_with(inp, preset)
        for y0, y1 in ((0, 40), (40, 41), (100, 160), (150, 200)):
            rows = dn.denoise_rows_with(inp.with_halo_of(y0, y1, preset), y0, y1, preset)
            assert same_bits(rows, whole[y0 * 30 * 4:y1 * 30 * 4]), (preset, y0, y1)
        lo = 100 - preset.halo() + 1
        short = band_input({k: v[lo:] for k, v in planes.items()}, spp)
        short = dn.DenoiseInput(**{**short.__dict__, "origin_y": lo, "height": 200})
        with pytest.raises(dn.DenoiseError, match="halo"):
            dn.denoise_rows_with(short, 100, 110, preset)
    assert dn.halo(0) == 72 and dn.halo(1) == 24 and dn.halo(2) == 0


def test_c_abi_counts_and_refusals(gpu):
    planes = orc.denoise_synthetic(40, 120, 16, 0.5, 2.0, 1.0)
    counts = np.where(planes["partId"].reshape(-1) == 2, 64.0, 16.0).astype(np.float32)
    got = dn.denoise(band_input(planes, 16, counts))
    st, want = orc.denoise(40, 120, 0, 120, 16, 0, 120, planes, orc.FULL, 1.0, counts)
    assert st == 0 and same_bits(got, want.reshape(-1))
    lib = gpu.library()
    data = band_input(planes, 16).planes()
    out = np.zeros(40 * 120 * 4, np.float32)
    assert lib.rt_denoise(40, 120, 0, 120, 16, 0, 120, data.ctypes.data, out.ctypes.data) == 0
    assert same_bits(out, dn.denoise(band_input(planes, 16)))
    assert lib.rt_denoise_with(7, 1.0, 40, 120, 0, 120, 16, 0, 120, data.ctypes.data, out.ctypes.data) == 5
    assert lib.rt_denoise(40, 119, 0, 120, 16, 0, 120, data.ctypes.data, out.ctypes.data) == 5
    assert lib.rt_denoise(40, 120, 0, 120, 16, 0, 120, None, out.ctypes.data) == 3
    with pytest.raises(dn.DenoiseError, match="lum2"):
        bad = band_input(planes, 16)
        dn.denoise(dn.DenoiseInput(**{**bad.__dict__, "lum2": bad.lum2[:10]}))


def test_large_render_denoise_equals_the_oracle_on_the_same_film(gpu):
    # The bench's 1024 x 1024 scene at 8 samples: the GPU film and passes fed to the oracle's filter
    # give the GPU denoiser's bits (the oracle does not render this image; it filters it).
    s = gpu.Scene(gpu.write_scene(2, 0))
    s.render_samples(0, 8)
    s.render_passes(0, 8)
    out = s.denoise(0)
    rgb, _, alpha, lum2 = s.film()
    ps = s.passes()
    planes = {"sums": np.concatenate([rgb, alpha[..., None]], axis=2), "lum2": lum2, "albedo": ps["albedo"],
              "normal": ps["normal"], "depth": ps["depth"], "partId": ps["partId"], "materialId": ps["materialId"]}
    st, want = orc.denoise(s.width, s.height, 0, s.height, 8, 0, s.height, planes, orc.FULL)
    assert st == 0
    assert same_bits(out, want)
    assert (ps["partId"] != 0).mean() > 0.2
