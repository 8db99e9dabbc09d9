"""The allocation-free tile entry point (rpx_render_tile_noalloc) against rt_render_tile through
the C ABI: the same film bits on two calls in a row, the second with a smaller tile (the shared
scratch region keeps the first call's contents), on scenes that use the textured-BSDF and media
parts of the region too; and the size limit refused with status 2."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def call(lib, fn, handle, x, y, w, h, sample):
    out = np.full((h, w, 3), np.nan, np.float32)
    st = getattr(lib, fn)(handle, x, y, w, h, sample, out.ctypes.data_as(ctypes.c_void_p))
    return st, out


def same_bits(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("scene", ["example", "media", "textured"])
def test_noalloc_tiles_equal_rt_render_tile(gpu, example_rsb, scene):
    lib = gpu.library()
    if scene == "example":
        rsb = example_rsb
    elif scene == "media":
        rsb = gpu.write_scene(5, 0)
    else:
        rsb = gpu.write_scene(4, 0)
    s = gpu.Scene(rsb)
    if scene == "textured":
        # The fit scene's floor (material 8) with a base-colour texture.
        tex = np.random.default_rng(3).uniform(0.1, 0.9, size=(64, 64, 3)).astype(np.float32)
        s.set_texture(8, tex, (-4000.0, -4000.0), (1.0 / 8000.0, 1.0 / 8000.0))
    W, H = s.width, s.height
    # Two calls in a row, the second smaller; then a tile over the image's edge.
    tiles = [((W // 2 - 64) & ~7, (H // 2 - 64) & ~7, 128, 128, 3), (W // 3, H // 4, 17, 9, 5),
             (W - 20, H - 12, 32, 32, 0)]
    for x, y, w, h, k in tiles:
        print("noalloc", scene, "tile", x, y, w, h, "sample", k, flush=True)
        st_a, got = call(lib, "rpx_render_tile_noalloc", s.handle, x, y, w, h, k)
        st_b, want = call(lib, "rt_render_tile", s.handle, x, y, w, h, k)
        assert st_a == 0 and st_b == 0
        assert same_bits(got, want), (scene, x, y, w, h, int(np.sum(got.view(np.uint32) != want.view(np.uint32))))
    if scene == "example":
        # The reference's pinned tile, through the allocation-free entry point.
        st, got = call(lib, "rpx_render_tile_noalloc", s.handle, 240, 176, 32, 32, 3)
        assert st == 0 and gpu.fnv1a(got) == 0xE2515C9A
        want, _ = orc.OracleScene(rsb).render_tile(240, 176, 32, 32, 3)
        assert same_bits(got, want)
    s.close()


def test_noalloc_refuses_tiles_beyond_its_limit(gpu, example_rsb):
    lib = gpu.library()
    s = gpu.Scene(example_rsb)
    st, _ = call(lib, "rpx_render_tile_noalloc", s.handle, 0, 0, 129, 128, 0)
    assert st == 2 and "16384" in gpu.last_error()
    st, _ = call(lib, "rpx_render_tile_noalloc", s.handle, 0, 0, 128, 128, 0)
    assert st == 0
    s.close()
