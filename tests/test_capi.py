"""The C ABI library without a GPU: it loads, exports every entry point of include/rpx.h, refuses
work it cannot do with the reference's status codes, and its scene generators write RSB bytes that
the oracle reads."""
import os
import re

import numpy as np
import pytest

from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "rpx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:uint32_t|uint64_t|void|uint8_t\s*\*)\s*(r[tp]x?_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol(rpx):
    lib = rpx.library()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(rpx.exported_symbols()) == names


def test_alloc_gives_zeroed_writable_buffers(rpx):
    # ffi.rs:27-48: zeroed bytes, never null (also for a zero length), released with their length.
    import ctypes
    lib = rpx.library()
    for n in (0, 1, 4096):
        p = lib.rt_alloc(n)
        assert p
        buf = (ctypes.c_uint8 * max(n, 1)).from_address(p)
        assert n == 0 or bytes(buf) == bytes(n)
        buf[0] = 7
        lib.rt_free(p, n)
    lib.rt_free(None, 8)


def test_status_codes_without_a_scene(rpx):
    lib = rpx.library()
    out = np.zeros(12, np.float32)
    assert lib.rt_render_tile(12345, 0, 0, 2, 2, 0, out.ctypes.data) == 1       # bad handle
    assert lib.rt_render_tile(12345, 0, 0, 2, 2, 0, None) == 3                  # null pointer
    assert lib.rt_scene_free(12345) == 1
    assert lib.rt_scene_load(None, 10) == 0
    assert "null pointer" in rpx.last_error()
    b = np.zeros(3, np.float32)
    assert lib.rpx_set_trip_budget(12345, 48) == 1                              # bad handle
    assert lib.rpx_trip_budget(12345, None) == 3                                # null pointer
    assert lib.rpx_tune_trip_budget(12345, 8, b.ctypes.data) == 1
    assert lib.rpx_queue_trip_budget(12345, b.ctypes.data) == 1


def test_load_without_a_gpu_fails_loudly(rpx, example_rsb):
    if rpx.device_count() != 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(rpx.RsbError, match="no HIP device"):
        rpx.Scene(example_rsb)


def test_generated_plain_scene_is_what_it_says(rpx):
    s = orc.OracleScene(rpx.write_scene(1, 0))
    g = s.geometry_info()
    assert (s.width, s.height, s.max_samples, s.depth) == (256, 256, 16, 8)
    assert g["triangles"] == 5120 + 2 and g["instances"] == 2


def test_scene_generators_are_deterministic(rpx):
    assert rpx.write_scene(1, 3) == rpx.write_scene(1, 3)
    assert rpx.write_scene(0) == rpx.write_scene(0)
