"""ctypes binding of librpx.so, the GPU tracer's C ABI (include/rpx.h), and a Python mirror of the
reference renderer's host-facing render API.

The reference's hosts drive its engine through the wasm C ABI (engine/crates/rt-wasm/src/ffi.rs:
rt_scene_load / rt_render_tile / rt_render_tile_rgba / rt_scene_update / rt_scene_free, with the
last load error as text) and accumulate each pixel's samples in f32 in sample order
(src/render/tiles.ts:154-219, rt-cli/src/lib.rs:266-366). `Scene` keeps those names, argument
meanings, status codes and error behaviour; `render_samples` / `film` are the GPU-native progressive
loop with the film resident in HBM.

There is no CPU fallback: without the built library, or without a GPU, the calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librpx.so")

STATUS = {
    0: "ok",
    1: "bad handle",
    2: "empty or oversized tile",
    3: "null pointer",
    4: "bad scene",
    5: "bad denoise input",
    16: "no GPU or GPU memory exhausted",
    17: "GPU kernel launch or execution failure",
}

# Every exported symbol of include/rpx.h, with its ctypes signature.
_SIGS = {
    "rt_version": (ctypes.c_uint32, []),
    "rt_features": (ctypes.c_uint32, []),
    "rt_alloc": (ctypes.c_void_p, [ctypes.c_uint32]),
    "rt_free": (None, [ctypes.c_void_p, ctypes.c_uint32]),
    "rt_scene_load": (ctypes.c_uint32, [ctypes.c_void_p, ctypes.c_uint32]),
    "rt_scene_update": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32]),
    "rt_scene_error_len": (ctypes.c_uint32, []),
    "rt_scene_error_copy": (None, [ctypes.c_void_p]),
    "rt_scene_free": (ctypes.c_uint32, [ctypes.c_uint32]),
    "rt_scene_caustics": (ctypes.c_uint32, [ctypes.c_uint32]),
    "rt_render_tile": (ctypes.c_uint32, [ctypes.c_uint32] * 6 + [ctypes.c_void_p]),
    "rt_render_tile_rgba": (ctypes.c_uint32, [ctypes.c_uint32] * 6 + [ctypes.c_void_p]),
    "rt_render_block_rgba": (ctypes.c_uint32, [ctypes.c_uint32] * 7 + [ctypes.c_void_p]),
    "rt_adaptive_decide": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                                             ctypes.c_uint32] + [ctypes.c_void_p] * 5),
    "rpx_render_samples": (ctypes.c_uint32, [ctypes.c_uint32] * 3),
    "rpx_film_reset": (ctypes.c_uint32, [ctypes.c_uint32]),
    "rpx_film_read": (ctypes.c_uint32, [ctypes.c_uint32] + [ctypes.c_void_p] * 4),
    "rpx_film_device": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_void_p]),
    "rpx_trace_rays": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32] + [ctypes.c_void_p] * 4),
    "rpx_trace_rays_wave": (ctypes.c_uint32,
                            [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32] + [ctypes.c_void_p] * 4),
    "rpx_scene_info": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_void_p]),
    "rpx_scene_stats": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_void_p]),
    "rpx_set_options": (ctypes.c_uint32, [ctypes.c_uint32] * 8),
    "rpx_autotune": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]),
    "rpx_set_trip_budget": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32]),
    "rpx_trip_budget": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_void_p]),
    "rpx_tune_trip_budget": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]),
    "rpx_queue_trip_budget": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_void_p]),
    "rpx_set_profiling": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32]),
    "rpx_last_timing": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_void_p]),
    "rpx_energy_tables": (ctypes.c_uint32, [ctypes.c_void_p] * 6),
    "rpx_write_scene": (ctypes.c_uint64, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64]),
    "rpx_write_queue_job": (ctypes.c_uint64, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64]),
    "rpx_material_params": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32]),
    "rpx_set_material_params": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32]),
    "rpx_set_texture": (ctypes.c_uint32, [ctypes.c_uint32] * 4 + [ctypes.c_void_p] + [ctypes.c_float] * 4),
    "rpx_adjoint": (ctypes.c_uint32, [ctypes.c_uint32] * 3 + [ctypes.c_void_p] * 3),
    "rpx_queue_create": (ctypes.c_uint32, []),
    "rpx_queue_add": (ctypes.c_uint32, [ctypes.c_uint32] * 4 + [ctypes.c_void_p]),
    "rpx_queue_run": (ctypes.c_uint32, [ctypes.c_uint32]),
    "rpx_queue_reset": (ctypes.c_uint32, [ctypes.c_uint32]),
    "rpx_queue_film_read": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32] + [ctypes.c_void_p] * 4),
    "rpx_queue_set_options": (ctypes.c_uint32, [ctypes.c_uint32] * 8),
    "rpx_queue_autotune": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]),
    "rpx_queue_set_profiling": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32]),
    "rpx_queue_last_timing": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_void_p]),
    "rpx_queue_stats": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_void_p]),
    "rpx_queue_free": (ctypes.c_uint32, [ctypes.c_uint32]),
    "rpx_render_adaptive": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]),
    "rpx_adaptive_counts": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32]),
    "rpx_adaptive_work": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_void_p]),
    "rt_render_tile_passes": (ctypes.c_uint32, [ctypes.c_uint32] * 6 + [ctypes.c_void_p]),
    "rt_render_tile_albedo": (ctypes.c_uint32, [ctypes.c_uint32] * 6 + [ctypes.c_void_p]),
    "rt_render_tile_alpha": (ctypes.c_uint32, [ctypes.c_uint32] * 6 + [ctypes.c_void_p]),
    "rt_denoise": (ctypes.c_uint32, [ctypes.c_uint32] * 7 + [ctypes.c_void_p] * 2),
    "rt_denoise_with": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_float] + [ctypes.c_uint32] * 7
                        + [ctypes.c_void_p] * 2),
    "rt_denoise_halo": (ctypes.c_uint32, [ctypes.c_uint32]),
    "rpx_render_passes": (ctypes.c_uint32, [ctypes.c_uint32] * 3),
    "rpx_passes_reset": (ctypes.c_uint32, [ctypes.c_uint32]),
    "rpx_passes_read": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]),
    "rpx_denoise_film": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint32,
                                           ctypes.c_void_p]),
    "rpx_denoise_timing": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_void_p]),
    "rpx_denoise_counted": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_float] + [ctypes.c_uint32] * 7
                            + [ctypes.c_void_p] * 2),
    "rpx_last_kernel_ms": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_void_p]),
    "rpx_render_tile_noalloc": (ctypes.c_uint32, [ctypes.c_uint32] * 6 + [ctypes.c_void_p]),
    "rpx_device_count": (ctypes.c_uint32, []),
}

KEEP = 0xFFFFFFFF

# The passes (rt-core/src/passes.rs:44-87): bit i of a mask is pass PASS_KEYS[i].
PASS_KEYS = ["albedo", "normal", "depth", "partId", "materialId", "coc", "alpha"]
PASS_CHANNELS = [3, 3, 1, 1, 1, 1, 1]
PASS_ALBEDO, PASS_NORMAL, PASS_DEPTH, PASS_PART_ID, PASS_MATERIAL_ID, PASS_COC, PASS_ALPHA = (1 << i for i in range(7))
PASS_ALL = (1 << 7) - 1
PASS_SAMPLED = PASS_ALBEDO | PASS_ALPHA
PASS_EXACT = PASS_ALL & ~PASS_SAMPLED
FLOOR_ID = 0xFFFFFFFF


def _planes_of(buf, mask, w, h):
    """Splits little-endian words of the planes of `mask` (bit order) into (h, w[, c]) arrays."""
    out, at = {}, 0
    for i, (k, c) in enumerate(zip(PASS_KEYS, PASS_CHANNELS)):
        if not mask & (1 << i):
            continue
        words = buf[at:at + w * h * c]
        at += w * h * c
        dt = np.uint32 if k in ("partId", "materialId") else np.float32
        a = words.view(dt).copy()
        out[k] = a.reshape((h, w, c) if c > 1 else (h, w))
    return out
FLAG_PER_PASS = 1  # one sample pass at a time (the pool drains after each pass)
FLAG_GRAPH = 2     # groups of iterations launched as captured HIP graphs
FLAG_NO_DRAIN = 4  # the iteration loop runs to the end (no drain launch for the last paths)
FLAG_ROUND_ONLY = 8  # adaptive rounds render only their own samples (nothing rendered ahead)

_lib = None


class RsbError(RuntimeError):
    """The scene bytes could not be loaded (the reference's RsbError text)."""


class RenderError(RuntimeError):
    """A render call returned a non-zero status."""


def library():
    """The loaded librpx.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "librpx.so is not built (run repackshot_amd/build.py or __graft_entry__.build())")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def exported_symbols():
    return list(_SIGS)


def device_count():
    return int(library().rpx_device_count())


def last_error():
    lib = library()
    n = lib.rt_scene_error_len()
    buf = ctypes.create_string_buffer(max(int(n), 1))
    lib.rt_scene_error_copy(buf)
    return buf.raw[:n].decode("utf-8", "replace")


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _check(status, what):
    if status != 0:
        raise RenderError("%s: status %d (%s) %s" % (what, status, STATUS.get(status, "?"), last_error()))


def write_scene(kind, seed=0):
    """RSB v1 bytes of one of this library's scenes (0 example, 1 icosphere on a floor, 2 eight
    displaced level-7 icospheres)."""
    lib = library()
    n = lib.rpx_write_scene(kind, seed, None, 0)
    if n == 0:
        raise ValueError("unknown scene kind %d" % kind)
    buf = np.zeros(n, np.uint8)
    got = lib.rpx_write_scene(kind, seed, _ptr(buf), n)
    if got != n:
        raise RuntimeError("scene generator returned %d bytes, expected %d" % (got, n))
    return buf.tobytes()


def write_queue_job(seed, job):
    """RSB v1 bytes of job `job` of the seeded mixed render queue."""
    lib = library()
    n = lib.rpx_write_queue_job(seed, job, None, 0)
    buf = np.zeros(n, np.uint8)
    got = lib.rpx_write_queue_job(seed, job, _ptr(buf), n)
    if got != n:
        raise RuntimeError("queue generator returned %d bytes, expected %d" % (got, n))
    return buf.tobytes()


def energy_tables():
    """The energy-compensation tables generated on the current GPU."""
    lib = library()
    a = np.zeros(1024, np.uint16)
    s = np.zeros(1024, np.uint16)
    ci = np.zeros(17408, np.uint16)
    co = np.zeros(17408, np.uint16)
    aa = np.zeros(32, np.float32)
    sa = np.zeros(32, np.float32)
    _check(lib.rpx_energy_tables(_ptr(a), _ptr(s), _ptr(ci), _ptr(co), _ptr(aa), _ptr(sa)), "energy tables")
    return {"albedo": a, "schlick": s, "clear_in": ci, "clear_out": co, "albedo_avg": aa, "schlick_avg": sa}


class Scene:
    """A scene loaded on the GPU from RSB v1 bytes (rt_scene_load)."""

    def __init__(self, rsb: bytes):
        lib = library()
        data = np.frombuffer(rsb, np.uint8)
        h = lib.rt_scene_load(_ptr(data), len(rsb))
        if h == 0:
            raise RsbError(last_error())
        self.handle = int(h)
        info = self.info()
        self.width, self.height = info["width"], info["height"]

    def close(self):
        if getattr(self, "handle", 0):
            library().rt_scene_free(self.handle)
            self.handle = 0

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update(self, rsb: bytes):
        """rt_scene_update: replaces the scene, keeping the meshes when the bytes carry none."""
        data = np.frombuffer(rsb, np.uint8)
        st = library().rt_scene_update(self.handle, _ptr(data), len(rsb))
        if st == 4:
            raise RsbError(last_error())
        _check(st, "rt_scene_update")
        info = self.info()
        self.width, self.height = info["width"], info["height"]

    def info(self):
        out = np.zeros(8, np.uint64)
        _check(library().rpx_scene_info(self.handle, _ptr(out)), "rpx_scene_info")
        keys = ["width", "height", "max_samples", "seed", "max_depth", "triangles", "nodes", "instances"]
        return {k: int(v) for k, v in zip(keys, out)}

    def caustics(self):
        """Caustic photons per sample (0: the scene has no caustics)."""
        return int(library().rt_scene_caustics(self.handle))

    def render_tile(self, x, y, w, h, sample_index):
        """One sample of every pixel of the tile: (h, w, 3) float32 linear RGB."""
        out = np.zeros((h, w, 3), np.float32)
        _check(library().rt_render_tile(self.handle, x, y, w, h, sample_index, _ptr(out)), "rt_render_tile")
        return out

    def kernel_times(self):
        """Per-kernel GPU ms of the last profiled render (rpx_last_kernel_ms)."""
        out = np.zeros(10, np.float32)
        _check(library().rpx_last_kernel_ms(self.handle, _ptr(out)), "rpx_last_kernel_ms")
        keys = ["extend_kernel_ms", "extend_launches", "shadow_kernel_ms", "shadow_launches", "phase_ms", "walk_ms",
                "drain_ms", "shade_ms", "photon_map_ms", "photon_maps_traced"]
        return {k: (int(v) if k.endswith("launches") or k.endswith("traced") else float(v)) for k, v in zip(keys, out)}

    def render_tile_noalloc(self, x, y, w, h, sample_index):
        """render_tile through the allocation-free entry point (tiles of at most 16384 pixels)."""
        out = np.zeros((h, w, 3), np.float32)
        _check(library().rpx_render_tile_noalloc(self.handle, x, y, w, h, sample_index, _ptr(out)),
               "rpx_render_tile_noalloc")
        return out

    def render_tile_rgba(self, x, y, w, h, sample_index):
        out = np.zeros((h, w, 4), np.float32)
        _check(library().rt_render_tile_rgba(self.handle, x, y, w, h, sample_index, _ptr(out)), "rt_render_tile_rgba")
        return out

    def render_block_rgba(self, x, y, w, h, sample_index, iteration=0):
        """rt_render_block_rgba: one sample of a block of an adaptive round, as render_tile_rgba."""
        out = np.zeros((h, w, 4), np.float32)
        _check(library().rt_render_block_rgba(self.handle, x, y, w, h, sample_index, iteration, _ptr(out)),
               "rt_render_block_rgba")
        return out

    def render_tile_passes(self, x, y, w, h, mask=PASS_EXACT):
        """rt_render_tile_passes: the exact passes of `mask` (normal, depth, partId, materialId, coc)
        of the tile, from one ray through each pixel's centre; a dict of planes."""
        mask &= PASS_EXACT
        words = sum(c for i, c in enumerate(PASS_CHANNELS) if mask & (1 << i)) * w * h
        buf = np.zeros(max(words, 1), np.uint32)
        _check(library().rt_render_tile_passes(self.handle, x, y, w, h, mask, _ptr(buf)), "rt_render_tile_passes")
        return _planes_of(buf, mask, w, h)

    def render_tile_albedo(self, x, y, w, h, sample_index):
        """rt_render_tile_albedo: (h, w, 3) base colour of what sample `sample_index`'s camera ray meets."""
        out = np.zeros((h, w, 3), np.float32)
        _check(library().rt_render_tile_albedo(self.handle, x, y, w, h, sample_index, _ptr(out)),
               "rt_render_tile_albedo")
        return out

    def render_tile_alpha(self, x, y, w, h, sample_index):
        """rt_render_tile_alpha: (h, w) coverage (1 where the sample's camera ray meets an instance)."""
        out = np.zeros((h, w), np.float32)
        _check(library().rt_render_tile_alpha(self.handle, x, y, w, h, sample_index, _ptr(out)),
               "rt_render_tile_alpha")
        return out

    def render_passes(self, first, count):
        """The exact passes of every pixel into the resident pass film, and the albedo and coverage of
        samples [first, first + count) added to its sums in sample order."""
        _check(library().rpx_render_passes(self.handle, first, count), "rpx_render_passes")

    def reset_passes(self):
        _check(library().rpx_passes_reset(self.handle), "rpx_passes_reset")

    def passes(self, mask=PASS_ALL):
        """The resolved planes of the resident pass film: a dict keyed as PASS_KEYS."""
        W, H = self.width, self.height
        words = sum(c for i, c in enumerate(PASS_CHANNELS) if mask & (1 << i)) * W * H
        buf = np.zeros(max(words, 1), np.uint32)
        _check(library().rpx_passes_read(self.handle, mask, _ptr(buf)), "rpx_passes_read")
        return _planes_of(buf, mask, W, H)

    def denoise(self, preset=0, strength=1.0, samples=0):
        """The resident film denoised with the resident passes (preset 0 full, 1 viewport): (H, W, 4)
        RGBA sums of `samples` samples (0: those rendered since the last film reset)."""
        out = np.zeros((self.height, self.width, 4), np.float32)
        _check(library().rpx_denoise_film(self.handle, preset, strength, samples, _ptr(out)), "rpx_denoise_film")
        return out

    def denoise_on_device(self, preset=0, strength=1.0, samples=0):
        """rpx_denoise_film with the result left in device memory (for timing)."""
        _check(library().rpx_denoise_film(self.handle, preset, strength, samples, None), "rpx_denoise_film")

    def denoise_timing(self):
        out = np.zeros(8, np.float32)
        _check(library().rpx_denoise_timing(self.handle, _ptr(out)), "rpx_denoise_timing")
        keys = ["total_ms", "setup_ms", "firefly_ms", "variance_ms", "atrous_ms", "final_ms", "levels"]
        return {k: float(v) for k, v in zip(keys, out)}

    def render_adaptive(self, cap=0):
        """Adaptive sampling in rounds into the resident film (cleared first): 8 x 8 blocks sampled
        until the scene's noise target or `cap` samples (0: the scene's max_samples). Returns the
        decisions, rounds, path samples, fewest / most samples of a block, the last decision's
        largest error and remaining work, and the GPU time (ms)."""
        out = np.zeros(8, np.uint64)
        _check(library().rpx_render_adaptive(self.handle, cap, _ptr(out)), "rpx_render_adaptive")
        bits = out[5:7].astype(np.uint32).view(np.float32)
        return {"decisions": int(out[0]), "rounds": int(out[1]), "paths": int(out[2]), "min_samples": int(out[3]),
                "max_samples": int(out[4]), "max_error": float(bits[0]), "remaining": float(bits[1]),
                "gpu_ms": int(out[7]) / 1.0e6}

    def adaptive_work(self):
        """The last adaptive render's work: path samples rendered, of them rendered ahead of their
        rounds, runs that rendered ahead, and samples rendered ahead that no round used."""
        out = np.zeros(4, np.uint64)
        _check(library().rpx_adaptive_work(self.handle, _ptr(out)), "rpx_adaptive_work")
        return {"rendered": int(out[0]), "rendered_ahead": int(out[1]), "runs_ahead": int(out[2]),
                "unused": int(out[3])}

    def adaptive_counts(self):
        """Samples of every 8 x 8 block of the last adaptive render, (rows, cols) uint32."""
        cols, rows = (self.width + 7) // 8, (self.height + 7) // 8
        out = np.zeros((rows, cols), np.uint32)
        _check(library().rpx_adaptive_counts(self.handle, _ptr(out), out.size), "rpx_adaptive_counts")
        return out

    def render_samples(self, first, count):
        """Accumulates samples [first, first + count) of every pixel into the resident film."""
        _check(library().rpx_render_samples(self.handle, first, count), "rpx_render_samples")

    def reset_film(self):
        _check(library().rpx_film_reset(self.handle), "rpx_film_reset")

    def film(self):
        """The film's sums: rgb (H, W, 3), weight, alpha, lum2 (H, W), float32."""
        H, W = self.height, self.width
        rgb = np.zeros((H, W, 3), np.float32)
        weight = np.zeros((H, W), np.float32)
        alpha = np.zeros((H, W), np.float32)
        lum2 = np.zeros((H, W), np.float32)
        _check(library().rpx_film_read(self.handle, _ptr(rgb), _ptr(weight), _ptr(alpha), _ptr(lum2)), "rpx_film_read")
        return rgb, weight, alpha, lum2

    def film_device_pointers(self):
        ptrs = (ctypes.c_void_p * 4)()
        _check(library().rpx_film_device(self.handle, ptrs), "rpx_film_device")
        return [int(p or 0) for p in ptrs]

    def resolve(self):
        """Film::resolve (image.rs:158-169): sums times the f32 reciprocal of the count."""
        rgb, weight, _, _ = self.film()
        return resolve(rgb, weight)

    def trace_rays(self, rays, wave_stack=0):
        """Nearest and any hit of rays (n, 8) float32: origin, direction, t_min, t_max. With
        `wave_stack` (1..1024), each ray is traced by a whole wave over a shared stack of that many
        entries (rpx_trace_rays_wave)."""
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        inst = np.zeros(n, np.uint32)
        tri = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        anyhit = np.zeros(n, np.uint8)
        if wave_stack:
            _check(library().rpx_trace_rays_wave(self.handle, _ptr(rays), n, wave_stack, _ptr(inst), _ptr(tri), _ptr(t),
                                                 _ptr(anyhit)), "rpx_trace_rays_wave")
        else:
            _check(library().rpx_trace_rays(self.handle, _ptr(rays), n, _ptr(inst), _ptr(tri), _ptr(t), _ptr(anyhit)),
                   "rpx_trace_rays")
        return inst, tri, t, anyhit

    def stats(self):
        out = np.zeros(4, np.uint64)
        _check(library().rpx_scene_stats(self.handle, _ptr(out)), "rpx_scene_stats")
        return {"rays": int(out[0]), "nearest": int(out[1]), "shadow": int(out[2]), "samples": int(out[3])}

    def set_options(self, pool=0, shape=KEEP, shadow_shape=KEEP, group=0, chunk=0, per_pass=False, graph=False,
                    drain=True, tail=0, adaptive_ahead=True):
        """Path-pool options (0 / KEEP keep the current value; the flags per_pass, graph, drain and
        adaptive_ahead are set by every call); the film does not depend on them. adaptive_ahead:
        the adaptive rounds render a candidate block's samples up to the cap in one run and keep
        them (False: each round renders only its own samples)."""
        flags = _flags(per_pass, graph, drain, adaptive_ahead)
        _check(library().rpx_set_options(self.handle, pool, shape, shadow_shape, group, chunk, flags, tail),
               "rpx_set_options")

    def material_params(self):
        """The differentiable parameters: (materials, 5) float32 — roughness, base r, g, b, IOR."""
        lib = library()
        n = lib.rpx_material_params(self.handle, None, 0)
        out = np.zeros(5 * n, np.float32)
        lib.rpx_material_params(self.handle, _ptr(out), out.size)
        return out.reshape(n, 5)

    def set_material_params(self, params):
        params = np.ascontiguousarray(params, np.float32).reshape(-1, 5)
        _check(library().rpx_set_material_params(self.handle, _ptr(params), params.shape[0]),
               "rpx_set_material_params")

    def set_texture(self, material, rgb, origin=(0.0, 0.0), scale=(1.0, 1.0)):
        """Gives `material` an RGB f32 base-colour texture (h, w, 3), planar in world x, y; None removes it."""
        if material is None:
            _check(library().rpx_set_texture(self.handle, KEEP, 0, 0, None, 0.0, 0.0, 1.0, 1.0), "rpx_set_texture")
            self.texture_shape = None
            return
        rgb = np.ascontiguousarray(rgb, np.float32)
        h, w = rgb.shape[:2]
        _check(library().rpx_set_texture(self.handle, material, w, h, _ptr(rgb), origin[0], origin[1], scale[0],
                                         scale[1]), "rpx_set_texture")
        self.texture_shape = (h, w)

    def adjoint(self, first, count, d_image):
        """d loss / d parameter ((materials, 5)) and d loss / d texel ((h, w, 3) or None) of a loss of
        the mean image of samples [first, first + count), given d loss / d image (H, W, 3)."""
        d_image = np.ascontiguousarray(d_image, np.float32)
        if d_image.shape != (self.height, self.width, 3):
            raise ValueError("d_image must be (height, width, 3)")
        n = library().rpx_material_params(self.handle, None, 0)
        gm = np.zeros((n, 5), np.float32)
        shape = getattr(self, "texture_shape", None)
        gt = np.zeros(shape + (3,), np.float32) if shape else None
        _check(library().rpx_adjoint(self.handle, first, count, _ptr(d_image), _ptr(gm), _ptr(gt)), "rpx_adjoint")
        return gm, gt

    def autotune(self, samples=32):
        """Runs the launch tuner; returns the chosen pool, shapes, trip budget and every shape
        candidate's ms."""
        out = np.zeros(60, np.float32)
        _check(library().rpx_autotune(self.handle, samples, _ptr(out)), "rpx_autotune")
        t = _tune(out)
        t["trip_budget"] = self.trip_budget()
        return t

    def set_trip_budget(self, trips):
        """Closest-hit loop trips before a walk is set aside for a second launch (0: one launch);
        the film does not depend on it."""
        _check(library().rpx_set_trip_budget(self.handle, trips), "rpx_set_trip_budget")

    def trip_budget(self):
        out = ctypes.c_uint32()
        _check(library().rpx_trip_budget(self.handle, ctypes.byref(out)), "rpx_trip_budget")
        return int(out.value)

    def tune_trip_budget(self, samples=32):
        """Times samples [0, samples) at the current launch options with each trip budget the
        tuner knows, keeps the faster; returns it and each candidate's ms."""
        out = np.zeros(3, np.float32)
        _check(library().rpx_tune_trip_budget(self.handle, samples, _ptr(out)), "rpx_tune_trip_budget")
        return {"trip_budget": int(out[0]), "ms": {0: round(float(out[1]), 3), 48: round(float(out[2]), 3)}}

    def set_profiling(self, on=True):
        _check(library().rpx_set_profiling(self.handle, 1 if on else 0), "rpx_set_profiling")

    def last_timing(self):
        """GPU time of the last render: total, per-kernel sums (ms) and launch counts."""
        out = np.zeros(12, np.float32)
        _check(library().rpx_last_timing(self.handle, _ptr(out)), "rpx_last_timing")
        return _timing(out)


ADAPTIVE_BLOCK = 8
ADAPTIVE_WARMUP = 16
DECISION_WORDS = 6


def adaptive_decide(width, height, ev, target, cap, counts, sums, lum2):
    """rt_adaptive_decide (rt-wasm/src/ffi.rs:237-296): the adaptive decision of a width x height
    image from its blocks' sample counts, RGBA sums (4 floats per pixel) and squared-luminance sums.
    Returns one uint8 per block (1: in the next round) and the six decision words (round, active
    blocks, blocks, then the f32 bits of the threshold, largest error and remaining work)."""
    cols, rows = (width + ADAPTIVE_BLOCK - 1) // ADAPTIVE_BLOCK, (height + ADAPTIVE_BLOCK - 1) // ADAPTIVE_BLOCK
    counts = np.ascontiguousarray(counts, np.uint32).reshape(-1)
    sums = np.ascontiguousarray(sums, np.float32).reshape(-1)
    lum2 = np.ascontiguousarray(lum2, np.float32).reshape(-1)
    if counts.size != cols * rows or sums.size != 4 * width * height or lum2.size != width * height:
        raise ValueError("rt_adaptive_decide: plane sizes do not fit the image")
    active = np.zeros(cols * rows, np.uint8)
    words = np.zeros(DECISION_WORDS, np.uint32)
    _check(library().rt_adaptive_decide(width, height, ev, target, cap, _ptr(counts), _ptr(sums), _ptr(lum2),
                                        _ptr(active), _ptr(words)), "rt_adaptive_decide")
    return active, words


class Adaptive:
    """The state of an adaptive render in rounds for a host that renders the blocks itself
    (rt-core/src/adaptive.rs:319-554): the samples of every 8 x 8 block, the active blocks and the
    current round, with each decision taken by rt_adaptive_decide on the GPU."""

    def __init__(self, width, height, target, ev, cap):
        self.width, self.height = width, height
        self.cols = (width + ADAPTIVE_BLOCK - 1) // ADAPTIVE_BLOCK
        self.rows = (height + ADAPTIVE_BLOCK - 1) // ADAPTIVE_BLOCK
        self.counts = np.zeros(self.cols * self.rows, np.uint32)
        self.active = np.ones(self.cols * self.rows, np.uint8)
        self.round = min(ADAPTIVE_WARMUP, cap) if cap > 0 else ADAPTIVE_WARMUP
        self.done = 0
        self.iteration = 0
        self.target, self.ev, self.cap = max(float(target), 0.0), float(ev), int(cap)

    def left(self):
        return self.round - self.done

    def converged(self):
        return self.round == 0

    def advance(self, count):
        self.done = min(self.done + count, self.round)

    def quotas(self):
        """adaptive.rs:421-430: each block's samples in the current round."""
        q = np.where(self.active != 0, self.round, 0).astype(np.int64)
        if self.cap > 0:
            q = np.minimum(q, np.maximum(self.cap - self.counts.astype(np.int64), 0))
        return q.astype(np.uint32)

    def samples_of_pixel(self, x, y):
        b = (y // ADAPTIVE_BLOCK) * self.cols + x // ADAPTIVE_BLOCK
        return int(self.counts[b]) + min(int(self.quotas()[b]), self.done)

    def decide(self, sums, stride, lum2):
        """Counts the round's samples and chooses the next round (adaptive.rs:517-539) from the
        image's sums (`stride` floats per pixel, RGB first). Returns the decision (round, active,
        blocks, threshold, largest error, remaining, as float32) and the chosen blocks."""
        self.counts += self.quotas()
        self.iteration += self.round
        self.done = 0
        n = self.width * self.height
        planes = np.asarray(sums, np.float32).reshape(n, stride)
        rgba = np.zeros((n, 4), np.float32)
        rgba[:, :min(stride, 4)] = planes[:, :4]
        active, words = adaptive_decide(self.width, self.height, self.ev, self.target, self.cap, self.counts, rgba,
                                        lum2)
        self.active = active
        self.round = int(words[0])
        out = np.zeros(6, np.float32)
        out[:3] = words[:3]
        out[3:] = words[3:].view(np.float32)
        return out, active


def _flags(per_pass, graph, drain, ahead=True):
    return ((FLAG_PER_PASS if per_pass else 0) | (FLAG_GRAPH if graph else 0) | (0 if drain else FLAG_NO_DRAIN)
            | (0 if ahead else FLAG_ROUND_ONLY))


def _tune(out):
    return {"pool": int(out[0]), "shape": int(out[1]), "shadow_shape": int(out[2]), "best_ms": float(out[3]),
            "ms": [[round(float(v), 3) for v in out[4 + 6 * s:10 + 6 * s]] for s in range(8)],
            "shadow_ms": [round(float(v), 3) for v in out[52:60]]}


def _timing(out):
    return {"total_ms": float(out[0]), "extend_ms": float(out[1]), "shade_ms": float(out[2]),
            "shadow_ms": float(out[3]), "regen_ms": float(out[4]), "extend_launches": int(out[5]),
            "shade_launches": int(out[6]), "shadow_launches": int(out[7]), "regen_launches": int(out[8]),
            "iterations": int(out[9]), "chunks": int(out[10]), "drain_ms": float(out[11])}


class Queue:
    """A render queue (rpx_queue_*): jobs of several scenes rendered by one persistent path pool,
    each into its own resident film."""

    def __init__(self):
        q = library().rpx_queue_create()
        if q == 0:
            raise RenderError(last_error())
        self.handle = int(q)
        self.jobs = []

    def add(self, scene, first, count):
        job = ctypes.c_uint32()
        _check(library().rpx_queue_add(self.handle, scene.handle, first, count, ctypes.byref(job)), "rpx_queue_add")
        self.jobs.append((scene.width, scene.height))
        return int(job.value)

    def run(self):
        _check(library().rpx_queue_run(self.handle), "rpx_queue_run")

    def reset(self):
        """Clears every job's film."""
        _check(library().rpx_queue_reset(self.handle), "rpx_queue_reset")

    def film(self, job):
        W, H = self.jobs[job]
        rgb = np.zeros((H, W, 3), np.float32)
        weight = np.zeros((H, W), np.float32)
        alpha = np.zeros((H, W), np.float32)
        lum2 = np.zeros((H, W), np.float32)
        _check(library().rpx_queue_film_read(self.handle, job, _ptr(rgb), _ptr(weight), _ptr(alpha), _ptr(lum2)),
               "rpx_queue_film_read")
        return rgb, weight, alpha, lum2

    def set_options(self, pool=0, shape=KEEP, shadow_shape=KEEP, group=0, chunk=0, per_pass=False, graph=False,
                    drain=True, tail=0):
        flags = _flags(per_pass, graph, drain)
        _check(library().rpx_queue_set_options(self.handle, pool, shape, shadow_shape, group, chunk, flags, tail),
               "rpx_queue_set_options")

    def autotune(self, paths=0):
        out = np.zeros(60, np.float32)
        _check(library().rpx_queue_autotune(self.handle, paths, _ptr(out)), "rpx_queue_autotune")
        t = _tune(out)
        b = ctypes.c_uint32()
        _check(library().rpx_queue_trip_budget(self.handle, ctypes.byref(b)), "rpx_queue_trip_budget")
        t["trip_budget"] = int(b.value)
        return t

    def set_profiling(self, on=True):
        _check(library().rpx_queue_set_profiling(self.handle, 1 if on else 0), "rpx_queue_set_profiling")

    def last_timing(self):
        out = np.zeros(12, np.float32)
        _check(library().rpx_queue_last_timing(self.handle, _ptr(out)), "rpx_queue_last_timing")
        return _timing(out)

    def stats(self):
        out = np.zeros(4, np.uint64)
        _check(library().rpx_queue_stats(self.handle, _ptr(out)), "rpx_queue_stats")
        return {"rays": int(out[0]), "nearest": int(out[1]), "shadow": int(out[2]), "samples": int(out[3])}

    def close(self):
        if getattr(self, "handle", 0):
            library().rpx_queue_free(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def resolve(rgb, weight):
    """Average radiance per pixel: rgb * f32(1 / weight), zero where no sample landed."""
    rgb = np.asarray(rgb, np.float32)
    weight = np.asarray(weight, np.float32)
    out = np.zeros_like(rgb)
    m = weight > 0
    inv = np.zeros_like(weight)
    inv[m] = np.float32(1.0) / weight[m]
    out[m] = rgb[m] * inv[m][..., None]
    return out


def fnv1a(values):
    """FNV-1a over the little-endian bytes of f32 values (render.rs:981-989)."""
    data = np.ascontiguousarray(values, dtype="<f4").view(np.uint8)
    h = np.uint32(0x811C9DC5)
    # vectorised in blocks is not possible (sequential); use Python ints over bytes
    hv = int(h)
    for b in data.tobytes():
        hv = ((hv ^ b) * 0x01000193) & 0xFFFFFFFF
    return hv
