"""The denoiser's host API, a mirror of the reference's rt-post/src/denoise.rs (the names, argument
meaning, bands with their halo and error behaviour of DenoiseInput, denoise, denoise_rows,
denoise_rows_with and denoise_rows_at), running on the GPU through the C ABI: rt_denoise_with
(rt-wasm/src/ffi.rs:638-687) for a band's 14 planes, rpx_denoise_counted when each pixel carries its
own sample count (an adaptive image, rt-wasm/src/lib.rs:369-475).

The filter is the reference's edge-avoiding a-trous transform guided by the exact passes and each
pixel's variance; every preset, band and strength gives the reference's bits. There is no CPU
fallback: without the built library or a GPU the calls raise.
"""
from dataclasses import dataclass, replace

import numpy as np

from . import native

GUIDE_PASSES = native.PASS_ALBEDO | native.PASS_NORMAL | native.PASS_DEPTH | native.PASS_PART_ID | \
    native.PASS_MATERIAL_ID
LEVELS = 5
MIN_SAMPLES_FOR_VARIANCE = 4


@dataclass(frozen=True)
class Preset:
    """How hard the filter works (denoise.rs:67-120): passes of the transform and the radius of the
    spatial variance estimate. The C ABI knows the two presets by number."""
    levels: int
    variance_radius: int
    id: int

    def halo(self):
        """Rows a band holds above and below the rows it computes: 2^(levels + 1) + 8."""
        return (2 << self.levels) + 8


FULL = Preset(5, 3, 0)
VIEWPORT = Preset(3, 1, 1)
HALO = FULL.halo()


class DenoiseError(ValueError):
    """Why the filter refused its input (denoise.rs:166-182)."""


def _shape_error(what):
    return DenoiseError("denoise: %s does not fit the image" % what)


ROWS_ERROR = "denoise: the band does not hold the rows' halo"


@dataclass(frozen=True)
class DenoiseInput:
    """A band of rows of the accumulated image and its passes (denoise.rs:132-164): every plane holds
    width x rows pixels from row origin_y. `sums` are RGBA sums of `samples` samples, `lum2` the
    sums of the samples' squared luminance, `counts` (or None) each pixel's own sample count."""
    width: int
    height: int
    origin_y: int
    rows: int
    samples: int
    sums: np.ndarray
    lum2: np.ndarray
    albedo: np.ndarray
    normal: np.ndarray
    depth: np.ndarray
    part_id: np.ndarray
    material_id: np.ndarray
    counts: np.ndarray = None

    @staticmethod
    def whole(samples, sums, lum2, passes):
        """The whole image: its RGBA sums and squared-luminance sums of `samples` samples and its
        resolved passes (a dict with the guide planes, as Scene.passes returns)."""
        for k in ("albedo", "normal", "depth", "partId", "materialId"):
            if k not in passes:
                raise _shape_error("the passes")
        h, w = np.asarray(passes["depth"]).shape[:2]
        inp = DenoiseInput(w, h, 0, h, int(samples), _f32(sums), _f32(lum2), _f32(passes["albedo"]),
                           _f32(passes["normal"]), _f32(passes["depth"]), _u32(passes["partId"]),
                           _u32(passes["materialId"]))
        inp.check()
        return inp

    def with_counts(self, counts):
        """This input with each pixel's own sample count (None: none)."""
        inp = replace(self, counts=None if counts is None else _f32(counts))
        inp.check()
        return inp

    def with_halo(self, y0, y1):
        return self.with_halo_of(y0, y1, FULL)

    def with_halo_of(self, y0, y1, preset):
        """The band this input needs for image rows y0..y1 with `preset`: those rows and its halo,
        within the image and this input (denoise.rs:235-262)."""
        halo = preset.halo()
        lo = max(max(y0 - halo, 0), self.origin_y)
        hi = max(min(y1 + halo, self.height, self.origin_y + self.rows), lo)
        w = self.width
        a, b = (lo - self.origin_y) * w, (hi - self.origin_y) * w

        def cut(plane, n):
            flat = plane.reshape(-1)
            return flat[min(a * n, flat.size):min(b * n, flat.size)]
        return replace(self, origin_y=lo, rows=hi - lo, sums=cut(self.sums, 4), lum2=cut(self.lum2, 1),
                       albedo=cut(self.albedo, 3), normal=cut(self.normal, 3), depth=cut(self.depth, 1),
                       part_id=cut(self.part_id, 1), material_id=cut(self.material_id, 1),
                       counts=None if self.counts is None else cut(self.counts, 1))

    def check(self):
        """denoise.rs:264-287: the band inside the image and every plane its size."""
        if self.width == 0 or self.origin_y + self.rows > self.height:
            raise _shape_error("the band")
        n = self.width * self.rows
        for plane, c, what in ((self.sums, 4, "sums"), (self.lum2, 1, "lum2"), (self.albedo, 3, "albedo"),
                               (self.normal, 3, "normal"), (self.depth, 1, "depth"), (self.part_id, 1, "part_id"),
                               (self.material_id, 1, "material_id")):
            if plane.size != n * c:
                raise _shape_error(what)
        if self.counts is not None and self.counts.size != n:
            raise _shape_error("counts")
        return n

    def planes(self):
        """The band as the C ABI's input: 14 (15 with counts) planes of little-endian words."""
        parts = [self.sums, self.lum2, self.albedo, self.normal, self.depth, self.part_id.view(np.float32),
                 self.material_id.view(np.float32)]
        if self.counts is not None:
            parts.append(self.counts)
        return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32, copy=False)


def _f32(a):
    return np.ascontiguousarray(a, np.float32).reshape(-1)


def _u32(a):
    return np.ascontiguousarray(a, np.uint32).reshape(-1)


def rgba_sums(rgb, alpha):
    """The RGBA sums of a film (denoise.rs:290-299): radiance and coverage, 4 per pixel."""
    rgb = np.asarray(rgb, np.float32)
    return np.concatenate([rgb.reshape(-1, 3), np.asarray(alpha, np.float32).reshape(-1, 1)], axis=1).reshape(-1)


def denoise(inp):
    """Every row of the input with the full filter: RGBA sums, width x rows pixels."""
    return denoise_with(inp, FULL)


def denoise_with(inp, preset):
    return denoise_rows_with(inp, inp.origin_y, inp.origin_y + inp.rows, preset)


def denoise_rows(inp, y0, y1):
    """Image rows y0..y1 with the full filter; the input must hold them with HALO rows around them
    (or up to the image's edge): exactly those rows of denoise over the whole image."""
    return denoise_rows_with(inp, y0, y1, FULL)


def denoise_rows_with(inp, y0, y1, preset):
    return denoise_rows_at(inp, y0, y1, preset, 1.0)


def denoise_rows_at(inp, y0, y1, preset, strength):
    """denoise_rows_with at `strength` (denoise.rs:386-407): 1 the filter's image, 0 the input's own
    rows (no filter runs), in between noisy x (1 - s) + denoised x s per colour channel."""
    inp.check()
    halo = preset.halo()
    band_end = inp.origin_y + inp.rows
    if (y0 > y1 or y0 < inp.origin_y or y1 > band_end or inp.origin_y > max(y0 - halo, 0)
            or band_end < min(y1 + halo, inp.height)):
        raise DenoiseError(ROWS_ERROR)
    out = np.zeros(max(inp.width * (y1 - y0) * 4, 1), np.float32)
    data = inp.planes()
    lib = native.library()
    fn = lib.rpx_denoise_counted if inp.counts is not None else lib.rt_denoise_with
    st = fn(preset.id, float(strength), inp.width, inp.height, inp.origin_y, inp.rows, inp.samples, y0, y1,
            native._ptr(data), native._ptr(out))
    if st == 5:
        raise DenoiseError(ROWS_ERROR)
    native._check(st, "rt_denoise_with")
    return out[:inp.width * (y1 - y0) * 4]


def halo(preset_id):
    """rt_denoise_halo: rows a band of preset `preset_id` (0 full, 1 viewport) holds; 0 if unknown."""
    return int(native.library().rt_denoise_halo(preset_id))


def means(sums, samples):
    """The RGB means (sum x f32(1 / n)) of RGBA sums of `samples` samples (denoise/tests.rs:361-368)."""
    inv = np.float32(1.0) / np.float32(samples)
    return (np.asarray(sums, np.float32).reshape(-1, 4)[:, :3] * inv).astype(np.float32).reshape(-1)
