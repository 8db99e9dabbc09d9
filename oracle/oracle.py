"""ORACLE — test infrastructure only.

Python binding of oracle/_build/liboracle.so, the CPU restatement of the reference renderer
(rt_core.hpp, rt_geometry.cpp, rt_bsdf.cpp, rt_render.cpp cite the reference lines each part
follows). Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the
checker; the product (repackshot_amd/) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
CACHE = os.path.join(HERE, "_build", "energy_tables.bin")

_lib = None


def build(force=False):
    """Compiles the oracle with its Makefile (g++, no FMA contraction)."""
    args = ["make", "-C", HERE]
    if force:
        args.append("-B")
    subprocess.run(args, check=True, capture_output=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, u32, u64, f32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float
        L.orc_set_energy_cache.argtypes = [ctypes.c_char_p]
        L.orc_scene_load.restype = vp
        L.orc_scene_load.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t]
        L.orc_scene_free.argtypes = [vp]
        L.orc_scene_info.argtypes = [vp, vp]
        L.orc_geometry_info.argtypes = [vp, vp]
        L.orc_render_tile.argtypes = [vp, u32, u32, u32, u32, u32, vp, vp, vp]
        L.orc_render_film.argtypes = [vp, u32, u32, ctypes.c_int, vp, vp, vp, vp, vp]
        L.orc_render_tiles.argtypes = [vp, vp, ctypes.c_size_t, u32, u32, ctypes.c_int, vp, vp, vp, vp, vp]
        L.orc_trace_rays.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int, vp, vp, vp, vp]
        L.orc_energy_tables.argtypes = [vp] * 6
        L.orc_fmath.restype = f32
        L.orc_fmath.argtypes = [ctypes.c_int, f32, f32]
        L.orc_pcg32.argtypes = [u64, u64, vp, ctypes.c_size_t]
        L.orc_fnv1a.restype = u32
        L.orc_fnv1a.argtypes = [vp, ctypes.c_size_t]
        L.orc_mix64.restype = u64
        L.orc_mix64.argtypes = [u64]
        L.orc_sobol.restype = u32
        L.orc_sobol.argtypes = [u32, u32]
        L.orc_owen.restype = u32
        L.orc_owen.argtypes = [u32, u32]
        L.orc_sampler_get1.restype = f32
        L.orc_sampler_get1.argtypes = [u32, u32, u32, u32, u32]
        L.orc_material_params.restype = u32
        L.orc_material_params.argtypes = [vp, vp, u32]
        L.orc_set_material_params.argtypes = [vp, vp, u32]
        L.orc_set_texture.argtypes = [vp, u32, u32, u32, vp, f32, f32, f32, f32]
        L.orc_adjoint.argtypes = [vp, u32, u32, ctypes.c_int, vp, vp, vp]
        L.orc_render_film_f64.argtypes = [vp, u32, u32, ctypes.c_int, vp, vp]
        L.orc_adaptive_new.restype = vp
        L.orc_adaptive_new.argtypes = [u32, u32, f32, f32, u32]
        L.orc_adaptive_free.argtypes = [vp]
        L.orc_adaptive_left.restype = u32
        L.orc_adaptive_left.argtypes = [vp]
        L.orc_adaptive_advance.argtypes = [vp, u32]
        L.orc_adaptive_samples_of_pixel.restype = u32
        L.orc_adaptive_samples_of_pixel.argtypes = [vp, u32, u32]
        L.orc_adaptive_blocks.restype = u32
        L.orc_adaptive_blocks.argtypes = [vp]
        L.orc_adaptive_decide.argtypes = [vp, vp, u32, vp, vp, vp]
        L.orc_adaptive_pixel_error.restype = f32
        L.orc_adaptive_pixel_error.argtypes = [f32, f32, f32, f32, u32, f32]
        L.orc_render_adaptive.restype = u32
        L.orc_render_adaptive.argtypes = [vp, u32, ctypes.c_int, vp, vp, vp, vp, vp]
        L.orc_render_passes.restype = u64
        L.orc_render_passes.argtypes = [vp, u32, u32, ctypes.c_int] + [vp] * 7
        L.orc_denoise.restype = ctypes.c_int
        L.orc_denoise.argtypes = [u32] * 9 + [f32] + [vp] * 9
        L.orc_denoise_synthetic.argtypes = [u32, u32, u32, f32, f32, f32] + [vp] * 7
        L.orc_bsdf_probe.restype = ctypes.c_int
        L.orc_bsdf_probe.argtypes = [vp, ctypes.c_int, vp, vp, f32, f32, ctypes.c_int, ctypes.c_int, vp]
        L.orc_fresnel_dielectric.restype = f32
        L.orc_fresnel_dielectric.argtypes = [f32, f32]
        L.orc_phase.restype = f32
        L.orc_phase.argtypes = [f32, f32]
        L.orc_sample_phase.restype = f32
        L.orc_sample_phase.argtypes = [vp, f32, f32, f32, vp]
        L.orc_lights_info.restype = u32
        L.orc_lights_info.argtypes = [vp, vp, vp, vp, u32]
        L.orc_lights_pmf.restype = f32
        L.orc_lights_pmf.argtypes = [vp, vp, u32]
        L.orc_lights_pick.argtypes = [vp, f32, vp, vp]
        L.orc_lights_pick_at.argtypes = [vp, vp, f32, vp, vp, vp]
        L.orc_set_techniques.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.orc_lights_sample.restype = ctypes.c_int
        L.orc_lights_sample.argtypes = [vp, vp, f32, f32, f32, vp]
        L.orc_lights_mis_probe.argtypes = [vp, vp, u32, vp]
        L.orc_lights_escaped.argtypes = [vp, vp, vp]
        L.orc_caustics_info.argtypes = [vp, vp]
        L.orc_caustics_radius.restype = f32
        L.orc_caustics_radius.argtypes = [vp, u32]
        L.orc_photons.restype = u32
        L.orc_photons.argtypes = [vp, u32, u32, u32, vp, u32, vp]
        L.orc_photon_gather.argtypes = [vp, u32, vp, u32, vp, u32, vp]
        L.orc_is_caster.restype = ctypes.c_int
        L.orc_is_caster.argtypes = [vp]
        L.orc_lights_emitter_weight.restype = f32
        L.orc_lights_emitter_weight.argtypes = [vp, vp, u32, f32, f32, f32]
        L.orc_soup.restype = u32
        L.orc_soup.argtypes = [vp, u32, vp]
        L.orc_random_rays.restype = ctypes.c_int
        L.orc_random_rays.argtypes = [vp, u32, u64, vp]
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        L.orc_set_energy_cache(CACHE.encode())
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class OracleScene:
    def __init__(self, rsb: bytes):
        err = ctypes.create_string_buffer(512)
        h = lib().orc_scene_load(rsb, len(rsb), err, 512)
        if not h:
            raise ValueError(err.value.decode())
        self.h = ctypes.c_void_p(h)
        info = np.zeros(5, np.uint32)
        lib().orc_scene_info(self.h, _p(info))
        self.width, self.height, self.max_samples, self.seed, self.depth = (int(v) for v in info)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_free(self.h)
            self.h = None

    def geometry_info(self):
        out = np.zeros(5, np.uint64)
        lib().orc_geometry_info(self.h, _p(out))
        return dict(zip(["meshes", "instances", "triangles", "nodes", "packets"], (int(v) for v in out)))

    def render_tile(self, x, y, w, h, sample_index):
        rgb = np.zeros((h, w, 3), np.float32)
        alpha = np.zeros((h, w), np.float32)
        rays = ctypes.c_uint64()
        lib().orc_render_tile(self.h, x, y, w, h, sample_index, _p(rgb), _p(alpha), ctypes.byref(rays))
        return rgb, alpha

    def render_film(self, first, count, threads=None, film=None):
        """Accumulates samples [first, first + count) into `film` (rgb, weight, alpha, lum2), or a new
        one; returns the film and the traversal counters."""
        if threads is None:
            threads = os.cpu_count() or 1
        H, W = self.height, self.width
        if film is None:
            film = (np.zeros((H, W, 3), np.float32), np.zeros((H, W), np.float32),
                    np.zeros((H, W), np.float32), np.zeros((H, W), np.float32))
        stats = np.zeros(7, np.uint64)
        lib().orc_render_film(self.h, first, count, threads, *(_p(a) for a in film), _p(stats))
        keys = ["rays", "nearest", "any", "nearest_nodes", "nearest_packets", "any_nodes", "any_packets"]
        return film, {k: int(v) for k, v in zip(keys, stats)}

    def render_tiles(self, rects, first, count, threads=None, film=None):
        """render_film over the listed (x, y, w, h) tiles only."""
        if threads is None:
            threads = os.cpu_count() or 1
        H, W = self.height, self.width
        if film is None:
            film = (np.zeros((H, W, 3), np.float32), np.zeros((H, W), np.float32),
                    np.zeros((H, W), np.float32), np.zeros((H, W), np.float32))
        r = np.ascontiguousarray(rects, np.uint32).reshape(-1, 4)
        stats = np.zeros(7, np.uint64)
        lib().orc_render_tiles(self.h, _p(r), r.shape[0], first, count, threads, *(_p(a) for a in film), _p(stats))
        keys = ["rays", "nearest", "any", "nearest_nodes", "nearest_packets", "any_nodes", "any_packets"]
        return film, {k: int(v) for k, v in zip(keys, stats)}

    def render_f64(self, first, count, threads=None):
        """The mean image of samples [first, first + count), radiance summed in float64: (H, W, 3)."""
        if threads is None:
            threads = os.cpu_count() or 1
        rgb = np.zeros((self.height, self.width, 3), np.float64)
        alpha = np.zeros((self.height, self.width), np.float64)
        lib().orc_render_film_f64(self.h, first, count, threads, _p(rgb), _p(alpha))
        return rgb / count

    def render_adaptive(self, cap=0, threads=None):
        """The adaptive render to `cap` samples (0: the scene's): film (rgb, weight, alpha, lum2),
        samples per block, number of decisions."""
        if threads is None:
            threads = os.cpu_count() or 1
        H, W = self.height, self.width
        film = (np.zeros((H, W, 3), np.float32), np.zeros((H, W), np.float32), np.zeros((H, W), np.float32),
                np.zeros((H, W), np.float32))
        blocks = ((W + 7) // 8) * ((H + 7) // 8)
        counts = np.zeros(blocks, np.uint32)
        n = lib().orc_render_adaptive(self.h, cap, threads, *(_p(a) for a in film), _p(counts))
        return film, counts, int(n)

    def render_passes(self, first, count, threads=None):
        """The resolved passes of the whole image (exact planes from the centre rays, albedo and
        alpha over samples [first, first + count)): a dict of planes in the reference's key order."""
        if threads is None:
            threads = os.cpu_count() or 1
        H, W = self.height, self.width
        planes = new_passes(W, H)
        lib().orc_render_passes(self.h, first, count, threads, *(_p(planes[k]) for k in PASS_KEYS))
        return planes

    def material_params(self):
        n = lib().orc_material_params(self.h, None, 0)
        out = np.zeros(5 * n, np.float32)
        lib().orc_material_params(self.h, _p(out), out.size)
        return out.reshape(n, 5)

    def set_material_params(self, params):
        params = np.ascontiguousarray(params, np.float32).reshape(-1, 5)
        if lib().orc_set_material_params(self.h, _p(params), params.shape[0]) != 0:
            raise ValueError("one parameter set per material expected")

    def set_texture(self, material, rgb, origin=(0.0, 0.0), scale=(1.0, 1.0)):
        self.texture = None
        if material is None:
            lib().orc_set_texture(self.h, 0xFFFFFFFF, 0, 0, None, 0, 0, 1, 1)
            return
        rgb = np.ascontiguousarray(rgb, np.float32)
        h, w = rgb.shape[:2]
        if lib().orc_set_texture(self.h, material, w, h, _p(rgb), origin[0], origin[1], scale[0], scale[1]) != 0:
            raise ValueError("bad texture")
        self.texture = (w, h)

    def adjoint(self, first, count, d_image, threads=None):
        """d loss / d parameter (materials x 5) and d loss / d texel (h, w, 3), in float64."""
        if threads is None:
            threads = os.cpu_count() or 1
        d_image = np.ascontiguousarray(d_image, np.float32)
        n = lib().orc_material_params(self.h, None, 0)
        gm = np.zeros(5 * n, np.float64)
        tex = getattr(self, "texture", None)
        gt = np.zeros((tex[1], tex[0], 3) if tex else (0,), np.float64)
        lib().orc_adjoint(self.h, first, count, threads, _p(d_image), _p(gm), _p(gt) if tex else None)
        return gm.reshape(n, 5), gt

    def lights(self):
        """(per_point, cdf, flux_pmf) of the prepared lights (light.rs:673-994)."""
        cdf = np.zeros(256, np.float32)
        pmf = np.zeros(256, np.float32)
        per = np.zeros(1, np.uint32)
        n = lib().orc_lights_info(self.h, _p(per), _p(cdf), _p(pmf), 256)
        return bool(per[0]), cdf[:n].copy(), pmf[:n].copy()

    def light_pmf(self, point, entry):
        """Probability that next-event estimation at `point` picks `entry` (light.rs:1076-1088)."""
        return float(lib().orc_lights_pmf(self.h, _p(np.asarray(point, np.float32)), entry))

    def light_pick(self, u):
        """pick() on the static cdf: (entry, rest) (light.rs:1059-1074)."""
        i = np.zeros(1, np.uint32)
        r = np.zeros(1, np.float32)
        lib().orc_lights_pick(self.h, u, _p(i), _p(r))
        return int(i[0]), float(r[0])

    def light_pick_at(self, point, u):
        """pick_at() on the selection at `point`: (entry, rest, probability) (light.rs:1118-1141)."""
        i = np.zeros(1, np.uint32)
        r = np.zeros(1, np.float32)
        q = np.zeros(1, np.float32)
        lib().orc_lights_pick_at(self.h, _p(np.asarray(point, np.float32)), u, _p(i), _p(r), _p(q))
        return int(i[0]), float(r[0]), float(q[0])

    def light_sample(self, point, u_select, u0, u1):
        """One light sample at `point` (light.rs:1145-1315), or None: a dict of wi, distance,
        radiance, pdf, delta, glossy, entry, kind (0 area, 1 sun, 2 sphere, 3 mesh, 4 environment)."""
        out = np.zeros(12, np.float32)
        if not lib().orc_lights_sample(self.h, _p(np.asarray(point, np.float32)), u_select, u0, u1, _p(out)):
            return None
        return {"wi": out[0:3].copy(), "distance": float(out[3]), "radiance": out[4:7].copy(), "pdf": float(out[7]),
                "delta": bool(out[8]), "glossy": bool(out[9]), "entry": int(out[10]), "kind": int(out[11])}

    def caustics(self):
        """The caustic setup (caustics.rs:229-345): has, photons, first radius, sources, caster
        spheres, most casters a photon crosses."""
        out = np.zeros(6, np.float32)
        lib().orc_caustics_info(self.h, _p(out))
        return {"on": bool(out[0]), "photons": int(out[1]), "radius": float(out[2]), "sources": int(out[3]),
                "spheres": int(out[4]), "max_bounces": int(out[5])}

    def caustic_radius(self, iteration):
        return float(lib().orc_caustics_radius(self.h, iteration))

    def photons(self, iteration, first=0, last=None):
        """The stored photons of photons [first, last) of pass `iteration`: an (n, 13) array of
        position, wi, power, normal, casters crossed, and the rays cast."""
        if last is None:
            last = self.caustics()["photons"]
        rays = np.zeros(1, np.uint64)
        n = lib().orc_photons(self.h, iteration, first, last, None, 0, _p(rays))
        out = np.zeros((max(n, 1), 13), np.float32)
        lib().orc_photons(self.h, iteration, first, last, _p(out), n, _p(rays))
        return out[:n], int(rays[0])

    def photon_gather(self, iteration, photons, queries):
        """Gathers (BSDF one) at queries (point, normal, most casters) from a map of `photons`."""
        ph = np.ascontiguousarray(photons, np.float32).reshape(-1, 13)
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 7)
        out = np.zeros((q.shape[0], 3), np.float32)
        lib().orc_photon_gather(self.h, iteration, _p(ph), ph.shape[0], _p(q), q.shape[0], _p(out))
        return out

    def light_escaped(self, direction):
        """What a camera ray leaving along `direction` sees of the sky (light.rs:1502-1520)."""
        out = np.zeros(3, np.float32)
        lib().orc_lights_escaped(self.h, _p(np.asarray(direction, np.float32)), _p(out))
        return out

    def emitter_weight(self, point, instance, bsdf_pdf, t, cos):
        """The MIS weight of an emitter found by a BSDF-sampled ray (light.rs:1537-1551)."""
        return float(lib().orc_lights_emitter_weight(self.h, _p(np.asarray(point, np.float32)), instance, bsdf_pdf,
                                                     t, cos))

    def light_mis_probe(self, tuples):
        """(kind, weight) per (point xyz, u_select, u0, u1): the MIS weight the BSDF side gives the
        light a sample found, at the sample's own density (kind -1: skipped)."""
        t = np.ascontiguousarray(tuples, np.float32).reshape(-1, 6)
        out = np.zeros((t.shape[0], 2), np.float32)
        lib().orc_lights_mis_probe(self.h, _p(t), t.shape[0], _p(out))
        return out[:, 0].astype(np.int32), out[:, 1].copy()

    def set_techniques(self, techniques, per_point=None):
        """0 MIS, 1 lights only, 2 BSDF only (the reference tests' switch); per_point forces the
        per-point selection on or off."""
        lib().orc_set_techniques(self.h, techniques, -1 if per_point is None else int(bool(per_point)))

    def random_rays(self, count, seed):
        """The reference tests' random rays around this scene's geometry (geometry/tests.rs:92-140),
        8 floats each: origin, direction, t_min, t_max."""
        out = np.zeros((count, 8), np.float32)
        if lib().orc_random_rays(self.h, count, seed, _p(out)) != 0:
            raise ValueError("the scene has no instances")
        return out

    def trace_rays(self, rays, brute=False):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        inst = np.zeros(n, np.uint32)
        tri = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        anyhit = np.zeros(n, np.uint8)
        lib().orc_trace_rays(self.h, _p(rays), n, 1 if brute else 0, _p(inst), _p(tri), _p(t), _p(anyhit))
        return inst, tri, t, anyhit


class Adaptive:
    """The state of an adaptive render in rounds (adaptive.rs Adaptive)."""

    def __init__(self, width, height, target, ev, cap):
        self.h = ctypes.c_void_p(lib().orc_adaptive_new(width, height, target, ev, cap))
        self.width, self.height = width, height
        self.n_blocks = int(lib().orc_adaptive_blocks(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_adaptive_free(self.h)
            self.h = None

    def left(self):
        return int(lib().orc_adaptive_left(self.h))

    def advance(self, n):
        lib().orc_adaptive_advance(self.h, n)

    def samples_of_pixel(self, x, y):
        return int(lib().orc_adaptive_samples_of_pixel(self.h, x, y))

    def decide(self, sums, stride, lum2):
        sums = np.ascontiguousarray(sums, np.float32)
        lum2 = np.ascontiguousarray(lum2, np.float32)
        out = np.zeros(6, np.float32)
        active = np.zeros(self.n_blocks, np.uint8)
        lib().orc_adaptive_decide(self.h, _p(sums), stride, _p(lum2), _p(out), _p(active))
        return out, active


def adaptive_pixel_error(r, g, b, lum2, n, floor):
    return np.float32(lib().orc_adaptive_pixel_error(r, g, b, lum2, n, floor))


def energy_tables():
    a = np.zeros(1024, np.uint16)
    s = np.zeros(1024, np.uint16)
    ci = np.zeros(17408, np.uint16)
    co = np.zeros(17408, np.uint16)
    aa = np.zeros(32, np.float32)
    sa = np.zeros(32, np.float32)
    lib().orc_energy_tables(_p(a), _p(s), _p(ci), _p(co), _p(aa), _p(sa))
    return {"albedo": a, "schlick": s, "clear_in": ci, "clear_out": co, "albedo_avg": aa, "schlick_avg": sa}


def fmath(which, x, y=0.0):
    return np.float32(lib().orc_fmath({"exp2": 0, "log2": 1, "powf": 2, "sin": 3, "cos": 4, "atan2": 5, "acos": 6}[which], x, y))


def pcg32(seed, stream, n):
    out = np.zeros(n, np.uint32)
    lib().orc_pcg32(seed, stream, _p(out), n)
    return out


def fnv1a(values):
    """FNV-1a over the little-endian bytes of an array (the reference's image and pass hashes)."""
    a = np.ascontiguousarray(values)
    return int(lib().orc_fnv1a(_p(a), a.nbytes))


def mix64(z):
    return int(lib().orc_mix64(z))


def sobol(dim, index):
    return int(lib().orc_sobol(dim, index))


def owen(x, seed):
    return int(lib().orc_owen(x, seed))


def sampler_get1(seed, x, y, index, dim):
    return np.float32(lib().orc_sampler_get1(seed, x, y, index, dim))


# The pass planes (rt-core/src/passes.rs:76-87), in the order of their bits.
PASS_KEYS = ["albedo", "normal", "depth", "partId", "materialId", "coc", "alpha"]
PASS_CHANNELS = [3, 3, 1, 1, 1, 1, 1]
FLOOR_ID = 0xFFFFFFFF


def new_passes(width, height):
    """Background passes of a width x height image (depth +inf, the rest zero)."""
    out = {}
    for k, c in zip(PASS_KEYS, PASS_CHANNELS):
        dt = np.uint32 if k in ("partId", "materialId") else np.float32
        shape = (height, width, c) if c > 1 else (height, width)
        out[k] = np.zeros(shape, dt)
    out["depth"][:] = np.inf
    return out


# Presets of the denoiser (rt-post/src/denoise.rs:76-90): (levels, variance radius).
FULL = (5, 3)
VIEWPORT = (3, 1)


def denoise(width, height, origin_y, rows, samples, y0, y1, planes, preset=FULL, strength=1.0, counts=None):
    """denoise_rows_at (rt-post/src/denoise.rs:386-407) over a band: `planes` holds sums (RGBA, 4
    per pixel), lum2, albedo, normal, depth, partId, materialId of `rows` rows. Returns (status,
    RGBA sums of rows y0..y1); status 0 ok, 1 shape, 2 rows."""
    def arr(k, dt):
        return np.ascontiguousarray(planes[k], dt)
    sums, lum2 = arr("sums", np.float32), arr("lum2", np.float32)
    albedo, normal, depth = arr("albedo", np.float32), arr("normal", np.float32), arr("depth", np.float32)
    part, material = arr("partId", np.uint32), arr("materialId", np.uint32)
    n = width * rows
    for a, c in ((sums, 4), (lum2, 1), (albedo, 3), (normal, 3), (depth, 1), (part, 1), (material, 1)):
        if a.size != n * c:
            return 1, None
    cnt = None if counts is None else np.ascontiguousarray(counts, np.float32)
    if cnt is not None and cnt.size != n:
        return 1, None
    out = np.zeros(max(width * max(y1 - y0, 0) * 4, 1), np.float32)
    st = lib().orc_denoise(width, height, origin_y, rows, samples, y0, y1, preset[0], preset[1], strength, _p(sums),
                           _p(lum2), _p(albedo), _p(normal), _p(depth), _p(part), _p(material), _p(cnt), _p(out))
    if st != 0:
        return int(st), None
    return 0, out[:width * (y1 - y0) * 4].reshape(y1 - y0, width, 4)


def denoise_synthetic(width, height, samples, left, right, spread):
    """The reference tests' synthetic film (rt-post/src/denoise/tests.rs:22-73) as denoiser planes."""
    n = width * height
    planes = {"sums": np.zeros((height, width, 4), np.float32), "lum2": np.zeros((height, width), np.float32),
              "albedo": np.zeros((height, width, 3), np.float32), "normal": np.zeros((height, width, 3), np.float32),
              "depth": np.zeros((height, width), np.float32), "partId": np.zeros((height, width), np.uint32),
              "materialId": np.zeros((height, width), np.uint32)}
    assert n > 0
    lib().orc_denoise_synthetic(width, height, samples, left, right, spread,
                                *(_p(planes[k]) for k in ("sums", "lum2", "albedo", "normal", "depth", "partId",
                                                          "materialId")))
    return planes


def phase(cos, g):
    """The Henyey-Greenstein phase function per steradian (integrator/medium.rs:143-150)."""
    return float(lib().orc_phase(cos, g))


def sample_phase(direction, g, u0, u1):
    """A direction around `direction` drawn from the phase function, and its density."""
    d = np.ascontiguousarray(direction, np.float32)
    out = np.zeros(3, np.float32)
    pdf = lib().orc_sample_phase(_p(d), g, u0, u1, _p(out))
    return out, float(pdf)


class Bsdf:
    """The oracle's BSDF of a material seen from outside (`entering`) or inside, in the local
    shading frame (z the normal): eval, pdf and sample (bsdf.rs:384-992)."""

    def __init__(self, kind=0, base=(0.8, 0.8, 0.8), metallic=0.0, roughness=0.5, ior=1.5, transmission=0.0,
                 coat=0.0, coat_roughness=0.0, entering=True):
        self.mat = np.array([kind, *base, metallic, roughness, ior, transmission, coat, coat_roughness], np.float32)
        self.entering = 1 if entering else 0

    def eval(self, wo, wi, all_lobes=True, single=False):
        """f(wo, wi); `single` leaves out the energy-compensation (multiple-scattering) lobes."""
        out = np.zeros(3, np.float32)
        mode = 2 if single else (1 if all_lobes else 0)
        lib().orc_bsdf_probe(_p(self.mat), self.entering, _p(np.asarray(wo, np.float32)),
                             _p(np.asarray(wi, np.float32)), 0.0, 0.0, 0, mode, _p(out))
        return out

    def pdf(self, wo, wi):
        out = np.zeros(1, np.float32)
        lib().orc_bsdf_probe(_p(self.mat), self.entering, _p(np.asarray(wo, np.float32)),
                             _p(np.asarray(wi, np.float32)), 0.0, 0.0, 1, 1, _p(out))
        return float(out[0])

    def param_derivs(self, wo, wi, all_lobes=True):
        """d f(wo, wi) / d (roughness, base r, g, b, IOR) as (5, 3): the BSDF rebuilt with each
        parameter as a dual number (the adjoint's per-vertex derivatives)."""
        out = np.zeros(15, np.float32)
        lib().orc_bsdf_probe(_p(self.mat), self.entering, _p(np.asarray(wo, np.float32)),
                             _p(np.asarray(wi, np.float32)), 0.0, 0.0, 3, 1 if all_lobes else 0, _p(out))
        return out.reshape(5, 3)

    def sample(self, wo, u0, u1):
        """(wi, weight, pdf) or None."""
        out = np.zeros(7, np.float32)
        z = np.zeros(3, np.float32)
        ok = lib().orc_bsdf_probe(_p(self.mat), self.entering, _p(np.asarray(wo, np.float32)), _p(z), u0, u1, 2, 1,
                                  _p(out))
        return (out[:3].copy(), out[3:6].copy(), float(out[6])) if ok else None


def soup():
    """The reference tests' degenerate triangle soup (geometry/tests.rs:48-88): positions (n, 3)
    and 15,000 indices."""
    pos = np.zeros((20000, 3), np.float32)
    idx = np.zeros(15000, np.uint32)
    n = lib().orc_soup(_p(pos), 20000, _p(idx))
    if n == 0:
        raise RuntimeError("soup positions overflow")
    return pos[:n].copy(), idx


def is_caster(kind=0, base=(0.8, 0.8, 0.8), metallic=0.0, roughness=0.5, ior=1.5, transmission=0.0, coat=0.0,
              coat_roughness=0.0):
    """caustics.rs:104-119: glass or metal sharp enough to focus light."""
    m = np.array([kind, *base, metallic, roughness, ior, transmission, coat, coat_roughness], np.float32)
    return bool(lib().orc_is_caster(_p(m)))


def fresnel_dielectric(cos_i, eta):
    return float(lib().orc_fresnel_dielectric(cos_i, eta))
