"""A writer of RSB v1 scene bytes (the reference's scene interchange format, scene/rsb.rs:1-130;
the interface's writer is src/io/rsb.ts) for the sections the GPU tracer reads: CAMR, MATL, MESH,
INST, LGHT, ENVR, FLOR, RNDR. Values are packed as given (little-endian f32/u32); no arithmetic is
done on them, so the bytes carry exactly the caller's floats.
"""
import struct

import numpy as np


def _f(*v):
    return struct.pack("<%df" % len(v), *[float(np.float32(x)) for x in v])


def _u(*v):
    return struct.pack("<%dI" % len(v), *[int(x) for x in v])


def _section(tag: bytes, payload: bytes) -> bytes:
    return tag + _u(len(payload)) + payload


class SceneWriter:
    def __init__(self):
        self.sections = []

    def camera(self, position, target, up=(0, 0, 1), focal_mm=50.0, sensor_mm=36.0, fstop=0.0,
               focus_dist=1000.0, ortho=False, ortho_width=1000.0):
        self.sections.append(_section(b"CAMR", _f(*position, *target, *up, focal_mm, sensor_mm, fstop, focus_dist)
                                      + _u(1 if ortho else 0) + _f(ortho_width)))
        return self

    def material(self, kind=0, base_color=(0.8, 0.8, 0.8), metallic=0.0, roughness=0.5, ior=1.5, transmission=0.0,
                 coat=0.0, coat_roughness=0.0, emission=(0, 0, 0), emission_power=0.0, flags=0, opacity=1.0,
                 anisotropy=0.0, absorption_mm=0.0, scatter_mm=0.0, scatter_color=(1.0, 1.0, 1.0), scatter_g=0.0):
        payload = (_u(kind) + _f(*base_color, metallic, roughness, ior, transmission, coat, coat_roughness,
                                 *emission, emission_power) + _u(flags)
                   + _f(opacity, anisotropy, absorption_mm, 0.0, scatter_mm, *scatter_color, scatter_g, 0.0))
        self.sections.append(_section(b"MATL", payload))
        return self

    def mesh(self, positions, indices, normals=None, material=0):
        p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        n = np.zeros_like(p) if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        i = np.ascontiguousarray(indices, np.uint32).reshape(-1)
        payload = _u(p.shape[0], i.size) + p.astype("<f4").tobytes() + n.astype("<f4").tobytes() \
            + i.astype("<u4").tobytes() + _u(material)
        self.sections.append(_section(b"MESH", payload))
        return self

    def instance(self, mesh, rows=((1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0)), material=0, flags=0):
        flat = [v for r in rows for v in r]
        self.sections.append(_section(b"INST", _u(mesh) + _f(*flat) + _u(material, flags)))
        return self

    def light(self, kind=0, position=(0, 0, 1000), direction=(0, 0, -1), up=(0, 1, 0), size_x=500.0, size_y=500.0,
              color=(1, 1, 1), power=1.0, spread_deg=180.0, angle_deg=0.0, flags=3):
        self.sections.append(_section(b"LGHT", _u(kind) + _f(*position, *direction, *up, size_x, size_y, *color, power,
                                                           spread_deg, angle_deg) + _u(flags)))
        return self

    def environment(self, kind=0, floor=0.1, horizon=0.3, zenith=0.6, strength=1.0, background_color=(0.05,) * 3,
                    background_mode=0):
        payload = (_u(kind) + _f(floor, horizon, zenith, strength, 0.0, *background_color) + _u(background_mode)
                   + _u(0xFFFFFFFF, 0xFFFFFFFF, 0) + _f(3.0, 0.0, 45.0, 0.0, 0.0, 1.0, 1.0) + _u(0))
        self.sections.append(_section(b"ENVR", payload))
        return self

    def floor(self, mode=0, z=0.0, color=(0.5, 0.5, 0.5)):
        self.sections.append(_section(b"FLOR", _u(mode) + _f(z, *color)))
        return self

    def render(self, width, height, max_samples=16, seed=0, max_depth=8, clamp_indirect=0.0, pixel_filter=0,
               caustic_photons=0, caustic_radius=0.0):
        payload = (_u(width, height, max_samples, seed) + _f(0.0, 0.0) + _u(1, max_depth) + _f(clamp_indirect)
                   + _u(0, 0, caustic_photons) + _f(caustic_radius, 1.0) + _u(pixel_filter, 1))
        self.sections.append(_section(b"RNDR", payload))
        return self

    def bytes(self) -> bytes:
        return b"RSB1" + _u(0) + b"".join(self.sections)


# The RNDR fields up to pixel_filter and edge_supersample with their defaults (scene/rsb.rs:46-57):
# width, height, max_samples, seed, max_seconds, exposure_ev, tonemap, max_depth, clamp_indirect,
# three u32 and a f32 off, denoise_strength 1, pixel_filter 0, edge_supersample 1.
_RNDR_DEFAULTS = (_u(0, 0, 0, 0) + _f(0.0, 0.0) + _u(1, 8) + _f(0.0) + _u(0, 0, 0) + _f(0.0, 1.0) + _u(0, 1))


def with_adaptive(rsb: bytes, noise_threshold: float, width=None, height=None) -> bytes:
    """The scene bytes with adaptive sampling asked for in the RNDR section (the E10 form: the
    section grows to 72 bytes with the adaptive bit and the noise target, scene/rsb.rs:57), and
    optionally another image size. Fields the section did not carry take their defaults."""
    out = bytearray(rsb)
    i = 8
    while i + 8 <= len(out):
        tag, n = bytes(out[i:i + 4]), struct.unpack("<I", out[i + 4:i + 8])[0]
        if tag == b"RNDR":
            body = bytearray(_RNDR_DEFAULTS)
            old = out[i + 8:i + 8 + n]
            body[:min(len(old), 64)] = old[:64]
            if width is not None:
                body[0:8] = _u(width, height)
            body += _u(1) + _f(noise_threshold)
            out[i:i + 8 + n] = _section(b"RNDR", bytes(body))
            return bytes(out)
        i += 8 + n
    raise ValueError("the scene bytes have no RNDR section")


def with_size(rsb: bytes, width: int, height: int, max_depth=None) -> bytes:
    """The scene bytes with another image size (and optionally path depth) in the RNDR section."""
    out = bytearray(rsb)
    i = 8
    while i + 8 <= len(out):
        tag, n = bytes(out[i:i + 4]), struct.unpack("<I", out[i + 4:i + 8])[0]
        if tag == b"RNDR":
            out[i + 8:i + 16] = _u(width, height)
            if max_depth is not None:
                out[i + 36:i + 40] = _u(max_depth)
            return bytes(out)
        i += 8 + n
    raise ValueError("the scene bytes have no RNDR section")


def box_mesh(lo, hi):
    """An axis-aligned box (positions, indices) wound outwards, without normals."""
    c = np.array([[hi[0] if (k & 1) else lo[0], hi[1] if (k & 2) else lo[1], hi[2] if (k & 4) else lo[2]]
                  for k in range(8)], np.float32)
    quads = [[0, 2, 3, 1], [4, 5, 7, 6], [0, 1, 5, 4], [2, 6, 7, 3], [1, 3, 7, 5], [0, 4, 6, 2]]
    idx = []
    for a, b, cc, d in quads:
        idx += [a, b, cc, a, cc, d]
    return c, np.array(idx, np.uint32)


def icosphere(center, radius, subdivisions=3):
    """A geodesic sphere (positions, indices) with flat facets, wound outwards: the regular
    icosahedron with each triangle split in four `subdivisions` times, new vertices pushed out to
    the sphere."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    dirs = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        mids = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in mids:
                m = dirs[a] + dirs[b]
                dirs.append(m / np.linalg.norm(m))
                mids[key] = len(dirs) - 1
            return mids[key]

        nxt = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nxt += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nxt
    p = np.asarray(center, np.float64) + radius * np.array(dirs)
    return p.astype(np.float32), np.array(faces, np.uint32).reshape(-1)


def uv_sphere(center, radius, lat=16, lon=32):
    """A UV sphere (positions, normals, indices) in float32."""
    th = np.linspace(0.0, np.pi, lat + 1)
    ph = np.linspace(0.0, 2 * np.pi, lon, endpoint=False)
    tt, pp = np.meshgrid(th, ph, indexing="ij")
    n = np.stack([np.sin(tt) * np.cos(pp), np.sin(tt) * np.sin(pp), np.cos(tt)], -1).reshape(-1, 3)
    p = np.asarray(center, np.float64) + radius * n
    idx = []
    for i in range(lat):
        for j in range(lon):
            a = i * lon + j
            b = i * lon + (j + 1) % lon
            c = a + lon
            d = b + lon
            idx += [a, c, b, b, c, d]
    return p.astype(np.float32), n.astype(np.float32), np.array(idx, np.uint32)
