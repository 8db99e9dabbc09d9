"""The oracle's adaptive sampling (adaptive.rs) against the reference's pinned decisions and the
properties its tests hold: six decisions over a synthetic image hash to PINNED_E10_DECISIONS
(adaptive/tests.rs:455-534), pixel errors of known sums, and an adaptive render whose every pixel is
the plain mean of its own samples (adaptive/tests.rs:390-434)."""
import struct

import numpy as np

from oracle import oracle as orc


PINNED_E10_DECISIONS = 0xB492454A
M32 = 0xFFFFFFFF
f32 = np.float32


def fnv1a_bytes(data):
    h = 0x811C9DC5
    for b in data:
        h = ((h ^ b) * 0x01000193) & M32
    return h


def mix(x):
    x ^= x >> 16
    x = (x * 0x7FEB352D) & M32
    x ^= x >> 15
    x = (x * 0x846CA68B) & M32
    x ^= x >> 16
    return x


def synthetic(ad, w, h, rnd):
    """Per pixel a colour and a spread from an integer hash of the pixel and the round: the sums of
    as many samples as the pixel has (four floats per pixel) and the squared-luminance sums."""
    sums = np.zeros(w * h * 4, np.float32)
    lum2 = np.zeros(w * h, np.float32)
    for y in range(h):
        for x in range(w):
            p = y * w + x
            v = mix(((p * 2654435761) & M32) ^ ((rnd * 0x9E3779B9) & M32))
            n = f32(ad.samples_of_pixel(x, y))
            c = [f32(v & 0xFF) / f32(255.0), f32((v >> 8) & 0xFF) / f32(255.0), f32((v >> 16) & 0xFF) / f32(255.0)]
            spread = f32((v >> 24) & 0xFF) / f32(255.0) * (f32(0.5) if x < w // 3 else f32(0.05)) \
                * (f32(2.0) if y < 9 else f32(1.0))
            l = f32(0.2126) * c[0] + f32(0.7152) * c[1] + f32(0.0722) * c[2]
            sums[4 * p:4 * p + 4] = [c[0] * n, c[1] * n, c[2] * n, n]
            lum2[p] = (l * l + spread * spread) * n
    return sums, lum2


def synthetic_decisions(rounds, make):
    """Six rounds of decisions on a 45 x 37 synthetic image (target 2 %, exposure +0.5, cap 200),
    hashed: each round's size, active blocks, block count, threshold, largest error and remaining
    work, and the chosen blocks. `make(w, h, target, ev, cap)` gives the adaptive state, whose
    `decide(sums, stride, lum2)` returns (out6, active)."""
    w, h = 45, 37
    ad = make(w, h, 0.02, 0.5, 200)
    data = bytearray()
    for rnd in range(rounds):
        ad.advance(ad.left())
        sums, lum2 = synthetic(ad, w, h, rnd)
        out, active = ad.decide(sums, 4, lum2)
        data += struct.pack("<3I", int(out[0]), int(out[1]), int(out[2]))
        data += np.asarray(out[3:6], "<f4").tobytes()
        data += bytes(active.astype(np.uint8))
        if int(out[0]) == 0:
            break
    return fnv1a_bytes(data)


def test_synthetic_decisions_have_the_pinned_hash():
    assert synthetic_decisions(6, orc.Adaptive) == PINNED_E10_DECISIONS


def test_pixel_error_of_known_sums():
    floor = f32(0.01)
    # Fewer than two samples, or non-finite sums: unmeasured.
    assert orc.adaptive_pixel_error(1, 1, 1, 1, 1, floor) == f32(1e30)
    assert orc.adaptive_pixel_error(np.inf, 1, 1, 1, 4, floor) == f32(1e30)
    # n grey samples of luminance m +- s (half each way): mean m, E[l^2] = m^2 + s^2, and the
    # error sqrt(s^2 / (n - 1)) / (m + floor) (the variance estimate the module states).
    for m, s, n in [(0.5, 0.1, 16), (0.2, 0.0, 8), (1.5, 0.7, 64)]:
        l2 = (m * m + s * s) * n
        got = orc.adaptive_pixel_error(m * n, m * n, m * n, l2, n, floor)
        want = np.sqrt(max(m * m + s * s - m * m, 0.0) / (n - 1)) / (m + 0.01)
        assert abs(float(got) - want) <= 1e-5 * max(want, 1e-6) + 1e-6, (m, s, n, got, want)


def adaptive_scene(target=0.05, width=64, height=48):
    """The convergence tests' glossy scene with adaptive sampling asked for (RNDR adaptive bit and a
    noise target)."""
    import test_gpu_convergence as conv
    from repackshot_amd.rsb import with_adaptive
    return with_adaptive(conv.scene(), target, width, height)


def test_adaptive_pixels_are_plain_means_of_their_own_samples():
    """Rendered adaptively to a cap, every pixel equals, bit for bit, the plain film's pixel of as
    many samples: a block that stops keeps its mean; no weights, no mixing."""
    o = orc.OracleScene(adaptive_scene())
    film, counts, decisions = o.render_adaptive(cap=64, threads=8)
    assert decisions >= 2
    assert counts.min() >= 16 and counts.max() <= 64 and counts.max() > counts.min(), counts
    rgb, weight, _, lum2 = film
    cols = (o.width + 7) // 8
    per_pixel = np.array([[counts[(y // 8) * cols + x // 8] for x in range(o.width)] for y in range(o.height)])
    assert np.array_equal(weight, per_pixel.astype(np.float32))
    checked = 0
    for n in np.unique(per_pixel):
        plain, _ = o.render_film(0, int(n), threads=8)
        m = per_pixel == n
        assert np.array_equal(rgb[m].view(np.uint32), plain[0][m].view(np.uint32))
        assert np.array_equal(lum2[m].view(np.uint32), plain[3][m].view(np.uint32))
        checked += int(m.sum())
    assert checked == o.width * o.height
