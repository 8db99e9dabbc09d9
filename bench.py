"""Benchmark of the GPU tracer on the single large render: 1024 x 1024, 64 spp, depth 12, eight
displaced level-7 icospheres (2,621,440 triangles), four rough conductors, two smooth glass, two
Lambert, three rect lights and the studio dome (repackshot_amd/csrc/rpx_generators.hip, kind 2).

A step renders the whole image at 64 samples per pixel into the film resident in HBM (the path
pool's launch shape and size chosen first by the library's tuner, outside the timed region).
Prints one JSON line: Mpaths/s of the timed steps, Mrays/s, the roofline of the traversal phase
(the closest-hit kernel with the shadow kernel beside it on a second stream), the CPU baseline (the oracle on this host's cores over a bounded sample of the
same workload), the parity of the GPU film against the oracle on that sample's tiles, and the
mixed render queue (64 seeded jobs of 256..1024 px, 10k..500k triangles, 16..64 spp) rendered by
one persistent pool beside the same jobs rendered one at a time in sample passes.

The adjoint record times the path-replay gradients of the inverse-material fit's scene against
central finite differences of the same forward renderer.

The denoise record times the guide passes and the reference's full denoising filter on the film of
the timed render.

The media record renders the caustics-and-media workload (a glass bottle holding a scattering
medium, lit by a sphere light; 1920 x 1080, 64 spp, depth 16) with the walks in their own kernel.

The caustics record renders the bottle of clear glass under a glossy disc light with 4,194,304
caustic photons per sample pass, the photon maps traced and built on the GPU.

The adaptive record renders the same scene adaptively (8 x 8 blocks in rounds to a 3 % noise target
or 64 samples) with the decisions on the GPU between rounds.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-queue] [--no-adjoint]
                       [--no-adaptive] [--no-denoise] [--no-media] [--no-caustics]
                       [--pool P --shape S --shadow-shape T]   (fixed launch options instead of the tuner)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpaths/s at 1024² 64spp & mixed queue; adjoint grads/s vs FD at fixed rRMSE"
HBM_PEAK_GBS = 8000.0  # MI355X datasheet: 8 TB/s HBM3E
# Algorithmic bytes of one node and one triangle packet of the reference's BVH8 layout
# (geometry/bvh.rs:95-102: bounds [3][2][8] f32 + child [8] u32 = 224 B, align 64 -> 256 B;
# geometry/triangle.rs:160-167: v [3][3][4] f32 + id [4] u32 = 160 B).
NODE_BYTES = 256
PACKET_BYTES = 160
# Memory-side bytes of one closest-hit launch per closest-hit ray, by trip budget (0: k_extend
# alone; 48: k_extend_budget + k_extend_resume): not measured in this run but taken from rocprofv3
# FETCH_SIZE and WRITE_SIZE runs of this workload (FETCH_SIZE scaled by 2.0, WRITE_SIZE by 1.0, the
# factors a 256 MiB device copy gives), recorded in TRAFFIC_SOURCE.
EXTEND_TRAFFIC_BYTES_PER_RAY = {0: 245.9, 48: 266.6}
TRAFFIC_SOURCE = ("profile-derived constant per closest-hit ray, not measured in this run: "
                  "profiles/hbm_traffic_large_render_1024_64spp.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(rsb, spp, budget_s=15.0):
    """The oracle (the CPU restatement of the reference renderer, its host schedule: 32-px tiles
    over an atomic tile queue on every hardware thread) timed on a bounded sample of the same
    workload: whole 32 x 32 tiles at all `spp` samples, batches of one tile per thread until about
    `budget_s` of wall time. Returns the baseline record, the rendered tiles (for the parity check
    of the GPU film) and the reference traversal's counters over them."""
    from oracle import oracle as orc

    threads = os.cpu_count() or 1
    t0 = time.perf_counter()
    o = orc.OracleScene(rsb)
    log("bench: oracle scene ready in %.1f s (%d threads)" % (time.perf_counter() - t0, threads))
    W, H = o.width, o.height
    order = [(tx * 32, ty * 32) for ty in range((H + 31) // 32) for tx in range((W + 31) // 32)]
    rng = np.random.default_rng(12345)
    rng.shuffle(order)
    film = None
    counts = {}
    rects = []
    elapsed = 0.0
    while elapsed < budget_s and len(rects) < len(order):
        batch = [(x, y, min(32, W - x), min(32, H - y)) for (x, y) in order[len(rects):len(rects) + threads]]
        t1 = time.perf_counter()
        film, st = o.render_tiles(batch, 0, spp, threads=threads, film=film)
        elapsed += time.perf_counter() - t1
        rects += batch
        for k, v in st.items():
            counts[k] = counts.get(k, 0) + v
        log("bench: cpu baseline %d tiles, %.1f s" % (len(rects), elapsed))
    paths = sum(w * h for (_, _, w, h) in rects) * spp
    record = {"value": round(paths / elapsed / 1e6, 4), "unit": "Mpaths/s", "cores": threads, "kind": "port",
              "sample": "%d random 32x32 tiles of the same image at all %d spp (%d paths, %d rays), %.1f s"
                        % (len(rects), spp, paths, counts["rays"], elapsed),
              "mrays_per_s": round(counts["rays"] / elapsed / 1e6, 4)}
    return record, rects, film, counts


def mixed_queue(native, seed=0, jobs=64):
    """The mixed render queue: every job loaded as its own scene; the persistent pool (one queue,
    tuned; and the same with captured graph launches) against the plain per-job tracer: each scene
    rendered alone through the same library at the queue's tuned options (whole sample range, the
    pool allocated by an untimed render first), and, as a second figure, alone in drained sample
    passes. Every job's film from the queue is compared bit for bit with the same scene alone."""
    t0 = time.perf_counter()
    scenes, spps, paths = [], [], 0
    for j in range(jobs):
        sc = native.Scene(native.write_queue_job(seed, j))
        info = sc.info()
        scenes.append(sc)
        spps.append(info["max_samples"])
        paths += info["width"] * info["height"] * info["max_samples"]
    log("bench: queue of %d jobs loaded in %.1f s (%d paths)" % (jobs, time.perf_counter() - t0, paths))
    q = native.Queue()
    ids = [q.add(sc, 0, spp) for sc, spp in zip(scenes, spps)]
    tune = q.autotune()
    opts = dict(pool=tune["pool"], shape=tune["shape"], shadow_shape=tune["shadow_shape"])
    log("bench: queue tuner chose pool %d shapes %d / %d trip budget %d"
        % (tune["pool"], tune["shape"], tune["shadow_shape"], tune["trip_budget"]))
    q.run()  # warm
    q.reset()
    st0 = q.stats()
    t1 = time.perf_counter()
    q.run()
    t_persistent = time.perf_counter() - t1
    rays = q.stats()["rays"] - st0["rays"]
    log("bench: persistent queue %.3f s" % t_persistent)
    films = [q.film(i) for i in ids]
    q.set_options(graph=True, **opts)
    q.reset()
    q.run()  # warm (capture, instantiation)
    q.reset()
    t1 = time.perf_counter()
    q.run()
    t_graph = time.perf_counter() - t1
    graph_same = all(all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(q.film(i), f))
                     for i, f in zip(ids, films))
    log("bench: persistent queue with graph launches %.3f s" % t_graph)
    q.close()
    # The plain per-job tracer: each scene alone, whole sample range, the queue's options; then the
    # same scene in drained sample passes. Each scene is closed after its job (its pool is freed).
    t_plain = t_pass = 0.0
    identical = True
    for sc, spp, film in zip(scenes, spps, films):
        sc.set_options(**opts)
        sc.set_trip_budget(tune["trip_budget"])
        sc.reset_film()
        sc.render_samples(0, spp)  # warm: the pool of this job's size
        sc.reset_film()
        t2 = time.perf_counter()
        sc.render_samples(0, spp)
        t_plain += time.perf_counter() - t2
        identical = identical and all(np.array_equal(a.view(np.uint32), b.view(np.uint32))
                                      for a, b in zip(film, sc.film()))
        sc.set_options(per_pass=True)
        sc.reset_film()
        t2 = time.perf_counter()
        sc.render_samples(0, spp)
        t_pass += time.perf_counter() - t2
        sc.close()
    log("bench: plain per-job %.3f s (in sample passes %.3f s)" % (t_plain, t_pass))
    return {"jobs": jobs, "paths": paths, "persistent_mpaths_per_s": round(paths / t_persistent / 1e6, 3),
            "persistent_mrays_per_s": round(rays / t_persistent / 1e6, 3),
            "plain_per_job_mpaths_per_s": round(paths / t_plain / 1e6, 3),
            "speedup": round(t_plain / t_persistent, 3),
            "speedup_basis": "each job alone through the same library at the queue's tuned options, whole sample "
                             "range per call, pool allocated beforehand",
            "persistent_s": round(t_persistent, 3), "plain_s": round(t_plain, 3),
            "graph_launches": {"persistent_s": round(t_graph, 3), "mpaths_per_s": round(paths / t_graph / 1e6, 3),
                               "films_equal_eager": graph_same},
            "per_pass_baseline": {"plain_s": round(t_pass, 3), "mpaths_per_s": round(paths / t_pass / 1e6, 3),
                                  "speedup": round(t_pass / t_persistent, 3),
                                  "what": "each job alone, one drained sample pass per call"},
            "tuned": {"pool": tune["pool"], "shape": tune["shape"], "shadow_shape": tune["shadow_shape"],
                      "trip_budget": tune["trip_budget"]},
            "films_bit_exact_vs_plain": identical}


FIT_PARAMS = [(m, 0) for m in range(6)] + [(m, 1) for m in range(8)] + [(4, 4), (5, 4)]

# The BSDFs read their energy-compensation tables linearly in roughness between nodes k / 31
# (bsdf/energy.rs, 32 nodes), and the compensation starts at alpha 2e-3 (energy.rs:58, roughness
# 0.0447): the image is piecewise smooth in roughness, with a kink at every node. A central
# difference whose step straddles a node measures the mean of two slopes, not the derivative, so the
# comparison is made at roughness values in the middle of their table cells (and away from 0.0447).
ROUGHNESS_NODES = 31
COMPENSATION_START = float(np.sqrt(np.float32(2.0e-3)))


def cell_centred(params, step):
    out = params.copy()
    for m in range(out.shape[0]):
        r = float(out[m, 0])
        if r >= 1.0:
            continue
        k = min(int(r * ROUGHNESS_NODES), ROUGHNESS_NODES - 1)
        c = (k + 0.5) / ROUGHNESS_NODES
        if abs(c - COMPENSATION_START) < 2.0 * step:
            c += 1.0 / ROUGHNESS_NODES
        out[m, 0] = np.float32(c)
    return out


def adjoint_bench(native, seed=0, ranges=48, steps=(1.0e-3, 1.0e-2)):
    """The inverse-material fit's workload (512x512, 32 spp, depth 6, 200,000 triangles, 8 materials x
    (roughness, rgb albedo, IOR) + a 256x256 rgb floor texture = 196,648 parameters): gradients per
    second of the path-replay adjoint against central finite differences of the same GPU forward
    renderer, and their agreement at 16 parameters over independent sample ranges (a fixed linear
    loss, so both estimate the same derivative), at the workload's difference step 1e-3 and at 1e-2,
    with the checked roughnesses in the middle of their energy-table cells. The adjoint and the
    differences of one range share its samples, so their per-range difference is what the check's
    error bars are taken from."""
    s = native.Scene(native.write_scene(4, seed))
    info = s.info()
    spp = info["max_samples"]
    rng = np.random.default_rng(3)
    tex = rng.uniform(0.2, 0.8, size=(256, 256, 3)).astype(np.float32)
    origin, scale = (-4000.0, -4000.0), (1.0 / 8000.0, 1.0 / 8000.0)
    s.set_texture(8, tex, origin, scale)
    params = cell_centred(s.material_params(), max(steps))
    s.set_material_params(params)
    n_params = params.size + tex.size
    weights = np.random.default_rng(5).uniform(0.5, 1.5, size=(s.height, s.width, 3))
    w32 = weights.astype(np.float32)

    def image(first):
        s.reset_film()
        s.render_samples(first, spp)
        rgb, wt, _, _ = s.film()
        return rgb.astype(np.float64) * (1.0 / wt.astype(np.float64))[..., None]

    s.adjoint(0, spp, w32)  # warm
    image(0)
    t0 = time.perf_counter()
    s.adjoint(0, spp, w32)
    t_adj = time.perf_counter() - t0
    t0 = time.perf_counter()
    s.reset_film()
    s.render_samples(0, spp)
    t_fwd = time.perf_counter() - t0
    A = []
    F = {h: [] for h in steps}
    for k in range(ranges):
        first = 1000 * k
        gm, _ = s.adjoint(first, spp, w32)
        A.append([gm[m, j] for m, j in FIT_PARAMS])
        for h in steps:
            row = []
            for m, j in FIT_PARAMS:
                up, dn = params.copy(), params.copy()
                up[m, j] += h
                dn[m, j] -= h
                s.set_material_params(up)
                lu = float(np.sum(weights * image(first)))
                s.set_material_params(dn)
                ld = float(np.sum(weights * image(first)))
                row.append((lu - ld) / (2 * h))
            s.set_material_params(params)
            F[h].append(row)
    s.close()
    A = np.array(A)
    am = A.mean(0)
    names = ["roughness %d" % m for m in range(6)] + ["base r %d" % m for m in range(8)] + ["ior 4", "ior 5"]

    def check(h):
        Fh = np.array(F[h])
        fm = Fh.mean(0)
        D = A - Fh
        sigma = D.std(0, ddof=1) / np.sqrt(ranges)  # paired: the per-range difference's error
        rel = np.abs(am - fm) / np.abs(fm)
        resolved = sigma <= 0.02 * np.abs(fm)
        return {"fd_step": h, "params": len(FIT_PARAMS), "ranges": ranges,
                "aggregate_rel_err": round(float(np.linalg.norm(am - fm) / np.linalg.norm(fm)), 4),
                "aggregate_rel_err_noise_floor": round(float(np.linalg.norm(sigma) / np.linalg.norm(fm)), 4),
                "max_abs_diff_over_sigma": round(float(np.max(np.abs(am - fm) / sigma)), 2),
                "resolved_to_2pct": int(resolved.sum()),
                "resolved_names": [n for n, r in zip(names, resolved) if r],
                "max_rel_err_resolved": round(float(np.max(rel[resolved])) if resolved.any() else -1.0, 4),
                "rel_sigma_per_param": {n: round(float(x), 4) for n, x in zip(names, sigma / np.abs(fm))}}

    adj_gps = n_params / t_adj
    fd_gps = 1.0 / (2.0 * t_fwd)
    rec = {"workload": "inverse-material fit: 512x512, 32 spp, depth 6, 199,714 tris, 8 materials x 5 + "
                       "256x256x3 texels", "parameters": int(n_params),
           "adjoint_s": round(t_adj, 4), "forward_s": round(t_fwd, 4),
           "adjoint_grads_per_s": round(adj_gps, 1), "fd_grads_per_s": round(fd_gps, 2),
           "grads_per_s_speedup": round(adj_gps / fd_gps, 1),
           "time_for_16_gradients": {"adjoint_s": round(t_adj, 4), "fd_s": round(32 * t_fwd, 4),
                                     "speedup": round(32 * t_fwd / t_adj, 2)},
           # Every material parameter (8 x 5): central differences need two renders each, the
           # adjoint's one pass gives them (and the texels) together.
           "time_for_40_material_gradients": {"adjoint_s": round(t_adj, 4), "fd_s": round(80 * t_fwd, 4),
                                              "speedup": round(80 * t_fwd / t_adj, 2)}}
    for h in steps:
        rec["fd_check" if h == max(steps) else "fd_check_step_%g" % h] = check(h)
    log("bench: adjoint %s" % rec)
    return rec


FIT_METALS = (0, 1, 2, 3)  # the fit scene's conductors (rpx_write_scene(4): materials 0..3)


def fit_texture():
    """The fit's true floor texture: a smooth 256 x 256 RGB pattern."""
    yy, xx = np.mgrid[0:256, 0:256] / 255.0
    return np.stack([0.5 + 0.3 * np.sin(6 * xx), 0.5 + 0.3 * np.cos(5 * yy), 0.4 + 0.2 * np.sin(4 * (xx + yy))],
                    -1).astype(np.float32)


def fit_bench(native, seed=0, steps=300, spp=16):
    """The inverse-material fit itself: the fit scene's target rendered at its true parameters
    (256 spp), the fit started from all 40 material parameters (8 materials x roughness, rgb, IOR)
    and every texel of the floor's 256 x 256 texture perturbed by up to +-30 % (seeded), Adam on the
    L2 image loss with the path-replay adjoint (repackshot_amd.fit): parameter, texture and image
    error before and after."""
    from repackshot_amd import fit as fitting

    s = native.Scene(native.write_scene(4, seed))
    origin, scale = (-4000.0, -4000.0), (1.0 / 8000.0, 1.0 / 8000.0)
    tex_true = fit_texture()
    p_true = s.material_params().copy()
    s.set_texture(8, tex_true, origin, scale)
    s.reset_film()
    s.render_samples(1_000_000, 256)
    target = s.resolve().astype(np.float64)
    rng = np.random.default_rng(11)
    p0 = np.clip(p_true * rng.uniform(0.7, 1.3, size=p_true.shape).astype(np.float32), fitting.PARAM_LO,
                 fitting.PARAM_HI)
    p0[8] = p_true[8]  # the floor: its base colour is the texture
    tex0 = np.clip(tex_true * rng.uniform(0.7, 1.3, size=tex_true.shape).astype(np.float32), fitting.TEXEL_LO,
                   fitting.TEXEL_HI)

    def image_err(p, tex, first):
        s.set_material_params(p)
        s.set_texture(8, tex, origin, scale)
        s.reset_film()
        s.render_samples(first, 32)
        return float(np.sqrt(np.mean((s.resolve() - target) ** 2) / np.mean(target ** 2)))

    mask = fitting.identifiable(p_true, FIT_METALS)[:8]

    def errors(p, tex):
        d, t = p[:8] - p_true[:8], p_true[:8]
        group = lambda cols: round(float(np.linalg.norm(d[:, cols][mask[:, cols]]) /
                                         np.linalg.norm(t[:, cols][mask[:, cols]])), 4)
        return {"all_40": round(float(np.linalg.norm(d) / np.linalg.norm(t)), 4),
                "identifiable": round(float(np.linalg.norm(d[mask]) / np.linalg.norm(t[mask])), 4),
                "roughness": group([0]), "base_colour": group([1, 2, 3]), "ior": group([4]),
                "texture": round(float(np.linalg.norm(tex - tex_true) / np.linalg.norm(tex_true)), 4)}

    before = errors(p0, tex0)
    img_before = image_err(p0, tex0, 2_000_000)
    noise_floor = image_err(p_true, tex_true, 2_100_000)
    t0 = time.perf_counter()
    p, tex, losses = fitting.fit(s, target, p0, tex0, 8, origin, scale, steps=steps, spp=spp, frozen=(8,))
    wall = time.perf_counter() - t0
    after = errors(p, tex)
    img_after = image_err(p, tex, 3_000_000)
    s.close()
    rec = {"workload": "inverse-material fit: 512x512, depth 6, 199,714 tris; 40 material parameters and "
                       "256x256x3 texels started up to +-30 % off, L2 loss against a 256-spp target",
           "steps": steps, "spp_per_step": spp, "adam_lr": {"params": 0.02, "texels": 0.004, "decay_to": 0.1},
           "seconds": round(wall, 2), "ms_per_step": round(1000.0 * wall / steps, 2),
           "param_rel_err": {"before": before, "after": after},
           "image_rel_rmse_32spp": {"before": round(img_before, 4), "after": round(img_after, 4),
                                    "at_the_true_parameters": round(noise_floor, 4)},
           "loss": {"first": round(losses[0], 6), "last_20_mean": round(float(np.mean(losses[-20:])), 6)},
           "identifiable_note": "a metal's IOR and a Lambert surface's roughness do not change the image"}
    log("bench: fit %s" % rec)
    return rec


# Algorithmic bytes of one a-trous pass per pixel: the guides (ids 8 B, normal + depth 16 B, albedo 16 B,
# depth gradient 8 B) and the illumination + variance (16 B) read once, the next illumination +
# variance (16 B) written; the 25 taps' re-reads come from the caches.
ATROUS_BYTES_PER_PIXEL = 80


def denoise_bench(scene, spp, repeats=5):
    """The passes and the denoiser (the reference's full filter: five a-trous passes guided by the
    albedo, normal, depth and ids) on the film of the timed render: GPU time of the pass rays (the
    centre ray and every sample's camera ray of each pixel) and of the filter, and the a-trous
    kernel's algorithmic bytes per second."""
    W, H = scene.width, scene.height
    scene.reset_passes()
    scene.render_passes(0, spp)  # warm
    scene.reset_passes()
    t0 = time.perf_counter()
    scene.render_passes(0, spp)
    t_passes = time.perf_counter() - t0
    scene.denoise_on_device(0, 1.0, spp)  # warm (work buffers)
    walls, gpu = [], []
    for _ in range(repeats):
        t1 = time.perf_counter()
        scene.denoise_on_device(0, 1.0, spp)
        walls.append(time.perf_counter() - t1)
        gpu.append(scene.denoise_timing())
    best = min(range(repeats), key=lambda i: gpu[i]["total_ms"])
    t = gpu[best]
    atrous_ms = t["atrous_ms"] / max(t["levels"], 1)
    achieved = ATROUS_BYTES_PER_PIXEL * W * H / (atrous_ms * 1e-3) / 1e9
    rec = {"workload": "full filter on the %dx%d film of %d spp" % (W, H, spp),
           "passes_ms": round(1000.0 * t_passes, 3),
           "pass_rays_per_s_m": round(W * H * (spp + 1) / t_passes / 1e6, 1),
           "denoise_ms": round(t["total_ms"], 4), "denoise_wall_ms": round(1000.0 * min(walls), 4),
           "mpixels_per_s": round(W * H / (t["total_ms"] * 1e-3) / 1e6, 1),
           "stage_ms": {k: round(t[k], 4) for k in ("setup_ms", "firefly_ms", "variance_ms", "atrous_ms", "final_ms")},
           "atrous_pass_ms": round(atrous_ms, 4),
           "atrous_roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": round(achieved / HBM_PEAK_GBS, 4),
                               "bytes_per_pixel": ATROUS_BYTES_PER_PIXEL}}
    log("bench: denoise %s" % rec)
    return rec


def adaptive_bench(native, rsb, tuned, target=0.03, cap=64):
    """Adaptive sampling on the same render: 8 x 8 blocks in rounds to a 3 % noise target (the
    reference's E10 example target) or 64 samples, the film resident in HBM and the decisions
    (per-pixel errors, smoothing, block maxima) on the GPU between rounds; the mean of three renders
    after a warm-up, at the large render's launch options and trip budget. Timed with candidate
    blocks' samples up to the cap rendered ahead in one run and kept for their rounds (the
    default), and with every round rendering only its own samples; both give the same film."""
    from repackshot_amd.rsb import with_adaptive

    scene = native.Scene(with_adaptive(rsb, target))
    W, H = scene.width, scene.height
    modes = {}
    films = {}
    for mode, ahead in (("rounds_only", False), ("ahead", True)):
        scene.set_options(pool=tuned["pool"], shape=tuned["shape"], shadow_shape=tuned["shadow_shape"],
                          adaptive_ahead=ahead)
        scene.set_trip_budget(tuned["trip_budget"])
        scene.render_adaptive(cap)  # warm-up (pool, planes and kept samples allocated)
        walls, gpu = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            st = scene.render_adaptive(cap)
            walls.append(time.perf_counter() - t0)
            gpu.append(st["gpu_ms"])
        wall = sum(walls) / len(walls)
        work = scene.adaptive_work()
        films[mode] = [f.copy() for f in scene.film()]
        modes[mode] = {"wall_s": round(wall, 4), "gpu_ms": round(sum(gpu) / len(gpu), 2), "renders_timed": 3,
                       "mpaths_per_s": round(st["paths"] / wall / 1e6, 3), "rendered_path_samples": work["rendered"],
                       "rendered_ahead": work["rendered_ahead"], "runs_ahead": work["runs_ahead"],
                       "unused_path_samples": work["unused"]}
    counts = scene.adaptive_counts()
    scene.close()
    uniform = W * H * cap
    same = all(np.array_equal(a.view(np.uint32), b.view(np.uint32))
               for a, b in zip(films["ahead"], films["rounds_only"]))
    rec = {"noise_target": target, "cap": cap, "decisions": st["decisions"], "rounds": st["rounds"],
           "path_samples": st["paths"], "uniform_path_samples": uniform,
           "fraction_of_uniform": round(st["paths"] / uniform, 4),
           "block_samples_min": st["min_samples"], "block_samples_max": st["max_samples"],
           "blocks_at_cap": int((counts >= cap).sum()), "blocks": int(counts.size),
           "last_max_error": round(st["max_error"], 5), "wall_s": modes["ahead"]["wall_s"],
           "gpu_ms": modes["ahead"]["gpu_ms"], "mpaths_per_s": modes["ahead"]["mpaths_per_s"],
           "modes": modes, "films_equal_across_modes": same}
    log("bench: adaptive %s" % rec)
    return rec


def media_bench(native, tuned, seed=0, tiles=((944, 320), (944, 480), (912, 640), (960, 760), (640, 800),
                                              (1280, 864))):
    """Caustics and media: the bottle workload (a 200,000-triangle glass bottle holding a scattering
    medium, one sphere light; 1920 x 1080, 64 spp, depth 16) rendered whole into the resident film
    with the large render's launch options, the random walks in their own kernel between the
    closest hits and the shade; and the GPU film against the oracle's on a few 32 x 32 tiles
    through the bottle, whose render time is the CPU baseline of this workload."""
    from oracle import oracle as orc

    rsb = native.write_scene(5, seed)
    scene = native.Scene(rsb)
    info = scene.info()
    W, H, spp = info["width"], info["height"], info["max_samples"]
    scene.set_options(pool=tuned["pool"], shape=tuned["shape"], shadow_shape=tuned["shadow_shape"])
    budget = scene.tune_trip_budget()
    scene.render_samples(0, spp)  # warm
    scene.reset_film()
    s0 = scene.stats()
    t0 = time.perf_counter()
    scene.render_samples(0, spp)
    wall = time.perf_counter() - t0
    s1 = scene.stats()
    scene.set_profiling(True)
    scene.reset_film()
    scene.render_samples(0, spp)
    timing = scene.last_timing()
    scene.set_profiling(False)
    rgb, weight, alpha, lum2 = scene.film()
    oracle = orc.OracleScene(rsb)
    rects = [(x, y, 32, 32) for x, y in tiles]
    threads = os.cpu_count() or 1
    t1 = time.perf_counter()
    ofilm, _ = oracle.render_tiles(rects, 0, spp, threads)
    cpu_s = time.perf_counter() - t1
    exact = all(np.array_equal(g[y:y + h, x:x + w].view(np.uint32), c[y:y + h, x:x + w].view(np.uint32))
                for (x, y, w, h) in rects for g, c in ((rgb, ofilm[0]), (weight, ofilm[1]), (alpha, ofilm[2]),
                                                      (lum2, ofilm[3])))
    paths = W * H * spp
    rec = {"workload": "caustics and media: 1920x1080, 64 spp, max depth 16, a 200,000-triangle glass bottle "
                       "(IOR 1.5) holding a medium of sigma_s 0.5 / sigma_a 0.05 per metre and g 0.7, Lambert "
                       "floor, one sphere light (radius 50 mm, radiance 500), no caustic photons (the reference's "
                       "sphere lights send none)",
           "width": W, "height": H, "spp": spp, "triangles": info["triangles"],
           "mpaths_per_s": round(paths / wall / 1e6, 3), "ms_per_step": round(1000.0 * wall, 2),
           "mrays_per_s": round((s1["rays"] - s0["rays"]) / wall / 1e6, 3),
           "rays_per_path": round((s1["rays"] - s0["rays"]) / paths, 4),
           "stage_ms_per_step": {k: round(timing[k], 2) for k in ("extend_ms", "shade_ms", "shadow_ms", "regen_ms",
                                                                   "total_ms")},
           "extend_ms_includes_walks": True,
           "trip_budget": budget,
           "parity": {"tiles": len(rects), "film_bit_exact_vs_oracle": bool(exact)},
           "cpu_baseline": {"value": round(len(rects) * 1024 * spp / cpu_s / 1e6, 4), "unit": "Mpaths/s",
                            "cores": threads, "kind": "port",
                            "sample": "%d tiles of 32 x 32 at %d spp through the bottle" % (len(rects), spp)}}
    scene.close()
    log("bench: media %s" % rec)
    return rec


def caustics_bench(native, tuned, seed=0, tiles=((864, 800), (832, 800), (800, 832), (768, 768))):
    """Caustics: the bottle of clear glass under a glossy disc light (rpx_write_scene(6); 1920 x 1080,
    64 spp, depth 16): before each chunk of up to 16 sample passes the GPU traces 4,194,304 photons
    per pass and builds each pass's hashed-grid map (radix sort by cell bucket); every receiver the
    camera paths meet gathers from its pass's map. Timed whole into the resident film with the
    large render's launch options. Parity: the GPU film of samples [0, 2) against the oracle's on
    32 x 32 tiles of the caustic under the bottle (the oracle traces the same two photon maps on
    this host's cores), whose time is the CPU baseline."""
    from oracle import oracle as orc

    rsb = native.write_scene(6, seed)
    scene = native.Scene(rsb)
    info = scene.info()
    W, H, spp = info["width"], info["height"], info["max_samples"]
    photons = scene.caustics()
    scene.set_options(pool=tuned["pool"], shape=tuned["shape"], shadow_shape=tuned["shadow_shape"])
    budget = scene.tune_trip_budget()
    scene.render_samples(spp, 16)  # warm: the path pool and the kept maps' memory
    scene.reset_film()
    s0 = scene.stats()
    t0 = time.perf_counter()
    scene.render_samples(0, spp)
    wall = time.perf_counter() - t0
    s1 = scene.stats()
    scene.set_profiling(True)
    scene.reset_film()
    scene.render_samples(2 * spp, spp)  # passes whose maps are not kept
    timing = scene.last_timing()
    kernels = scene.kernel_times()
    scene.set_profiling(False)
    scene.reset_film()
    scene.render_samples(0, 2)
    rgb, weight, alpha, lum2 = scene.film()
    oracle = orc.OracleScene(rsb)
    rects = [(x, y, 32, 32) for x, y in tiles]
    threads = os.cpu_count() or 1
    t1 = time.perf_counter()
    ofilm, _ = oracle.render_tiles(rects, 0, 2, threads)
    cpu_s = time.perf_counter() - t1
    exact = all(np.array_equal(g[y:y + h, x:x + w].view(np.uint32), c[y:y + h, x:x + w].view(np.uint32))
                for (x, y, w, h) in rects for g, c in ((rgb, ofilm[0]), (weight, ofilm[1]), (alpha, ofilm[2]),
                                                      (lum2, ofilm[3])))
    lit = float(np.mean([rgb[y:y + h, x:x + w].mean() for (x, y, w, h) in rects]))
    paths = W * H * spp
    map_ms = kernels["photon_map_ms"]
    rec = {"workload": "caustics: 1920x1080, 64 spp, max depth 16, a 200,000-triangle bottle of clear glass "
                       "(IOR 1.5) on a Lambert floor under a glossy disc light (100 mm, radiance 500), %d caustic "
                       "photons per sample pass (the most a scene may ask for), first gather radius 3 pixels"
                       % photons,
           "width": W, "height": H, "spp": spp, "triangles": info["triangles"], "photons_per_pass": photons,
           "mpaths_per_s": round(paths / wall / 1e6, 3), "ms_per_step": round(1000.0 * wall, 2),
           "mrays_per_s": round((s1["rays"] - s0["rays"]) / wall / 1e6, 3),
           "photon_maps_ms_per_step": round(map_ms, 2),
           "mphotons_per_s": round(photons * spp / (map_ms * 1e-3) / 1e6, 1) if map_ms > 0 else None,
           "photon_map_share": round(map_ms / timing["total_ms"], 3) if timing["total_ms"] > 0 else None,
           "stage_ms_per_step": {k: round(timing[k], 2) for k in ("extend_ms", "shade_ms", "shadow_ms", "regen_ms",
                                                                   "drain_ms", "total_ms")},
           "shade_includes_gather": True,
           "trip_budget": budget,
           "parity": {"tiles": len(rects), "samples": 2, "film_bit_exact_vs_oracle": bool(exact),
                      "tiles_mean_rgb": round(lit, 4)},
           "cpu_baseline": {"value": round(len(rects) * 1024 * 2 / cpu_s / 1e6, 5), "unit": "Mpaths/s",
                            "cores": threads, "kind": "port",
                            "sample": "%d tiles of 32 x 32 at 2 spp under the bottle, the two passes' photon maps "
                                      "(%d photons each) traced first" % (len(rects), photons)}}
    scene.close()
    log("bench: caustics %s" % rec)
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-queue", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-adaptive", action="store_true")
    ap.add_argument("--no-denoise", action="store_true")
    ap.add_argument("--no-media", action="store_true")
    ap.add_argument("--no-caustics", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--pool", type=int, default=0)
    ap.add_argument("--shape", type=int, default=-1)
    ap.add_argument("--shadow-shape", type=int, default=0)
    ap.add_argument("--trip-budget", type=int, default=48, help="closest-hit trip budget with --shape")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this tracer runs on one GPU; --gpus %d is not supported (one process, one GPU)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    from repackshot_amd import native

    if native.device_count() == 0:
        raise SystemExit("bench.py needs a GPU")
    t0 = time.perf_counter()
    rsb = native.write_scene(2, args.seed)
    log("bench: scene generated (%.1f MB) in %.1f s" % (len(rsb) / 1e6, time.perf_counter() - t0))
    t0 = time.perf_counter()
    scene = native.Scene(rsb)
    info = scene.info()
    load_s = time.perf_counter() - t0
    log("bench: scene loaded on the GPU in %.1f s: %s" % (load_s, info))
    spp = info["max_samples"]
    W, H = info["width"], info["height"]
    paths_per_step = W * H * spp

    if args.shape >= 0:
        scene.set_options(pool=args.pool or (1 << 21), shape=args.shape, shadow_shape=args.shadow_shape)
        scene.set_trip_budget(args.trip_budget)
        tuned = {"pool": args.pool or (1 << 21), "shape": args.shape, "shadow_shape": args.shadow_shape,
                 "trip_budget": scene.trip_budget(), "by": "fixed"}
    else:
        t = scene.autotune(spp)
        tuned = {"pool": t["pool"], "shape": t["shape"], "shadow_shape": t["shadow_shape"],
                 "trip_budget": t["trip_budget"], "by": "tuner", "candidates_ms": t["ms"],
                 "shadow_candidates_ms": t["shadow_ms"]}
        log("bench: tuner chose pool %d shapes %d / %d trip budget %d (%s %s)"
            % (t["pool"], t["shape"], t["shadow_shape"], t["trip_budget"], t["ms"], t["shadow_ms"]))
    scene.set_profiling(False)
    for i in range(args.warmup):
        scene.reset_film()
        scene.render_samples(0, spp)
        log("bench: warmup step %d done" % i)
    # Timed steps: the library call is synchronous (it waits for its stream), so the wall clock
    # brackets the GPU work exactly.
    stats0 = scene.stats()
    times = []
    for i in range(args.steps):
        scene.reset_film()
        t1 = time.perf_counter()
        scene.render_samples(0, spp)
        times.append(time.perf_counter() - t1)
        log("bench: step %d: %.4f s" % (i, times[-1]))
    stats1 = scene.stats()
    total = sum(times)
    mpaths = paths_per_step * args.steps / total / 1e6
    rays = stats1["rays"] - stats0["rays"]
    mrays = rays / total / 1e6

    # Per-kernel timing on the tracer's stream (HIP events) for one more step.
    scene.set_profiling(True)
    scene.reset_film()
    s_before = scene.stats()
    scene.render_samples(0, spp)
    s_after = scene.stats()
    timing = scene.last_timing()
    kernels = scene.kernel_times()
    scene.set_profiling(False)
    nearest_prof = s_after["nearest"] - s_before["nearest"]
    shadow_prof = s_after["shadow"] - s_before["shadow"]
    phase_avg_ms = timing["extend_ms"] / max(timing["extend_launches"], 1)
    extend_avg_ms = kernels["extend_kernel_ms"] / max(kernels["extend_launches"], 1)
    shadow_avg_ms = kernels["shadow_kernel_ms"] / max(kernels["shadow_launches"], 1)

    # The same step with each group of iterations launched as a captured HIP graph (same options).
    graph_ms = None
    if not args.no_graph:
        scene.set_options(pool=tuned["pool"], shape=tuned["shape"], shadow_shape=tuned["shadow_shape"], graph=True)
        scene.reset_film()
        scene.render_samples(0, spp)  # warm (capture and instantiation)
        gt = []
        for _ in range(max(args.steps, 1)):
            scene.reset_film()
            t1 = time.perf_counter()
            scene.render_samples(0, spp)
            gt.append(time.perf_counter() - t1)
        graph_ms = 1000.0 * sum(gt) / len(gt)
        scene.set_options(pool=tuned["pool"], shape=tuned["shape"], shadow_shape=tuned["shadow_shape"])
        scene.reset_film()
        scene.render_samples(0, spp)
        log("bench: graph launches %.3f ms per step" % graph_ms)

    result = {
        "metric": METRIC,
        "value": round(mpaths, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator; no external assets)",
        "config": {"workload": "single large render: 1024x1024, 64 spp, max depth 12, 8 displaced level-7 "
                               "icospheres (2,621,440 tris), 4 rough conductors, 2 smooth glass, 2 Lambert, "
                               "3 rect lights + studio dome, MIS NEE, f32 radiance",
                   "width": W, "height": H, "spp": spp, "max_depth": info["max_depth"],
                   "triangles": info["triangles"], "seed": args.seed},
        "mrays_per_s": round(mrays, 3),
        "rays_per_path": round(rays / (paths_per_step * args.steps), 4),
        "launch": tuned,
        "stage_ms_per_step": {k: round(timing[k], 2) for k in ("extend_ms", "shade_ms", "shadow_ms", "regen_ms",
                                                                "drain_ms", "total_ms")},
        "iterations_per_step": timing["iterations"],
        "graph_launches": None if graph_ms is None else {
            "ms_per_step": round(graph_ms, 3), "mpaths_per_s": round(paths_per_step / (graph_ms * 1e-3) / 1e6, 3),
            "note": "each group of 8 pool iterations launched as one captured HIP graph, same pool and shapes"},
        # RSB parse, host BVH build (binned SAH, 4-wide, all host threads) and upload, outside the steps.
        "scene_load_s": round(load_s, 3),
    }

    if not args.no_cpu_baseline:
        cpu, rects, ofilm, counts = cpu_baseline(rsb, spp)
        result["cpu_baseline"] = cpu
        # Parity of the GPU film (samples 0..spp-1 of the profiled step) on the baseline's tiles.
        rgb, weight, alpha, lum2 = scene.film()
        exact = True
        max_rel = 0.0
        for (x, y, w, h) in rects:
            for g, c in ((rgb, ofilm[0]), (weight, ofilm[1]), (alpha, ofilm[2]), (lum2, ofilm[3])):
                gs, cs = g[y:y + h, x:x + w], c[y:y + h, x:x + w]
                if not np.array_equal(gs.view(np.uint32), cs.view(np.uint32)):
                    exact = False
                    max_rel = max(max_rel, float(np.max(np.abs(gs - cs) / np.maximum(np.abs(cs), 1e-6))))
        # Image-level error of the resolved tiles (the bar is rRMSE <= 1e-3 at equal spp).
        num = den = 0.0
        for (x, y, w, h) in rects:
            gi = rgb[y:y + h, x:x + w].astype(np.float64) / weight[y:y + h, x:x + w, None]
            ci = ofilm[0][y:y + h, x:x + w].astype(np.float64) / ofilm[1][y:y + h, x:x + w, None]
            num += float(np.sum((gi - ci) ** 2))
            den += float(np.sum(ci ** 2))
        result["parity"] = {"tiles": len(rects), "pixels": int(sum(w * h for (_, _, w, h) in rects)),
                            "film_bit_exact_vs_oracle": exact, "max_rel_err": max_rel,
                            "image_rrmse": float(np.sqrt(num / max(den, 1e-30)))}
        # Algorithmic bytes of one closest-hit and one any-hit ray: the reference traversal's BVH8
        # nodes and packets visited per query, counted by the oracle over the baseline's rays. The
        # measured span is one traversal phase: the closest-hit kernel and, beside it on a second
        # stream, the previous iteration's shadow rays.
        b_ray = (NODE_BYTES * counts["nearest_nodes"] + PACKET_BYTES * counts["nearest_packets"]) / max(
            counts["nearest"], 1)
        b_any = (NODE_BYTES * counts["any_nodes"] + PACKET_BYTES * counts["any_packets"]) / max(counts["any"], 1)
        launches = max(kernels["extend_launches"], 1)
        traffic_per_ray = EXTEND_TRAFFIC_BYTES_PER_RAY.get(tuned["trip_budget"])
        ext_bytes = b_ray * nearest_prof / launches
        shadow_bytes = b_any * shadow_prof / max(kernels["shadow_launches"], 1)
        achieved = ext_bytes / (extend_avg_ms * 1e-3) / 1e9
        shadow_achieved = shadow_bytes / (shadow_avg_ms * 1e-3) / 1e9
        phase_achieved = (ext_bytes + shadow_bytes) / (phase_avg_ms * 1e-3) / 1e9
        result["roofline"] = {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None if traffic_per_ray is None else round(traffic_per_ray * nearest_prof / launches),
            "traffic_source": TRAFFIC_SOURCE,
            # What the kernel moves from memory per launch over its time: the L2 and the Infinity
            # Cache serve most node fetches, so the kernel is held by divergence and dependent-fetch
            # latency, well under the HBM roof it is priced against.
            "traffic_frac": None if traffic_per_ray is None else round(
                traffic_per_ray * nearest_prof / launches / (extend_avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "limiter": "divergence and dependent-fetch latency (17 of 64 lanes active, VALU busy 72 %)",
            "kernel": ("k_extend (closest hit, shape %d): algorithmic bytes of its rays per launch over its own "
                       "average launch time (HIP events on its stream)" % tuned["shape"]) if not tuned["trip_budget"]
                      else ("k_extend_budget + k_extend_resume (the closest hit, shape %d, walks past %d trips "
                            "finished by the second kernel): algorithmic bytes of its rays per launch pair over the "
                            "pair's average time (HIP events on its stream)" % (tuned["shape"], tuned["trip_budget"])),
            "bytes_per_ray": round(b_ray, 1),
            "closest_hit_rays_per_launch": round(nearest_prof / launches, 1),
            "extend_avg_ms": round(extend_avg_ms, 4),
            "extend_launches_per_step": kernels["extend_launches"],
            "other_kernels": {
                "k_shadow": {"achieved": round(shadow_achieved, 1), "frac": round(shadow_achieved / HBM_PEAK_GBS, 4),
                             "bytes_per_ray": round(b_any, 1),
                             "rays_per_launch": round(shadow_prof / max(kernels["shadow_launches"], 1), 1),
                             "avg_ms": round(shadow_avg_ms, 4), "stream": "second (beside k_extend)"},
                "traversal_phase": {"achieved": round(phase_achieved, 1),
                                    "frac": round(phase_achieved / HBM_PEAK_GBS, 4), "avg_ms": round(phase_avg_ms, 4),
                                    "what": "k_extend until k_shadow beside it has joined; both kernels' bytes"}},
        }
    if not args.no_denoise:
        result["denoise"] = denoise_bench(scene, spp)
    # Each later workload loads its own scene after this one is closed (its pool, results and
    # continuation buffers freed): a scene loaded while they were held measured its adaptive render
    # 11 % slower for its whole life (61.5 against 55.2 ms), a placement effect in device memory.
    scene.close()
    if not args.no_adaptive:
        result["adaptive"] = adaptive_bench(native, rsb, tuned)
    if not args.no_media:
        result["media"] = media_bench(native, tuned, args.seed)
    if not args.no_caustics:
        result["caustics"] = caustics_bench(native, tuned, args.seed)
    if not args.no_queue:
        result["mixed_queue"] = mixed_queue(native, args.seed)
    if not args.no_adjoint:
        result["adjoint"] = adjoint_bench(native, args.seed)
        result["fit"] = fit_bench(native, args.seed)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
