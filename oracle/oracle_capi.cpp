// ORACLE — test infrastructure only. The C ABI through which tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg drive the CPU restatement of the reference renderer. The product
// (repackshot_amd/) never loads this library.
//
// The film entry point follows the reference host's schedule (rt-cli/src/lib.rs:266-366): 32-pixel
// tiles claimed from one atomic counter by `threads` workers; each pixel's samples are added to
// its f32 film sums in sample-index order (image.rs:81-94), so the film does not depend on the
// number of threads.
#include <atomic>
#include <cstring>
#include <thread>
#include "rt_adaptive.hpp"
#include "rt_bsdf_grad.hpp"
#include "rt_denoise.hpp"
#include "rt_passes.hpp"
#include "rt_caustics.hpp"
#include "rt_handle.hpp"

using namespace orc;

namespace {
void set_err(char* err, size_t cap, const std::string& msg) {
  if (!err || cap == 0) return;
  size_t n = std::min(cap - 1, msg.size());
  std::memcpy(err, msg.data(), n);
  err[n] = 0;
}
}  // namespace

extern "C" {

void orc_set_energy_cache(const char* path) { energy::set_cache_path(path); }

void* orc_scene_load(const uint8_t* bytes, size_t len, char* err, size_t err_cap) {
  auto* h = new Handle();
  std::string e;
  if (!scene_from_rsb(bytes, len, h->scene, e) || !h->prep.build(h->scene, e)) {
    set_err(err, err_cap, e);
    delete h;
    return nullptr;
  }
  return h;
}

void orc_scene_free(void* p) { delete static_cast<Handle*>(p); }

// width, height, max_samples, seed, depth
void orc_scene_info(void* p, uint32_t out[5]) {
  const Scene& s = static_cast<Handle*>(p)->scene;
  out[0] = s.render.width;
  out[1] = s.render.height;
  out[2] = s.render.max_samples;
  out[3] = s.render.seed;
  out[4] = s.render.depth();
}

// Geometry counts: meshes, instances, triangles (each mesh once), wide nodes, packets.
void orc_geometry_info(void* p, uint64_t out[5]) {
  const Geometry& g = static_cast<Handle*>(p)->scene.geometry;
  uint64_t tris = 0, nodes = g.tlas.nodes.size(), packets = 0;
  for (const Mesh& m : g.meshes) {
    tris += m.triangle_count();
    nodes += m.bvh.nodes.size();
    packets += m.packets.size();
  }
  out[0] = g.meshes.size();
  out[1] = g.instances.size();
  out[2] = tris;
  out[3] = nodes;
  out[4] = packets;
}

// One sample (`sample_index`) of every pixel of the tile: w*h*3 radiance floats and w*h alphas,
// row major from the tile's top-left (the value rt_render_tile_rgba splats).
int orc_render_tile(void* p, uint32_t x, uint32_t y, uint32_t w, uint32_t h, uint32_t sample_index, float* rgb,
                    float* alpha, uint64_t* rays) {
  auto* hd = static_cast<Handle*>(p);
  const Scene& s = hd->scene;
  RenderStats st;
  PhotonMap map;
  if (hd->prep.has_caustics) map = trace_photon_map(s, hd->prep, sample_index, (int)std::thread::hardware_concurrency());
  const PhotonMap* photons = hd->prep.has_caustics ? &map : nullptr;
  for (uint32_t j = 0; j < h; ++j)
    for (uint32_t i = 0; i < w; ++i) {
      uint32_t px = x + i, py = y + j;
      size_t k = (size_t)j * w + i;
      if (px >= s.render.width || py >= s.render.height) {
        rgb[3 * k] = rgb[3 * k + 1] = rgb[3 * k + 2] = 0.0f;
        if (alpha) alpha[k] = 0.0f;
        continue;
      }
      PixelSampler sampler = PixelSampler::make(s.render.seed, px, py, sample_index);
      Ray ray;
      camera_sample(s, sampler, px, py, ray);
      Rgb c;
      float a;
      radiance(s, hd->prep, sampler, ray, c, a, &st, nullptr, photons);
      rgb[3 * k] = c.r;
      rgb[3 * k + 1] = c.g;
      rgb[3 * k + 2] = c.b;
      if (alpha) alpha[k] = a;
    }
  if (rays) *rays = st.rays;
  return 0;
}

// Accumulates samples [first, first + count) of every pixel into the whole-image film arrays
// (rgb: 3*W*H, weight/alpha/lum2: W*H; caller-initialised), 32-pixel tiles over `threads` workers.
// stats (may be null): rays, nearest queries, any queries, nearest nodes, nearest packets,
// any nodes, any packets.
static int render_tiles(Handle* hd, const std::vector<TileRect>& tiles, uint32_t first, uint32_t count, int threads,
                        float* rgb, float* weight, float* alpha, float* lum2, uint64_t* stats_out);

int orc_render_film(void* p, uint32_t first, uint32_t count, int threads, float* rgb, float* weight, float* alpha,
                    float* lum2, uint64_t* stats_out) {
  auto* hd = static_cast<Handle*>(p);
  const Scene& s = hd->scene;
  return render_tiles(hd, tile_grid(s.render.width, s.render.height, 32), first, count, threads, rgb, weight, alpha,
                      lum2, stats_out);
}

// orc_render_film restricted to the listed tiles (x, y, w, h quadruples, inside the image).
int orc_render_tiles(void* p, const uint32_t* rects, size_t n, uint32_t first, uint32_t count, int threads, float* rgb,
                     float* weight, float* alpha, float* lum2, uint64_t* stats_out) {
  std::vector<TileRect> tiles(n);
  for (size_t i = 0; i < n; ++i) tiles[i] = TileRect{rects[4 * i], rects[4 * i + 1], rects[4 * i + 2], rects[4 * i + 3]};
  return render_tiles(static_cast<Handle*>(p), tiles, first, count, threads, rgb, weight, alpha, lum2, stats_out);
}

}  // extern "C"

static int render_tiles(Handle* hd, const std::vector<TileRect>& tiles, uint32_t first, uint32_t count, int threads,
                        float* rgb, float* weight, float* alpha, float* lum2, uint64_t* stats_out) {
  const Scene& s = hd->scene;
  uint32_t W = s.render.width;
  // The photon maps of the samples (one per sample index), traced before the tiles.
  std::vector<PhotonMap> maps;
  if (hd->prep.has_caustics)
    for (uint32_t k = 0; k < count; ++k) maps.push_back(trace_photon_map(s, hd->prep, first + k, threads));
  std::atomic<size_t> next{0};
  std::vector<RenderStats> per(std::max(threads, 1));
  auto worker = [&](int tid) {
    for (;;) {
      size_t t = next.fetch_add(1);
      if (t >= tiles.size()) return;
      TileRect tile = tiles[t];
      Film film;
      film.init(tile.x, tile.y, tile.w, tile.h);
      for (uint32_t j = 0; j < tile.h; ++j)
        for (uint32_t i = 0; i < tile.w; ++i) {
          size_t src = (size_t)(tile.y + j) * W + tile.x + i, dst = (size_t)j * tile.w + i;
          for (int c = 0; c < 3; ++c) film.rgb[3 * dst + c] = rgb[3 * src + c];
          film.weight[dst] = weight[src];
          film.alpha[dst] = alpha[src];
          film.lum2[dst] = lum2[src];
        }
      for (uint32_t k = 0; k < count; ++k)
        render_tile(s, hd->prep, tile, first + k, film, &per[tid], maps.empty() ? nullptr : &maps[k]);
      for (uint32_t j = 0; j < tile.h; ++j)
        for (uint32_t i = 0; i < tile.w; ++i) {
          size_t dst = (size_t)(tile.y + j) * W + tile.x + i, src = (size_t)j * tile.w + i;
          for (int c = 0; c < 3; ++c) rgb[3 * dst + c] = film.rgb[3 * src + c];
          weight[dst] = film.weight[src];
          alpha[dst] = film.alpha[src];
          lum2[dst] = film.lum2[src];
        }
    }
  };
  int n = std::max(threads, 1);
  std::vector<std::thread> pool;
  for (int t = 0; t < n; ++t) pool.emplace_back(worker, t);
  for (auto& t : pool) t.join();
  if (stats_out) {
    uint64_t acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (const RenderStats& r : per) {
      acc[0] += r.rays;
      acc[1] += r.nearest;
      acc[2] += r.any;
      acc[3] += r.nearest_trav.nodes;
      acc[4] += r.nearest_trav.packets;
      acc[5] += r.any_trav.nodes;
      acc[6] += r.any_trav.packets;
    }
    std::memcpy(stats_out, acc, sizeof(acc));
  }
  return 0;
}

extern "C" {

// Nearest and any hit of `n` rays given as 8 floats each (origin, direction, t_min, t_max).
// Outputs per ray: instance, triangle (0xffffffff on a miss), t, any (0/1). `brute` selects the
// brute-force loop instead of the tree.
int orc_trace_rays(void* p, const float* rays, size_t n, int brute, uint32_t* inst, uint32_t* tri, float* t,
                   uint8_t* any) {
  const Geometry& g = static_cast<Handle*>(p)->scene.geometry;
  for (size_t i = 0; i < n; ++i) {
    const float* r = rays + 8 * i;
    Ray ray(Vec3(r[0], r[1], r[2]), Vec3(r[3], r[4], r[5]));
    ray.t_min = r[6];
    ray.t_max = r[7];
    RayHit h;
    bool hit, a;
    if (brute) {
      hit = g.brute_force(ray, h, a);
    } else {
      hit = g.trace(ray, false, h);
      a = g.occluded(ray);
    }
    inst[i] = hit ? h.instance : 0xffffffffu;
    tri[i] = hit ? h.triangle : 0xffffffffu;
    t[i] = hit ? h.t : INF;
    any[i] = a ? 1 : 0;
  }
  return 0;
}

// Energy tables (regenerated by the restated recipe, or read from the cache file).
void orc_energy_tables(uint16_t* albedo, uint16_t* schlick, uint16_t* clear_in, uint16_t* clear_out,
                       float* albedo_avg, float* schlick_avg) {
  const energy::Tables& t = energy::tables();
  std::memcpy(albedo, t.albedo.data(), t.albedo.size() * 2);
  std::memcpy(schlick, t.schlick.data(), t.schlick.size() * 2);
  std::memcpy(clear_in, t.clear_in.data(), t.clear_in.size() * 2);
  std::memcpy(clear_out, t.clear_out.data(), t.clear_out.size() * 2);
  std::memcpy(albedo_avg, t.albedo_avg.data(), t.albedo_avg.size() * 4);
  std::memcpy(schlick_avg, t.schlick_avg.data(), t.schlick_avg.size() * 4);
}

// The BSDF of a material seen from one side, for its closed-form tests (bsdf.rs:384-992):
// `mat` = kind, base r, g, b, metallic, roughness, ior, transmission, coat, coat roughness;
// op 0 eval (out rgb; `all` 0 for the diffuse lobes only, 2 without the energy-compensation
// lobes), 1 pdf (out[0]), 2 sample (out: wi xyz, weight rgb, pdf; returns 0 when it gives none),
// 3 d eval / d (roughness, base r, g, b, ior) (out: 5 rgb triples, `all` as for op 0 but 0 or 1).
int orc_bsdf_probe(const float* mat, int entering, const float* wo, const float* wi, float u0, float u1, int op,
                   int all, float* out) {
  Material m;
  m.kind = (u32)mat[0];
  m.base_color = Rgb(mat[1], mat[2], mat[3]);
  m.metallic = mat[4];
  m.roughness = mat[5];
  m.ior = mat[6];
  m.transmission = mat[7];
  m.coat = mat[8];
  m.coat_roughness = mat[9];
  Vec3 o(wo[0], wo[1], wo[2]);
  if (op == 3) {
    Rgb d[5];
    bsdf_param_derivs_exact(m, entering != 0, o, Vec3(wi[0], wi[1], wi[2]), all ? Lobes::All : Lobes::Diffuse, d);
    for (int k = 0; k < 5; ++k) {
      out[3 * k] = d[k].r;
      out[3 * k + 1] = d[k].g;
      out[3 * k + 2] = d[k].b;
    }
    return 1;
  }
  Bsdf b = Bsdf::surface(m, entering != 0);
  if (op == 0) {
    if (all == 2) {
      // The single-scattering lobes alone: the energy-compensation lobes switched off.
      b.ms_spec_on = false;
      b.ms_coat_on = false;
    }
    Rgb f = b.eval(o, Vec3(wi[0], wi[1], wi[2]), all ? Lobes::All : Lobes::Diffuse);
    out[0] = f.r;
    out[1] = f.g;
    out[2] = f.b;
    return 1;
  }
  if (op == 1) {
    out[0] = b.pdf(o, Vec3(wi[0], wi[1], wi[2]));
    return 1;
  }
  BsdfSample s;
  if (!b.sample(o, u0, u1, s)) return 0;
  out[0] = s.wi.x;
  out[1] = s.wi.y;
  out[2] = s.wi.z;
  out[3] = s.weight.r;
  out[4] = s.weight.g;
  out[5] = s.weight.b;
  out[6] = s.pdf;
  return 1;
}
float orc_fresnel_dielectric(float cos_i, float eta) { return fresnel_dielectric(cos_i, eta); }

// The prepared lights of a scene, for the reference's selection tests (light.rs:1026-1141,
// 1708-1750): entries, whether the selection is per point, the static cdf and per-entry pmf.
uint32_t orc_lights_info(void* p, uint32_t* per_point, float* cdf, float* flux_pmf, uint32_t cap) {
  const Lights& l = static_cast<Handle*>(p)->prep.lights;
  *per_point = l.per_point ? 1u : 0u;
  for (size_t k = 0; k < l.cdf.size() && k < cap; ++k) {
    cdf[k] = l.cdf[k];
    flux_pmf[k] = l.flux_pmf[k];
  }
  return (uint32_t)l.entries.size();
}

// The probability with which next-event estimation at `point` picks `entry` (Lights::pmf of
// Lights::select there).
float orc_lights_pmf(void* p, const float* point, uint32_t entry) {
  const Lights& l = static_cast<Handle*>(p)->prep.lights;
  Selection s = l.select(Vec3(point[0], point[1], point[2]));
  return l.pmf(entry, s);
}

// pick() on the static cdf: the entry and u rescaled inside its bucket.
void orc_lights_pick(void* p, float u, uint32_t* index, float* rest) {
  const Lights& l = static_cast<Handle*>(p)->prep.lights;
  size_t i = 0;
  l.pick_static(u, i, *rest);
  *index = (uint32_t)i;
}

// pick_at() on the selection at `point`: entry, rescaled u and probability.
void orc_lights_pick_at(void* p, const float* point, float u, uint32_t* index, float* rest, float* prob) {
  const Lights& l = static_cast<Handle*>(p)->prep.lights;
  Selection s = l.select(Vec3(point[0], point[1], point[2]));
  size_t i = 0;
  l.pick_at(s, u, i, *rest, *prob);
  *index = (uint32_t)i;
}

// One light sample at `point` with the selection there (Lights::sample, light.rs:1145-1315):
// returns 0 when it gives none; out: wi xyz, distance, radiance rgb, pdf, delta, glossy, the entry
// picked, its kind (0 area, 1 sun, 2 sphere, 3 mesh, 4 environment).
int orc_lights_sample(void* p, const float* point, float u_select, float u0, float u1, float* out) {
  const Lights& l = static_cast<Handle*>(p)->prep.lights;
  Vec3 x(point[0], point[1], point[2]);
  Selection sel = l.select(x);
  LightSample s;
  if (l.entries.empty() || !l.sample_selected(sel, u_select, u0, u1, s)) return 0;
  size_t index = 0;
  float rest, prob;
  if (sel.by_flux) l.pick_static(u_select, index, rest);
  else l.pick_at(sel, u_select, index, rest, prob);
  float v[12] = {s.wi.x, s.wi.y, s.wi.z, s.distance, s.radiance.r, s.radiance.g, s.radiance.b, s.pdf,
                 s.delta ? 1.0f : 0.0f, s.glossy ? 1.0f : 0.0f, (float)index, (float)l.entries[index].kind};
  std::memcpy(out, v, sizeof(v));
  return 1;
}

// Radiance a camera ray leaving the scene along `dir` sees of the sky (Lights::escaped without a
// BSDF: the environment and the suns flagged camera, light.rs:1502-1520).
void orc_lights_escaped(void* p, const float* dir, float* out3) {
  const Lights& l = static_cast<Handle*>(p)->prep.lights;
  Rgb r = l.escaped(Vec3(dir[0], dir[1], dir[2]), false, 0.0f, nullptr);
  out3[0] = r.r;
  out3[1] = r.g;
  out3[2] = r.b;
}

// Lights::emitter_weight of `instance` for a BSDF-sampled ray of density bsdf_pdf that left
// `point` (its selection there) and met the instance at distance t with `cos` (light.rs:1537-1551).
float orc_lights_emitter_weight(void* p, const float* point, uint32_t instance, float bsdf_pdf, float t, float cos) {
  const Lights& l = static_cast<Handle*>(p)->prep.lights;
  Selection sel = l.select(Vec3(point[0], point[1], point[2]));
  return l.emitter_weight(instance, bsdf_pdf, sel, t, cos);
}

// ---------------------------------------------------------------- caustics (caustics.rs)
// has caustics, photons per sample, first radius (mm), sources, caster spheres, most casters a
// photon crosses.
void orc_caustics_info(void* p, float* out6) {
  const Prepared& pr = static_cast<Handle*>(p)->prep;
  const CausticSetup& c = pr.caustics;
  out6[0] = pr.has_caustics ? 1.0f : 0.0f;
  out6[1] = (float)c.photons;
  out6[2] = c.radius;
  out6[3] = (float)c.sources.size();
  out6[4] = (float)c.sphere_c.size();
  out6[5] = (float)c.max_bounces;
}
float orc_caustics_radius(void* p, uint32_t iteration) {
  return static_cast<Handle*>(p)->prep.caustics.radius_of(iteration);
}

// The stored photons of photons [first, last) of pass `iteration`, 13 floats each (position, wi,
// power, normal, casters crossed), up to `cap`; returns how many were stored (even beyond cap).
uint32_t orc_photons(void* p, uint32_t iteration, uint32_t first, uint32_t last, float* out, uint32_t cap,
                     uint64_t* rays) {
  auto* hd = static_cast<Handle*>(p);
  std::vector<Photon> ph;
  u64 r = 0;
  if (hd->prep.has_caustics) trace_photons(hd->scene, hd->prep, hd->prep.caustics, iteration, first, last, ph, r);
  for (size_t i = 0; i < ph.size() && i < cap; ++i) {
    const Photon& q = ph[i];
    float v[13] = {q.position.x, q.position.y, q.position.z, q.wi.x, q.wi.y, q.wi.z, q.power.r, q.power.g,
                   q.power.b, q.normal.x, q.normal.y, q.normal.z, (float)q.bounces};
    std::memcpy(out + 13 * i, v, sizeof(v));
  }
  if (rays) *rays = r;
  return (uint32_t)ph.size();
}

// The map of pass `iteration` built from n given photons (13 floats each), gathered with a BSDF
// of one at nq queries (point xyz, normal xyz, most casters): rgb per query.
void orc_photon_gather(void* p, uint32_t iteration, const float* photons, uint32_t n, const float* queries, uint32_t nq,
                       float* out3) {
  auto* hd = static_cast<Handle*>(p);
  std::vector<Photon> ph(n);
  for (uint32_t i = 0; i < n; ++i) {
    const float* v = photons + 13 * (size_t)i;
    ph[i] = Photon{Vec3(v[0], v[1], v[2]), Vec3(v[3], v[4], v[5]), Rgb(v[6], v[7], v[8]), Vec3(v[9], v[10], v[11]),
                   (u32)v[12]};
  }
  PhotonMap m = build_photon_map(hd->prep.caustics, iteration, std::move(ph), 0);
  for (uint32_t k = 0; k < nq; ++k) {
    const float* q = queries + 7 * (size_t)k;
    Rgb c = m.gather(Vec3(q[0], q[1], q[2]), Vec3(q[3], q[4], q[5]), (u32)q[6], [](Vec3) { return Rgb(1, 1, 1); });
    out3[3 * k] = c.r;
    out3[3 * k + 1] = c.g;
    out3[3 * k + 2] = c.b;
  }
}

// caustics.rs:104-119 for a material given as orc_bsdf_probe's `mat` (kind, base rgb, metallic,
// roughness, ior, transmission, coat, coat roughness).
int orc_is_caster(const float* mat) {
  Material m;
  m.kind = (u32)mat[0];
  m.base_color = Rgb(mat[1], mat[2], mat[3]);
  m.metallic = mat[4];
  m.roughness = mat[5];
  m.ior = mat[6];
  m.transmission = mat[7];
  m.coat = mat[8];
  m.coat_roughness = mat[9];
  return is_caster(m) ? 1 : 0;
}

// What the BSDF side gives a light that next-event estimation sampled (light.rs:2396-2470): for
// each of n tuples (point xyz, u_select, u0, u1), a light sample at the point and, when it is not
// a delta, the MIS weight of the same light found by a BSDF-sampled ray along the sample's
// direction whose density is the sample's (crossed for areas and spheres crossed alone, the sky
// beyond the environment for a sun, the environment's own, emitter_weight for the mesh of
// instance 0). out per tuple: the entry's kind (-1: none, a delta or crossing several lights) and
// the weight, which the power heuristic makes 1/2.
void orc_lights_mis_probe(void* p, const float* tuples, uint32_t n, float* out) {
  const Lights& l = static_cast<Handle*>(p)->prep.lights;
  for (uint32_t k = 0; k < n; ++k) {
    const float* t = tuples + 6 * (size_t)k;
    out[2 * k] = -1.0f;
    out[2 * k + 1] = 0.0f;
    Vec3 x(t[0], t[1], t[2]);
    Selection sel = l.select(x);
    LightSample s;
    if (l.entries.empty() || !l.sample_selected(sel, t[3], t[4], t[5], s) || s.delta) continue;
    size_t index = 0;
    float rest, prob;
    if (sel.by_flux) l.pick_static(t[3], index, rest);
    else l.pick_at(sel, t[3], index, rest, prob);
    Ray ray(x, s.wi);
    int crossings = 0;
    for (const AreaLight& a : l.areas) {
      float th, c;
      if (a.intersect(ray, INF, th, c)) ++crossings;
    }
    for (const SphereLight& sp : l.spheres) {
      float th;
      if (sp.intersect(ray, INF, th)) ++crossings;
    }
    const LightRef ref = l.entries[index];
    Rgb env = Rgb::splat(l.env.value(s.wi.z));
    Rgb env_part = l.has_env_entry ? env * l.found_weight(s.pdf, l.pmf(l.env_entry, sel) * l.env.pdf(s.wi)) : env;
    float w;
    if ((ref.kind == REF_AREA || ref.kind == REF_SPHERE) && crossings != 1) continue;
    if (ref.kind == REF_AREA) {
      const AreaLight& a = l.areas[ref.index];
      float cos = -s.wi.dot(a.frame.z);
      w = l.crossed_from(ray, INF, true, s.pdf, &sel, 0.0f).g / a.radiance_at(cos).g;
    } else if (ref.kind == REF_SPHERE) {
      const SphereLight& sp = l.spheres[ref.index];
      w = l.crossed_from(ray, INF, true, s.pdf, &sel, 0.0f).g / (sp.radiance.g * sp.mask(x));
    } else if (ref.kind == REF_SUN) {
      w = (l.escaped(s.wi, true, s.pdf, &sel).g - env_part.g) / l.suns[ref.index].radiance.g;
    } else if (ref.kind == REF_ENV) {
      w = env_part.g / env.g;
    } else {
      const MeshEmitter& m = l.meshes[ref.index];
      w = l.emitter_weight(0, s.pdf, sel, s.distance, s.wi.dot(m.normals[0]));
    }
    out[2 * k] = (float)ref.kind;
    out[2 * k + 1] = w;
  }
}

// The reference tests' switch (light.rs:550-558, integrator/tests.rs:2574-2583): techniques
// 0 MIS, 1 lights only, 2 BSDF only; per_point 0 or 1 sets the per-point selection, other values
// keep it.
void orc_set_techniques(void* p, int techniques, int per_point) {
  Lights& l = static_cast<Handle*>(p)->prep.lights;
  l.techniques = techniques;
  if (per_point == 0 || per_point == 1) l.per_point = per_point == 1 && !l.entries.empty();
}

// Scalar numerics for the golden-vector tests: 0 exp2, 1 log2, 2 powf, 3 sin, 4 cos, 5 atan2,
// 6 acos.
float orc_fmath(int which, float x, float y) {
  switch (which) {
    case 0: return exp2f_(x);
    case 1: return log2f_(x);
    case 2: return powf_(x, y);
    case 3: return fsin(x);
    case 4: return fcos(x);
    case 5: return atan2f_(x, y);
    default: return acosf_(x);
  }
}

void orc_pcg32(uint64_t seed, uint64_t stream, uint32_t* out, size_t n) {
  Pcg32 r = Pcg32::seeded(seed, stream);
  for (size_t i = 0; i < n; ++i) out[i] = r.next_u32();
}
uint64_t orc_mix64(uint64_t z) { return mix64(z); }
// FNV-1a over bytes (render.rs:981-989 hashes the little-endian bytes of f32 planes this way).
uint32_t orc_fnv1a(const uint8_t* bytes, size_t n) {
  uint32_t h = 0x811c9dc5u;
  for (size_t i = 0; i < n; ++i) h = (h ^ bytes[i]) * 0x01000193u;
  return h;
}
uint32_t orc_sobol(uint32_t dim, uint32_t index) { return sobol(dim, index); }
uint32_t orc_owen(uint32_t x, uint32_t seed) { return owen_scramble(x, seed); }
float orc_sampler_get1(uint32_t seed, uint32_t x, uint32_t y, uint32_t index, uint32_t dim) {
  return PixelSampler::make(seed, x, y, index).get1(dim);
}

// ---------------------------------------------------------------- differentiable-render extension
// Parameters per material: roughness, base r, g, b, IOR. Returns the number of materials.
uint32_t orc_material_params(void* p, float* out, uint32_t cap) {
  const Scene& s = static_cast<Handle*>(p)->scene;
  for (size_t m = 0; m < s.materials.size() && out && (m + 1) * 5 <= cap; ++m) {
    const Material& mm = s.materials[m];
    float* o = out + 5 * m;
    o[0] = mm.roughness;
    o[1] = mm.base_color.r;
    o[2] = mm.base_color.g;
    o[3] = mm.base_color.b;
    o[4] = mm.ior;
  }
  return (uint32_t)s.materials.size();
}

int orc_set_material_params(void* p, const float* params, uint32_t n) {
  Scene& s = static_cast<Handle*>(p)->scene;
  if (n != s.materials.size()) return 1;
  for (uint32_t m = 0; m < n; ++m) {
    Material& mm = s.materials[m];
    mm.roughness = params[5 * m];
    mm.base_color = Rgb(params[5 * m + 1], params[5 * m + 2], params[5 * m + 3]);
    mm.ior = params[5 * m + 4];
  }
  // Media absorb by the base colour (integrator/mod.rs:300-311).
  std::string err;
  return static_cast<Handle*>(p)->prep.build(s, err) ? 0 : 2;
}

int orc_set_texture(void* p, uint32_t material, uint32_t w, uint32_t h, const float* rgb, float ox, float oy, float sx,
                    float sy) {
  Scene& s = static_cast<Handle*>(p)->scene;
  s.texture = Texture{};
  if (material == 0xffffffffu) return 0;
  if (material >= s.materials.size() || w == 0 || h == 0) return 1;
  s.texture.material = material;
  s.texture.w = w;
  s.texture.h = h;
  s.texture.ox = ox;
  s.texture.oy = oy;
  s.texture.sx = sx;
  s.texture.sy = sy;
  s.texture.rgb.assign(rgb, rgb + (size_t)w * h * 3);
  return 0;
}

// Path replay backpropagation on the CPU (the GPU adjoint's reference): for samples
// [first, first + count) of every pixel, the forward radiance, then the replay adding
// d loss / d parameter (d_image: d loss / d image, 3*W*H; the image is the samples' mean) into
// grad_mat (5 per material) and grad_tex (3 per texel), summed in double.
int orc_adjoint(void* p, uint32_t first, uint32_t count, int threads, const float* d_image, double* grad_mat,
                double* grad_tex) {
  auto* hd = static_cast<Handle*>(p);
  const Scene& s = hd->scene;
  u32 W = s.render.width, H = s.render.height;
  size_t nm = s.materials.size() * 5, nt = (size_t)s.texture.w * s.texture.h * 3;
  int n = std::max(threads, 1);
  std::vector<std::vector<double>> gm(n, std::vector<double>(nm, 0.0)), gt(n, std::vector<double>(nt, 0.0));
  std::atomic<u32> next_row{0};
  float scale = 1.0f / (float)count;
  auto worker = [&](int tid) {
    for (;;) {
      u32 y = next_row.fetch_add(1);
      if (y >= H) return;
      for (u32 x = 0; x < W; ++x) {
        const float* d = d_image + 3 * ((size_t)y * W + x);
        for (u32 k = 0; k < count; ++k) {
          PixelSampler sampler = PixelSampler::make(s.render.seed, x, y, first + k);
          Ray ray;
          camera_sample(s, sampler, x, y, ray);
          Rgb total;
          float a;
          radiance(s, hd->prep, sampler, ray, total, a, nullptr);
          AdjointPath ap;
          ap.adj = Rgb(d[0] * scale, d[1] * scale, d[2] * scale);
          ap.total = total;
          ap.grad_mat = gm[tid].data();
          ap.grad_tex = gt[tid].empty() ? nullptr : gt[tid].data();
          Rgb again;
          radiance(s, hd->prep, sampler, ray, again, a, nullptr, &ap);
        }
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < n; ++t) pool.emplace_back(worker, t);
  for (auto& t : pool) t.join();
  for (size_t i = 0; i < nm; ++i) {
    double acc = 0.0;
    for (int t = 0; t < n; ++t) acc += gm[t][i];
    grad_mat[i] = acc;
  }
  for (size_t i = 0; i < nt && grad_tex; ++i) {
    double acc = 0.0;
    for (int t = 0; t < n; ++t) acc += gt[t][i];
    grad_tex[i] = acc;
  }
  return 0;
}

// The convergence reference: samples [first, first + count) of every pixel with the radiance summed
// in double (rgb: 3*W*H, alpha: W*H, caller-zeroed), rows over `threads` workers.
int orc_render_film_f64(void* p, uint32_t first, uint32_t count, int threads, double* rgb, double* alpha) {
  auto* hd = static_cast<Handle*>(p);
  const Scene& s = hd->scene;
  u32 W = s.render.width, H = s.render.height;
  // Scenes with caustics go sample by sample (each sample index has its photon map); every
  // pixel's samples are summed in index order either way.
  bool by_sample = hd->prep.has_caustics;
  std::vector<double> acc(by_sample ? (size_t)W * H * 4 : 0, 0.0);
  PhotonMap map;
  u32 k0 = 0, k1 = count;
  auto worker = [&](std::atomic<u32>* next_row) {
    for (;;) {
      u32 y = next_row->fetch_add(1);
      if (y >= H) return;
      for (u32 x = 0; x < W; ++x) {
        size_t i = (size_t)y * W + x;
        double r = 0, g = 0, b = 0, a = 0;
        if (by_sample) {
          r = acc[4 * i];
          g = acc[4 * i + 1];
          b = acc[4 * i + 2];
          a = acc[4 * i + 3];
        }
        for (u32 k = k0; k < k1; ++k) {
          PixelSampler sampler = PixelSampler::make(s.render.seed, x, y, first + k);
          Ray ray;
          camera_sample(s, sampler, x, y, ray);
          Rgb c;
          float al;
          radiance(s, hd->prep, sampler, ray, c, al, nullptr, nullptr, by_sample ? &map : nullptr);
          r += c.r;
          g += c.g;
          b += c.b;
          a += al;
        }
        if (by_sample && k1 < count) {
          acc[4 * i] = r;
          acc[4 * i + 1] = g;
          acc[4 * i + 2] = b;
          acc[4 * i + 3] = a;
          continue;
        }
        rgb[3 * i] += r;
        rgb[3 * i + 1] += g;
        rgb[3 * i + 2] += b;
        alpha[i] += a;
      }
    }
  };
  auto run = [&]() {
    std::atomic<u32> next_row{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < std::max(threads, 1); ++t) pool.emplace_back(worker, &next_row);
    for (auto& t : pool) t.join();
  };
  if (!by_sample) {
    run();
    return 0;
  }
  for (u32 k = 0; k < count; ++k) {
    map = trace_photon_map(s, hd->prep, first + k, threads);
    k0 = k;
    k1 = k + 1;
    run();
  }
  return 0;
}

// ---------------------------------------------------------------- adaptive sampling (adaptive.rs)
void* orc_adaptive_new(uint32_t width, uint32_t height, float target, float ev, uint32_t cap) {
  return new adaptive::Adaptive(width, height, target, ev, cap);
}
void orc_adaptive_free(void* a) { delete static_cast<adaptive::Adaptive*>(a); }
uint32_t orc_adaptive_left(void* a) { return static_cast<adaptive::Adaptive*>(a)->left(); }
void orc_adaptive_advance(void* a, uint32_t n) { static_cast<adaptive::Adaptive*>(a)->advance(n); }
uint32_t orc_adaptive_samples_of_pixel(void* a, uint32_t x, uint32_t y) {
  auto* ad = static_cast<adaptive::Adaptive*>(a);
  return ad->samples_of(ad->block_of(x, y));
}
uint32_t orc_adaptive_blocks(void* a) { return static_cast<adaptive::Adaptive*>(a)->blocks(); }
// The round's decision from whole-image planes: out6 = round, active, blocks (as floats), threshold,
// max error, remaining; active_out one byte per block.
void orc_adaptive_decide(void* a, const float* sums, uint32_t stride, const float* lum2, float* out6,
                         uint8_t* active_out) {
  auto* ad = static_cast<adaptive::Adaptive*>(a);
  adaptive::Decision d = ad->decide(sums, stride, lum2);
  out6[0] = (float)d.round;
  out6[1] = (float)d.active;
  out6[2] = (float)d.blocks;
  out6[3] = d.threshold;
  out6[4] = d.max_error;
  out6[5] = d.remaining;
  std::memcpy(active_out, ad->active().data(), ad->active().size());
}
float orc_adaptive_pixel_error(float r, float g, float b, float lum2, uint32_t n, float floor) {
  return adaptive::pixel_error(r, g, b, lum2, n, floor);
}

// The scene rendered adaptively to the cap (0: the scene's max_samples) with the host loop of the
// reference's tests (adaptive/tests.rs:294-356): whole rounds, the first with the plain renderer,
// later ones block by block, each pixel's samples in index order. Film arrays as orc_render_film
// (caller-zeroed); block_counts gets each block's samples; returns the number of decisions.
uint32_t orc_render_adaptive(void* p, uint32_t cap, int threads, float* rgb, float* weight, float* alpha, float* lum2,
                             uint32_t* block_counts) {
  auto* hd = static_cast<Handle*>(p);
  const Scene& s = hd->scene;
  u32 W = s.render.width, H = s.render.height;
  if (cap == 0) cap = s.render.max_samples;
  adaptive::Adaptive ad(W, H, s.render.noise_threshold, s.render.exposure_ev, cap);
  Film film;
  film.init(0, 0, W, H);
  u32 decisions = 0;
  int n = std::max(threads, 1);
  for (;;) {
    if (ad.left() == 0) {
      ad.decide(film.rgb.data(), 3, film.lum2.data());
      ++decisions;
      if (ad.converged()) break;
    }
    u32 count = ad.left();
    // The round's blocks (all of them in the first), each its quota from its own count.
    std::vector<std::pair<u32, u32>> work;  // (block, samples)
    for (u32 b = 0; b < ad.blocks(); ++b) {
      u32 q = ad.quota(b);
      if (q > 0) work.push_back({b, std::min(q, count)});
    }
    std::atomic<size_t> next{0};
    auto worker = [&]() {
      for (;;) {
        size_t k = next.fetch_add(1);
        if (k >= work.size()) return;
        u32 b = work[k].first, bx = b % ad.cols(), by = b / ad.cols();
        TileRect t{bx * adaptive::BLOCK, by * adaptive::BLOCK, std::min(adaptive::BLOCK, W - bx * adaptive::BLOCK),
                   std::min(adaptive::BLOCK, H - by * adaptive::BLOCK)};
        u32 first = ad.counts()[b];
        for (u32 j = 0; j < work[k].second; ++j) render_tile(s, hd->prep, t, first + j, film, nullptr);
      }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < n; ++t) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
    ad.advance(count);
  }
  std::memcpy(rgb, film.rgb.data(), film.rgb.size() * 4);
  std::memcpy(weight, film.weight.data(), film.weight.size() * 4);
  std::memcpy(alpha, film.alpha.data(), film.alpha.size() * 4);
  std::memcpy(lum2, film.lum2.data(), film.lum2.size() * 4);
  std::memcpy(block_counts, ad.counts().data(), ad.counts().size() * 4);
  return decisions;
}

// The passes of the whole image as the reference's tests assemble them
// (rt-core/src/render/passes_tests.rs:18-37): per 32-pixel tile, the exact passes from the centre
// rays, then the albedo and alpha of samples [first, first + count) in index order, resolved.
// Planes (caller-sized, any may be null): albedo 3N, normal 3N, depth N, part N, material N, coc N,
// alpha N. `threads` workers claim tiles from one counter. Returns the rays cast.
uint64_t orc_render_passes(void* p, uint32_t first, uint32_t count, int threads, float* albedo, float* normal,
                           float* depth, uint32_t* part, uint32_t* material, float* coc, float* alpha) {
  auto* hd = static_cast<Handle*>(p);
  const Scene& s = hd->scene;
  u32 W = s.render.width;
  std::vector<TileRect> tiles = tile_grid(s.render.width, s.render.height, 32);
  std::atomic<size_t> next{0};
  std::atomic<uint64_t> rays{0};
  auto worker = [&]() {
    for (;;) {
      size_t t = next.fetch_add(1);
      if (t >= tiles.size()) return;
      TileRect tile = tiles[t];
      PassFilm film;
      film.init(tile.x, tile.y, tile.w, tile.h, PASS_ALL);
      uint64_t r = render_tile_exact(s, hd->prep, tile, film);
      for (u32 k = 0; k < count; ++k) r += render_tile_albedo(s, hd->prep, tile, first + k, film);
      Passes ps = film.resolve();
      for (u32 j = 0; j < tile.h; ++j)
        for (u32 i = 0; i < tile.w; ++i) {
          size_t src = (size_t)j * tile.w + i, dst = (size_t)(tile.y + j) * W + tile.x + i;
          for (int c = 0; c < 3; ++c) {
            if (albedo) albedo[3 * dst + c] = ps.albedo[3 * src + c];
            if (normal) normal[3 * dst + c] = ps.normal[3 * src + c];
          }
          if (depth) depth[dst] = ps.depth[src];
          if (part) part[dst] = ps.part_id[src];
          if (material) material[dst] = ps.material_id[src];
          if (coc) coc[dst] = ps.coc[src];
          if (alpha) alpha[dst] = ps.alpha[src];
        }
      rays += r;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < std::max(threads, 1); ++t) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
  return rays.load();
}

// The denoiser (rt-post/src/denoise.rs:386-407, denoise_rows_at) on a band of `rows` rows from
// `origin_y` of a width x height image: rows y0..y1 of RGBA sums into `out` (width * (y1 - y0) * 4
// floats). `levels` and `radius` give the preset (clamped as the reference clamps); `counts` may be
// null. Returns 0, 1 (shape) or 2 (rows).
int orc_denoise(uint32_t width, uint32_t height, uint32_t origin_y, uint32_t rows, uint32_t samples, uint32_t y0,
                uint32_t y1, uint32_t levels, uint32_t radius, float strength, const float* sums, const float* lum2,
                const float* albedo, const float* normal, const float* depth, const uint32_t* part,
                const uint32_t* material, const float* counts, float* out) {
  denoise::Input in;
  in.width = width;
  in.height = height;
  in.origin_y = origin_y;
  in.rows = rows;
  in.samples = samples;
  in.counts = counts;
  in.sums = sums;
  in.lum2 = lum2;
  in.albedo = albedo;
  in.normal = normal;
  in.depth = depth;
  in.part_id = part;
  in.material_id = material;
  std::vector<float> v;
  denoise::Status st = denoise::denoise_rows_at(in, y0, y1, denoise::Preset{levels, radius}, strength, v);
  if (st != denoise::OK) return (int)st;
  std::memcpy(out, v.data(), v.size() * 4);
  return 0;
}

// The Henyey-Greenstein phase function and its sampling (integrator/medium.rs:143-169), for the
// tests of its density.
float orc_phase(float cos, float g);
float orc_sample_phase(const float* dir, float g, float u0, float u1, float* out) {
  float density;
  Vec3 d = sample_phase_for_tests(Vec3(dir[0], dir[1], dir[2]), g, u0, u1, density);
  out[0] = d.x;
  out[1] = d.y;
  out[2] = d.z;
  return density;
}
float orc_phase(float cos, float g) { return phase_for_tests(cos, g); }

// The reference tests' synthetic film (rt-post/src/denoise/tests.rs:22-73): the top row
// background (a gradient), the left half part 1 around `left`, the right half part 2 around
// `right`, exponential-like noise of relative `spread` from Pcg32(7, 11). Planes as orc_denoise.
void orc_denoise_synthetic(uint32_t width, uint32_t height, uint32_t samples, float left, float right, float spread,
                           float* sums, float* lum2, float* albedo, float* normal, float* depth, uint32_t* part,
                           uint32_t* material) {
  size_t n = (size_t)width * height;
  std::memset(sums, 0, n * 16);
  std::memset(lum2, 0, n * 4);
  std::memset(albedo, 0, n * 12);
  std::memset(normal, 0, n * 12);
  std::memset(part, 0, n * 4);
  std::memset(material, 0, n * 4);
  for (size_t p = 0; p < n; ++p) depth[p] = INF;
  Pcg32 rng = Pcg32::seeded(7, 11);
  for (u32 y = 0; y < height; ++y)
    for (u32 x = 0; x < width; ++x) {
      size_t p = (size_t)y * width + x;
      if (y == 0) {
        for (int k = 0; k < 3; ++k) sums[p * 4 + k] = ((float)x * 0.01f + (float)k) * (float)samples;
        sums[p * 4 + 3] = (float)samples;
        continue;
      }
      bool left_side = x < width / 2;
      u32 id = left_side ? 1 : 2;
      float base = left_side ? left : right;
      part[p] = id;
      material[p] = id;
      for (int k = 0; k < 3; ++k) albedo[p * 3 + k] = 0.8f;
      normal[p * 3 + 2] = 1.0f;
      depth[p] = 500.0f;
      for (u32 s = 0; s < samples; ++s) {
        float u = fmax_r(rng.next_f32(), 1.0e-6f);
        float v = base * fmax_r(1.0f + spread * (-lnf_(u) - 1.0f), 0.0f);
        for (int k = 0; k < 3; ++k) sums[p * 4 + k] += v;
        sums[p * 4 + 3] += 1.0f;
        float l = 0.2126f * v + 0.7152f * v + 0.0722f * v;
        lum2[p] += l * l;
      }
    }
}

}  // extern "C"
