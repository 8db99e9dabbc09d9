// ORACLE — test infrastructure only (see rt_core.hpp). The exact passes; reference lines per
// function.
#include "rt_passes.hpp"

namespace orc {

// integrator/mod.rs:911-990 (trace_camera is trace here: no particles), light.rs:1479-1488
FirstHit first_hit(const Scene& scene, const Prepared& prep, const Ray& ray) {
  FirstHit out;
  RayHit hit;
  bool has_hit = scene.geometry.trace(ray, true, hit);
  float t_geometry = has_hit ? hit.t : INF;
  float t_floor = prep.floor_t(ray, true);
  float t_nearest = fmin_r(t_geometry, t_floor);
  if (t_nearest == INF || prep.lights.camera_sees_panel(ray, t_nearest)) return out;
  if (t_floor < t_geometry) {
    out.kind = FIRST_FLOOR;
    out.t = t_floor;
    out.normal = Vec3(0.0f, 0.0f, ray.dir.z < 0.0f ? 1.0f : -1.0f);
    return out;
  }
  Hit details;
  if (!scene.geometry.hit_details(ray, hit, details)) return out;
  const Material& material = prep.material(scene, details.instance);
  Rgb albedo = material.base_color;
  u32 mat_id = details.instance < scene.geometry.instances.size()
                   ? scene.geometry.instances[details.instance].material_id
                   : 0xffffffffu;
  if (scene.texture.on() && mat_id == scene.texture.material && mat_id < scene.materials.size())
    albedo = scene.texture.value(details.point.x, details.point.y);
  out.kind = FIRST_SURFACE;
  out.instance = details.instance;
  out.t = details.t;
  out.normal = details.normal;  // bevel.rs:372-377 without a bevel: the hit's shading normal
  out.albedo = albedo;
  return out;
}

// passes.rs:144-153
float coc_diameter(const Camera& camera, float depth, u32 long_side) {
  if (camera.lens_radius() <= 0.0f || !(depth > 0.0f && is_finite(depth))) return 0.0f;
  float aperture = camera.focal_mm / camera.fstop;
  float on_focus_plane = aperture * std::fabs(depth - camera.focus_dist) / depth;
  float pixel_on_focus_plane =
      camera.sensor_mm / (float)std::max<u32>(long_side, 1) * (camera.focus_dist / camera.focal_mm);
  return on_focus_plane / pixel_on_focus_plane;
}

// passes.rs:177-206
Exact exact_of(const Scene& scene, const Ray& ray, const FirstHit& hit) {
  Exact e;
  float t;
  Vec3 normal;
  u32 part, material;
  if (hit.kind == FIRST_BACKGROUND) return e;
  if (hit.kind == FIRST_FLOOR) {
    t = hit.t;
    normal = hit.normal;
    part = FLOOR_ID;
    material = FLOOR_ID;
  } else {
    u32 m = hit.instance < scene.geometry.instances.size() ? scene.geometry.instances[hit.instance].material_id
                                                           : 0xffffffffu;
    material = m < scene.materials.size() ? m + 1 : 0;
    t = hit.t;
    normal = hit.normal;
    part = hit.instance + 1;
  }
  Vec3 forward, right, up;
  scene.camera.basis(forward, right, up);
  float depth = t * ray.dir.dot(forward) + (ray.origin - scene.camera.position).dot(forward);
  u32 long_side = std::max(scene.render.width, scene.render.height);
  e.normal = normal;
  e.depth = depth;
  e.part_id = part;
  e.material_id = material;
  e.coc = coc_diameter(scene.camera, depth, long_side);
  return e;
}

// render.rs:500-509
Ray centre_ray(const Scene& scene, u32 x, u32 y) {
  u32 width = std::max<u32>(scene.render.width, 1), height = std::max<u32>(scene.render.height, 1);
  float aspect = (float)width / (float)height;
  float u = ((float)x + 0.5f) / (float)width;
  float v = ((float)y + 0.5f) / (float)height;
  return scene.camera.generate_ray(aspect, u, v, 0.0f, 0.0f);
}

// passes.rs:252-330
void PassFilm::init(u32 ox, u32 oy, u32 w, u32 h, u32 m) {
  origin_x = ox;
  origin_y = oy;
  width = w;
  height = h;
  mask = m & PASS_ALL;
  size_t n = (size_t)w * h;
  auto has = [&](u32 bit) { return (mask & bit) != 0; };
  albedo.assign(has(PASS_ALBEDO) ? n * 3 : 0, 0.0f);
  albedo_weight.assign(has(PASS_ALBEDO) ? n : 0, 0.0f);
  normal.assign(has(PASS_NORMAL) ? n * 3 : 0, 0.0f);
  depth.assign(has(PASS_DEPTH) ? n : 0, INF);
  part_id.assign(has(PASS_PART_ID) ? n : 0, 0u);
  material_id.assign(has(PASS_MATERIAL_ID) ? n : 0, 0u);
  coc.assign(has(PASS_COC) ? n : 0, 0.0f);
  alpha.assign(has(PASS_ALPHA) ? n : 0, 0.0f);
  alpha_weight.assign(has(PASS_ALPHA) ? n : 0, 0.0f);
}

bool PassFilm::index(u32 x, u32 y, size_t& i) const {
  if (x < origin_x || y < origin_y) return false;
  u32 lx = x - origin_x, ly = y - origin_y;
  if (lx >= width || ly >= height) return false;
  i = (size_t)ly * width + lx;
  return true;
}

// passes.rs:333-361
void PassFilm::add_albedo(u32 x, u32 y, Rgb c) {
  size_t i;
  if (!(mask & PASS_ALBEDO) || !index(x, y, i)) return;
  albedo[i * 3] += c.r;
  albedo[i * 3 + 1] += c.g;
  albedo[i * 3 + 2] += c.b;
  albedo_weight[i] += 1.0f;
}
void PassFilm::add_alpha(u32 x, u32 y, float a) {
  size_t i;
  if (!(mask & PASS_ALPHA) || !index(x, y, i)) return;
  alpha[i] += a;
  alpha_weight[i] += 1.0f;
}

// passes.rs:364-386
void PassFilm::set_exact(u32 x, u32 y, const Exact& e) {
  size_t i;
  if (!index(x, y, i)) return;
  if (mask & PASS_NORMAL) {
    normal[i * 3] = e.normal.x;
    normal[i * 3 + 1] = e.normal.y;
    normal[i * 3 + 2] = e.normal.z;
  }
  if (mask & PASS_DEPTH) depth[i] = e.depth;
  if (mask & PASS_PART_ID) part_id[i] = e.part_id;
  if (mask & PASS_MATERIAL_ID) material_id[i] = e.material_id;
  if (mask & PASS_COC) coc[i] = e.coc;
}

// passes.rs:388-420: the sums times the f32 reciprocal of their counts
Passes PassFilm::resolve() const {
  Passes p;
  p.width = width;
  p.height = height;
  p.mask = mask;
  p.albedo.assign(albedo.size(), 0.0f);
  for (size_t i = 0; i < albedo_weight.size(); ++i) {
    float w = albedo_weight[i];
    if (w > 0.0f) {
      float inv = 1.0f / w;
      for (int c = 0; c < 3; ++c) p.albedo[i * 3 + c] = albedo[i * 3 + c] * inv;
    }
  }
  p.alpha.assign(alpha.size(), 0.0f);
  for (size_t i = 0; i < alpha_weight.size(); ++i) {
    float w = alpha_weight[i];
    if (w > 0.0f) p.alpha[i] = alpha[i] * (1.0f / w);
  }
  p.normal = normal;
  p.depth = depth;
  p.part_id = part_id;
  p.material_id = material_id;
  p.coc = coc;
  return p;
}

// render.rs:515-556
u64 render_tile_exact(const Scene& scene, const Prepared& prep, TileRect tile, PassFilm& film) {
  if (!(film.mask & PASS_EXACT)) return 0;
  u32 x_end = std::min<u32>(tile.x + tile.w, scene.render.width);
  u32 y_end = std::min<u32>(tile.y + tile.h, scene.render.height);
  u64 rays = 0;
  for (u32 y = tile.y; y < y_end; ++y)
    for (u32 x = tile.x; x < x_end; ++x) {
      Ray ray = centre_ray(scene, x, y);
      ++rays;
      film.set_exact(x, y, exact_of(scene, ray, first_hit(scene, prep, ray)));
    }
  return rays;
}

// render.rs:558-640 (passes.rs:209-235 albedo_of, coverage_of)
u64 render_tile_albedo(const Scene& scene, const Prepared& prep, TileRect tile, u32 sample_index, PassFilm& film) {
  if (!(film.mask & PASS_SAMPLED)) return 0;
  u32 x_end = std::min<u32>(tile.x + tile.w, scene.render.width);
  u32 y_end = std::min<u32>(tile.y + tile.h, scene.render.height);
  u64 rays = 0;
  for (u32 y = tile.y; y < y_end; ++y)
    for (u32 x = tile.x; x < x_end; ++x) {
      PixelSampler sampler = PixelSampler::make(scene.render.seed, x, y, sample_index);
      Ray ray;
      camera_sample(scene, sampler, x, y, ray);
      ++rays;
      FirstHit hit = first_hit(scene, prep, ray);
      Rgb albedo = hit.kind == FIRST_BACKGROUND ? Rgb(0.0f, 0.0f, 0.0f)
                   : hit.kind == FIRST_FLOOR    ? scene.floor.color
                                                : hit.albedo;
      film.add_albedo(x, y, albedo);
      film.add_alpha(x, y, hit.kind == FIRST_SURFACE ? 1.0f : 0.0f);
    }
  return rays;
}

}  // namespace orc
