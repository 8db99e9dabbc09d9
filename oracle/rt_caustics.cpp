// ORACLE — test infrastructure only (see rt_core.hpp and rt_caustics.hpp).
#include "rt_caustics.hpp"
#include <algorithm>
#include <thread>

namespace orc {
namespace {

constexpr float MM2_TO_M2 = 1.0e-6f;
constexpr float MM_TO_M = 1.0e-3f;

inline bool positive(float x) { return x > 0.0f; }

// light.rs:636-654
std::vector<float> cdf_of(const std::vector<float>& weights) {
  float total = -0.0f;
  for (float w : weights) total = total + w;
  if (!(total > 0.0f && std::isfinite(total))) return {};
  std::vector<float> cdf;
  float acc = 0.0f;
  for (float w : weights) {
    acc += w;
    cdf.push_back(acc / total);
  }
  cdf.back() = 1.0f;
  return cdf;
}
size_t pick_index(const std::vector<float>& cdf, float u) {
  size_t i = 0;
  while (i < cdf.size() && cdf[i] <= u) ++i;
  return std::min(i, cdf.empty() ? (size_t)0 : cdf.size() - 1);
}
float bucket_width(const std::vector<float>& cdf, size_t i) { return cdf[i] - (i == 0 ? 0.0f : cdf[i - 1]); }

// warp.rs:48-54
Vec3 uniform_sphere(float u0, float u1) {
  float z = 1.0f - 2.0f * u0;
  float r = std::sqrt(fmax_r(1.0f - z * z, 0.0f));
  float s, c;
  sin_cos(TAU * u1, s, c);
  return Vec3(r * c, r * s, z);
}

// caustics.rs:171-178
float cone_one_minus_cos(float r, float d) {
  float sin2 = (r * r) / (d * d);
  float cos = std::sqrt(fmax_r(1.0f - sin2, 0.0f));
  return sin2 / (1.0f + cos);
}

void sphere_of(const Aabb& b, Vec3& c, float& r) {
  c = b.centroid();
  r = fmax_r(0.5f * (b.max - b.min).length(), 1.0e-3f);
}

// caustics.rs:122-143: the world box of instance i's mesh box corners.
Aabb instance_bounds(const Scene& scene, size_t i) {
  const Geometry& g = scene.geometry;
  const Instance& inst = g.instances[i];
  if (inst.mesh >= g.meshes.size()) return Aabb();
  Aabb b = g.meshes[inst.mesh].bounds;
  if (b.is_empty()) return b;
  Aabb out;
  for (u32 k = 0; k < 8; ++k) {
    Vec3 corner((k & 1) == 0 ? b.min.x : b.max.x, (k & 2) == 0 ? b.min.y : b.max.y,
                (k & 4) == 0 ? b.min.z : b.max.z);
    out = out.extend(inst.transform.point(corner));
  }
  return out;
}

// caustics.rs:145-169
void instance_sphere(const Scene& scene, size_t i, Vec3& centre, float& radius) {
  Aabb b = instance_bounds(scene, i);
  centre = b.centroid();
  const Instance& inst = scene.geometry.instances[i];
  if (inst.mesh >= scene.geometry.meshes.size()) {
    sphere_of(b, centre, radius);
    return;
  }
  const Mesh& mesh = scene.geometry.meshes[inst.mesh];
  float r2 = 0.0f;
  for (const Vec3& p : mesh.positions) r2 = fmax_r(r2, (inst.transform.point(p) - centre).length_squared());
  float r = std::sqrt(r2) * 1.0001f + 1.0e-3f;
  if (r > 0.0f && std::isfinite(r)) radius = r;
  else sphere_of(b, centre, radius);
}

// caustics.rs:180-197
float pixel_footprint(const Scene& scene) {
  const Camera& c = scene.camera;
  u32 w = std::max<u32>(scene.render.width, 1), h = std::max<u32>(scene.render.height, 1);
  if (c.ortho) return std::fabs(c.ortho_width) / (float)w;
  float distance = (c.target - c.position).length();
  return distance * c.sensor_mm / (fmax_r(c.focal_mm, 1.0e-3f) * (float)std::max(w, h));
}

// caustics.rs:495-559: a direction from x into the casters' cones and its density under the
// mixture of every cone (`normal`: the emitting side of a one-sided light, or null).
bool sample_towards(const CausticSetup& s, Vec3 x, const Vec3* normal, float u_pick, float u0, float u1, Vec3& dir,
                    float& pdf_out) {
  size_t count = s.sphere_c.size();
  Vec3 axes[CAUSTIC_MAX_SPHERES];
  float omcs[CAUSTIC_MAX_SPHERES];
  std::vector<float> weights(count);
  for (size_t i = 0; i < count; ++i) {
    Vec3 d = s.sphere_c[i] - x;
    float dist = d.length();
    Vec3 axis = dist > 0.0f ? d / dist : Vec3(0.0f, 0.0f, 1.0f);
    float omc, omega;
    if (dist <= s.sphere_r[i]) {
      omc = 0.0f;
      omega = PI;
    } else {
      omc = cone_one_minus_cos(s.sphere_r[i], dist);
      omega = TAU * omc;
    }
    float facing = normal ? 0.05f + fmax_r(normal->dot(axis), 0.0f) : 1.0f;
    axes[i] = axis;
    omcs[i] = omc;
    weights[i] = omega * facing;
  }
  std::vector<float> cdf = cdf_of(weights);
  if (cdf.empty()) return false;
  size_t k = pick_index(cdf, u_pick);
  if (omcs[k] > 0.0f) {
    dir = Frame::from_normal(axes[k]).to_world(uniform_cone(u0, u1, omcs[k])).normalize();
  } else if (normal) {
    dir = Frame::from_normal(*normal).to_world(cosine_hemisphere(u0, u1)).normalize();
  } else {
    dir = uniform_sphere(u0, u1);
  }
  float pdf = 0.0f;
  for (size_t i = 0; i < count; ++i) {
    float p = bucket_width(cdf, i);
    if (!positive(p)) continue;
    if (omcs[i] > 0.0f) {
      if (1.0f - dir.dot(axes[i]) <= omcs[i]) pdf += p / (TAU * omcs[i]);
    } else {
      pdf += p * (normal ? fmax_r(dir.dot(*normal), 0.0f) * FRAC_1_PI : 0.25f * FRAC_1_PI);
    }
  }
  if (!(pdf > 0.0f && std::isfinite(pdf))) return false;
  pdf_out = pdf;
  return true;
}

// caustics.rs:369-491: where a photon leaves its light, its direction and its flux (W).
bool emit(const CausticSetup& s, const Lights& lights, Pcg32& rng, Ray& ray, Rgb& flux) {
  float n = (float)s.photons;
  size_t j = pick_index(s.cdf, rng.next_f32());
  float p_light = bucket_width(s.cdf, j);
  if (!positive(p_light)) return false;
  float up0 = rng.next_f32();
  float up1 = rng.next_f32();
  float u_pick = rng.next_f32();
  float ud0 = rng.next_f32();
  float ud1 = rng.next_f32();
  const CausticSetup::Source src = s.sources[j];
  if (src.kind == 0) {
    const AreaLight& light = lights.areas[src.index];
    float x0, y0;
    if (light.rect) {
      x0 = (2.0f * up0 - 1.0f) * light.half_x;
      y0 = (2.0f * up1 - 1.0f) * light.half_y;
    } else {
      float dx, dy;
      concentric_disk(up0, up1, dx, dy);
      x0 = dx * light.radius;
      y0 = dy * light.radius;
    }
    Vec3 x = light.centre + light.frame.x * x0 + light.frame.y * y0;
    Vec3 dir;
    float pdf_dir;
    if (!sample_towards(s, x, &light.frame.z, u_pick, ud0, ud1, dir, pdf_dir)) return false;
    float cos = dir.dot(light.frame.z);
    if (!positive(cos)) return false;
    float pdf_area = 1.0f / (light.area * MM2_TO_M2);
    flux = light.radiance_at(cos) * (cos / (pdf_area * pdf_dir * p_light * n));
    ray = Ray(x, dir);
    return true;
  }
  if (src.kind == 2) {
    const MeshEmitter& m = lights.meshes[src.index];
    size_t t = pick_index(m.cdf, up0);
    float lo = t == 0 ? 0.0f : m.cdf[t - 1];
    float width = m.cdf[t] - lo;
    float rest = width > 0.0f ? clampf((up0 - lo) / width, 0.0f, 0.99999994f) : 0.0f;
    float w[3];
    uniform_triangle(rest, up1, w);
    Vec3 x = m.a[t] * w[0] + m.b[t] * w[1] + m.c[t] * w[2];
    Vec3 normal = m.normals[t];
    Vec3 dir;
    float pdf_dir;
    if (!sample_towards(s, x, m.two_sided ? nullptr : &normal, u_pick, ud0, ud1, dir, pdf_dir)) return false;
    float facing = dir.dot(normal);
    if (!m.two_sided && facing <= 0.0f) return false;
    float cos = std::fabs(facing);
    if (!positive(cos)) return false;
    float pdf_area = 1.0f / (m.area * MM2_TO_M2);
    flux = m.radiance * (cos / (pdf_area * pdf_dir * p_light * n));
    ray = Ray(offset_origin(x, facing > 0.0f ? normal : -normal), dir);
    return true;
  }
  const SunLight& sun = lights.suns[src.index];
  Vec3 to_sun = sun.frame.z;
  Vec3 travel;
  float cos, pdf_dir;
  if (sun.delta) {
    travel = -to_sun;
    cos = 1.0f;
    pdf_dir = 1.0f;
  } else {
    Vec3 local = uniform_cone(ud0, ud1, sun.one_minus_cos_max);
    Vec3 d = sun.frame.to_world(local).normalize();
    travel = -d;
    cos = local.z;
    pdf_dir = 1.0f / sun.solid_angle;
  }
  Vec3 plane = s.scene_centre + to_sun * (2.0f * s.scene_radius);
  float tilt = 0.0f;
  if (!sun.delta) {
    float s2 = 2.0f * sun.one_minus_cos_max - sun.one_minus_cos_max * sun.one_minus_cos_max;
    tilt = std::sqrt(fmax_r(s2, 0.0f)) * 4.0f * s.scene_radius;
  }
  std::vector<float> areas;
  for (float r : s.sphere_r) areas.push_back(PI * (r + tilt) * (r + tilt));
  std::vector<float> cdf = cdf_of(areas);
  if (cdf.empty()) return false;
  size_t k = pick_index(cdf, u_pick);
  Frame frame = Frame::from_normal(to_sun);
  auto project = [&](Vec3 c) { return c - to_sun * (c - plane).dot(to_sun); };
  float dx, dy;
  concentric_disk(up0, up1, dx, dy);
  float rk = s.sphere_r[k] + tilt;
  Vec3 x = project(s.sphere_c[k]) + frame.x * (dx * rk) + frame.y * (dy * rk);
  float pdf_area = 0.0f;
  for (size_t i = 0; i < s.sphere_c.size(); ++i) {
    float ri = s.sphere_r[i] + tilt;
    if ((x - project(s.sphere_c[i])).length_squared() <= ri * ri)
      pdf_area += bucket_width(cdf, i) / (PI * ri * ri * MM2_TO_M2);
  }
  if (!positive(pdf_area)) return false;
  flux = sun.radiance * (cos / (pdf_area * pdf_dir * p_light * n));
  ray = Ray(x, travel);
  return true;
}

// caustics.rs:581-735
u64 trace_photon(const Scene& scene, const Prepared& prep, const CausticSetup& s, u32 iteration, u32 index,
                 std::vector<Photon>& out) {
  u64 key = ((u64)s.seed << 32) | (u64)iteration;
  Pcg32 rng = Pcg32::seeded(mix64(key ^ 0x0e5b0001c0decafeull), (u64)index);
  Ray ray;
  Rgb power;
  if (!emit(s, prep.lights, rng, ray, power)) return 0;
  if (!(peak(power) > 0.0f && std::isfinite(peak(power)))) return 0;
  const Geometry& geo = scene.geometry;
  float throughput = 1.0f;
  u32 bounces = 0;
  bool in_medium = false;
  Rgb sigma;
  u64 rays = 0;
  for (u32 step = 0; step < s.max_bounces + 64 + 64; ++step) {
    ++rays;
    RayHit hit;
    bool has_hit = geo.trace(ray, false, hit);
    float t_geometry = has_hit ? hit.t : INF;
    float t_floor = prep.floor_t(ray, true);
    float t_nearest = fmin_r(t_geometry, t_floor);
    if (in_medium && std::isfinite(t_nearest)) {
      Rgb bl(expf_(-sigma.r * t_nearest), expf_(-sigma.g * t_nearest), expf_(-sigma.b * t_nearest));
      power = power * bl;
    }
    if (t_nearest == INF) break;
    if (t_floor < t_geometry) {
      if (bounces > 0) {
        Vec3 p = ray.at(t_floor);
        out.push_back(Photon{Vec3(p.x, p.y, prep.floor_z), -ray.dir, power,
                             Vec3(0.0f, 0.0f, ray.dir.z < 0.0f ? 1.0f : -1.0f), bounces});
      }
      break;
    }
    Hit details;
    if (!geo.hit_details(ray, hit, details)) break;
    const Material& m = prep.material(scene, details.instance);
    if (m.is_emissive()) break;
    Vec3 ng = details.geometric_normal;
    bool caster = s.is_caster_instance(details.instance);
    Bsdf bsdf = Bsdf::surface(m, details.outside);
    if (!caster) {
      if (bounces > 0) out.push_back(Photon{details.point, -ray.dir, power, ng, bounces});
      break;  // no thin walls in the supported scenes: a receiver ends the photon
    }
    if (bounces >= s.max_bounces) break;
    // The shading point the camera paths use (bevel.rs:232-252 without a bevel).
    Vec3 normal = details.normal;
    if (normal.dot(ng) < 0.0f) normal = -normal;
    if (normal.dot(ray.dir) >= -1.0e-4f) normal = ng;
    Vec3 point = geo.terminator_point(details);
    Frame frame = Frame::from_normal(normal);
    Vec3 wo_world = -ray.dir;
    Vec3 wo = frame.to_local(wo_world);
    float u0 = rng.next_f32();
    float u1 = rng.next_f32();
    BsdfSample bs;
    if (!bsdf.sample(wo, u0, u1, bs)) break;
    bool transmitted = bs.wi.z < 0.0f;
    Vec3 wi = frame.to_world(bs.wi).normalize();
    if (transmitted) {
      if (wi.dot(ng) >= 0.0f) break;
    } else if (wi.dot(ng) <= 0.0f) {
      break;
    }
    Rgb weight = bs.weight;
    if (transmitted) weight = weight * (bsdf.eta * bsdf.eta);
    float num = std::fabs(wo_world.dot(normal)) * std::fabs(wi.dot(ng));
    float den = std::fabs(wo_world.dot(ng)) * std::fabs(wi.dot(normal));
    if (den > 0.0f) weight = weight * fmin_r(num / den, 10.0f);
    ++bounces;
    power = power * weight;
    throughput *= peak(weight);
    if (bounces >= 3) {
      float survive = clampf(throughput, 0.05f, 1.0f);
      if (rng.next_f32() >= survive) break;
      power = power * (1.0f / survive);
      throughput /= survive;
    }
    if (!(peak(power) > 0.0f && std::isfinite(peak(power)))) break;
    Vec3 next_origin = transmitted ? offset_origin(details.point, -ng) : offset_origin(point, ng);
    if (transmitted) {
      in_medium = false;
      if (details.outside) {
        u32 id = geo.instances[details.instance].material_id;
        if (id < prep.has_medium.size() && prep.has_medium[id]) {
          in_medium = true;
          sigma = prep.media[id];
        }
      }
    }
    ray = Ray(next_origin, wi);
  }
  return rays;
}

}  // namespace

// caustics.rs:104-119
bool is_caster(const Material& m) {
  float r = m.roughness;  // anisotropic materials are refused at load
  if (m.is_emissive() || std::isnan(r) || r > CAUSTIC_MAX_ROUGHNESS) return false;
  ScatterMedium unused;
  if (scatter_medium_of(m, unused)) return false;
  bool glass = m.kind == MATERIAL_GLASS || (m.transmission >= 0.5f && m.metallic < 0.5f);
  bool metal = m.kind != MATERIAL_GLASS && m.metallic >= 0.5f;
  return glass || metal;
}

float CausticSetup::radius_of(u32 iteration) const {
  float r2 = radius * radius;
  for (u32 k = 1; k <= iteration; ++k) r2 *= ((float)k + CAUSTIC_RADIUS_ALPHA) / ((float)k + 1.0f);
  return std::sqrt(r2);
}

// caustics.rs:229-345
bool caustic_setup(const Scene& scene, const Lights& lights, CausticSetup& out) {
  out = CausticSetup();
  u32 photons = scene.render.caustic_photons;
  if (photons == 0) return false;
  const Geometry& g = scene.geometry;
  out.casters.assign(g.instances.size(), 0);
  std::vector<size_t> parts;
  for (size_t i = 0; i < g.instances.size(); ++i) {
    u32 id = g.instances[i].material_id;
    out.casters[i] = id < scene.materials.size() && is_caster(scene.materials[id]);
    if (out.casters[i] && !instance_bounds(scene, i).is_empty()) parts.push_back(i);
  }
  if (parts.empty()) return false;
  if (parts.size() > CAUSTIC_MAX_SPHERES) {
    Aabb all;
    for (size_t i : parts) all = all.unite(instance_bounds(scene, i));
    Vec3 c;
    float r;
    sphere_of(all, c, r);
    out.sphere_c.push_back(c);
    out.sphere_r.push_back(r);
  } else {
    for (size_t i : parts) {
      Vec3 c;
      float r;
      instance_sphere(scene, i, c, r);
      out.sphere_c.push_back(c);
      out.sphere_r.push_back(r);
    }
  }
  auto towards = [&](Vec3 x, const Vec3* normal) {
    float sum = -0.0f;
    for (size_t i = 0; i < out.sphere_c.size(); ++i) {
      Vec3 d = out.sphere_c[i] - x;
      float dist = d.length();
      float facing = normal ? 0.05f + fmax_r(normal->dot(d / fmax_r(dist, 1.0e-6f)), 0.0f) : 1.0f;
      float omega = dist <= out.sphere_r[i] ? PI : TAU * cone_one_minus_cos(out.sphere_r[i], dist);
      sum = sum + omega * facing;
    }
    return sum;
  };
  std::vector<float> weights;
  for (size_t i = 0; i < lights.areas.size(); ++i) {
    const AreaLight& a = lights.areas[i];
    if (!a.glossy) continue;
    float w = a.radiance.luminance() * a.area * MM2_TO_M2 * towards(a.centre, &a.frame.z);
    if (w > 0.0f && std::isfinite(w)) {
      out.sources.push_back({0, (u32)i});
      weights.push_back(w);
    }
  }
  for (size_t i = 0; i < lights.suns.size(); ++i) {
    const SunLight& sn = lights.suns[i];
    if (!sn.glossy) continue;
    Rgb irradiance = sn.delta ? sn.radiance : sn.radiance * (sn.solid_angle * (1.0f - 0.5f * sn.one_minus_cos_max));
    float area = -0.0f;
    for (float r : out.sphere_r) area = area + PI * r * r * MM2_TO_M2;
    float w = irradiance.luminance() * area;
    if (w > 0.0f && std::isfinite(w)) {
      out.sources.push_back({1, (u32)i});
      weights.push_back(w);
    }
  }
  for (size_t i = 0; i < lights.meshes.size(); ++i) {
    const MeshEmitter& m = lights.meshes[i];
    Vec3 acc(0.0f, 0.0f, 0.0f);
    for (size_t t = 0; t < m.a.size(); ++t) acc = acc + (m.a[t] + m.b[t] + m.c[t]) * (1.0f / 3.0f);
    Vec3 centroid = acc * (1.0f / (float)std::max<size_t>(m.a.size(), 1));
    float w = m.radiance.luminance() * m.area * MM2_TO_M2 * 0.5f * towards(centroid, nullptr);
    if (w > 0.0f && std::isfinite(w)) {
      out.sources.push_back({2, (u32)i});
      weights.push_back(w);
    }
  }
  out.cdf = cdf_of(weights);
  if (out.cdf.empty()) return false;
  Aabb bounds = g.bounds;
  if (bounds.is_empty()) {
    out.scene_centre = Vec3(0.0f, 0.0f, 0.0f);
    out.scene_radius = 1.0f;
  } else {
    sphere_of(bounds, out.scene_centre, out.scene_radius);
  }
  float radius_px = scene.render.caustic_radius > 0.0f ? scene.render.caustic_radius : CAUSTIC_DEFAULT_RADIUS_PX;
  out.radius = fmax_r(radius_px * pixel_footprint(scene), 1.0e-4f);
  out.photons = photons;
  out.seed = scene.render.seed;
  u32 depth = scene.render.depth();
  out.max_bounces = depth > 0 ? depth - 1 : 0;
  return true;
}

void trace_photons(const Scene& scene, const Prepared& prep, const CausticSetup& setup, u32 iteration, u32 first,
                   u32 last, std::vector<Photon>& out, u64& rays) {
  rays = 0;
  for (u32 i = first; i < std::min(last, setup.photons); ++i) rays += trace_photon(scene, prep, setup, iteration, i, out);
}

// caustics.rs:807-860
PhotonMap build_photon_map(const CausticSetup& setup, u32 iteration, std::vector<Photon> photons, u64 rays) {
  PhotonMap m;
  m.iteration = iteration;
  m.radius = setup.radius_of(iteration);
  float r_m = m.radius * MM_TO_M;
  m.inv_cell = 1.0f / (2.0f * m.radius);
  size_t want = std::max<size_t>(photons.size(), 1) * 2;
  size_t buckets = 1;
  while (buckets < want) buckets <<= 1;
  buckets = std::min<size_t>(buckets, (size_t)1 << 24);
  m.mask = (u32)buckets - 1;
  std::vector<u32> keys(photons.size());
  for (size_t i = 0; i < photons.size(); ++i) {
    int c[3];
    caustic_cell(photons[i].position, m.inv_cell, c);
    keys[i] = caustic_bucket(c, m.mask);
  }
  m.starts.assign(buckets + 1, 0);
  for (u32 k : keys) m.starts[k + 1]++;
  for (size_t i = 0; i < buckets; ++i) m.starts[i + 1] += m.starts[i];
  std::vector<u32> fill(m.starts.begin(), m.starts.end());
  m.order.assign(photons.size(), 0);
  for (size_t i = 0; i < keys.size(); ++i) m.order[fill[keys[i]]++] = (u32)i;
  m.inv_r2 = 1.0f / (m.radius * m.radius);
  m.norm = 2.0f / (PI * r_m * r_m);
  m.photons = std::move(photons);
  m.rays = rays;
  return m;
}

PhotonMap trace_photon_map(const Scene& scene, const Prepared& prep, u32 iteration, int threads) {
  const CausticSetup& s = prep.caustics;
  if (!prep.has_caustics) return PhotonMap();
  int n = std::max(threads, 1);
  std::vector<std::vector<Photon>> parts(n);
  std::vector<u64> rays(n, 0);
  u32 per = (s.photons + n - 1) / n;
  std::vector<std::thread> pool;
  for (int t = 0; t < n; ++t)
    pool.emplace_back([&, t]() {
      u32 a = std::min(s.photons, per * (u32)t), b = std::min(s.photons, per * (u32)(t + 1));
      trace_photons(scene, prep, s, iteration, a, b, parts[t], rays[t]);
    });
  for (auto& th : pool) th.join();
  std::vector<Photon> all;
  u64 total = 0;
  for (int t = 0; t < n; ++t) {
    all.insert(all.end(), parts[t].begin(), parts[t].end());
    total += rays[t];
  }
  return build_photon_map(s, iteration, std::move(all), total);
}

}  // namespace orc
