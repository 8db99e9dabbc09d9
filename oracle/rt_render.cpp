// ORACLE — test infrastructure only (see rt_core.hpp). Citations in rt_render.hpp.
#include "rt_render.hpp"
#include "rt_caustics.hpp"
#include "rt_bsdf_grad.hpp"
#include <algorithm>
#include <cstring>

namespace orc {

constexpr float MM2_TO_M2 = 1.0e-6f;
constexpr float MM_TO_M = 1.0e-3f;
// light.rs:118, 323-325: the closest a light's selection weight takes it; below this 1 - cos of
// its cone a sphere light is a point.
constexpr float MIN_DISTANCE2_M2 = 1.0e-6f;
constexpr float POINT_CONE = 1.0e-7f;
// medium.rs:69-87: events of a walk before its roulette, the largest anisotropy, and the offsets
// inside one event's dimensions.
constexpr u32 MEDIUM_ROULETTE_AFTER = 32;
constexpr float MAX_G = 0.9f;
constexpr u32 MEDIUM_OFFSET_DISTANCE = 0, MEDIUM_OFFSET_PHASE = 1, MEDIUM_OFFSET_ROULETTE = 3,
              MEDIUM_OFFSET_FRESNEL = 4, MEDIUM_OFFSET_HERO = 5;
constexpr size_t MAX_LOCAL_LIGHTS = 64;

// ------------------------------------------------------------------ scene/mod.rs:160-357
void Camera::basis(Vec3& forward, Vec3& right, Vec3& upv) const {
  forward = (target - position).normalize();
  if (forward.length_squared() == 0.0f) forward = Vec3(0, 1, 0);
  right = forward.cross(up).normalize();
  if (right.length_squared() < 0.5f) {
    Vec3 axis = std::fabs(forward.z) < 0.9f ? Vec3(0, 0, 1) : Vec3(1, 0, 0);
    right = forward.cross(axis).normalize();
  }
  upv = right.cross(forward);
}

float Camera::lens_radius() const {
  bool through_lens = !ortho && projection == 0;
  if (fstop > 0.0f && focus_dist > 0.0f && through_lens) return focal_mm / fstop * 0.5f;
  return 0.0f;
}

Ray Camera::generate_ray(float aspect, float u, float v, float lx, float ly) const {
  Vec3 forward, right, upv;
  basis(forward, right, upv);
  if (ortho) {
    float w = ortho_width;
    Vec3 origin = position + right * ((u - 0.5f) * w) + upv * ((0.5f - v) * w / aspect);
    return Ray(origin, forward);
  }
  float sensor_w, sensor_h;
  if (aspect >= 1.0f) { sensor_w = sensor_mm; sensor_h = sensor_mm / aspect; }
  else { sensor_w = sensor_mm * aspect; sensor_h = sensor_mm; }
  float x = (u - 0.5f) * sensor_w;
  float y = (0.5f - v) * sensor_h;
  Vec3 dir = (forward * focal_mm + right * x + upv * y).normalize();
  float radius = lens_radius();
  if (radius == 0.0f) return Ray(position, dir);
  Vec3 focus_point = position + dir * (focus_dist / dir.dot(forward));
  Vec3 origin = position + (right * lx + upv * ly) * radius;
  return Ray(origin, (focus_point - origin).normalize());
}

// ------------------------------------------------------------------ rsb.rs:1197-1610 (subset)
namespace {
struct Reader {
  const uint8_t* p;
  size_t len, pos = 0;
  bool ok = true;
  size_t remaining() const { return len - pos; }
  u32 u32_() {
    if (remaining() < 4) { ok = false; return 0; }
    u32 v;
    std::memcpy(&v, p + pos, 4);
    pos += 4;
    return v;
  }
  float f32_() { return u2f(u32_()); }
  u32 opt_u32(u32 d) { return remaining() >= 4 ? u32_() : d; }
  float opt_f32(float d) { return remaining() >= 4 ? f32_() : d; }
  Vec3 vec3() { float a = f32_(), b = f32_(), c = f32_(); return Vec3(a, b, c); }
  Rgb rgb() { float a = f32_(), b = f32_(), c = f32_(); return Rgb(a, b, c); }
};
}  // namespace

bool scene_from_rsb(const uint8_t* bytes, size_t len, Scene& scene, std::string& err) {
  if (len < 4 || std::memcmp(bytes, "RSB1", 4) != 0) { err = "bad magic"; return false; }
  if (len < 8) { err = "truncated header"; return false; }
  u32 header_len;
  std::memcpy(&header_len, bytes + 4, 4);
  size_t pos = 8 + (size_t)header_len;
  if (pos > len) { err = "truncated header"; return false; }
  scene = Scene();
  bool have_cam = false, have_env = false, have_floor = false, have_render = false;
  struct RawMesh { std::vector<Vec3> pos, nrm; std::vector<u32> idx; u32 material; };
  std::vector<RawMesh> meshes;
  std::vector<Instance> instances;
  while (pos < len) {
    if (pos + 8 > len) { err = "truncated section header"; return false; }
    char tag[5] = {0};
    std::memcpy(tag, bytes + pos, 4);
    u32 n;
    std::memcpy(&n, bytes + pos + 4, 4);
    if (pos + 8 + (size_t)n > len) { err = "truncated section payload"; return false; }
    Reader r{bytes + pos + 8, n};
    pos += 8 + (size_t)n;
    std::string t(tag);
    if (t == "CAMR") {
      if (have_cam) { err = "duplicate CAMR"; return false; }
      have_cam = true;
      Camera& c = scene.camera;
      c.position = r.vec3();
      c.target = r.vec3();
      c.up = r.vec3();
      c.focal_mm = r.f32_();
      c.sensor_mm = r.f32_();
      c.fstop = r.f32_();
      c.focus_dist = r.f32_();
      c.ortho = r.u32_() != 0;
      c.ortho_width = r.f32_();
      c.projection = r.opt_u32(0);
      c.fov_x_deg = r.opt_f32(360);
      c.fov_y_deg = r.opt_f32(180);
      c.exterior = r.opt_f32(1);
      if (!r.ok) { err = "CAMR too short"; return false; }
      if (c.projection == 1 || c.projection == 2) { err = "panoramic cameras are not supported"; return false; }
    } else if (t == "MATL") {
      Material m;
      m.kind = r.u32_();
      m.base_color = r.rgb();
      m.metallic = r.f32_();
      m.roughness = r.f32_();
      m.ior = r.f32_();
      m.transmission = r.f32_();
      m.coat = r.f32_();
      m.coat_roughness = r.f32_();
      m.emission = r.rgb();
      m.emission_power = r.f32_();
      m.flags = r.opt_u32(0);
      m.opacity = r.opt_f32(1);
      m.anisotropy = r.opt_f32(0);
      m.absorption_mm = r.opt_f32(0);
      m.bevel_mm = r.opt_f32(0);
      m.scatter_mm = r.opt_f32(0);
      float sr = r.opt_f32(1), sg = r.opt_f32(1), sb = r.opt_f32(1);
      m.scatter_color = Rgb(sr, sg, sb);
      m.scatter_g = r.opt_f32(0);
      m.abbe = r.opt_f32(0);
      if (!r.ok) { err = "MATL too short"; return false; }
      if (r.remaining() >= 4) { err = "material texture blocks are not supported"; return false; }
      scene.materials.push_back(m);
    } else if (t == "MESH") {
      RawMesh rm;
      u32 vc = r.u32_(), ic = r.u32_();
      if (!r.ok || ((u64)vc * 6 + ic + 1) * 4 > r.remaining()) { err = "MESH too short"; return false; }
      if (ic % 3 != 0) { err = "index_count is not a multiple of 3"; return false; }
      rm.pos.resize(vc);
      rm.nrm.resize(vc);
      for (u32 i = 0; i < vc; ++i) rm.pos[i] = r.vec3();
      for (u32 i = 0; i < vc; ++i) rm.nrm[i] = r.vec3();
      rm.idx.resize(ic);
      for (u32 i = 0; i < ic; ++i) {
        rm.idx[i] = r.u32_();
        if (rm.idx[i] >= vc) { err = "vertex index out of range"; return false; }
      }
      rm.material = r.u32_();
      if (r.remaining() >= 4 && r.u32_() != 0) { err = "UV sets are not supported"; return false; }
      meshes.push_back(std::move(rm));
    } else if (t == "INST") {
      Instance inst;
      inst.mesh = r.u32_();
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 4; ++j) inst.transform.rows[i][j] = r.f32_();
      inst.material_id = r.u32_();
      inst.flags = r.u32_();
      if (!r.ok) { err = "INST too short"; return false; }
      instances.push_back(inst);
    } else if (t == "LGHT") {
      Light l;
      l.kind = r.u32_();
      if (l.kind > 5) { err = "unknown light kind"; return false; }
      l.position = r.vec3();
      l.direction = r.vec3();
      l.up = r.vec3();
      l.size_x = r.f32_();
      l.size_y = r.f32_();
      l.color = r.rgb();
      l.power = r.f32_();
      l.spread_deg = r.f32_();
      l.angle_deg = r.f32_();
      l.flags = r.u32_();
      if (!r.ok) { err = "LGHT too short"; return false; }
      scene.lights.push_back(l);
    } else if (t == "ENVR") {
      if (have_env) { err = "duplicate ENVR"; return false; }
      have_env = true;
      Environment& e = scene.environment;
      e.kind = r.u32_();
      if (e.kind > 3) { err = "unknown environment kind"; return false; }
      e.floor = r.f32_();
      e.horizon = r.f32_();
      e.zenith = r.f32_();
      e.strength = r.f32_();
      e.rotation_deg = r.f32_();
      e.background_color = r.rgb();
      e.background_mode = r.u32_();
      if (!r.ok) { err = "ENVR too short"; return false; }
      u32 hdri = r.opt_u32(0xffffffffu);
      if (e.kind == 3 && hdri != 0xffffffffu) { err = "HDRI environments are not supported"; return false; }
      if (e.kind == 2) { err = "sky environments are not supported"; return false; }
      if (e.kind == 3) e.kind = 0;  // an HDRI without its bitmap renders the dome
    } else if (t == "FLOR") {
      if (have_floor) { err = "duplicate FLOR"; return false; }
      have_floor = true;
      scene.floor.mode = r.u32_();
      scene.floor.z = r.f32_();
      scene.floor.color = r.rgb();
      if (!r.ok) { err = "FLOR too short"; return false; }
      if (scene.floor.mode > 2) { err = "unknown floor mode"; return false; }
      if (scene.floor.mode == 2) { err = "the shadow catcher is not supported"; return false; }
    } else if (t == "RNDR") {
      if (have_render) { err = "duplicate RNDR"; return false; }
      have_render = true;
      RenderSettings& s = scene.render;
      s.width = r.u32_();
      s.height = r.u32_();
      s.max_samples = r.u32_();
      s.seed = r.u32_();
      s.max_seconds = r.f32_();
      s.exposure_ev = r.f32_();
      s.tonemap = r.u32_();
      if (!r.ok) { err = "RNDR too short"; return false; }
      s.max_depth = r.opt_u32(8);
      s.clamp_indirect = r.opt_f32(0);
      (void)r.opt_u32(0);   // passes
      (void)r.opt_u32(0);   // denoise
      s.caustic_photons = r.opt_u32(0);
      s.caustic_radius = r.opt_f32(0);
      (void)r.opt_f32(1);   // denoise strength
      s.pixel_filter = r.opt_u32(0);
      s.edge_supersample = r.opt_u32(1);
      // E10 (rsb.rs:1105-1110): the adaptive bit asks the host for rounds; the renderer reads it not.
      s.adaptive = r.opt_u32(0);
      s.noise_threshold = r.opt_f32(0);
      if (s.caustic_photons > MAX_CAUSTIC_PHOTONS) { err = "caustic_photons must be at most 4194304"; return false; }
      if (s.caustic_radius < 0.0f) { err = "caustic_radius must not be negative"; return false; }
      if (!(s.noise_threshold >= 0.0f && s.noise_threshold <= 1.0f)) {
        err = "noise_threshold must be between 0 and 1";
        return false;
      }
      if (s.pixel_filter == 3) { err = "the Blackman-Harris filter is not supported"; return false; }
      if (s.edge_supersample > 1) { err = "edge supersampling is not supported"; return false; }
      if (s.width < 1 || s.width > 16384 || s.height < 1 || s.height > 16384) {
        err = "width and height must be between 1 and 16384";
        return false;
      }
    } else if (t == "LENS" || t == "BMAP" || t == "PATN" || t == "IMGS" || t == "PROG" || t == "IESP" ||
               t == "FXAM" || t == "SECT" || t == "GLAS" || t == "WETP" || t == "WTP2" || t == "WETF" ||
               t == "WETG" || t == "NEST") {
      err = "section " + t + " is not supported";
      return false;
    }
  }
  // Materials: refuse what the restated integrator does not follow.
  for (const Material& m : scene.materials) {
    if (m.anisotropy > 0.0f) { err = "anisotropic materials are not supported"; return false; }
    if (m.opacity < 1.0f) { err = "cut-outs are not supported"; return false; }
    if (m.flags & MATERIAL_FLAG_THIN) { err = "thin walls are not supported"; return false; }
    if (m.abbe > 0.0f) { err = "dispersion is not supported"; return false; }
    if (m.bevel_mm > 0.0f) { err = "edge bevels are not supported"; return false; }
  }
  // A scene with a scattering medium and a thin wall (integrator/mod.rs:289-299: smooth, wholly clear
  // glass without a medium of its own), which shadow rays from the medium's exits would cross, is
  // outside the restated branches.
  bool any_scatter = false, any_wall = false;
  for (const Material& m : scene.materials) {
    ScatterMedium sm;
    bool scattering = scatter_medium_of(m, sm);
    any_scatter = any_scatter || scattering;
    any_wall = any_wall || (is_clear_solid(m) && !scattering && (m.kind == MATERIAL_GLASS || m.metallic <= 0.0f) &&
                            m.transmission >= 1.0f && m.roughness <= 0.05f && m.coat <= 0.0f && m.opacity >= 1.0f);
  }
  if (any_scatter && any_wall) { err = "thin walls beside a scattering medium are not supported"; return false; }
  for (const Light& l : scene.lights)
    if (l.kind == LIGHT_IES) { err = "IES lights are not supported"; return false; }
  std::vector<Mesh> built(meshes.size());
  for (size_t i = 0; i < meshes.size(); ++i) {
    if (meshes[i].material >= scene.materials.size()) { err = "material_id does not refer to a MATL section"; return false; }
    if (!Mesh::create(meshes[i].pos, meshes[i].nrm, meshes[i].idx, meshes[i].material, built[i], err)) return false;
  }
  if (instances.empty()) {
    for (size_t i = 0; i < built.size(); ++i) {
      Instance inst;
      inst.mesh = (u32)i;
      inst.material_id = built[i].material_id;
      instances.push_back(inst);
    }
  }
  for (const Instance& inst : instances) {
    if (inst.mesh >= built.size()) { err = "mesh_id does not refer to a MESH section"; return false; }
    if (inst.material_id >= scene.materials.size()) { err = "material_id does not refer to a MATL section"; return false; }
  }
  return scene.geometry.build(std::move(built), std::move(instances), err);
}

// ------------------------------------------------------------------ environment.rs:275-431
float EnvRamp::value(float z) const {
  z = clampf(z, -1.0f, 1.0f);
  if (z < 0.0f) return floor * -z + horizon * (1.0f + z);
  return horizon * (1.0f - z) + zenith * z;
}
float EnvRamp::pdf(Vec3 dir) const {
  if (total_mass <= 0.0f) return 0.0f;
  return value(dir.z) / (TAU * total_mass);
}
static float invert_linear(float a, float b, float u) {
  float c = u * 0.5f * (a + b);
  float denom = a + std::sqrt(fmax_r(a * a + 2.0f * (b - a) * c, 0.0f));
  if (denom > 0.0f) return clampf(2.0f * c / denom, 0.0f, 1.0f);
  return 0.0f;
}
bool EnvRamp::sample(float u0, float u1, Vec3& dir, float& pdf_out) const {
  if (total_mass <= 0.0f) return false;
  float target = u0 * total_mass;
  float z;
  if (target < lower_mass) {
    z = -1.0f + invert_linear(floor, horizon, target / lower_mass);
  } else {
    float upper = total_mass - lower_mass;
    float fraction = clampf((target - lower_mass) / upper, 0.0f, 1.0f);
    z = invert_linear(horizon, zenith, fraction);
  }
  z = clampf(z, -1.0f, 1.0f);
  float r = std::sqrt(fmax_r((1.0f - z) * (1.0f + z), 0.0f));
  float s, c;
  sin_cos(TAU * u1, s, c);
  dir = Vec3(r * c, r * s, z);
  pdf_out = pdf(dir);
  return pdf_out > 0.0f;
}
static EnvRamp make_ramp(float f, float h, float z, float strength) {
  EnvRamp r;
  r.floor = fmax_r(f * strength, 0.0f);
  r.horizon = fmax_r(h * strength, 0.0f);
  r.zenith = fmax_r(z * strength, 0.0f);
  r.lower_mass = 0.5f * (r.floor + r.horizon);
  float upper = 0.5f * (r.horizon + r.zenith);
  r.total_mass = r.lower_mass + upper;
  return r;
}

// ------------------------------------------------------------------ light.rs:120-1551
static inline bool positive(float x) { return x > 0.0f; }
static inline bool usable_total(float t) { return t > 0.0f && std::isfinite(t); }
static float power_heuristic(float pa, float pb) {
  if (!positive(pa)) return 0.0f;
  float r = pb / pa;
  return 1.0f / (1.0f + r * r);
}
static float spread_exponent(float spread_deg) {
  if (spread_deg >= 180.0f) return 0.0f;
  float half = 0.5f * fmax_r(spread_deg, 1.0f) * DEG_TO_RAD;
  return -1.0f / log2f_(fcos(half));
}
static Frame light_frame(Vec3 normal, Vec3 up) {
  Vec3 z = normal.normalize();
  Vec3 y = (up - z * z.dot(up)).normalize();
  if (y.length_squared() < 0.5f) return Frame::from_normal(z);
  Frame f;
  f.x = y.cross(z);
  f.y = y;
  f.z = z;
  return f;
}
static float orientation_bound(Vec3 v, float d2, Vec3 axis, float cos_o, float sin_o, float bound) {
  if (d2 <= bound * bound) return 1.0f;
  float inv_d = 1.0f / std::sqrt(d2);
  float sin_u = bound * inv_d;
  float cos_u = std::sqrt(fmax_r(1.0f - sin_u * sin_u, 0.0f));
  float cos_s = cos_o * cos_u - sin_o * sin_u;
  float sin_s = sin_o * cos_u + cos_o * sin_u;
  if (sin_s < 0.0f) return 1.0f;
  float cos_t = v.dot(axis) * inv_d;
  if (cos_t >= cos_s) return 1.0f;
  float sin_t = std::sqrt(fmax_r(1.0f - cos_t * cos_t, 0.0f));
  return fmax_r(cos_t * cos_s + sin_t * sin_s, 0.0f);
}

bool AreaLight::intersect(const Ray& ray, float t_max, float& t_out, float& cos_out) const {
  Vec3 n = frame.z;
  float denom = ray.dir.dot(n);
  if (denom >= 0.0f) return false;
  float t = (centre - ray.origin).dot(n) / denom;
  if (!(t > ray.t_min && t < t_max)) return false;
  Vec3 q = ray.at(t) - centre;
  float x = q.dot(frame.x), y = q.dot(frame.y);
  bool inside = rect ? (std::fabs(x) <= half_x && std::fabs(y) <= half_y) : (x * x + y * y <= radius * radius);
  if (!inside) return false;
  t_out = t;
  cos_out = -denom;
  return true;
}

float AreaLight::weight(Vec3 x) const {
  Vec3 v = x - centre;
  float d2 = v.length_squared();
  float orientation = orientation_bound(v, d2, frame.z, 1.0f, 0.0f, bound);
  if (orientation <= 0.0f) return 0.0f;
  return base * orientation / (fmax_r(d2, area) * MM2_TO_M2);
}

// light.rs:364-420
float SphereLight::mask(Vec3 x) const {
  if (cos_outer < -1.0f) return 1.0f;
  Vec3 d = x - centre;
  float len2 = d.length_squared();
  if (!positive(len2)) return 0.0f;
  float c = d.dot(axis) / std::sqrt(len2);
  if (c <= cos_outer) return 0.0f;
  if (c >= cos_inner) return 1.0f;
  float t = (c - cos_outer) / (cos_inner - cos_outer);
  return t * t * (3.0f - 2.0f * t);
}

bool SphereLight::cone(float dist2, float& omc) const {
  float r2 = radius * radius;
  if (!positive(dist2 - r2)) return false;
  float sin2 = r2 / dist2;
  float v = sin2 / (1.0f + std::sqrt(fmax_r(1.0f - sin2, 0.0f)));
  omc = v < POINT_CONE ? 0.0f : v;
  return true;
}

bool SphereLight::intersect(const Ray& ray, float t_max, float& t_out) const {
  if (!positive(radius)) return false;
  Vec3 oc = ray.origin - centre;
  float r2 = radius * radius;
  if (oc.length_squared() <= r2) return false;
  float b = oc.dot(ray.dir);
  float disc = r2 - (oc - ray.dir * b).length_squared();
  if (disc < 0.0f) return false;
  float t = -b - std::sqrt(disc);
  if (!(t > ray.t_min && t < t_max)) return false;
  t_out = t;
  return true;
}

float SphereLight::weight(Vec3 x) const {
  float m = mask(x);
  if (m <= 0.0f) return 0.0f;
  float d2 = (x - centre).length_squared() * MM2_TO_M2;
  return base * m / fmax_r(d2, near2);
}

static std::vector<float> build_cdf(const std::vector<float>& weights) {
  float total = -0.0f;
  for (float w : weights) total = total + w;
  if (!(total > 0.0f && std::isfinite(total))) return {};
  std::vector<float> cdf;
  float acc = 0.0f;
  for (float w : weights) {
    acc += w;
    cdf.push_back(acc / total);
  }
  if (!cdf.empty()) cdf.back() = 1.0f;
  return cdf;
}

static void pick(const std::vector<float>& cdf, float u, size_t& index, float& rest) {
  size_t i = 0;
  while (i < cdf.size() && cdf[i] <= u) ++i;
  i = std::min(i, cdf.empty() ? 0 : cdf.size() - 1);
  float lo = i == 0 ? 0.0f : cdf[i - 1];
  float width = cdf[i] - lo;
  rest = width > 0.0f ? clampf((u - lo) / width, 0.0f, 0.99999994f) : 0.0f;
  index = i;
}

bool Lights::build(const Scene& scene, std::string& err) {
  Aabb bounds = scene.geometry.bounds;
  Vec3 centre;
  float radius;
  if (bounds.is_empty()) {
    centre = Vec3(0, 0, 0);
    radius = 1000.0f;
  } else {
    centre = bounds.centroid();
    radius = fmax_r(0.5f * (bounds.max - bounds.min).length(), 1.0f);
  }
  float radius_m = radius * MM_TO_M;
  auto reach = [&](Vec3 p, bool has_normal, Vec3 normal) -> float {
    Vec3 to_centre = centre - p;
    float d = to_centre.length() * MM_TO_M;
    float facing = has_normal ? 0.05f + 0.95f * fmax_r(normal.dot(to_centre.normalize()), 0.0f) : 1.0f;
    d = fmax_r(d, radius_m);
    return facing / (d * d);
  };
  std::vector<float> weights;
  areas.clear();
  spheres.clear();
  entries.clear();
  for (const Light& light : scene.lights) {
    bool glossy = (light.flags & LIGHT_FLAG_GLOSSY) != 0;
    bool camera = (light.flags & LIGHT_FLAG_CAMERA) != 0;
    bool shadow = (light.flags & LIGHT_FLAG_NO_SHADOW) == 0;
    float power = fmax_r(light.power, 0.0f);
    u32 entry = (u32)entries.size();
    if (light.kind == LIGHT_POINT || light.kind == LIGHT_SPOT) {
      // light.rs:793-863 (no IES profile)
      Rgb color(fmax_r(light.color.r, 0.0f), fmax_r(light.color.g, 0.0f), fmax_r(light.color.b, 0.0f));
      Rgb intensity = color * (power / (4.0f * PI));
      float radius = 0.5f * std::fabs(light.size_x);
      if (!std::isfinite(radius)) radius = 0.0f;
      SphereLight sp;
      if (light.kind == LIGHT_SPOT) {
        Vec3 axis = light.direction.normalize();
        if (axis.length_squared() < 0.5f) continue;
        float half = 0.5f * clampf(light.spread_deg, 1.0f, 180.0f) * DEG_TO_RAD;
        float soft = fmax_r(light.angle_deg, 0.0f) * DEG_TO_RAD;
        float inner = fmax_r(half - soft, 0.0f);
        sp.axis = axis;
        sp.cos_outer = fcos(half);
        sp.cos_inner = fcos(inner);
      } else {
        sp.axis = Vec3(0.0f, 0.0f, -1.0f);
        sp.cos_outer = -2.0f;
        sp.cos_inner = -2.0f;
      }
      float flux = color.luminance() * power;
      if (!positive(flux)) continue;
      float r_m = radius * MM_TO_M;
      sp.radiance = radius > 0.0f ? intensity * (1.0f / (PI * r_m * r_m)) : Rgb(0.0f, 0.0f, 0.0f);
      sp.centre = light.position;
      sp.radius = radius;
      sp.intensity = intensity;
      sp.glossy = glossy;
      sp.camera = camera;
      sp.shadow = shadow;
      sp.entry = entry;
      sp.base = intensity.luminance();
      sp.near2 = fmax_r(r_m * r_m, MIN_DISTANCE2_M2);
      float toward = 0.05f + 0.95f * sp.mask(centre);
      weights.push_back(flux * toward * reach(light.position, false, Vec3(0, 0, 0)));
      entries.push_back(LightRef{REF_SPHERE, (u32)spheres.size()});
      spheres.push_back(sp);
      continue;
    }
    if (light.kind == LIGHT_SUN) {
      // light.rs:755-791
      Vec3 to_sun = -light.direction.normalize();
      if (to_sun.length_squared() < 0.5f) continue;
      float half = fmin_r(0.5f * fmax_r(light.angle_deg, 0.0f) * DEG_TO_RAD, FRAC_PI_2);
      Rgb color(fmax_r(light.color.r, 0.0f), fmax_r(light.color.g, 0.0f), fmax_r(light.color.b, 0.0f));
      Rgb irradiance = color * power;
      if (!positive(irradiance.luminance())) continue;
      SunLight sun;
      sun.delta = half < 1.0e-5f;
      float sh = fsin(0.5f * half);
      sun.one_minus_cos_max = 2.0f * sh * sh;
      float sin_max = fsin(half);
      sun.radiance = sun.delta ? irradiance : irradiance * (1.0f / (PI * sin_max * sin_max));
      weights.push_back(PI * irradiance.luminance());
      entries.push_back(LightRef{REF_SUN, (u32)suns.size()});
      sun.frame = Frame::from_normal(to_sun);
      sun.cos_max = 1.0f - sun.one_minus_cos_max;
      sun.solid_angle = TAU * sun.one_minus_cos_max;
      sun.glossy = glossy;
      sun.camera = camera;
      sun.shadow = shadow;
      sun.entry = entry;
      sun.base = irradiance.luminance();
      suns.push_back(sun);
      continue;
    }
    if (light.kind != LIGHT_RECT && light.kind != LIGHT_DISK) { err = "unsupported light kind"; return false; }
    AreaLight a;
    a.rect = light.kind == LIGHT_RECT;
    a.half_x = 0.5f * std::fabs(light.size_x);
    a.half_y = 0.5f * std::fabs(light.size_y);
    a.radius = 0.5f * std::fabs(light.size_x);
    if (a.rect) {
      a.area = 4.0f * a.half_x * a.half_y;
      a.bound = std::sqrt(a.half_x * a.half_x + a.half_y * a.half_y);
    } else {
      a.area = PI * a.radius * a.radius;
      a.bound = a.radius;
    }
    a.frame = light_frame(light.direction, light.up);
    if (!positive(a.area) || a.frame.z.length_squared() < 0.5f) continue;
    Rgb color(fmax_r(light.color.r, 0.0f), fmax_r(light.color.g, 0.0f), fmax_r(light.color.b, 0.0f));
    a.radiance = color * (power / (PI * a.area * MM2_TO_M2));
    a.exponent = spread_exponent(light.spread_deg);
    float flux = color.luminance() * power * 2.0f / (a.exponent + 2.0f);
    if (!positive(flux)) continue;
    weights.push_back(flux * reach(light.position, true, a.frame.z));
    entries.push_back(LightRef{REF_AREA, (u32)areas.size()});
    a.centre = light.position;
    a.glossy = glossy;
    a.camera = camera;
    a.shadow = shadow;
    a.entry = entry;
    a.base = a.radiance.luminance() * a.area * MM2_TO_M2;
    areas.push_back(a);
  }
  // light.rs:866-953: emissive instances whose material is flagged a light.
  const Geometry& geo = scene.geometry;
  meshes.clear();
  instance_emitter.assign(geo.instances.size(), -1);
  for (size_t i = 0; i < geo.instances.size(); ++i) {
    const Instance& inst = geo.instances[i];
    if (inst.material_id >= scene.materials.size()) continue;
    const Material& mat = scene.materials[inst.material_id];
    Rgb radiance = mat.radiance();
    if ((mat.flags & MATERIAL_FLAG_LIGHT) == 0 || !positive(radiance.luminance())) continue;
    if (inst.mesh >= geo.meshes.size()) continue;
    Transform inv;
    if (!inst.transform.inverse(inv)) continue;
    const Mesh& mesh = geo.meshes[inst.mesh];
    MeshEmitter m;
    std::vector<float> tri_areas;
    Vec3 centroid(0.0f, 0.0f, 0.0f);
    for (size_t t = 0; t < mesh.triangle_count(); ++t) {
      Vec3 p0, p1, p2;
      mesh.triangle(t, p0, p1, p2);
      Vec3 w0 = inst.transform.point(p0), w1 = inst.transform.point(p1), w2 = inst.transform.point(p2);
      Vec3 cross = (w1 - w0).cross(w2 - w0);
      float area = 0.5f * cross.length();
      if (!(area > 0.0f && std::isfinite(area))) continue;
      centroid = centroid + (w0 + w1 + w2) * (area / 3.0f);
      m.a.push_back(w0);
      m.b.push_back(w1);
      m.c.push_back(w2);
      m.normals.push_back(cross.normalize());
      tri_areas.push_back(area);
    }
    float area = -0.0f;
    for (float x : tri_areas) area = area + x;
    m.cdf = build_cdf(tri_areas);
    if (m.cdf.empty()) continue;
    m.two_sided = (inst.flags & INSTANCE_ONE_SIDED) == 0;
    float sides = m.two_sided ? 2.0f : 1.0f;
    float flux = radiance.luminance() * PI * area * MM2_TO_M2 * sides;
    m.centroid = centroid / area;
    float bound = 0.0f;
    for (size_t t = 0; t < m.a.size(); ++t) {
      bound = fmax_r(bound, (m.a[t] - m.centroid).length());
      bound = fmax_r(bound, (m.b[t] - m.centroid).length());
      bound = fmax_r(bound, (m.c[t] - m.centroid).length());
    }
    m.bound = bound;
    Vec3 mean(0.0f, 0.0f, 0.0f);
    for (size_t t = 0; t < m.normals.size(); ++t) mean = mean + m.normals[t] * tri_areas[t];
    float length = mean.length();
    if (m.two_sided || length <= 1.0e-3f * area) {
      m.axis = Vec3(0.0f, 0.0f, 1.0f);
      m.cos_o = -1.0f;
      m.sin_o = 0.0f;
    } else {
      m.axis = mean / length;
      float cos_o = 1.0f;
      for (const Vec3& n : m.normals) cos_o = fmin_r(cos_o, n.dot(m.axis));
      cos_o = clampf(cos_o, -1.0f, 1.0f);
      m.cos_o = cos_o;
      m.sin_o = std::sqrt(fmax_r(1.0f - cos_o * cos_o, 0.0f));
    }
    weights.push_back(flux * reach(m.centroid, false, Vec3(0, 0, 0)));
    m.entry = (u32)entries.size();
    entries.push_back(LightRef{REF_MESH, (u32)meshes.size()});
    instance_emitter[i] = (int)meshes.size();
    m.area = area;
    m.radiance = radiance;
    m.base = radiance.luminance() * area * MM2_TO_M2;
    meshes.push_back(m);
  }
  const Environment& e = scene.environment;
  float strength = fmax_r(e.strength, 0.0f);
  env = e.kind == 1 ? make_ramp(e.horizon, e.horizon, e.horizon, strength) : make_ramp(e.floor, e.horizon, e.zenith, strength);
  float mean = 0.5f * env.total_mass;
  env_base = PI * mean;
  has_env_entry = false;
  if (!(env.total_mass <= 0.0f)) {
    weights.push_back(4.0f * PI * PI * mean);
    env_entry = (u32)entries.size();
    has_env_entry = true;
    entries.push_back(LightRef{REF_ENV, 0});
  }
  cdf = build_cdf(weights);
  if (cdf.empty()) {
    entries.clear();
    has_env_entry = false;
  }
  flux_pmf.clear();
  for (size_t k = 0; k < cdf.size(); ++k) flux_pmf.push_back(cdf[k] - (k == 0 ? 0.0f : cdf[k - 1]));
  per_point = entries.size() >= 2 && entries.size() <= MAX_LOCAL_LIGHTS;
  return true;
}

float Lights::weight(size_t entry, Vec3 x) const {
  const LightRef& r = entries[entry];
  switch (r.kind) {
    case REF_AREA: return areas[r.index].weight(x);
    case REF_SUN: return suns[r.index].base;
    case REF_SPHERE: return spheres[r.index].weight(x);
    case REF_MESH: return meshes[r.index].weight(x);
    default: return env_base;
  }
}

// light.rs:465-481
float MeshEmitter::weight(Vec3 x) const {
  Vec3 v = x - centroid;
  float d2 = v.length_squared();
  float orientation = cos_o <= -1.0f ? 1.0f : orientation_bound(v, d2, axis, cos_o, sin_o, bound);
  if (orientation <= 0.0f) return 0.0f;
  return base * orientation / (fmax_r(d2, area) * MM2_TO_M2);
}

float Lights::emitter_weight(u32 instance, float bsdf_pdf, const Selection& from, float t, float cos) const {
  if (!is_mesh_light(instance)) return 1.0f;
  const MeshEmitter& m = meshes[instance_emitter[instance]];
  float light_pdf = pmf(m.entry, from) * t * t / (m.area * fmax_r(std::fabs(cos), 1.0e-7f));
  return found_weight(bsdf_pdf, light_pdf);
}

Selection Lights::select(Vec3 x) const {
  size_t n = entries.size();
  if (!per_point || n == 0) return select_by_flux(x);
  Selection s;
  s.point = x;
  s.total = 0.0f;
  s.by_flux = false;
  for (size_t e = 0; e < n; ++e) {
    float w = weight(e, x);
    if (e < (size_t)KEPT_WEIGHTS) s.weights[e] = w;
    s.total += w;
  }
  return s;
}

float Lights::weight_at(size_t entry, const Selection& s) const {
  return entry < (size_t)KEPT_WEIGHTS ? s.weights[entry] : weight(entry, s.point);
}

// light.rs:1118-1141: pick() on the cumulative distribution of the weights at the selection's
// vertex, built entry by entry up to the bucket that holds u: the entry, u rescaled inside its
// bucket, and weight / total.
void Lights::pick_at(const Selection& from, float u, size_t& index, float& rest, float& p) const {
  float total = from.total;
  size_t last = entries.size() - 1;
  float acc = 0.0f, lo = 0.0f;
  index = last;
  rest = 0.0f;
  p = 0.0f;
  for (size_t e = 0; e <= last; ++e) {
    float w = weight_at(e, from);
    acc += w;
    float hi = e == last ? 1.0f : acc / total;
    if (hi > u || e == last) {
      float width = hi - lo;
      rest = width > 0.0f ? clampf((u - lo) / width, 0.0f, 0.99999994f) : 0.0f;
      index = e;
      p = w / total;
      return;
    }
    lo = hi;
  }
}

void Lights::pick_static(float u, size_t& index, float& rest) const { pick(cdf, u, index, rest); }

// light.rs:1086-1111: the MIS weights of a light found by a BSDF-sampled ray and of a light
// sample, with the reference tests' switch to one technique alone.
float Lights::found_weight(float bsdf_pdf, float light_pdf) const {
  if (techniques == TECH_LIGHTS_ONLY) return positive(light_pdf) ? 0.0f : 1.0f;
  if (techniques == TECH_BSDF_ONLY) return 1.0f;
  return power_heuristic(bsdf_pdf, light_pdf);
}

float Lights::sampled_weight(float light_pdf, float bsdf_pdf) const {
  if (techniques == TECH_LIGHTS_ONLY) return 1.0f;
  if (techniques == TECH_BSDF_ONLY) return 0.0f;
  return power_heuristic(light_pdf, bsdf_pdf);
}

float Lights::pmf(u32 entry, const Selection& from) const {
  if (entry >= entries.size()) return 0.0f;
  if (from.by_flux) return flux_pmf[entry];
  if (!usable_total(from.total)) return 0.0f;
  return weight_at(entry, from) / from.total;
}

bool Lights::sample_selected(const Selection& from, float u_select, float u0, float u1, LightSample& out) const {
  size_t n = entries.size();
  if (n == 0) return false;
  Vec3 x = from.point;
  size_t index;
  float rest, p;
  if (from.by_flux) {
    pick(cdf, u_select, index, rest);
    p = pmf((u32)index, from);
  } else {
    if (!usable_total(from.total)) return false;
    pick_at(from, u_select, index, rest, p);
  }
  if (!positive(p)) return false;
  LightSample s;
  const LightRef ref = entries[index];
  if (ref.kind == REF_SUN) {
    // light.rs:1197-1223
    const SunLight& sun = suns[ref.index];
    if (sun.delta) {
      s.wi = sun.frame.z;
      s.pdf = p;
    } else {
      Vec3 local = uniform_cone(u0, u1, sun.one_minus_cos_max);
      s.wi = sun.frame.to_world(local).normalize();
      s.pdf = p / sun.solid_angle;
    }
    s.distance = INF;
    s.radiance = sun.radiance;
    s.delta = sun.delta;
    s.glossy = sun.glossy;
    s.photon = sun.glossy;
    s.shadow = sun.shadow;
  } else if (ref.kind == REF_MESH) {
    // light.rs:1270-1297: the triangle by area with the selection's remainder, a point on it.
    const MeshEmitter& m = meshes[ref.index];
    size_t t;
    float unused;
    pick(m.cdf, rest, t, unused);
    float w[3];
    uniform_triangle(u0, u1, w);
    Vec3 pt = m.a[t] * w[0] + m.b[t] * w[1] + m.c[t] * w[2];
    Vec3 d = pt - x;
    float dist2 = d.length_squared();
    float distance = std::sqrt(dist2);
    Vec3 wi = d / distance;
    float facing = -wi.dot(m.normals[t]);
    if (!positive(distance) || (!m.two_sided && facing <= 0.0f)) return false;
    float cos = std::fabs(facing);
    if (!positive(cos)) return false;
    s.wi = wi;
    s.distance = distance;
    s.radiance = m.radiance;
    s.pdf = p * dist2 / (m.area * cos);
    s.delta = false;
    s.glossy = true;
    s.photon = true;
    s.shadow = true;
  } else if (ref.kind == REF_SPHERE) {
    // light.rs:1224-1268
    const SphereLight& light = spheres[ref.index];
    float mask = light.mask(x);
    Vec3 d = light.centre - x;
    float dist2 = d.length_squared();
    if (!(mask > 0.0f && dist2 > 0.0f)) return false;
    float dist = std::sqrt(dist2);
    Vec3 towards = d / dist;
    float omc = 0.0f;
    if (light.radius > 0.0f && !light.cone(dist2, omc)) return false;
    if (omc == 0.0f) {
      s.wi = towards;
      s.distance = dist;
      s.radiance = light.intensity * (mask / (dist2 * MM2_TO_M2));
      s.pdf = p;
      s.delta = true;
    } else {
      Vec3 local = uniform_cone(u0, u1, omc);
      Vec3 wi = Frame::from_normal(towards).to_world(local).normalize();
      float b = wi.dot(d);
      float r2 = light.radius * light.radius;
      float near = b - std::sqrt(fmax_r(r2 - (d - wi * b).length_squared(), 0.0f));
      s.wi = wi;
      s.distance = fmax_r(near, 0.0f);
      s.radiance = light.radiance * mask;
      s.pdf = p / (TAU * omc);
      s.delta = false;
    }
    s.glossy = light.glossy;
    s.photon = false;
    s.shadow = light.shadow;
  } else if (ref.kind == REF_AREA) {
    const AreaLight& light = areas[ref.index];
    float x0, y0;
    if (light.rect) {
      x0 = (2.0f * u0 - 1.0f) * light.half_x;
      y0 = (2.0f * u1 - 1.0f) * light.half_y;
    } else {
      float dx, dy;
      concentric_disk(u0, u1, dx, dy);
      x0 = dx * light.radius;
      y0 = dy * light.radius;
    }
    Vec3 pt = light.centre + light.frame.x * x0 + light.frame.y * y0;
    Vec3 d = pt - x;
    float dist2 = d.length_squared();
    float distance = std::sqrt(dist2);
    Vec3 wi = d / distance;
    float cos = -wi.dot(light.frame.z);
    if (!(cos > 0.0f && distance > 0.0f)) return false;
    s.wi = wi;
    s.distance = distance;
    s.radiance = light.radiance_at(cos);
    s.pdf = p * dist2 / (light.area * cos);
    s.delta = false;
    s.glossy = light.glossy;
    s.photon = light.glossy;
    s.shadow = light.shadow;
  } else {
    Vec3 wi;
    float pdf;
    if (!env.sample(u0, u1, wi, pdf)) return false;
    s.wi = wi;
    s.distance = INF;
    s.radiance = Rgb::splat(env.value(wi.z));
    s.pdf = p * pdf;
    s.delta = false;
    s.glossy = true;
    s.photon = false;
    s.shadow = true;
  }
  if (!(s.pdf > 0.0f && std::isfinite(s.pdf) && s.radiance.luminance() > 0.0f)) return false;
  out = s;
  return true;
}

Rgb Lights::crossed_from(const Ray& ray, float t_max, bool has_bsdf, float bsdf_pdf, const Selection* from,
                         float offset, bool with_areas) const {
  Rgb sum = BLACK;
  // Without the areas (their light is the photon map's): light.rs:1344-1352.
  for (const AreaLight& light : areas) {
    if (!with_areas) break;
    bool visible = has_bsdf ? light.glossy : light.camera;
    if (!visible) continue;
    float t, cos;
    if (!light.intersect(ray, t_max, t, cos)) continue;
    float weight = 1.0f;
    if (has_bsdf) {
      float d = offset + t;
      float light_pdf = pmf(light.entry, *from) * d * d / (light.area * cos);
      weight = found_weight(bsdf_pdf, light_pdf);
    }
    sum += light.radiance_at(cos) * weight;
  }
  // light.rs:1386-1425: the spheres after the areas, seen from the vertex the ray left.
  for (const SphereLight& light : spheres) {
    bool visible = has_bsdf ? light.glossy : light.camera;
    if (!visible) continue;
    float t;
    if (!light.intersect(ray, t_max, t)) continue;
    Vec3 vertex = ray.origin - ray.dir * offset;
    float mask = light.mask(vertex);
    if (!positive(mask)) continue;
    float weight = 1.0f;
    if (has_bsdf) {
      float omc;
      if (light.cone((light.centre - vertex).length_squared(), omc) && omc > 0.0f)
        weight = found_weight(bsdf_pdf, pmf(light.entry, *from) / (TAU * omc));
      else
        weight = 0.0f;
    }
    sum += light.radiance * (mask * weight);
  }
  return sum;
}

bool Lights::camera_sees_panel(const Ray& ray, float t_max) const {
  for (const AreaLight& light : areas) {
    float t, cos;
    if (light.camera && light.intersect(ray, t_max, t, cos)) return true;
  }
  for (const SphereLight& light : spheres) {
    float t;
    if (light.camera && light.intersect(ray, t_max, t) && light.mask(ray.origin) > 0.0f) return true;
  }
  return false;
}

// light.rs:1502-1520: the environment, then the sun discs (a camera ray sees the suns flagged
// camera, a BSDF-sampled ray those flagged glossy, weighted against next-event estimation).
Rgb Lights::escaped(Vec3 dir, bool has_bsdf, float bsdf_pdf, const Selection* from, bool with_suns) const {
  Rgb sum = Rgb::splat(env.value(dir.z));
  if (has_bsdf && has_env_entry) sum = sum * found_weight(bsdf_pdf, pmf(env_entry, *from) * env.pdf(dir));
  if (!with_suns) return sum;  // light.rs:1522-1533: the environment alone
  for (const SunLight& sun : suns) {
    bool visible = has_bsdf ? sun.glossy : sun.camera;
    if (visible && sun.contains(dir)) {
      float weight = has_bsdf ? found_weight(bsdf_pdf, pmf(sun.entry, *from) / sun.solid_angle) : 1.0f;
      sum += sun.radiance * weight;
    }
  }
  return sum;
}

// ------------------------------------------------------------------ integrator/mod.rs:237-864
bool Prepared::build(const Scene& scene, std::string& err) {
  if (!lights.build(scene, err)) return false;
  u32 mode = scene.environment.background_mode;
  background = mode == 1 ? 1 : (mode == 3 ? 1 : (mode == 2 ? 2 : 0));
  Rgb bc = scene.environment.background_color;
  background_color = Rgb(fmax_r(bc.r, 0.0f), fmax_r(bc.g, 0.0f), fmax_r(bc.b, 0.0f));
  floor_mode = scene.floor.mode;
  floor_z = scene.floor.z;
  floor_bsdf = Bsdf::lambert(scene.floor.color);
  max_depth = scene.render.depth();
  clamp_indirect = scene.render.clamp_indirect;
  fallback = Material();
  has_medium.clear();
  media.clear();
  has_scatter.clear();
  scatter.clear();
  for (const Material& m : scene.materials) {
    ScatterMedium sm{};
    bool scattering = scatter_medium_of(m, sm);
    has_scatter.push_back(scattering);
    scatter.push_back(sm);
    if (scattering) {
      // integrator/mod.rs:419: a scattering medium folds the absorption into its walk.
      has_medium.push_back(false);
      media.push_back(Rgb());
      continue;
    }
    bool clear = m.transmission > 0.0f && m.metallic < 1.0f && (m.flags & MATERIAL_FLAG_THIN) == 0;
    bool on = false;
    Rgb sigma;
    if (clear && m.absorption_mm > 0.0f && !m.is_emissive()) {
      Rgb c = m.base_color;
      sigma = Rgb(-lnf_(clampf(c.r, 1.0e-4f, 1.0f)) / m.absorption_mm,
                  -lnf_(clampf(c.g, 1.0e-4f, 1.0f)) / m.absorption_mm,
                  -lnf_(clampf(c.b, 1.0e-4f, 1.0f)) / m.absorption_mm);
      on = sigma.r > 0.0f || sigma.g > 0.0f || sigma.b > 0.0f;
    }
    has_medium.push_back(on);
    media.push_back(sigma);
  }
  has_caustics = caustic_setup(scene, lights, caustics);
  return true;
}

float Prepared::floor_t(const Ray& ray, bool camera_ray) const {
  bool present = floor_mode == 1 || (floor_mode == 2 && camera_ray);
  if (!present || ray.dir.z == 0.0f) return INF;
  float t = (floor_z - ray.origin.z) / ray.dir.z;
  return (t > ray.t_min && t < ray.t_max) ? t : INF;
}

const Material& Prepared::material(const Scene& scene, u32 instance) const {
  if (instance < scene.geometry.instances.size()) {
    u32 id = scene.geometry.instances[instance].material_id;
    if (id < scene.materials.size()) return scene.materials[id];
  }
  return fallback;
}

void Prepared::add(Rgb& sum, Rgb c, bool indirect) const {
  if (!(std::isfinite(c.r) && std::isfinite(c.g) && std::isfinite(c.b))) return;
  float pk = fmax_r(fmax_r(c.r, c.g), c.b);
  if (indirect && clamp_indirect > 0.0f && pk > clamp_indirect) sum += c * (clamp_indirect / pk);
  else sum += c;
}

// ------------------------------------------------------------------ integrator/medium.rs
// medium.rs:99-110
bool is_clear_solid(const Material& m) {
  bool metal = m.kind != MATERIAL_GLASS && m.metallic >= 1.0f;
  return m.transmission > 0.0f && !metal && (m.flags & MATERIAL_FLAG_THIN) == 0 && !m.is_emissive() &&
         m.ior >= 1.0001f;
}

// medium.rs:113-141
bool scatter_medium_of(const Material& m, ScatterMedium& out) {
  if (!(m.scatter_mm > 0.0f && std::isfinite(m.scatter_mm) && is_clear_solid(m))) return false;
  float sigma = 1.0f / m.scatter_mm;
  Rgb albedo(clampf(m.scatter_color.r, 0.0f, 1.0f), clampf(m.scatter_color.g, 0.0f, 1.0f),
             clampf(m.scatter_color.b, 0.0f, 1.0f));
  Rgb sigma_a((1.0f - albedo.r) * sigma, (1.0f - albedo.g) * sigma, (1.0f - albedo.b) * sigma);
  if (m.absorption_mm > 0.0f) {
    Rgb c = m.base_color;
    sigma_a += Rgb(-lnf_(clampf(c.r, 1.0e-4f, 1.0f)) / m.absorption_mm,
                   -lnf_(clampf(c.g, 1.0e-4f, 1.0f)) / m.absorption_mm,
                   -lnf_(clampf(c.b, 1.0e-4f, 1.0f)) / m.absorption_mm);
  }
  Rgb sigma_s = albedo * sigma;
  out.sigma_s = sigma_s;
  out.sigma_a = sigma_a;
  out.sigma_t = sigma_s + sigma_a;
  out.g = std::isfinite(m.scatter_g) ? clampf(m.scatter_g, -MAX_G, MAX_G) : 0.0f;
  out.ior = m.ior;
  return true;
}

namespace {
// medium.rs:143-150
float phase(float cos, float g) {
  float d = 1.0f + g * g - 2.0f * g * cos;
  return (1.0f - g * g) / (4.0f * PI * d * std::sqrt(fmax_r(d, 1.0e-12f)));
}

// medium.rs:152-169
Vec3 sample_phase(Vec3 dir, float g, float u0, float u1, float& density) {
  float cos;
  if (std::fabs(g) < 1.0e-3f) {
    cos = 1.0f - 2.0f * u0;
  } else {
    float s = (1.0f - g * g) / (1.0f - g + 2.0f * g * u0);
    cos = clampf((1.0f + g * g - s * s) / (2.0f * g), -1.0f, 1.0f);
  }
  float sin = std::sqrt(fmax_r(1.0f - cos * cos, 0.0f));
  float sp, cp;
  sin_cos(TAU * u1, sp, cp);
  Frame frame = Frame::from_normal(dir);
  density = phase(cos, g);
  return frame.to_world(Vec3(sin * cp, sin * sp, cos)).normalize();
}

inline float channel(Rgb c, u32 k) { return k == 0 ? c.r : (k == 1 ? c.g : c.b); }
inline Rgb transmittance(Rgb sigma, float t) {
  return Rgb(expf_(-sigma.r * t), expf_(-sigma.g * t), expf_(-sigma.b * t));
}

enum WalkEnd { WALK_CROSSED, WALK_EXIT, WALK_ESCAPED, WALK_ABSORBED };

// medium.rs:226-360: the random walk inside `m` from `ray` until the path leaves the solid.
WalkEnd walk(const Scene& scene, const PixelSampler& sampler, const ScatterMedium& m, Ray& ray, Rgb& throughput,
             u32& path_events, u32 bounce, RenderStats* stats, RayHit& hit_out, float& eta2_out, float& pdf_out) {
  const Geometry& geo = scene.geometry;
  u32 first = medium_dim(bounce, 0);
  float uh = sampler.get1(first + MEDIUM_OFFSET_HERO) * 3.0f;
  u32 hero = (u32)std::min<u64>(f_to_usize(uh), 2);
  float sigma_hero = channel(m.sigma_t, hero);
  Rgb w = WHITE, ratio = WHITE;
  u32 events = 0, steps = 0;
  float pdf = 0.0f;
  Rgb entry = throughput;
  auto effective = [](Rgb w_, Rgb r_) { return w_ * (3.0f / (r_.r + r_.g + r_.b)); };
  for (;;) {
    if (stats) { stats->rays++; stats->nearest++; }
    RayHit hit;
    bool has_hit = geo.trace(ray, false, hit, stats ? &stats->nearest_trav : nullptr);
    float t_hit = has_hit ? hit.t : INF;
    bool capped = path_events >= MEDIUM_EVENTS || std::isnan(sigma_hero) || sigma_hero <= 0.0f;
    u32 dims = medium_dim(bounce, steps);
    steps += 1;
    float t = INF;
    if (!capped) {
      float u = sampler.get1(dims + MEDIUM_OFFSET_DISTANCE);
      t = -lnf_(1.0f - u) / sigma_hero;
    }
    if (t < t_hit) {
      // A scattering event.
      Rgb tr = transmittance(m.sigma_t, t);
      Rgb p = m.sigma_t * tr;
      Rgb f = m.sigma_s * tr;
      float ph = channel(p, hero);
      if (std::isnan(ph) || ph <= 0.0f) return WALK_ABSORBED;
      w = w * (f / ph);
      ratio = ratio * (p / ph);
      float mx = peak(ratio);
      if (mx > 0.0f) {
        w = w / mx;
        ratio = ratio / mx;
      }
      events += 1;
      path_events += 1;
      Rgb current = entry * effective(w, ratio);
      if (!(peak(current) > 0.0f && std::isfinite(peak(current)))) return WALK_ABSORBED;
      if (events > MEDIUM_ROULETTE_AFTER) {
        float survive = clampf(peak(current) / peak(entry), 0.05f, 1.0f);
        if (sampler.get1(dims + MEDIUM_OFFSET_ROULETTE) >= survive) return WALK_ABSORBED;
        w = w / survive;
      }
      Vec3 x = ray.at(t);
      float u0, u1;
      sampler.get2(dims + MEDIUM_OFFSET_PHASE, u0, u1);
      float density;
      Vec3 dir = sample_phase(ray.dir, m.g, u0, u1, density);
      pdf = density;
      ray = Ray(x, dir);
      continue;
    }
    // No scattering before the surface (or the step cap is reached: the path then goes to the
    // surface with only the absorption applied).
    if (std::isfinite(t_hit)) {
      if (capped) {
        w = w * transmittance(m.sigma_a, t_hit);
      } else {
        Rgb p = transmittance(m.sigma_t, t_hit);
        float ph = channel(p, hero);
        if (std::isnan(ph) || ph <= 0.0f) return WALK_ABSORBED;
        w = w * (p / ph);
        ratio = ratio * (p / ph);
      }
    }
    throughput = entry * effective(w, ratio);
    if (!(peak(throughput) > 0.0f && std::isfinite(peak(throughput)))) return WALK_ABSORBED;
    if (!has_hit) return WALK_ESCAPED;
    hit_out = hit;
    if (events == 0) return WALK_CROSSED;
    Hit details;
    if (!geo.hit_details(ray, hit, details)) return WALK_ABSORBED;
    Vec3 ng = details.geometric_normal;
    float cos = clampf(-ray.dir.dot(ng), 0.0f, 1.0f);
    // medium.rs:218-224: leave with the probability of the Fresnel transmittance out.
    bool leave = capped || sampler.get1(dims + MEDIUM_OFFSET_FRESNEL) >= fresnel_dielectric(cos, 1.0f / fmax_r(m.ior, 1.0001f));
    if (leave) {
      eta2_out = m.ior * m.ior;
      pdf_out = pdf;
      return WALK_EXIT;
    }
    // Fresnel sent it inside again: this counts against the walk's step cap like a scattering event.
    path_events += 1;
    Vec3 dir = (ray.dir + ng * (2.0f * cos)).normalize();
    ray = Ray(offset_origin(details.point, ng), dir);
  }
}
}  // namespace

float phase_for_tests(float cos, float g) { return phase(cos, g); }
Vec3 sample_phase_for_tests(Vec3 dir, float g, float u0, float u1, float& density) {
  return sample_phase(dir, g, u0, u1, density);
}

static Ray shadow_ray(Vec3 origin, Vec3 dir, float distance) {
  Ray r(origin, dir);
  if (std::isfinite(distance)) r.t_max = distance * (1.0f - 1.0e-4f);
  return r;
}

// ------------------------------------------------------------------ integrator/mod.rs:1140-1891
void Texture::tap(float x, float y, u32 tx[4], u32 ty[4], float wt[4], float& ax, float& ay) const {
  float u = (x - ox) * sx, v = (y - oy) * sy;
  if (!std::isfinite(u)) u = 0.0f;
  if (!std::isfinite(v)) v = 0.0f;
  u = u - std::floor(u);
  v = v - std::floor(v);
  float fx = u * (float)w - 0.5f, fy = v * (float)h - 0.5f;
  float x0f = std::floor(fx), y0f = std::floor(fy);
  ax = fx - x0f;
  ay = fy - y0f;
  int ix = (int)x0f, iy = (int)y0f;
  u32 x0 = (u32)(ix < 0 ? ix + (int)w : (ix >= (int)w ? ix - (int)w : ix));
  u32 y0 = (u32)(iy < 0 ? iy + (int)h : (iy >= (int)h ? iy - (int)h : iy));
  u32 x1 = x0 + 1 == w ? 0 : x0 + 1, y1 = y0 + 1 == h ? 0 : y0 + 1;
  tx[0] = x0; ty[0] = y0;
  tx[1] = x1; ty[1] = y0;
  tx[2] = x0; ty[2] = y1;
  tx[3] = x1; ty[3] = y1;
  wt[0] = (1.0f - ax) * (1.0f - ay);
  wt[1] = ax * (1.0f - ay);
  wt[2] = (1.0f - ax) * ay;
  wt[3] = ax * ay;
}

Rgb Texture::value(float x, float y) const {
  u32 tx[4], ty[4];
  float wt[4], ax, ay;
  tap(x, y, tx, ty, wt, ax, ay);
  float c[3];
  for (int ch = 0; ch < 3; ++ch) {
    float a = rgb[3 * ((size_t)ty[0] * w + tx[0]) + ch], b = rgb[3 * ((size_t)ty[1] * w + tx[1]) + ch];
    float d = rgb[3 * ((size_t)ty[2] * w + tx[2]) + ch], e = rgb[3 * ((size_t)ty[3] * w + tx[3]) + ch];
    c[ch] = ((a * wt[0] + b * wt[1]) + d * wt[2]) + e * wt[3];
  }
  return Rgb(c[0], c[1], c[2]);
}

namespace {
float dot3(Rgb a, Rgb b) { return a.r * b.r + a.g * b.g + a.b * b.b; }

// Where a vertex's gradient goes: material `mat` (or none) and, for a textured one, its texels.
struct GradTarget {
  bool on = false;
  u32 mat = 0;
  bool tex = false;
  u32 tx[4], ty[4];
  float wt[4];
};

void commit(const Scene& scene, const AdjointPath& a, const GradTarget& t, const float g[5]) {
  if (!t.on) return;
  for (int k = 0; k < 5; ++k) {
    if (t.tex && k >= 1 && k <= 3) continue;
    a.grad_mat[5 * (size_t)t.mat + k] += (double)g[k];
  }
  if (t.tex) {
    const Texture& tx = scene.texture;
    for (int q = 0; q < 4; ++q)
      for (int c = 0; c < 3; ++c)
        a.grad_tex[3 * ((size_t)t.ty[q] * tx.w + t.tx[q]) + c] += (double)(g[1 + c] * t.wt[q]);
  }
}
}  // namespace

void radiance(const Scene& scene, const Prepared& prep, const PixelSampler& sampler, Ray camera_ray, Rgb& out_sum,
              float& out_alpha, RenderStats* stats, const AdjointPath* adjoint, const PhotonMap* photons) {
  // Caustics (integrator/mod.rs:1084-1113): where the path is in the split between the photon map
  // and the tracer: before its first receiver, after one, or after one and casters since.
  enum { EYE, AFTER_RECEIVER, AFTER_RECEIVER_CASTERS };
  const PhotonMap* map = prep.has_caustics ? photons : nullptr;
  int cstate = EYE;
  // The adjoint's pending term: d log f / d parameter of the last BSDF sample, and its target.
  GradTarget pending_target;
  Rgb pending_dl[5];
  const Lights& lights = prep.lights;
  const Geometry& geo = scene.geometry;
  Ray ray = camera_ray;
  Rgb throughput = WHITE;
  Rgb sum = BLACK;
  float alpha = 1.0f;
  bool has_pdf = false;
  float bsdf_pdf = 0.0f;
  Selection sampled_at = lights.select_by_flux(camera_ray.origin);
  u32 bounce = 0;
  bool primary = true;
  float offset = 0.0f;
  bool in_medium = false;
  Rgb sigma;
  // The scattering medium the ray is inside (a material index, -1 for none), the steps its walks
  // took, and whether the last vertex was a medium's exit (integrator/mod.rs:1177-1185).
  int scatter = -1;
  u32 medium_events = 0;
  bool exit_pdf = false;
  for (;;) {
    // A scattering walk that reached its solid's surface leaves by the exit lobe, whose albedo is carried here.
    bool exit = false;
    float exit_eta2 = 0.0f;
    RayHit hit;
    bool has_hit;
    if (scatter < 0) {
      if (stats) { stats->rays++; stats->nearest++; }
      has_hit = geo.trace(ray, primary, hit, stats ? &stats->nearest_trav : nullptr);
    } else {
      // integrator/mod.rs:1209-1242
      float eta2 = 0.0f, pdf = 0.0f;
      WalkEnd end = walk(scene, sampler, prep.scatter[scatter], ray, throughput, medium_events, bounce, stats, hit,
                         eta2, pdf);
      if (end == WALK_ABSORBED) break;
      has_hit = end != WALK_ESCAPED;
      if (end == WALK_EXIT) {
        scatter = -1;
        exit = true;
        exit_eta2 = eta2;
        has_pdf = true;
        bsdf_pdf = pdf;
        primary = false;
        offset = 0.0f;
      }
    }
    float t_geometry = has_hit ? hit.t : INF;
    float t_floor = exit ? INF : prep.floor_t(ray, primary);
    float t_nearest = fmin_r(t_geometry, t_floor);
    if (adjoint && pending_target.on) {
      // The radiance gathered after the last vertex: the forward's total minus what is summed.
      Rgb rest(adjoint->total.r - sum.r, adjoint->total.g - sum.g, adjoint->total.b - sum.b);
      float g[5];
      for (int k = 0; k < 5; ++k) g[k] = dot3(adjoint->adj, rest * pending_dl[k]);
      commit(scene, *adjoint, pending_target, g);
    }
    pending_target = GradTarget{};
    if (in_medium && std::isfinite(t_nearest)) {
      Rgb bl(expf_(-sigma.r * t_nearest), expf_(-sigma.g * t_nearest), expf_(-sigma.b * t_nearest));
      throughput = throughput * bl;
    }
    // Light of photon lights arriving now is the photon map's (integrator/mod.rs:1261-1272): the
    // area lights crossed are left out, the spheres stay.
    bool dropped = map && cstate == AFTER_RECEIVER_CASTERS;
    // The walk's last segment ends at its exit vertex inside the solid, so no light panel lies on it.
    Rgb crossed = exit ? BLACK : lights.crossed_from(ray, t_nearest, has_pdf, bsdf_pdf, &sampled_at, offset, !dropped);
    prep.add(sum, throughput * crossed, bounce >= 2);
    if (t_nearest == INF) {
      if (primary) {
        if (prep.background == 0) {
          Rgb seen = lights.escaped(ray.dir, false, 0.0f, nullptr);
          if (scene.camera.exterior != 1.0f) seen = seen * scene.camera.exterior;
          sum += throughput * seen;
        } else if (prep.background == 1) {
          sum += throughput * prep.background_color;
        } else {
          alpha = peak(crossed) > 0.0f ? 1.0f : 0.0f;
        }
      } else {
        Rgb sky = lights.escaped(ray.dir, has_pdf, bsdf_pdf, &sampled_at, !(dropped && has_pdf));
        prep.add(sum, throughput * sky, bounce >= 2);
      }
      break;
    }
    // The surface.
    Vec3 s_point, s_hit_point, s_normal, s_geometric;
    bool s_has_instance;
    u32 s_instance = 0;
    bool s_front = true, s_outside = true;
    if (t_floor < t_geometry) {
      Vec3 p = ray.at(t_floor);
      Vec3 point(p.x, p.y, prep.floor_z);
      Vec3 n(0.0f, 0.0f, ray.dir.z < 0.0f ? 1.0f : -1.0f);
      s_point = point;
      s_hit_point = point;
      s_normal = n;
      s_geometric = n;
      s_has_instance = false;
      s_outside = true;
    } else {
      Hit details;
      if (!geo.hit_details(ray, hit, details)) break;
      // bevel::shading_point without a bevel (bevel.rs:232-252)
      Vec3 ng = details.geometric_normal;
      Vec3 normal = details.normal;
      if (normal.dot(ng) < 0.0f) normal = -normal;
      if (normal.dot(ray.dir) >= -1.0e-4f) normal = ng;
      s_point = geo.terminator_point(details);
      s_hit_point = details.point;
      s_normal = normal;
      s_geometric = ng;
      s_has_instance = true;
      s_instance = details.instance;
      s_front = details.front_face;
      s_outside = details.outside;
    }
    const Material* material = s_has_instance ? &prep.material(scene, s_instance) : nullptr;
    if (material && material->is_emissive()) {
      bool one_sided = (geo.instances[s_instance].flags & INSTANCE_ONE_SIDED) != 0;
      if ((s_front || !one_sided) && !(dropped && lights.is_mesh_light(s_instance))) {
        // integrator/mod.rs:1548-1568: a BSDF-sampled ray weighs an emitter sampled as a light
        // against next-event estimation; others have a weight of one.
        float weight = 1.0f;
        if (has_pdf) {
          float cos = s_geometric.dot(ray.dir);
          weight = lights.emitter_weight(s_instance, bsdf_pdf, sampled_at, offset + t_geometry, cos);
        }
        prep.add(sum, throughput * material->radiance() * weight, bounce >= 2);
      }
      break;
    }
    if (bounce >= prep.max_depth) break;
    // A textured material's base colour comes from its texture at the hit point.
    u32 mat_id = s_has_instance ? geo.instances[s_instance].material_id : 0xffffffffu;
    bool has_mat = s_has_instance && mat_id < scene.materials.size();
    Material shaded = material ? *material : Material{};
    GradTarget target;
    if (has_mat) {
      target.on = true;
      target.mat = mat_id;
      if (scene.texture.on() && mat_id == scene.texture.material) {
        float ax, ay;
        scene.texture.tap(s_hit_point.x, s_hit_point.y, target.tx, target.ty, target.wt, ax, ay);
        shaded.base_color = scene.texture.value(s_hit_point.x, s_hit_point.y);
        target.tex = true;
      }
    }
    Bsdf bsdf = exit ? Bsdf::medium_exit(exit_eta2) : (material ? Bsdf::surface(shaded, s_outside) : prep.floor_bsdf);
    Frame frame = Frame::from_normal(s_normal);
    Vec3 wo = frame.to_local(-ray.dir);
    Vec3 ng = s_geometric;
    Vec3 origin = offset_origin(s_point, ng);
    u32 dims = bounce_dim(bounce);
    // Caustics (integrator/mod.rs:1672-1694): every receiver gathers the photons (the first one's
    // light direct, a later one's indirect); the state after this vertex says whether its light
    // of photon lights is the map's.
    int next_cstate = cstate;
    if (map) {
      bool caster = s_has_instance && prep.caustics.is_caster_instance(s_instance);
      if (!caster) {
        u32 budget = prep.max_depth > bounce + 1 ? prep.max_depth - (bounce + 1) : 0u;
        Vec3 side = exit ? -s_geometric : s_geometric;
        Rgb c = map->gather(s_hit_point, side, budget,
                            [&](Vec3 wi) { return bsdf.eval(wo, frame.to_local(wi), Lobes::All); });
        prep.add(sum, throughput * c, cstate != EYE);
        next_cstate = AFTER_RECEIVER;
      } else if (cstate != EYE) {
        next_cstate = AFTER_RECEIVER_CASTERS;
      }
    }
    bool drop_nee = map && next_cstate == AFTER_RECEIVER_CASTERS;
    Selection selection = lights.select(s_point);
    LightSample ls;
    float ul0, ul1;
    sampler.get2(dims + OFFSET_LIGHT, ul0, ul1);
    if (lights.sample_selected(selection, sampler.get1(dims + OFFSET_LIGHT_SELECT), ul0, ul1, ls) &&
        !(drop_nee && ls.photon)) {
      Vec3 wi = frame.to_local(ls.wi);
      Rgb f = bsdf.eval(wo, wi, ls.glossy ? Lobes::All : Lobes::Diffuse);
      if (peak(f) > 0.0f) {
        Vec3 from = wi.z < 0.0f ? offset_origin(s_hit_point, -ng) : origin;
        float distance = ls.distance;
        bool arrives = true;
        if (ls.shadow) {
          if (stats) { stats->rays++; stats->any++; }
          Ray sr = shadow_ray(from, ls.wi, distance);
          bool blocked = (prep.floor_mode == 1 && prep.floor_t(sr, false) < INF) ||
                         geo.occluded(sr, stats ? &stats->any_trav : nullptr);
          arrives = !blocked;
        }
        if (arrives) {
          float weight = (ls.delta || !ls.glossy) ? 1.0f : prep.lights.sampled_weight(ls.pdf, bsdf.pdf(wo, wi));
          Rgb c = throughput * f * ls.radiance * (std::fabs(wi.z) * weight / ls.pdf);
          if (adjoint && has_mat && std::isfinite(c.r) && std::isfinite(c.g) && std::isfinite(c.b)) {
            Rgb d[5];
            bsdf_param_derivs_exact(shaded, s_outside, wo, wi, ls.glossy ? Lobes::All : Lobes::Diffuse, d);
            float pk = fmax_r(fmax_r(c.r, c.g), c.b);
            float scale = (bounce >= 1 && prep.clamp_indirect > 0.0f && pk > prep.clamp_indirect)
                              ? prep.clamp_indirect / pk : 1.0f;
            float g[5];
            for (int k = 0; k < 5; ++k)
              g[k] = dot3(adjoint->adj, throughput * d[k] * ls.radiance * (std::fabs(wi.z) * weight / ls.pdf * scale));
            commit(scene, *adjoint, target, g);
          }
          prep.add(sum, c, bounce >= 1);
        }
      }
    }
    BsdfSample bs;
    float ub0, ub1;
    sampler.get2(dims + OFFSET_BSDF, ub0, ub1);
    if (!bsdf.sample(wo, ub0, ub1, bs)) break;
    bool transmitted = bs.wi.z < 0.0f;
    Vec3 wi = bs.straight ? ray.dir : frame.to_world(bs.wi).normalize();
    if (transmitted) {
      if (wi.dot(ng) >= 0.0f) break;
    } else if (wi.dot(ng) <= 0.0f) {
      break;
    }
    throughput = throughput * bs.weight;
    if (bounce >= 2) {
      float survive = clampf(peak(throughput), 0.05f, 1.0f);
      if (sampler.get1(dims + OFFSET_ROULETTE) >= survive) break;
      throughput = throughput / survive;
    }
    if (!(peak(throughput) > 0.0f && std::isfinite(peak(throughput)))) break;
    if (adjoint && has_mat) {
      Rgb f_s = bsdf.eval(wo, bs.wi, Lobes::All);
      Rgb d[5];
      bsdf_param_derivs_exact(shaded, s_outside, wo, bs.wi, Lobes::All, d);
      for (int k = 0; k < 5; ++k)
        pending_dl[k] = Rgb(f_s.r > 0.0f ? d[k].r / f_s.r : 0.0f, f_s.g > 0.0f ? d[k].g / f_s.g : 0.0f,
                            f_s.b > 0.0f ? d[k].b / f_s.b : 0.0f);
      pending_target = target;
    }
    Vec3 next_origin = transmitted ? offset_origin(s_hit_point, -ng) : origin;
    if (bs.straight) {
      offset += (next_origin - ray.origin).dot(ray.dir);
    } else {
      has_pdf = true;
      bsdf_pdf = bs.pdf;
      sampled_at = selection;
      primary = false;
      offset = 0.0f;
      cstate = next_cstate;
      // integrator/mod.rs:1822-1857: a BSDF sample transmitted into a solid right after an exit
      // vertex is probed for a thin wall (one ray counted; the scenes with media have no walls).
      bool next_exit = false;
      if (exit) {
        next_exit = true;
      } else if (exit_pdf && s_has_instance && transmitted && s_outside) {
        if (stats) stats->rays++;
      }
      exit_pdf = next_exit;
      if (transmitted) {
        if (s_outside && material && s_has_instance) {
          u32 id = geo.instances[s_instance].material_id;
          in_medium = id < prep.has_medium.size() && prep.has_medium[id];
          if (in_medium) sigma = prep.media[id];
          scatter = id < prep.has_scatter.size() && prep.has_scatter[id] ? (int)id : -1;
        } else {
          in_medium = false;
          scatter = -1;
        }
      }
    }
    ray = Ray(next_origin, wi);
    bounce += 1;
  }
  out_sum = sum;
  out_alpha = alpha;
}

// ------------------------------------------------------------------ render.rs:44-497, pixel_filter.rs
static float filter_offset(u32 filter, float u) {
  if (filter == 1) return 0.5f;
  if (filter == 2) return 0.5f + (u < 0.5f ? std::sqrt(2.0f * u) - 1.0f : 1.0f - std::sqrt(2.0f - 2.0f * u));
  return u;
}

void camera_sample(const Scene& scene, const PixelSampler& sampler, u32 x, u32 y, Ray& ray) {
  u32 width = std::max<u32>(scene.render.width, 1), height = std::max<u32>(scene.render.height, 1);
  float aspect = (float)width / (float)height;
  float du, dv;
  sampler.get2(DIM_CAMERA, du, dv);
  du = filter_offset(scene.render.pixel_filter, du);
  dv = filter_offset(scene.render.pixel_filter, dv);
  float u = ((float)x + du) / (float)width;
  float v = ((float)y + dv) / (float)height;
  float lx = 0.0f, ly = 0.0f;
  if (scene.camera.lens_radius() > 0.0f) {
    float a, b;
    sampler.get2(DIM_LENS, a, b);
    concentric_disk(a, b, lx, ly);
  }
  ray = scene.camera.generate_ray(aspect, u, v, lx, ly);
}

std::vector<TileRect> tile_grid(u32 width, u32 height, u32 size) {
  size = std::max<u32>(size, 1);
  std::vector<TileRect> tiles;
  for (u32 y = 0; y < height;) {
    u32 h = std::min(size, height - y);
    for (u32 x = 0; x < width;) {
      u32 w = std::min(size, width - x);
      tiles.push_back(TileRect{x, y, w, h});
      x += w;
    }
    y += h;
  }
  return tiles;
}

void Film::init(u32 ox, u32 oy, u32 w, u32 h) {
  origin_x = ox;
  origin_y = oy;
  width = w;
  height = h;
  size_t n = (size_t)w * h;
  rgb.assign(n * 3, 0.0f);
  weight.assign(n, 0.0f);
  alpha.assign(n, 0.0f);
  lum2.assign(n, 0.0f);
}

void Film::splat_alpha(u32 x, u32 y, Rgb c, float a) {
  if (x < origin_x || y < origin_y) return;
  u32 lx = x - origin_x, ly = y - origin_y;
  if (lx >= width || ly >= height) return;
  size_t i = (size_t)ly * width + lx;
  rgb[i * 3] += c.r;
  rgb[i * 3 + 1] += c.g;
  rgb[i * 3 + 2] += c.b;
  weight[i] += 1.0f;
  alpha[i] += a;
  float l = c.luminance();
  lum2[i] += l * l;
}

std::vector<float> Film::resolve() const {
  std::vector<float> out(rgb.size(), 0.0f);
  for (size_t i = 0; i < weight.size(); ++i) {
    float w = weight[i];
    if (w > 0.0f) {
      float inv = 1.0f / w;
      for (int c = 0; c < 3; ++c) out[i * 3 + c] = rgb[i * 3 + c] * inv;
    }
  }
  return out;
}

void render_tile(const Scene& scene, const Prepared& prep, TileRect tile, u32 sample_index, Film& film,
                 RenderStats* stats, const PhotonMap* photons) {
  u32 width = scene.render.width, height = scene.render.height;
  u32 x_end = std::min<u32>(tile.x + tile.w, width);
  u32 y_end = std::min<u32>(tile.y + tile.h, height);
  for (u32 y = tile.y; y < y_end; ++y) {
    for (u32 x = tile.x; x < x_end; ++x) {
      PixelSampler sampler = PixelSampler::make(scene.render.seed, x, y, sample_index);
      Ray ray;
      camera_sample(scene, sampler, x, y, ray);
      Rgb c;
      float a;
      radiance(scene, prep, sampler, ray, c, a, stats, nullptr, photons);
      film.splat_alpha(x, y, c * WHITE, a);
    }
  }
}

}  // namespace orc
