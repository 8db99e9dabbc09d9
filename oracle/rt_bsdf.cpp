// ORACLE — test infrastructure only (see rt_core.hpp). Citations in rt_bsdf.hpp.
#include "rt_bsdf.hpp"
#include <cstdio>
#include <atomic>
#include <mutex>
#include <string>
#include <thread>

namespace orc {

// ------------------------------------------------------------------ bsdf.rs:89-300
float Ggx::d(Vec3 h) const {
  // isotropic only (ggx_d, bsdf.rs:239-247)
  float t = a2 * h.z * h.z + (h.x * h.x + h.y * h.y);
  if (t <= 0.0f) return 0.0f;
  return a2 / (PI * t * t);
}
float Ggx::lambda(Vec3 w) const {
  float cos2 = w.z * w.z;
  if (cos2 <= 0.0f) return INF;
  float tan2 = (w.x * w.x + w.y * w.y) / cos2;
  return (std::sqrt(1.0f + a2 * tan2) - 1.0f) * 0.5f;
}

static inline float schlick(float f0, float f90, float cos) {
  float m = clampf(1.0f - cos, 0.0f, 1.0f);
  float m2 = m * m;
  return f0 + (f90 - f0) * (m2 * m2 * m);
}

float fresnel_dielectric(float cos_i, float eta) {
  cos_i = clampf(cos_i, 0.0f, 1.0f);
  float sin2_t = (1.0f - cos_i * cos_i) / (eta * eta);
  if (sin2_t >= 1.0f) return 1.0f;
  float cos_t = std::sqrt(1.0f - sin2_t);
  float rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
  float rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
  return 0.5f * (rs * rs + rp * rp);
}

static inline float thin_reflectance(float cos, float eta) {
  float r = fresnel_dielectric(cos, eta);
  return 2.0f * r / (1.0f + r);
}

bool refract(Vec3 wo, Vec3 h, float eta, Vec3& out) {
  float cos_i = wo.dot(h);
  float sin2_t = (1.0f - cos_i * cos_i) / (eta * eta);
  if (sin2_t >= 1.0f) return false;
  float cos_t = std::sqrt(1.0f - sin2_t);
  out = (h * (cos_i / eta - cos_t) - wo / eta).normalize();
  return true;
}

Vec3 sample_vndf(Vec3 wo, float ax, float ay, float u0, float u1) {
  Vec3 vh = Vec3(ax * wo.x, ay * wo.y, wo.z).normalize();
  float lensq = vh.x * vh.x + vh.y * vh.y;
  Vec3 t1 = lensq > 0.0f ? Vec3(-vh.y, vh.x, 0.0f) / std::sqrt(lensq) : Vec3(1.0f, 0.0f, 0.0f);
  Vec3 t2 = vh.cross(t1);
  float r = std::sqrt(u0);
  float sp, cp;
  sin_cos(TAU * u1, sp, cp);
  float p1 = r * cp;
  float s = 0.5f * (1.0f + vh.z);
  float p2 = (1.0f - s) * std::sqrt(fmax_r(1.0f - p1 * p1, 0.0f)) + s * (r * sp);
  Vec3 nh = t1 * p1 + t2 * p2 + vh * std::sqrt(fmax_r(1.0f - p1 * p1 - p2 * p2, 0.0f));
  return Vec3(ax * nh.x, ay * nh.y, fmax_r(nh.z, 1.0e-7f)).normalize();
}

// ------------------------------------------------------------------ energy.rs:55-554
namespace energy {
namespace {
constexpr u64 SEED = 0xe9ae9a;
constexpr u32 STRATA = 64;
constexpr float MIN_MU = 1.0f / 2048.0f;
constexpr u32 AVG_STEPS = 4096;
constexpr float UNIT = 1.0f / 65535.0f;
const float MAX_X = (MAX_IOR - 1.0f) / (MAX_IOR + 1.0f);

float node_mu(int i) {
  float t = (float)i / (float)(NODES_MU - 1);
  return fmax_r(t * t, MIN_MU);
}
float node_alpha(int j) {
  float r = (float)j / (float)(NODES_ROUGHNESS - 1);
  return r * r;
}
float node_ior(int k) {
  float x = MAX_X * (float)k / (float)(NODES_IOR - 1);
  return (1.0f + x) / (1.0f - x);
}
Vec3 outgoing(float mu) { return Vec3(std::sqrt(fmax_r(1.0f - mu * mu, 0.0f)), 0.0f, mu); }
void jitter(Pcg32& rng, u32 s, u32 strata, float& a, float& b) {
  u32 i = s % strata, j = s / strata;
  float inv = 1.0f / (float)strata;
  a = fmin_r(((float)i + rng.next_f32()) * inv, 0.99999994f);
  b = fmin_r(((float)j + rng.next_f32()) * inv, 0.99999994f);
}
double g2_over_g1(const Ggx& g, Vec3 wo, Vec3 wi) {
  float lo = g.lambda(wo), li = g.lambda(wi);
  return (double)((1.0f + lo) / (1.0f + lo + li));
}
double integrate_albedo(float mu, float alpha, u64 stream) {
  if (alpha <= 0.0f) return 1.0;
  Ggx g = Ggx::make(alpha);
  Vec3 wo = outgoing(mu);
  Pcg32 rng = Pcg32::seeded(SEED, stream);
  double sum = 0.0;
  u32 n = STRATA * STRATA;
  for (u32 s = 0; s < n; ++s) {
    float a, b;
    jitter(rng, s, STRATA, a, b);
    Vec3 h = sample_vndf(wo, alpha, alpha, a, b);
    Vec3 wi = h * (2.0f * wo.dot(h)) - wo;
    if (wi.z > 0.0f) sum += g2_over_g1(g, wo, wi);
  }
  return sum / (double)n;
}
double integrate_schlick(float mu, float alpha, u64 stream) {
  if (alpha <= 0.0f) {
    double m = (double)(1.0f - mu);
    return m * m * m * m * m;
  }
  Ggx g = Ggx::make(alpha);
  Vec3 wo = outgoing(mu);
  Pcg32 rng = Pcg32::seeded(SEED ^ 0x5c41, stream);
  double sum = 0.0;
  u32 n = STRATA * STRATA;
  for (u32 s = 0; s < n; ++s) {
    float a, b;
    jitter(rng, s, STRATA, a, b);
    Vec3 h = sample_vndf(wo, alpha, alpha, a, b);
    float c = wo.dot(h);
    Vec3 wi = h * (2.0f * c) - wo;
    if (wi.z > 0.0f) {
      double m = (double)clampf(1.0f - c, 0.0f, 1.0f);
      sum += m * m * m * m * m * g2_over_g1(g, wo, wi);
    }
  }
  return sum / (double)n;
}
double integrate_clear(float mu, float alpha, float eta, u64 stream) {
  if (alpha <= 0.0f || std::fabs(eta - 1.0f) < 1.0e-6f) return 1.0;
  Ggx g = Ggx::make(alpha);
  Vec3 wo = outgoing(mu);
  Pcg32 rng = Pcg32::seeded(SEED ^ 0xc1ea, stream);
  double sum = 0.0;
  u32 n = STRATA * STRATA;
  for (u32 s = 0; s < n; ++s) {
    float a, b;
    jitter(rng, s, STRATA, a, b);
    Vec3 h = sample_vndf(wo, alpha, alpha, a, b);
    float cos = wo.dot(h);
    float f = fresnel_dielectric(cos, eta);
    Vec3 wr = h * (2.0f * cos) - wo;
    if (wr.z > 0.0f) sum += (double)f * g2_over_g1(g, wo, wr);
    if (f < 1.0f) {
      Vec3 wt;
      if (refract(wo, h, eta, wt) && wt.z < 0.0f) sum += (double)(1.0f - f) * g2_over_g1(g, wo, wt);
    }
  }
  return sum / (double)n;
}
uint16_t quantise(double v) {
  double c = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
  return (uint16_t)std::round(c * 65535.0);
}
void locate(float x, int n, int& i, float& w) {
  float t = clampf(x, 0.0f, 1.0f) * (float)(n - 1);
  u64 ti = f_to_usize(t);
  i = (int)std::min<u64>(ti, (u64)(n - 2));
  w = t - (float)i;
}
float lookup(const uint16_t* table, float mu, float alpha) {
  int i, j;
  float u, v;
  locate(std::sqrt(fmax_r(mu, 0.0f)), NODES_MU, i, u);
  locate(std::sqrt(fmax_r(alpha, 0.0f)), NODES_ROUGHNESS, j, v);
  auto at = [&](int jj, int ii) { return (float)table[jj * NODES_MU + ii] * UNIT; };
  float low = at(j, i) + (at(j, i + 1) - at(j, i)) * u;
  float high = at(j + 1, i) + (at(j + 1, i + 1) - at(j + 1, i)) * u;
  return low + (high - low) * v;
}
std::vector<float> albedo_avg_of(const std::vector<uint16_t>& table) {
  std::vector<float> out(NODES_ROUGHNESS);
  for (int j = 0; j < NODES_ROUGHNESS; ++j) {
    float alpha = node_alpha(j);
    double sum = 0.0;
    for (u32 m = 0; m < AVG_STEPS; ++m) {
      float mu = ((float)m + 0.5f) / (float)AVG_STEPS;
      sum += (double)lookup(table.data(), mu, alpha) * (double)mu;
    }
    out[j] = (float)(2.0 * sum / (double)AVG_STEPS);
  }
  return out;
}

std::string g_cache_path;
std::once_flag g_once;
Tables g_tables;

bool load_cache(Tables& t) {
  if (g_cache_path.empty()) return false;
  FILE* f = std::fopen(g_cache_path.c_str(), "rb");
  if (!f) return false;
  const size_t slice = NODES_MU * NODES_ROUGHNESS, clear = slice * NODES_IOR;
  t.albedo.resize(slice);
  t.schlick.resize(slice);
  t.clear_in.resize(clear);
  t.clear_out.resize(clear);
  t.albedo_avg.resize(NODES_ROUGHNESS);
  t.schlick_avg.resize(NODES_ROUGHNESS);
  bool ok = std::fread(t.albedo.data(), 2, slice, f) == slice && std::fread(t.schlick.data(), 2, slice, f) == slice &&
            std::fread(t.clear_in.data(), 2, clear, f) == clear && std::fread(t.clear_out.data(), 2, clear, f) == clear &&
            std::fread(t.albedo_avg.data(), 4, NODES_ROUGHNESS, f) == (size_t)NODES_ROUGHNESS &&
            std::fread(t.schlick_avg.data(), 4, NODES_ROUGHNESS, f) == (size_t)NODES_ROUGHNESS;
  std::fclose(f);
  return ok;
}
void save_cache(const Tables& t) {
  if (g_cache_path.empty()) return;
  std::string tmp = g_cache_path + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return;
  std::fwrite(t.albedo.data(), 2, t.albedo.size(), f);
  std::fwrite(t.schlick.data(), 2, t.schlick.size(), f);
  std::fwrite(t.clear_in.data(), 2, t.clear_in.size(), f);
  std::fwrite(t.clear_out.data(), 2, t.clear_out.size(), f);
  std::fwrite(t.albedo_avg.data(), 4, t.albedo_avg.size(), f);
  std::fwrite(t.schlick_avg.data(), 4, t.schlick_avg.size(), f);
  std::fclose(f);
  std::rename(tmp.c_str(), g_cache_path.c_str());
}
}  // namespace

void set_cache_path(const char* path) { g_cache_path = path ? path : ""; }

void generate(Tables& out, int threads) {
  const int slice = NODES_MU * NODES_ROUGHNESS;
  out.albedo.assign(slice, 0);
  out.schlick.assign(slice, 0);
  out.clear_in.assign(slice * NODES_IOR, 0);
  out.clear_out.assign(slice * NODES_IOR, 0);
  // Work items: albedo nodes, schlick nodes, clear-in nodes, clear-out nodes.
  const int total = 2 * slice + 2 * slice * NODES_IOR;
  std::atomic<int> next{0};
  auto worker = [&]() {
    for (;;) {
      int n = next.fetch_add(1);
      if (n >= total) return;
      if (n < slice) {
        int j = n / NODES_MU, i = n % NODES_MU;
        out.albedo[n] = quantise(integrate_albedo(node_mu(i), node_alpha(j), (u64)n));
      } else if (n < 2 * slice) {
        int m = n - slice;
        int j = m / NODES_MU, i = m % NODES_MU;
        out.schlick[m] = quantise(integrate_schlick(node_mu(i), node_alpha(j), (u64)m + (1ull << 24)));
      } else {
        int m = n - 2 * slice;
        bool entering = m < slice * NODES_IOR;
        if (!entering) m -= slice * NODES_IOR;
        int k = m / slice, rest = m % slice;
        int j = rest / NODES_MU, i = rest % NODES_MU;
        float ior = node_ior(k);
        float eta = entering ? ior : 1.0f / ior;
        u64 stream = (u64)m + (entering ? 0ull : (1ull << 20));
        uint16_t q = quantise(integrate_clear(node_mu(i), node_alpha(j), eta, stream));
        (entering ? out.clear_in : out.clear_out)[m] = q;
      }
    }
  };
  if (threads < 1) threads = 1;
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
  out.albedo_avg = albedo_avg_of(out.albedo);
  out.schlick_avg = albedo_avg_of(out.schlick);
}

const Tables& tables() {
  std::call_once(g_once, []() {
    if (!load_cache(g_tables)) {
      unsigned hw = std::thread::hardware_concurrency();
      generate(g_tables, hw ? (int)hw : 4);
      save_cache(g_tables);
    }
  });
  return g_tables;
}

float albedo(float mu, float alpha) { return lookup(tables().albedo.data(), mu, alpha); }
float schlick_albedo(float mu, float alpha) { return lookup(tables().schlick.data(), mu, alpha); }
float albedo_avg(float alpha) {
  int j;
  float v;
  locate(std::sqrt(fmax_r(alpha, 0.0f)), NODES_ROUGHNESS, j, v);
  const auto& a = tables().albedo_avg;
  return a[j] + (a[j + 1] - a[j]) * v;
}
float schlick_avg(float alpha) {
  int j;
  float v;
  locate(std::sqrt(fmax_r(alpha, 0.0f)), NODES_ROUGHNESS, j, v);
  const auto& a = tables().schlick_avg;
  return a[j] + (a[j + 1] - a[j]) * v;
}
float clear_albedo(float mu, float alpha, float eta) {
  const Tables& t = tables();
  const std::vector<uint16_t>* table;
  float ior;
  if (eta >= 1.0f) { table = &t.clear_in; ior = eta; }
  else { table = &t.clear_out; ior = 1.0f / eta; }
  float x = (ior - 1.0f) / (ior + 1.0f);
  int k;
  float w;
  locate(x / MAX_X, NODES_IOR, k, w);
  const int slice = NODES_MU * NODES_ROUGHNESS;
  float a = lookup(table->data() + k * slice, mu, alpha);
  float b = lookup(table->data() + (k + 1) * slice, mu, alpha);
  return a + (b - a) * w;
}

Layer Layer::make(float f0, float f90, float strength, float alpha, float ms) {
  if (std::isnan(strength) || strength <= 0.0f || (f90 <= 0.0f && f0 <= 0.0f)) return Layer();
  Layer l;
  l.f0 = f0;
  l.f90 = f90;
  l.strength = strength;
  l.rough = alpha >= MIN_COMPENSATED_ALPHA;
  l.rough_alpha = alpha;
  l.rough_ms = ms;
  float avg;
  if (l.rough) {
    float e_avg = albedo_avg(alpha);
    avg = f0 * e_avg + (f90 - f0) * schlick_avg(alpha) + ms * (1.0f - e_avg);
  } else {
    avg = schlick_average(f0, f90);
  }
  avg = avg * strength;
  l.norm = 1.0f / (1.0f - fmin_r(avg, 0.999f));
  return l;
}
float Layer::albedo_at(float mu) const {
  float e;
  if (rough) {
    float ee = albedo(mu, rough_alpha);
    e = f0 * ee + (f90 - f0) * schlick_albedo(mu, rough_alpha) + rough_ms * (1.0f - ee);
  } else {
    float m = clampf(1.0f - mu, 0.0f, 1.0f);
    float m2 = m * m;
    e = f0 + (f90 - f0) * (m2 * m2 * m);
  }
  return clampf(e * strength, 0.0f, 1.0f);
}

static const float GAUSS_8[8][2] = {
    {0.019855072f, 0.05061427f}, {0.10166676f, 0.11119052f}, {0.2372338f, 0.15685332f},
    {0.40828268f, 0.18134189f},  {0.5917173f, 0.18134189f},  {0.7627662f, 0.15685332f},
    {0.8983332f, 0.11119052f},   {0.9801449f, 0.05061427f},
};
float joint_norm(const Layer& below, const Layer& above) {
  if (above.is_clear()) return below.norm;
  if (below.is_clear()) return above.norm;
  float sum = -0.0f;
  for (int k = 0; k < 8; ++k) {
    float mu = GAUSS_8[k][0], w = GAUSS_8[k][1];
    sum = sum + w * 2.0f * mu * below.passes(mu) * above.passes(mu);
  }
  return 1.0f / fmax_r(sum, 0.001f);
}

bool Lobe::make(float alpha, Lobe& out) {
  if (std::isnan(alpha) || alpha < MIN_COMPENSATED_ALPHA) return false;
  float e_avg = albedo_avg(alpha);
  float lost = 1.0f - e_avg;
  if (!(lost > 1.0e-6f)) return false;
  out.alpha = alpha;
  out.norm = FRAC_1_PI / lost;
  out.e_avg = e_avg;
  return true;
}
}  // namespace energy

// ------------------------------------------------------------------ bsdf.rs:384-996
static inline Rgb clamp01(Rgb c) { return Rgb(clampf(c.r, 0, 1), clampf(c.g, 0, 1), clampf(c.b, 0, 1)); }
static inline float peak_of(Rgb c) { return fmax_r(fmax_r(c.r, c.g), c.b); }
constexpr float COAT_F0 = 0.04f;
constexpr float MAX_REFLECT_PROBABILITY = 0.99f;

Bsdf Bsdf::surface_at(const Material& m, bool entering, float ior) {
  Rgb base = clamp01(m.base_color);
  float metallic = m.kind == MATERIAL_GLASS ? 0.0f : clampf(m.metallic, 0.0f, 1.0f);
  float transmission = clampf(m.transmission, 0.0f, 1.0f);
  ior = fmax_r(ior, 1.0f);
  float f0 = ((ior - 1.0f) / (ior + 1.0f)) * ((ior - 1.0f) / (ior + 1.0f));
  float roughness = clampf(m.roughness, 0.0f, 1.0f);
  float coat_roughness = clampf(m.coat_roughness, 0.0f, 1.0f);
  float clear = (1.0f - metallic) * transmission;
  bool thin = (m.flags & MATERIAL_FLAG_THIN) != 0 || ior < 1.0001f;
  float eta = (thin || entering) ? ior : 1.0f / ior;
  Rgb tint;
  if (thin) tint = base;
  else if (m.absorption_mm > 0.0f) tint = WHITE;
  else tint = Rgb(std::sqrt(base.r), std::sqrt(base.g), std::sqrt(base.b));
  float coat = (clear > 0.0f && !thin && !entering) ? 0.0f : clampf(m.coat, 0.0f, 1.0f);
  Bsdf b;
  b.diffuse = base * ((1.0f - metallic) * (1.0f - transmission));
  b.metal_f0 = base;
  b.metal_f90 = Rgb(fmin_r(50.0f * base.r, 1.0f), fmin_r(50.0f * base.g, 1.0f), fmin_r(50.0f * base.b, 1.0f));
  b.metallic = metallic;
  b.dielectric = (1.0f - metallic) * (1.0f - transmission);
  b.f0 = f0;
  b.f90 = fmin_r(50.0f * f0, 1.0f);
  b.spec = Ggx::make(fmax_r(roughness * roughness, MIN_ALPHA));
  b.coat = coat;
  b.coat_ggx = Ggx::make(fmax_r(coat_roughness * coat_roughness, MIN_ALPHA));
  b.clear = clear;
  b.eta = eta;
  b.tint = tint;
  b.thin = thin;
  b.base_layer = energy::Layer::clear();
  b.coat_layer = energy::Layer::clear();
  b.diffuse_norm = 1.0f;
  b.multiple_scattering();
  b.layers();
  return b;
}

Bsdf Bsdf::lambert(Rgb albedo) {
  Bsdf b;
  b.diffuse = clamp01(albedo);
  b.metal_f0 = BLACK;
  b.metal_f90 = BLACK;
  b.metallic = 0.0f;
  b.dielectric = 1.0f;
  b.f0 = 0.0f;
  b.f90 = 0.0f;
  b.spec = Ggx::make(1.0f);
  b.coat = 0.0f;
  b.coat_ggx = Ggx::make(1.0f);
  b.clear = 0.0f;
  b.eta = 1.0f;
  b.tint = WHITE;
  b.thin = false;
  b.base_layer = energy::Layer::clear();
  b.coat_layer = energy::Layer::clear();
  b.diffuse_norm = 1.0f;
  return b;
}

Bsdf Bsdf::medium_exit(float albedo) {
  Bsdf b = lambert(BLACK);
  b.exit = fmax_r(albedo, 0.0f);
  b.diffuse = BLACK;
  b.clear = 1.0f;
  return b;
}

void Bsdf::layers() {
  float alpha = spec.ax;  // isotropic
  if (diffuse.r > 0.0f || diffuse.g > 0.0f || diffuse.b > 0.0f) {
    float ms = ms_spec_on ? energy::multiple_fresnel(energy::schlick_average(f0, f90), ms_spec_lobe.e_avg) : 0.0f;
    base_layer = energy::Layer::make(f0, f90, 1.0f, alpha, ms);
  }
  if (coat > 0.0f) {
    float ms = ms_coat_on ? energy::multiple_fresnel(energy::schlick_average(COAT_F0, 1.0f), ms_coat_lobe.e_avg) : 0.0f;
    coat_layer = energy::Layer::make(COAT_F0, 1.0f, coat, coat_ggx.ax, ms);
  }
  diffuse_norm = energy::joint_norm(base_layer, coat_layer);
}

void Bsdf::multiple_scattering() {
  float alpha = spec.ax;
  energy::Lobe lobe;
  ms_spec_on = false;
  if (energy::Lobe::make(alpha, lobe) && has_specular()) {
    auto ms = [&](float a0, float a90) {
      return energy::multiple_fresnel(energy::schlick_average(a0, a90), lobe.e_avg);
    };
    float diel = dielectric > 0.0f ? dielectric * ms(f0, f90) : 0.0f;
    Rgb colour;
    if (metallic > 0.0f) {
      colour = Rgb(ms(metal_f0.r, metal_f90.r), ms(metal_f0.g, metal_f90.g), ms(metal_f0.b, metal_f90.b)) * metallic +
               Rgb::splat(diel);
    } else {
      colour = Rgb::splat(diel);
    }
    if (peak_of(colour) > 0.0f) {
      ms_spec_on = true;
      ms_spec_lobe = lobe;
      ms_spec_colour = colour;
    }
  }
  ms_coat_on = false;
  if (coat > 0.0f) {
    energy::Lobe cl;
    if (energy::Lobe::make(coat_ggx.ax, cl)) {
      float c = coat * energy::multiple_fresnel(energy::schlick_average(COAT_F0, 1.0f), cl.e_avg);
      if (c > 0.0f) {
        ms_coat_on = true;
        ms_coat_lobe = cl;
        ms_coat_c = c;
      }
    }
  }
  ms_clear_on = clear > 0.0f && !thin && alpha >= energy::MIN_COMPENSATED_ALPHA;
  ms_clear_alpha = alpha;
}

bool Bsdf::clear_compensation(float mu_o, float& k) const {
  if (!ms_clear_on) return false;
  float e = energy::clear_albedo(mu_o, ms_clear_alpha, eta);
  k = 1.0f / fmax_r(e, 0.05f);
  return true;
}

Rgb Bsdf::straight_transmittance(float cos) const {
  if (!is_thin()) return BLACK;
  cos = fmin_r(std::fabs(cos), 1.0f);
  float t = 1.0f - thin_reflectance(cos, eta);
  return tint * (clear * (1.0f - coat_fresnel(cos)) * t);
}

Rgb Bsdf::specular_fresnel(float cos) const {
  float d = schlick(f0, f90, cos);
  if (metallic <= 0.0f && clear <= 0.0f) return Rgb::splat(d);
  Rgb metal(schlick(metal_f0.r, metal_f90.r, cos), schlick(metal_f0.g, metal_f90.g, cos),
            schlick(metal_f0.b, metal_f90.b, cos));
  return Rgb::splat(d) * dielectric + metal * metallic;
}
float Bsdf::coat_fresnel(float cos) const { return coat <= 0.0f ? 0.0f : coat * schlick(COAT_F0, 1.0f, cos); }
float Bsdf::clear_fresnel(float cos) const { return thin ? thin_reflectance(cos, eta) : fresnel_dielectric(cos, eta); }
float Bsdf::reflect_probability(float cos) const { return fmin_r(clear_fresnel(cos), MAX_REFLECT_PROBABILITY); }
float Bsdf::reflect_chance(float mu_o, float cos_h) const {
  return ms_clear_on ? clear_fresnel(cos_h) : reflect_probability(mu_o);
}

LobeWeights Bsdf::lobe_weights(Vec3 wo) const {
  float c = coat_fresnel(wo.z);
  float below = 1.0f - c;
  float specular = has_specular() ? below * fmax_r(specular_fresnel(wo.z).luminance(), 0.0f) : 0.0f;
  float diff = below * diffuse.luminance() * base_layer.passes(wo.z);
  float clr = below * clear;
  float ms = 0.0f;
  if (ms_spec_on) ms += below * fmax_r(ms_spec_colour.luminance(), 0.0f) * ms_spec_lobe.lost(wo.z);
  if (ms_coat_on) ms += ms_coat_c * ms_coat_lobe.lost(wo.z);
  float total = c + specular + diff + clr + ms;
  if (total > 0.0f) {
    float inv = 1.0f / total;
    return LobeWeights{c * inv, specular * inv, diff * inv, clr * inv, ms * inv};
  }
  return LobeWeights{0, 0, 0, 0, 0};
}

Rgb Bsdf::eval(Vec3 wo, Vec3 wi, Lobes lobes) const {
  if (wo.z <= 0.0f) return BLACK;
  if (exit > 0.0f) return wi.z < 0.0f ? Rgb::splat(exit * FRAC_1_PI) : BLACK;
  if (wi.z <= 0.0f) {
    if (wi.z < 0.0f && lobes == Lobes::All && clear > 0.0f && !thin) return eval_refraction(wo, wi);
    return BLACK;
  }
  float attenuation = coat_layer.passes(wo.z) * coat_layer.passes(wi.z);
  Rgb diff = diffuse * (FRAC_1_PI * base_layer.passes(wo.z) * base_layer.passes(wi.z) * diffuse_norm);
  if (lobes == Lobes::Diffuse) return diff * attenuation;
  Vec3 h = (wo + wi).normalize();
  float cos_h = fmax_r(wo.dot(h), 0.0f);
  Rgb below = diff;
  if (has_specular()) below += specular_fresnel(cos_h) * spec.microfacet(wo, wi, h);
  if (ms_spec_on) below += ms_spec_colour * ms_spec_lobe.shape(wo.z, wi.z);
  if (clear > 0.0f) {
    float c = clear * clear_fresnel(cos_h) * spec.microfacet(wo, wi, h);
    float k;
    if (clear_compensation(wo.z, k)) c *= k;
    below += Rgb::splat(c);
  }
  Rgb f = below * attenuation;
  if (coat > 0.0f) f += Rgb::splat(coat_fresnel(cos_h) * coat_ggx.microfacet(wo, wi, h));
  if (ms_coat_on) f += Rgb::splat(ms_coat_c * ms_coat_lobe.shape(wo.z, wi.z));
  return f;
}

bool Bsdf::refraction_half(Vec3 wo, Vec3 wi, Vec3& h, float& co, float& ci) const {
  h = (wo + wi * eta).normalize();
  if (h.z < 0.0f) h = -h;
  co = wo.dot(h);
  ci = wi.dot(h);
  return co > 0.0f && ci < 0.0f && h.length_squared() > 0.5f;
}

Rgb Bsdf::eval_refraction(Vec3 wo, Vec3 wi) const {
  Vec3 h;
  float co, ci;
  if (!refraction_half(wo, wi, h, co, ci)) return BLACK;
  float f = fresnel_dielectric(co, eta);
  float denom = co + eta * ci;
  float g2 = 1.0f / (1.0f + spec.lambda(wo) + spec.lambda(wi));
  float value = (1.0f - f) * spec.d(h) * g2 * (-ci) * co / (-wi.z * wo.z * denom * denom);
  float k;
  if (clear_compensation(wo.z, k)) value *= k;
  float attenuation = coat_layer.passes(wo.z);
  return tint * (clear * attenuation * value);
}

float Bsdf::pdf(Vec3 wo, Vec3 wi) const {
  if (wo.z <= 0.0f) return 0.0f;
  if (exit > 0.0f) return fmax_r(-wi.z, 0.0f) * FRAC_1_PI;
  if (wi.z <= 0.0f) {
    if (wi.z < 0.0f && clear > 0.0f && !thin) return refraction_pdf(lobe_weights(wo), wo, wi);
    return 0.0f;
  }
  return pdf_with(lobe_weights(wo), wo, wi);
}

float Bsdf::pdf_with(const LobeWeights& w, Vec3 wo, Vec3 wi) const {
  float p = (w.diffuse + w.ms) * wi.z * FRAC_1_PI;
  if (w.specular > 0.0f || w.coat > 0.0f || w.clear > 0.0f) {
    Vec3 h = (wo + wi).normalize();
    if (wo.dot(h) > 0.0f) {
      if (w.specular > 0.0f) p += w.specular * spec.reflection_pdf(wo, h);
      if (w.coat > 0.0f) p += w.coat * coat_ggx.reflection_pdf(wo, h);
      if (w.clear > 0.0f) p += w.clear * reflect_chance(wo.z, wo.dot(h)) * spec.reflection_pdf(wo, h);
    }
  }
  return p;
}

float Bsdf::refraction_pdf(const LobeWeights& w, Vec3 wo, Vec3 wi) const {
  if (std::isnan(w.clear) || w.clear <= 0.0f) return 0.0f;
  Vec3 h;
  float co, ci;
  if (!refraction_half(wo, wi, h, co, ci)) return 0.0f;
  float g1 = 1.0f / (1.0f + spec.lambda(wo));
  float visible = g1 * spec.d(h) * co / wo.z;
  float denom = co + eta * ci;
  float jacobian = eta * eta * (-ci) / (denom * denom);
  return w.clear * (1.0f - reflect_chance(wo.z, co)) * visible * jacobian;
}

static inline bool pdf_ok(float p) { return p > 0.0f && std::isfinite(p); }

bool Bsdf::sample(Vec3 wo, float u0, float u1, BsdfSample& out) const {
  if (wo.z <= 0.0f) return false;
  if (exit > 0.0f) {
    // bsdf.rs:837-847: a cosine-distributed direction through the surface.
    Vec3 d = cosine_hemisphere(u0, u1);
    float pdf = d.z * FRAC_1_PI;
    if (!(pdf > 0.0f)) return false;
    out.wi = Vec3(d.x, d.y, -d.z);
    out.weight = Rgb::splat(exit);
    out.pdf = pdf;
    out.straight = false;
    return true;
  }
  LobeWeights w = lobe_weights(wo);
  Vec3 wi;
  if (u0 < w.coat) {
    Vec3 h = sample_vndf(wo, coat_ggx.ax, coat_ggx.ay, u0 / w.coat, u1);
    wi = h * (2.0f * wo.dot(h)) - wo;
  } else if (u0 < w.coat + w.specular) {
    Vec3 h = sample_vndf(wo, spec.ax, spec.ay, (u0 - w.coat) / w.specular, u1);
    wi = h * (2.0f * wo.dot(h)) - wo;
  } else if (w.diffuse + w.ms > 0.0f && (w.clear <= 0.0f || u0 < w.coat + w.specular + (w.diffuse + w.ms))) {
    float start = w.coat + w.specular;
    float uu = clampf((u0 - start) / (w.diffuse + w.ms), 0.0f, 0.99999994f);
    wi = cosine_hemisphere(uu, u1);
  } else if (w.clear > 0.0f) {
    return sample_clear(w, wo, u0, u1, out);
  } else {
    return false;
  }
  if (wi.z <= 0.0f) return false;
  float p = pdf_with(w, wo, wi);
  if (!pdf_ok(p)) return false;
  out.wi = wi;
  out.weight = eval(wo, wi, Lobes::All) * (wi.z / p);
  out.pdf = p;
  out.straight = false;
  return true;
}

bool Bsdf::sample_clear(const LobeWeights& w, Vec3 wo, float u0, float u1, BsdfSample& out) const {
  float start = w.coat + w.specular + (w.diffuse + w.ms);
  float uc = clampf((u0 - start) / w.clear, 0.0f, 0.99999994f);
  if (ms_clear_on) return sample_rough_clear(w, wo, uc, u1, out);
  float pr = reflect_probability(wo.z);
  if (uc < pr) {
    Vec3 h = sample_vndf(wo, spec.ax, spec.ay, uc / pr, u1);
    Vec3 wi = h * (2.0f * wo.dot(h)) - wo;
    if (wi.z <= 0.0f) return false;
    float p = pdf_with(w, wo, wi);
    if (!pdf_ok(p)) return false;
    out.wi = wi;
    out.weight = eval(wo, wi, Lobes::All) * (wi.z / p);
    out.pdf = p;
    out.straight = false;
    return true;
  }
  if (thin) {
    float chance = w.clear * (1.0f - pr);
    if (std::isnan(chance) || chance <= 0.0f) return false;
    out.wi = -wo;
    out.weight = straight_transmittance(wo.z) / chance;
    out.pdf = chance;
    out.straight = true;
    return true;
  }
  float ut = clampf((uc - pr) / (1.0f - pr), 0.0f, 0.99999994f);
  Vec3 h = sample_vndf(wo, spec.ax, spec.ay, ut, u1);
  Vec3 wi;
  if (!refract(wo, h, eta, wi)) return false;
  if (wi.z >= 0.0f) return false;
  float p = refraction_pdf(w, wo, wi);
  if (!pdf_ok(p)) return false;
  out.wi = wi;
  out.weight = eval_refraction(wo, wi) * (-wi.z / p);
  out.pdf = p;
  out.straight = false;
  return true;
}

bool Bsdf::sample_rough_clear(const LobeWeights& w, Vec3 wo, float u0, float u1, BsdfSample& out) const {
  Vec3 h = sample_vndf(wo, spec.ax, spec.ay, u0, u1);
  float fresnel = clear_fresnel(wo.dot(h));
  u64 key = ((u64)f2u(u0) << 32) | (u64)f2u(u1);
  float choice = (float)(mix64(key ^ 0x9e3779b97f4a7c15ull) >> 40) * (1.0f / 16777216.0f);
  if (choice < fresnel) {
    Vec3 wi = h * (2.0f * wo.dot(h)) - wo;
    if (wi.z <= 0.0f) return false;
    float p = pdf_with(w, wo, wi);
    if (!pdf_ok(p)) return false;
    out.wi = wi;
    out.weight = eval(wo, wi, Lobes::All) * (wi.z / p);
    out.pdf = p;
    out.straight = false;
    return true;
  }
  Vec3 wi;
  if (!refract(wo, h, eta, wi)) return false;
  if (wi.z >= 0.0f) return false;
  float p = refraction_pdf(w, wo, wi);
  if (!pdf_ok(p)) return false;
  out.wi = wi;
  out.weight = eval_refraction(wo, wi) * (-wi.z / p);
  out.pdf = p;
  out.straight = false;
  return true;
}

}  // namespace orc
