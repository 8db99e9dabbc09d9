// ORACLE — test infrastructure only (see rt_core.hpp).
// Restates the principled BSDF and its energy compensation (paths inside engine/crates/rt-core/src):
//   bsdf.rs:89-996          Ggx, schlick, fresnel_dielectric, refract, sample_vndf, Bsdf::surface_at,
//                           lobe_weights, eval_surface, eval_refraction, pdf_with, refraction_pdf,
//                           sample_surface, sample_clear, sample_rough_clear
//   bsdf/energy.rs:55-554   the table recipe (regenerated here from its description), lookups,
//                           Layer, joint_norm, Lobe, clear_albedo, multiple_fresnel
//   bsdf/forward.rs:138-190 eval/pdf/sample wrappers (no forward lobe on the configs' surfaces)
// The anisotropic branch (anisotropy > 0) is not restated: scenes that set it are refused.
#pragma once
#include <vector>
#include "rt_core.hpp"

namespace orc {

struct Material {
  u32 kind = 0;
  Rgb base_color = Rgb::splat(0.8f);
  float metallic = 0.0f, roughness = 0.5f, ior = 1.5f, transmission = 0.0f;
  float coat = 0.0f, coat_roughness = 0.0f;
  Rgb emission = Rgb(0, 0, 0);
  float emission_power = 0.0f;
  u32 flags = 0;
  float opacity = 1.0f, anisotropy = 0.0f, absorption_mm = 0.0f, bevel_mm = 0.0f;
  float scatter_mm = 0.0f;
  Rgb scatter_color = Rgb(1, 1, 1);
  float scatter_g = 0.0f, abbe = 0.0f;
  bool is_emissive() const { return kind == 2; }
  Rgb radiance() const {
    if (!is_emissive()) return Rgb(0, 0, 0);
    Rgb e = emission * emission_power;
    return Rgb(fmax_r(e.r, 0.0f), fmax_r(e.g, 0.0f), fmax_r(e.b, 0.0f));
  }
};
constexpr u32 MATERIAL_GLASS = 1;
constexpr u32 MATERIAL_FLAG_LIGHT = 1;
constexpr u32 MATERIAL_FLAG_THIN = 2;

// ---------------------------------------------------------------- energy tables
namespace energy {
constexpr float MIN_COMPENSATED_ALPHA = 2.0e-3f;
constexpr int NODES_MU = 32, NODES_ROUGHNESS = 32, NODES_IOR = 17;
constexpr float MAX_IOR = 3.0f;
struct Tables {
  std::vector<uint16_t> albedo, schlick, clear_in, clear_out;
  std::vector<float> albedo_avg, schlick_avg;
};
// Regenerates every table by the recipe of energy.rs (seed 0xe9ae9a, 64x64 jittered strata per
// node); `threads` workers split the nodes, the values do not depend on it.
void generate(Tables& out, int threads);
const Tables& tables();  // generated once (or loaded from the cache file) on first use
void set_cache_path(const char* path);

float albedo(float mu, float alpha);
float albedo_avg(float alpha);
float schlick_albedo(float mu, float alpha);
float schlick_avg(float alpha);
float clear_albedo(float mu, float alpha, float eta);
inline float schlick_average(float f0, float f90) { return f0 + (f90 - f0) * (1.0f / 21.0f); }
inline float multiple_fresnel(float f_avg, float e_avg) {
  float d = 1.0f - f_avg * (1.0f - e_avg);
  return d > 0.0f ? f_avg * f_avg * e_avg / d : 0.0f;
}

struct Layer {
  float f0 = 0, f90 = 0, strength = 0;
  bool rough = false;
  float rough_alpha = 0, rough_ms = 0;
  float norm = 1.0f;
  static Layer clear() { return Layer(); }
  static Layer make(float f0, float f90, float strength, float alpha, float ms);
  float albedo_at(float mu) const;
  float passes(float mu) const { return 1.0f - albedo_at(mu); }
  bool is_clear() const { return strength <= 0.0f; }
};
float joint_norm(const Layer& below, const Layer& above);

struct Lobe {
  float alpha, norm, e_avg;
  static bool make(float alpha, Lobe& out);
  float lost(float mu) const { return fmax_r(1.0f - energy::albedo(mu, alpha), 0.0f); }
  float shape(float mo, float mi) const { return lost(mo) * lost(mi) * norm; }
};
}  // namespace energy

// ---------------------------------------------------------------- bsdf
constexpr float MIN_ALPHA = 1.0e-3f;

struct Ggx {
  float ax, ay, a2;
  bool iso;
  static Ggx make(float alpha) { return Ggx{alpha, alpha, alpha * alpha, true}; }
  float d(Vec3 h) const;
  float lambda(Vec3 w) const;
  float reflection_pdf(Vec3 wo, Vec3 h) const {
    float g1 = 1.0f / (1.0f + lambda(wo));
    return g1 * d(h) / (4.0f * wo.z);
  }
  float microfacet(Vec3 wo, Vec3 wi, Vec3 h) const {
    float g2 = 1.0f / (1.0f + lambda(wo) + lambda(wi));
    return d(h) * g2 / (4.0f * wo.z * wi.z);
  }
};

float fresnel_dielectric(float cos_i, float eta);
bool refract(Vec3 wo, Vec3 h, float eta, Vec3& out);
Vec3 sample_vndf(Vec3 wo, float ax, float ay, float u0, float u1);

enum class Lobes { All, Diffuse };

struct BsdfSample {
  Vec3 wi;
  Rgb weight;
  float pdf;
  bool straight;
};

struct LobeWeights { float coat, specular, diffuse, clear, ms; };

struct Bsdf {
  Rgb diffuse, metal_f0, metal_f90;
  float metallic, dielectric, f0, f90;
  Ggx spec;
  float coat;
  Ggx coat_ggx;
  float clear, eta;
  Rgb tint;
  bool thin;
  // multiple scattering (MultiScatter)
  bool ms_spec_on = false;
  energy::Lobe ms_spec_lobe{};
  Rgb ms_spec_colour;
  bool ms_coat_on = false;
  energy::Lobe ms_coat_lobe{};
  float ms_coat_c = 0;
  bool ms_clear_on = false;
  float ms_clear_alpha = 0;
  energy::Layer base_layer, coat_layer;
  float diffuse_norm = 1.0f;
  // Above zero, the surface is the exit of a scattering medium: a Lambertian transmission lobe of
  // this albedo and nothing else (bsdf.rs:343-345, 549-556).
  float exit = 0.0f;

  static Bsdf surface(const Material& m, bool entering) { return surface_at(m, entering, m.ior); }
  static Bsdf medium_exit(float albedo);
  static Bsdf surface_at(const Material& m, bool entering, float ior);
  static Bsdf lambert(Rgb albedo);

  Rgb eval(Vec3 wo, Vec3 wi, Lobes lobes) const;
  float pdf(Vec3 wo, Vec3 wi) const;
  bool sample(Vec3 wo, float u0, float u1, BsdfSample& out) const;

  bool is_thin() const { return thin && clear > 0.0f; }
  Rgb straight_transmittance(float cos) const;

  // internals
  bool has_specular() const { return (f90 > 0.0f && dielectric > 0.0f) || metallic > 0.0f; }
  Rgb specular_fresnel(float cos) const;
  float coat_fresnel(float cos) const;
  float clear_fresnel(float cos) const;
  float reflect_probability(float cos) const;
  float reflect_chance(float mu_o, float cos_h) const;
  bool clear_compensation(float mu_o, float& k) const;
  LobeWeights lobe_weights(Vec3 wo) const;
  bool refraction_half(Vec3 wo, Vec3 wi, Vec3& h, float& co, float& ci) const;
  Rgb eval_refraction(Vec3 wo, Vec3 wi) const;
  float pdf_with(const LobeWeights& w, Vec3 wo, Vec3 wi) const;
  float refraction_pdf(const LobeWeights& w, Vec3 wo, Vec3 wi) const;
  bool sample_clear(const LobeWeights& w, Vec3 wo, float u0, float u1, BsdfSample& out) const;
  bool sample_rough_clear(const LobeWeights& w, Vec3 wo, float u0, float u1, BsdfSample& out) const;
  void multiple_scattering();
  void layers();
};

}  // namespace orc
