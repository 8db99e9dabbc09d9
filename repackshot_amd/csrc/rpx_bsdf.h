// The principled BSDF of the reference renderer (Lambert + GGX conductor / dielectric, clear
// transmission, clear coat, Kulla–Conty and Turquin energy compensation), as host/device code.
// The host builds one `Bsdf` per (material, side) when a scene loads; the shade kernel reads it
// and evaluates / samples it per vertex.
//
// Reference (paths inside engine/crates/rt-core/src): bsdf.rs:89-996 (Ggx, schlick,
// fresnel_dielectric, refract, sample_vndf, Bsdf::surface_at, lobe_weights, eval_surface,
// eval_refraction, pdf_with, refraction_pdf, sample_surface, sample_clear, sample_rough_clear),
// bsdf/energy.rs:55-554 (table recipe, lookups, Layer, joint_norm, Lobe, clear_albedo),
// bsdf/forward.rs:138-190 (eval/pdf/sample without a forward lobe).
// Isotropic lobes only: scenes with anisotropy > 0 are refused at load.
#pragma once
#include "rpx_math.h"

namespace rpx {

constexpr int E_NODES_MU = 32, E_NODES_ROUGHNESS = 32, E_NODES_IOR = 17;
constexpr int E_SLICE = E_NODES_MU * E_NODES_ROUGHNESS;
constexpr float E_MAX_IOR = 3.0f;
constexpr float MIN_COMPENSATED_ALPHA = 2.0e-3f;
constexpr float MIN_ALPHA = 1.0e-3f;

// Energy tables: E(mu, alpha), E5(mu, alpha), E_t in/out (16-bit fractions) and the averages.
struct EnergyTables {
  const uint16_t* albedo;     // [32 roughness][32 mu]
  const uint16_t* schlick;    // [32][32]
  const uint16_t* clear_in;   // [17 ior][32][32]
  const uint16_t* clear_out;  // [17][32][32]
  const float* albedo_avg;    // [32]
  const float* schlick_avg;   // [32]
};

RPX_HD void e_locate(float x, int n, int& i, float& w) {
  float t = clampf(x, 0.0f, 1.0f) * (float)(n - 1);
  u32 ti = f_to_u32(t);
  i = (int)(ti < (u32)(n - 2) ? ti : (u32)(n - 2));
  w = t - (float)i;
}
RPX_HD float e_lookup(const uint16_t* table, float mu, float alpha) {
  int i, j;
  float u, v;
  e_locate(sqrt_(fmax_r(mu, 0.0f)), E_NODES_MU, i, u);
  e_locate(sqrt_(fmax_r(alpha, 0.0f)), E_NODES_ROUGHNESS, j, v);
  const float UNIT = 1.0f / 65535.0f;
  float a00 = (float)table[j * E_NODES_MU + i] * UNIT, a01 = (float)table[j * E_NODES_MU + i + 1] * UNIT;
  float a10 = (float)table[(j + 1) * E_NODES_MU + i] * UNIT, a11 = (float)table[(j + 1) * E_NODES_MU + i + 1] * UNIT;
  float low = a00 + (a01 - a00) * u;
  float high = a10 + (a11 - a10) * u;
  return low + (high - low) * v;
}
RPX_HD float e_avg_lookup(const float* a, float alpha) {
  int j;
  float v;
  e_locate(sqrt_(fmax_r(alpha, 0.0f)), E_NODES_ROUGHNESS, j, v);
  return a[j] + (a[j + 1] - a[j]) * v;
}
RPX_HD float e_clear_albedo(const EnergyTables& t, float mu, float alpha, float eta) {
  const uint16_t* table;
  float ior;
  if (eta >= 1.0f) {
    table = t.clear_in;
    ior = eta;
  } else {
    table = t.clear_out;
    ior = 1.0f / eta;
  }
  float x = (ior - 1.0f) / (ior + 1.0f);
  const float MAX_X = (E_MAX_IOR - 1.0f) / (E_MAX_IOR + 1.0f);
  int k;
  float w;
  e_locate(x / MAX_X, E_NODES_IOR, k, w);
  float a = e_lookup(table + k * E_SLICE, mu, alpha);
  float b = e_lookup(table + (k + 1) * E_SLICE, mu, alpha);
  return a + (b - a) * w;
}
RPX_HD float schlick_average(float f0, float f90) { return f0 + (f90 - f0) * (1.0f / 21.0f); }
RPX_HD float multiple_fresnel(float f_avg, float e_avg) {
  float d = 1.0f - f_avg * (1.0f - e_avg);
  return d > 0.0f ? f_avg * f_avg * e_avg / d : 0.0f;
}

// Layer (energy.rs:364-455).
struct Layer {
  float f0, f90, strength;
  u32 rough;
  float alpha, ms, norm;
};
RPX_HD Layer layer_clear() { return Layer{0, 0, 0, 0, 0, 0, 1.0f}; }
RPX_HD float layer_albedo(const EnergyTables& t, const Layer& l, float mu) {
  float e;
  if (l.rough) {
    float ee = e_lookup(t.albedo, mu, l.alpha);
    e = l.f0 * ee + (l.f90 - l.f0) * e_lookup(t.schlick, mu, l.alpha) + l.ms * (1.0f - ee);
  } else {
    float m = clampf(1.0f - mu, 0.0f, 1.0f);
    float m2 = m * m;
    e = l.f0 + (l.f90 - l.f0) * (m2 * m2 * m);
  }
  return clampf(e * l.strength, 0.0f, 1.0f);
}
RPX_HD float layer_passes(const EnergyTables& t, const Layer& l, float mu) { return 1.0f - layer_albedo(t, l, mu); }
RPX_HD Layer layer_make(const EnergyTables& t, float f0, float f90, float strength, float alpha, float ms) {
  if (is_nan(strength) || strength <= 0.0f || (f90 <= 0.0f && f0 <= 0.0f)) return layer_clear();
  Layer l;
  l.f0 = f0;
  l.f90 = f90;
  l.strength = strength;
  l.rough = alpha >= MIN_COMPENSATED_ALPHA ? 1u : 0u;
  l.alpha = alpha;
  l.ms = ms;
  float avg;
  if (l.rough) {
    float e_avg = e_avg_lookup(t.albedo_avg, alpha);
    avg = f0 * e_avg + (f90 - f0) * e_avg_lookup(t.schlick_avg, alpha) + ms * (1.0f - e_avg);
  } else {
    avg = schlick_average(f0, f90);
  }
  avg = avg * strength;
  l.norm = 1.0f / (1.0f - fmin_r(avg, 0.999f));
  return l;
}
RPX_HD float joint_norm(const EnergyTables& t, const Layer& below, const Layer& above) {
  if (above.strength <= 0.0f) return below.norm;
  if (below.strength <= 0.0f) return above.norm;
  // 8-point Gauss-Legendre nodes and weights on [0, 1] (energy.rs:458-467).
  const float G_MU[8] = {0.019855072f, 0.10166676f, 0.2372338f, 0.40828268f,
                         0.5917173f,   0.7627662f,  0.8983332f, 0.9801449f};
  const float G_W[8] = {0.05061427f, 0.11119052f, 0.15685332f, 0.18134189f,
                        0.18134189f, 0.15685332f, 0.11119052f, 0.05061427f};
  float sum = -0.0f;
  for (int k = 0; k < 8; ++k)
    sum = sum + G_W[k] * 2.0f * G_MU[k] * layer_passes(t, below, G_MU[k]) * layer_passes(t, above, G_MU[k]);
  return 1.0f / fmax_r(sum, 0.001f);
}

// Lobe (energy.rs:508-554).
struct MsLobe {
  float alpha, norm, e_avg;
};
RPX_HD bool ms_lobe_make(const EnergyTables& t, float alpha, MsLobe& out) {
  if (is_nan(alpha) || alpha < MIN_COMPENSATED_ALPHA) return false;
  float e_avg = e_avg_lookup(t.albedo_avg, alpha);
  float lost = 1.0f - e_avg;
  if (!(lost > 1.0e-6f)) return false;
  out.alpha = alpha;
  out.norm = FRAC_1_PI / lost;
  out.e_avg = e_avg;
  return true;
}
RPX_HD float ms_lost(const EnergyTables& t, const MsLobe& l, float mu) {
  return fmax_r(1.0f - e_lookup(t.albedo, mu, l.alpha), 0.0f);
}
RPX_HD float ms_shape(const EnergyTables& t, const MsLobe& l, float mo, float mi) {
  return ms_lost(t, l, mo) * ms_lost(t, l, mi) * l.norm;
}

// ---------------------------------------------------------------- GGX (isotropic)
RPX_HD float ggx_d(V3 h, float a2) {
  float t = a2 * h.z * h.z + (h.x * h.x + h.y * h.y);
  if (t <= 0.0f) return 0.0f;
  return a2 / (PI * t * t);
}
RPX_HD float ggx_lambda(V3 w, float a2) {
  float cos2 = w.z * w.z;
  if (cos2 <= 0.0f) return RPX_INF;
  float tan2 = (w.x * w.x + w.y * w.y) / cos2;
  return (sqrt_(1.0f + a2 * tan2) - 1.0f) * 0.5f;
}
RPX_HD float ggx_reflection_pdf(float a2, V3 wo, V3 h) {
  float g1 = 1.0f / (1.0f + ggx_lambda(wo, a2));
  return g1 * ggx_d(h, a2) / (4.0f * wo.z);
}
RPX_HD float ggx_microfacet(float a2, V3 wo, V3 wi, V3 h) {
  float g2 = 1.0f / (1.0f + ggx_lambda(wo, a2) + ggx_lambda(wi, a2));
  return ggx_d(h, a2) * g2 / (4.0f * wo.z * wi.z);
}
RPX_HD float schlick_f(float f0, float f90, float cos) {
  float m = clampf(1.0f - cos, 0.0f, 1.0f);
  float m2 = m * m;
  return f0 + (f90 - f0) * (m2 * m2 * m);
}
RPX_HD float fresnel_dielectric(float cos_i, float eta) {
  cos_i = clampf(cos_i, 0.0f, 1.0f);
  float sin2_t = (1.0f - cos_i * cos_i) / (eta * eta);
  if (sin2_t >= 1.0f) return 1.0f;
  float cos_t = sqrt_(1.0f - sin2_t);
  float rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
  float rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
  return 0.5f * (rs * rs + rp * rp);
}
RPX_HD float thin_reflectance(float cos, float eta) {
  float r = fresnel_dielectric(cos, eta);
  return 2.0f * r / (1.0f + r);
}
RPX_HD bool refract_dir(V3 wo, V3 h, float eta, V3& out) {
  float cos_i = dot(wo, h);
  float sin2_t = (1.0f - cos_i * cos_i) / (eta * eta);
  if (sin2_t >= 1.0f) return false;
  float cos_t = sqrt_(1.0f - sin2_t);
  out = normalize(h * (cos_i / eta - cos_t) - wo / eta);
  return true;
}
RPX_HD V3 sample_vndf(V3 wo, float ax, float ay, float u0, float u1) {
  V3 vh = normalize(v3(ax * wo.x, ay * wo.y, wo.z));
  float lensq = vh.x * vh.x + vh.y * vh.y;
  V3 t1 = lensq > 0.0f ? v3(-vh.y, vh.x, 0.0f) / sqrt_(lensq) : v3(1.0f, 0.0f, 0.0f);
  V3 t2 = cross(vh, t1);
  float r = sqrt_(u0);
  float sp, cp;
  sin_cos(TAU * u1, sp, cp);
  float p1 = r * cp;
  float s = 0.5f * (1.0f + vh.z);
  float p2 = (1.0f - s) * sqrt_(fmax_r(1.0f - p1 * p1, 0.0f)) + s * (r * sp);
  V3 nh = t1 * p1 + t2 * p2 + vh * sqrt_(fmax_r(1.0f - p1 * p1 - p2 * p2, 0.0f));
  return normalize(v3(ax * nh.x, ay * nh.y, fmax_r(nh.z, 1.0e-7f)));
}

// ---------------------------------------------------------------- the BSDF of one surface side
struct Bsdf {
  C3 diffuse, metal_f0, metal_f90;
  float metallic, dielectric, f0, f90;
  float spec_alpha, spec_a2;
  float coat, coat_alpha, coat_a2;
  float clear, eta;
  C3 tint;
  u32 thin;
  u32 ms_spec_on;
  MsLobe ms_spec_lobe;
  C3 ms_spec_colour;
  u32 ms_coat_on;
  MsLobe ms_coat_lobe;
  float ms_coat_c;
  u32 ms_clear_on;
  float ms_clear_alpha;
  Layer base_layer, coat_layer;
  float diffuse_norm;
};

struct LobeWeights {
  float coat, specular, diffuse, clear, ms;
};

constexpr float COAT_F0 = 0.04f;
constexpr float MAX_REFLECT_PROBABILITY = 0.99f;
constexpr u32 MATERIAL_GLASS = 1;
constexpr u32 MATERIAL_FLAG_LIGHT = 1;
constexpr u32 MATERIAL_FLAG_THIN = 2;

RPX_HD C3 clamp01(C3 c) { return c3(clampf(c.r, 0, 1), clampf(c.g, 0, 1), clampf(c.b, 0, 1)); }
RPX_HD bool has_specular(const Bsdf& b) { return (b.f90 > 0.0f && b.dielectric > 0.0f) || b.metallic > 0.0f; }

// bsdf.rs:384-438 with multiple_scattering (:451-493) and layers (:426-449).
RPX_HD Bsdf bsdf_surface(const EnergyTables& t, u32 kind, C3 base_color, float metallic_in, float roughness_in,
                         float ior, float transmission_in, float coat_in, float coat_roughness_in, u32 flags,
                         float absorption_mm, bool entering) {
  C3 base = clamp01(base_color);
  float metallic = kind == MATERIAL_GLASS ? 0.0f : clampf(metallic_in, 0.0f, 1.0f);
  float transmission = clampf(transmission_in, 0.0f, 1.0f);
  ior = fmax_r(ior, 1.0f);
  float f0 = ((ior - 1.0f) / (ior + 1.0f)) * ((ior - 1.0f) / (ior + 1.0f));
  float roughness = clampf(roughness_in, 0.0f, 1.0f);
  float coat_roughness = clampf(coat_roughness_in, 0.0f, 1.0f);
  float clear = (1.0f - metallic) * transmission;
  bool thin = (flags & MATERIAL_FLAG_THIN) != 0 || ior < 1.0001f;
  float eta = (thin || entering) ? ior : 1.0f / ior;
  C3 tint;
  if (thin) tint = base;
  else if (absorption_mm > 0.0f) tint = splat(1.0f);
  else tint = c3(sqrt_(base.r), sqrt_(base.g), sqrt_(base.b));
  float coat = (clear > 0.0f && !thin && !entering) ? 0.0f : clampf(coat_in, 0.0f, 1.0f);
  Bsdf b;
  b.diffuse = base * ((1.0f - metallic) * (1.0f - transmission));
  b.metal_f0 = base;
  b.metal_f90 = c3(fmin_r(50.0f * base.r, 1.0f), fmin_r(50.0f * base.g, 1.0f), fmin_r(50.0f * base.b, 1.0f));
  b.metallic = metallic;
  b.dielectric = (1.0f - metallic) * (1.0f - transmission);
  b.f0 = f0;
  b.f90 = fmin_r(50.0f * f0, 1.0f);
  b.spec_alpha = fmax_r(roughness * roughness, MIN_ALPHA);
  b.spec_a2 = b.spec_alpha * b.spec_alpha;
  b.coat = coat;
  b.coat_alpha = fmax_r(coat_roughness * coat_roughness, MIN_ALPHA);
  b.coat_a2 = b.coat_alpha * b.coat_alpha;
  b.clear = clear;
  b.eta = eta;
  b.tint = tint;
  b.thin = thin ? 1u : 0u;
  b.base_layer = layer_clear();
  b.coat_layer = layer_clear();
  b.diffuse_norm = 1.0f;
  // multiple scattering
  float alpha = b.spec_alpha;
  MsLobe lobe;
  b.ms_spec_on = 0;
  b.ms_spec_colour = splat(0.0f);
  b.ms_spec_lobe = MsLobe{0, 0, 0};
  if (ms_lobe_make(t, alpha, lobe) && has_specular(b)) {
    float diel = b.dielectric > 0.0f ? b.dielectric * multiple_fresnel(schlick_average(b.f0, b.f90), lobe.e_avg) : 0.0f;
    C3 colour;
    if (b.metallic > 0.0f) {
      C3 m = c3(multiple_fresnel(schlick_average(b.metal_f0.r, b.metal_f90.r), lobe.e_avg),
                multiple_fresnel(schlick_average(b.metal_f0.g, b.metal_f90.g), lobe.e_avg),
                multiple_fresnel(schlick_average(b.metal_f0.b, b.metal_f90.b), lobe.e_avg));
      colour = m * b.metallic + splat(diel);
    } else {
      colour = splat(diel);
    }
    if (peak(colour) > 0.0f) {
      b.ms_spec_on = 1;
      b.ms_spec_lobe = lobe;
      b.ms_spec_colour = colour;
    }
  }
  b.ms_coat_on = 0;
  b.ms_coat_lobe = MsLobe{0, 0, 0};
  b.ms_coat_c = 0.0f;
  if (coat > 0.0f) {
    MsLobe cl;
    if (ms_lobe_make(t, b.coat_alpha, cl)) {
      float c = coat * multiple_fresnel(schlick_average(COAT_F0, 1.0f), cl.e_avg);
      if (c > 0.0f) {
        b.ms_coat_on = 1;
        b.ms_coat_lobe = cl;
        b.ms_coat_c = c;
      }
    }
  }
  b.ms_clear_on = (clear > 0.0f && !thin && alpha >= MIN_COMPENSATED_ALPHA) ? 1u : 0u;
  b.ms_clear_alpha = alpha;
  // layers
  if (b.diffuse.r > 0.0f || b.diffuse.g > 0.0f || b.diffuse.b > 0.0f) {
    float ms = b.ms_spec_on ? multiple_fresnel(schlick_average(b.f0, b.f90), b.ms_spec_lobe.e_avg) : 0.0f;
    b.base_layer = layer_make(t, b.f0, b.f90, 1.0f, alpha, ms);
  }
  if (coat > 0.0f) {
    float ms = b.ms_coat_on ? multiple_fresnel(schlick_average(COAT_F0, 1.0f), b.ms_coat_lobe.e_avg) : 0.0f;
    b.coat_layer = layer_make(t, COAT_F0, 1.0f, coat, b.coat_alpha, ms);
  }
  b.diffuse_norm = joint_norm(t, b.base_layer, b.coat_layer);
  return b;
}

// bsdf.rs:494-518: the Lambertian BSDF of the visible floor.
RPX_HD Bsdf bsdf_lambert(C3 albedo) {
  Bsdf b;
  b.diffuse = clamp01(albedo);
  b.metal_f0 = splat(0.0f);
  b.metal_f90 = splat(0.0f);
  b.metallic = 0.0f;
  b.dielectric = 1.0f;
  b.f0 = 0.0f;
  b.f90 = 0.0f;
  b.spec_alpha = 1.0f;
  b.spec_a2 = 1.0f;
  b.coat = 0.0f;
  b.coat_alpha = 1.0f;
  b.coat_a2 = 1.0f;
  b.clear = 0.0f;
  b.eta = 1.0f;
  b.tint = splat(1.0f);
  b.thin = 0;
  b.ms_spec_on = 0;
  b.ms_spec_lobe = MsLobe{0, 0, 0};
  b.ms_spec_colour = splat(0.0f);
  b.ms_coat_on = 0;
  b.ms_coat_lobe = MsLobe{0, 0, 0};
  b.ms_coat_c = 0.0f;
  b.ms_clear_on = 0;
  b.ms_clear_alpha = 1.0f;
  b.base_layer = layer_clear();
  b.coat_layer = layer_clear();
  b.diffuse_norm = 1.0f;
  return b;
}

RPX_HD C3 specular_fresnel(const Bsdf& b, float cos) {
  float d = schlick_f(b.f0, b.f90, cos);
  if (b.metallic <= 0.0f && b.clear <= 0.0f) return splat(d);
  C3 metal = c3(schlick_f(b.metal_f0.r, b.metal_f90.r, cos), schlick_f(b.metal_f0.g, b.metal_f90.g, cos),
                schlick_f(b.metal_f0.b, b.metal_f90.b, cos));
  return splat(d) * b.dielectric + metal * b.metallic;
}
RPX_HD float coat_fresnel(const Bsdf& b, float cos) { return b.coat <= 0.0f ? 0.0f : b.coat * schlick_f(COAT_F0, 1.0f, cos); }
RPX_HD float clear_fresnel(const Bsdf& b, float cos) {
  return b.thin ? thin_reflectance(cos, b.eta) : fresnel_dielectric(cos, b.eta);
}
RPX_HD float reflect_probability(const Bsdf& b, float cos) {
  return fmin_r(clear_fresnel(b, cos), MAX_REFLECT_PROBABILITY);
}
RPX_HD float reflect_chance(const Bsdf& b, float mu_o, float cos_h) {
  return b.ms_clear_on ? clear_fresnel(b, cos_h) : reflect_probability(b, mu_o);
}
RPX_HD bool clear_compensation(const EnergyTables& t, const Bsdf& b, float mu_o, float& k) {
  if (!b.ms_clear_on) return false;
  float e = e_clear_albedo(t, mu_o, b.ms_clear_alpha, b.eta);
  k = 1.0f / fmax_r(e, 0.05f);
  return true;
}

RPX_HD LobeWeights lobe_weights(const EnergyTables& t, const Bsdf& b, V3 wo) {
  float c = coat_fresnel(b, wo.z);
  float below = 1.0f - c;
  float specular = has_specular(b) ? below * fmax_r(luminance(specular_fresnel(b, wo.z)), 0.0f) : 0.0f;
  float diff = below * luminance(b.diffuse) * layer_passes(t, b.base_layer, wo.z);
  float clr = below * b.clear;
  float ms = 0.0f;
  if (b.ms_spec_on) ms += below * fmax_r(luminance(b.ms_spec_colour), 0.0f) * ms_lost(t, b.ms_spec_lobe, wo.z);
  if (b.ms_coat_on) ms += b.ms_coat_c * ms_lost(t, b.ms_coat_lobe, wo.z);
  float total = c + specular + diff + clr + ms;
  if (total > 0.0f) {
    float inv = 1.0f / total;
    return LobeWeights{c * inv, specular * inv, diff * inv, clr * inv, ms * inv};
  }
  return LobeWeights{0, 0, 0, 0, 0};
}

RPX_HD bool refraction_half(const Bsdf& b, V3 wo, V3 wi, V3& h, float& co, float& ci) {
  h = normalize(wo + wi * b.eta);
  if (h.z < 0.0f) h = -h;
  co = dot(wo, h);
  ci = dot(wi, h);
  return co > 0.0f && ci < 0.0f && length_sq(h) > 0.5f;
}

RPX_HD C3 eval_refraction(const EnergyTables& t, const Bsdf& b, V3 wo, V3 wi) {
  V3 h;
  float co, ci;
  if (!refraction_half(b, wo, wi, h, co, ci)) return splat(0.0f);
  float f = fresnel_dielectric(co, b.eta);
  float denom = co + b.eta * ci;
  float g2 = 1.0f / (1.0f + ggx_lambda(wo, b.spec_a2) + ggx_lambda(wi, b.spec_a2));
  float value = (1.0f - f) * ggx_d(h, b.spec_a2) * g2 * (-ci) * co / (-wi.z * wo.z * denom * denom);
  float k;
  if (clear_compensation(t, b, wo.z, k)) value *= k;
  float attenuation = layer_passes(t, b.coat_layer, wo.z);
  return b.tint * (b.clear * attenuation * value);
}

// `diffuse_only`: the diffuse lobe alone (lights without the glossy flag).
RPX_HD C3 bsdf_eval(const EnergyTables& t, const Bsdf& b, V3 wo, V3 wi, bool diffuse_only) {
  if (wo.z <= 0.0f) return splat(0.0f);
  if (wi.z <= 0.0f) {
    if (wi.z < 0.0f && !diffuse_only && b.clear > 0.0f && !b.thin) return eval_refraction(t, b, wo, wi);
    return splat(0.0f);
  }
  float attenuation = layer_passes(t, b.coat_layer, wo.z) * layer_passes(t, b.coat_layer, wi.z);
  C3 diff = b.diffuse * (FRAC_1_PI * layer_passes(t, b.base_layer, wo.z) * layer_passes(t, b.base_layer, wi.z) *
                         b.diffuse_norm);
  if (diffuse_only) return diff * attenuation;
  V3 h = normalize(wo + wi);
  float cos_h = fmax_r(dot(wo, h), 0.0f);
  C3 below = diff;
  if (has_specular(b)) below = below + specular_fresnel(b, cos_h) * ggx_microfacet(b.spec_a2, wo, wi, h);
  if (b.ms_spec_on) below = below + b.ms_spec_colour * ms_shape(t, b.ms_spec_lobe, wo.z, wi.z);
  if (b.clear > 0.0f) {
    float c = b.clear * clear_fresnel(b, cos_h) * ggx_microfacet(b.spec_a2, wo, wi, h);
    float k;
    if (clear_compensation(t, b, wo.z, k)) c *= k;
    below = below + splat(c);
  }
  C3 f = below * attenuation;
  if (b.coat > 0.0f) f = f + splat(coat_fresnel(b, cos_h) * ggx_microfacet(b.coat_a2, wo, wi, h));
  if (b.ms_coat_on) f = f + splat(b.ms_coat_c * ms_shape(t, b.ms_coat_lobe, wo.z, wi.z));
  return f;
}

RPX_HD float pdf_with(const EnergyTables& t, const Bsdf& b, const LobeWeights& w, V3 wo, V3 wi) {
  (void)t;
  float p = (w.diffuse + w.ms) * wi.z * FRAC_1_PI;
  if (w.specular > 0.0f || w.coat > 0.0f || w.clear > 0.0f) {
    V3 h = normalize(wo + wi);
    if (dot(wo, h) > 0.0f) {
      if (w.specular > 0.0f) p += w.specular * ggx_reflection_pdf(b.spec_a2, wo, h);
      if (w.coat > 0.0f) p += w.coat * ggx_reflection_pdf(b.coat_a2, wo, h);
      if (w.clear > 0.0f) p += w.clear * reflect_chance(b, wo.z, dot(wo, h)) * ggx_reflection_pdf(b.spec_a2, wo, h);
    }
  }
  return p;
}

RPX_HD float refraction_pdf(const Bsdf& b, const LobeWeights& w, V3 wo, V3 wi) {
  if (is_nan(w.clear) || w.clear <= 0.0f) return 0.0f;
  V3 h;
  float co, ci;
  if (!refraction_half(b, wo, wi, h, co, ci)) return 0.0f;
  float g1 = 1.0f / (1.0f + ggx_lambda(wo, b.spec_a2));
  float visible = g1 * ggx_d(h, b.spec_a2) * co / wo.z;
  float denom = co + b.eta * ci;
  float jacobian = b.eta * b.eta * (-ci) / (denom * denom);
  return w.clear * (1.0f - reflect_chance(b, wo.z, co)) * visible * jacobian;
}

RPX_HD float bsdf_pdf(const EnergyTables& t, const Bsdf& b, V3 wo, V3 wi) {
  if (wo.z <= 0.0f) return 0.0f;
  if (wi.z <= 0.0f) {
    if (wi.z < 0.0f && b.clear > 0.0f && !b.thin) return refraction_pdf(b, lobe_weights(t, b, wo), wo, wi);
    return 0.0f;
  }
  return pdf_with(t, b, lobe_weights(t, b, wo), wo, wi);
}

RPX_HD bool pdf_ok(float p) { return p > 0.0f && is_finite(p); }

struct BsdfSample {
  V3 wi;
  C3 weight;
  float pdf;
};

RPX_HD bool finish_reflection(const EnergyTables& t, const Bsdf& b, const LobeWeights& w, V3 wo, V3 wi,
                              BsdfSample& out) {
  if (wi.z <= 0.0f) return false;
  float p = pdf_with(t, b, w, wo, wi);
  if (!pdf_ok(p)) return false;
  out.wi = wi;
  out.weight = bsdf_eval(t, b, wo, wi, false) * (wi.z / p);
  out.pdf = p;
  return true;
}
RPX_HD bool finish_refraction(const EnergyTables& t, const Bsdf& b, const LobeWeights& w, V3 wo, V3 h,
                              BsdfSample& out) {
  V3 wi;
  if (!refract_dir(wo, h, b.eta, wi)) return false;
  if (wi.z >= 0.0f) return false;
  float p = refraction_pdf(b, w, wo, wi);
  if (!pdf_ok(p)) return false;
  out.wi = wi;
  out.weight = eval_refraction(t, b, wo, wi) * (-wi.z / p);
  out.pdf = p;
  return true;
}

// bsdf.rs:833-992. Thin walls (straight samples) are refused at scene load.
RPX_HD bool bsdf_sample(const EnergyTables& t, const Bsdf& b, V3 wo, float u0, float u1, BsdfSample& out) {
  if (wo.z <= 0.0f) return false;
  LobeWeights w = lobe_weights(t, b, wo);
  V3 wi;
  if (u0 < w.coat) {
    V3 h = sample_vndf(wo, b.coat_alpha, b.coat_alpha, u0 / w.coat, u1);
    wi = h * (2.0f * dot(wo, h)) - wo;
  } else if (u0 < w.coat + w.specular) {
    V3 h = sample_vndf(wo, b.spec_alpha, b.spec_alpha, (u0 - w.coat) / w.specular, u1);
    wi = h * (2.0f * dot(wo, h)) - wo;
  } else if (w.diffuse + w.ms > 0.0f && (w.clear <= 0.0f || u0 < w.coat + w.specular + (w.diffuse + w.ms))) {
    float start = w.coat + w.specular;
    float uu = clampf((u0 - start) / (w.diffuse + w.ms), 0.0f, 0.99999994f);
    wi = cosine_hemisphere(uu, u1);
  } else if (w.clear > 0.0f) {
    float start = w.coat + w.specular + (w.diffuse + w.ms);
    float uc = clampf((u0 - start) / w.clear, 0.0f, 0.99999994f);
    if (b.ms_clear_on) {
      V3 h = sample_vndf(wo, b.spec_alpha, b.spec_alpha, uc, u1);
      float fresnel = clear_fresnel(b, dot(wo, h));
      u64 key = ((u64)f2u(uc) << 32) | (u64)f2u(u1);
      float choice = (float)(mix64(key ^ 0x9e3779b97f4a7c15ull) >> 40) * (1.0f / 16777216.0f);
      if (choice < fresnel) return finish_reflection(t, b, w, wo, h * (2.0f * dot(wo, h)) - wo, out);
      return finish_refraction(t, b, w, wo, h, out);
    }
    float pr = reflect_probability(b, wo.z);
    if (uc < pr) {
      V3 h = sample_vndf(wo, b.spec_alpha, b.spec_alpha, uc / pr, u1);
      return finish_reflection(t, b, w, wo, h * (2.0f * dot(wo, h)) - wo, out);
    }
    float ut = clampf((uc - pr) / (1.0f - pr), 0.0f, 0.99999994f);
    V3 h = sample_vndf(wo, b.spec_alpha, b.spec_alpha, ut, u1);
    return finish_refraction(t, b, w, wo, h, out);
  } else {
    return false;
  }
  return finish_reflection(t, b, w, wo, wi, out);
}

}  // namespace rpx
