// Derivatives of the BSDF's value with respect to one material parameter (roughness, base colour
// r, g or b, IOR) at fixed directions, by forward-mode differentiation: every quantity the BSDF
// builds from the parameters carries its value and its derivative (a dual number) through the same
// steps as bsdf_surface and bsdf_eval (rpx_bsdf.h). The energy tables are read as the float code
// reads them (bilinear in sqrt(mu) and sqrt(alpha), linear between IOR slices), so their
// derivative is that of the interpolation inside the cell the value falls in; a clamp passes the
// derivative inside its range and stops it outside; min and max take the chosen operand's.
// Branches are decided on the values, as the float code decides them.
#pragma once
#include "rpx_bsdf.h"

namespace rpx {

struct Dual {
  float v, d;
};
RPX_HD Dual dc(float v) { return Dual{v, 0.0f}; }
RPX_HD Dual operator+(Dual a, Dual b) { return Dual{a.v + b.v, a.d + b.d}; }
RPX_HD Dual operator-(Dual a, Dual b) { return Dual{a.v - b.v, a.d - b.d}; }
RPX_HD Dual operator-(Dual a) { return Dual{-a.v, -a.d}; }
RPX_HD Dual operator*(Dual a, Dual b) { return Dual{a.v * b.v, a.d * b.v + a.v * b.d}; }
RPX_HD Dual operator/(Dual a, Dual b) {
  float q = a.v / b.v;
  return Dual{q, (a.d - q * b.d) / b.v};
}
RPX_HD Dual operator+(Dual a, float b) { return Dual{a.v + b, a.d}; }
RPX_HD Dual operator+(float a, Dual b) { return Dual{a + b.v, b.d}; }
RPX_HD Dual operator-(Dual a, float b) { return Dual{a.v - b, a.d}; }
RPX_HD Dual operator-(float a, Dual b) { return Dual{a - b.v, -b.d}; }
RPX_HD Dual operator*(Dual a, float b) { return Dual{a.v * b, a.d * b}; }
RPX_HD Dual operator*(float a, Dual b) { return Dual{a * b.v, a * b.d}; }
RPX_HD Dual operator/(Dual a, float b) { return Dual{a.v / b, a.d / b}; }
RPX_HD Dual operator/(float a, Dual b) {
  float q = a / b.v;
  return Dual{q, -q * b.d / b.v};
}
RPX_HD Dual dsqrt(Dual a) {
  float s = sqrt_(a.v);
  return Dual{s, s > 0.0f ? a.d * (0.5f / s) : 0.0f};
}
RPX_HD Dual dmax(Dual a, Dual b) { return a.v >= b.v ? a : b; }
RPX_HD Dual dmin(Dual a, Dual b) { return a.v <= b.v ? a : b; }
RPX_HD Dual dmax(Dual a, float b) { return a.v >= b ? a : dc(b); }
RPX_HD Dual dmin(Dual a, float b) { return a.v <= b ? a : dc(b); }
RPX_HD Dual dclamp(Dual x, float lo, float hi) {
  if (x.v < lo) return dc(lo);
  if (x.v > hi) return dc(hi);
  return x;
}

struct DC3 {
  Dual r, g, b;
};
RPX_HD DC3 dsplat(Dual v) { return DC3{v, v, v}; }
RPX_HD DC3 operator+(DC3 a, DC3 b) { return DC3{a.r + b.r, a.g + b.g, a.b + b.b}; }
RPX_HD DC3 operator*(DC3 a, DC3 b) { return DC3{a.r * b.r, a.g * b.g, a.b * b.b}; }
RPX_HD DC3 operator*(DC3 a, Dual s) { return DC3{a.r * s, a.g * s, a.b * s}; }
RPX_HD DC3 operator*(DC3 a, float s) { return DC3{a.r * s, a.g * s, a.b * s}; }
RPX_HD Dual dluminance(DC3 c) { return 0.2126f * c.r + 0.7152f * c.g + 0.0722f * c.b; }
RPX_HD C3 dvalue(DC3 c) { return c3(c.r.v, c.g.v, c.b.v); }
RPX_HD C3 dderiv(DC3 c) { return c3(c.r.d, c.g.d, c.b.d); }

struct DV3 {
  Dual x, y, z;
};
RPX_HD DV3 dv3(V3 a) { return DV3{dc(a.x), dc(a.y), dc(a.z)}; }
RPX_HD Dual ddot(V3 a, DV3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// ---------------------------------------------------------------- energy tables
RPX_HD void de_locate(Dual x, int n, int& i, Dual& w) {
  Dual t = dclamp(x, 0.0f, 1.0f) * (float)(n - 1);
  u32 ti = f_to_u32(t.v);
  i = (int)(ti < (u32)(n - 2) ? ti : (u32)(n - 2));
  w = t - (float)i;
}
RPX_HD Dual de_lookup(const uint16_t* table, float mu, Dual alpha) {
  int i, j;
  float u;
  Dual v;
  e_locate(sqrt_(fmax_r(mu, 0.0f)), E_NODES_MU, i, u);
  de_locate(dsqrt(dmax(alpha, 0.0f)), E_NODES_ROUGHNESS, j, v);
  const float UNIT = 1.0f / 65535.0f;
  float a00 = (float)table[j * E_NODES_MU + i] * UNIT, a01 = (float)table[j * E_NODES_MU + i + 1] * UNIT;
  float a10 = (float)table[(j + 1) * E_NODES_MU + i] * UNIT, a11 = (float)table[(j + 1) * E_NODES_MU + i + 1] * UNIT;
  float low = a00 + (a01 - a00) * u;
  float high = a10 + (a11 - a10) * u;
  return low + (high - low) * v;
}
RPX_HD Dual de_avg_lookup(const float* a, Dual alpha) {
  int j;
  Dual v;
  de_locate(dsqrt(dmax(alpha, 0.0f)), E_NODES_ROUGHNESS, j, v);
  return a[j] + (a[j + 1] - a[j]) * v;
}
RPX_HD Dual de_clear_albedo(const EnergyTables& t, float mu, Dual alpha, Dual eta) {
  const uint16_t* table;
  Dual ior;
  if (eta.v >= 1.0f) {
    table = t.clear_in;
    ior = eta;
  } else {
    table = t.clear_out;
    ior = 1.0f / eta;
  }
  Dual x = (ior - 1.0f) / (ior + 1.0f);
  const float MAX_X = (E_MAX_IOR - 1.0f) / (E_MAX_IOR + 1.0f);
  int k;
  Dual w;
  de_locate(x / MAX_X, E_NODES_IOR, k, w);
  Dual a = de_lookup(table + k * E_SLICE, mu, alpha);
  Dual b = de_lookup(table + (k + 1) * E_SLICE, mu, alpha);
  return a + (b - a) * w;
}
RPX_HD Dual dschlick_average(Dual f0, Dual f90) { return f0 + (f90 - f0) * (1.0f / 21.0f); }
RPX_HD Dual dmultiple_fresnel(Dual f_avg, Dual e_avg) {
  Dual d = 1.0f - f_avg * (1.0f - e_avg);
  return d.v > 0.0f ? f_avg * f_avg * e_avg / d : dc(0.0f);
}

struct DLayer {
  Dual f0, f90;
  float strength;
  u32 rough;
  Dual alpha, ms, norm;
};
RPX_HD DLayer dlayer_clear() { return DLayer{dc(0), dc(0), 0.0f, 0, dc(0), dc(0), dc(1.0f)}; }
RPX_HD Dual dlayer_albedo(const EnergyTables& t, const DLayer& l, float mu) {
  Dual e;
  if (l.rough) {
    Dual ee = de_lookup(t.albedo, mu, l.alpha);
    e = l.f0 * ee + (l.f90 - l.f0) * de_lookup(t.schlick, mu, l.alpha) + l.ms * (1.0f - ee);
  } else {
    float m = clampf(1.0f - mu, 0.0f, 1.0f);
    float m2 = m * m;
    e = l.f0 + (l.f90 - l.f0) * (m2 * m2 * m);
  }
  return dclamp(e * l.strength, 0.0f, 1.0f);
}
RPX_HD Dual dlayer_passes(const EnergyTables& t, const DLayer& l, float mu) { return 1.0f - dlayer_albedo(t, l, mu); }
RPX_HD DLayer dlayer_make(const EnergyTables& t, Dual f0, Dual f90, float strength, Dual alpha, Dual ms) {
  if (is_nan(strength) || strength <= 0.0f || (f90.v <= 0.0f && f0.v <= 0.0f)) return dlayer_clear();
  DLayer l;
  l.f0 = f0;
  l.f90 = f90;
  l.strength = strength;
  l.rough = alpha.v >= MIN_COMPENSATED_ALPHA ? 1u : 0u;
  l.alpha = alpha;
  l.ms = ms;
  Dual avg;
  if (l.rough) {
    Dual e_avg = de_avg_lookup(t.albedo_avg, alpha);
    avg = f0 * e_avg + (f90 - f0) * de_avg_lookup(t.schlick_avg, alpha) + ms * (1.0f - e_avg);
  } else {
    avg = dschlick_average(f0, f90);
  }
  avg = avg * strength;
  l.norm = 1.0f / (1.0f - dmin(avg, 0.999f));
  return l;
}
RPX_HD Dual djoint_norm(const EnergyTables& t, const DLayer& below, const DLayer& above) {
  if (above.strength <= 0.0f) return below.norm;
  if (below.strength <= 0.0f) return above.norm;
  const float G_MU[8] = {0.019855072f, 0.10166676f, 0.2372338f, 0.40828268f,
                         0.5917173f,   0.7627662f,  0.8983332f, 0.9801449f};
  const float G_W[8] = {0.05061427f, 0.11119052f, 0.15685332f, 0.18134189f,
                        0.18134189f, 0.15685332f, 0.11119052f, 0.05061427f};
  Dual sum = dc(-0.0f);
  for (int k = 0; k < 8; ++k)
    sum = sum + G_W[k] * 2.0f * G_MU[k] * dlayer_passes(t, below, G_MU[k]) * dlayer_passes(t, above, G_MU[k]);
  return 1.0f / dmax(sum, 0.001f);
}

struct DMsLobe {
  Dual alpha, norm, e_avg;
};
RPX_HD bool dms_lobe_make(const EnergyTables& t, Dual alpha, DMsLobe& out) {
  if (is_nan(alpha.v) || alpha.v < MIN_COMPENSATED_ALPHA) return false;
  Dual e_avg = de_avg_lookup(t.albedo_avg, alpha);
  Dual lost = 1.0f - e_avg;
  if (!(lost.v > 1.0e-6f)) return false;
  out.alpha = alpha;
  out.norm = FRAC_1_PI / lost;
  out.e_avg = e_avg;
  return true;
}
RPX_HD Dual dms_lost(const EnergyTables& t, const DMsLobe& l, float mu) {
  return dmax(1.0f - de_lookup(t.albedo, mu, l.alpha), 0.0f);
}
RPX_HD Dual dms_shape(const EnergyTables& t, const DMsLobe& l, float mo, float mi) {
  return dms_lost(t, l, mo) * dms_lost(t, l, mi) * l.norm;
}

// ---------------------------------------------------------------- GGX, Fresnel
RPX_HD Dual dggx_d(DV3 h, Dual a2) {
  Dual t = a2 * h.z * h.z + (h.x * h.x + h.y * h.y);
  if (t.v <= 0.0f) return dc(0.0f);
  return a2 / (PI * t * t);
}
RPX_HD Dual dggx_lambda(V3 w, Dual a2) {
  float cos2 = w.z * w.z;
  if (cos2 <= 0.0f) return dc(RPX_INF);
  float tan2 = (w.x * w.x + w.y * w.y) / cos2;
  return (dsqrt(1.0f + a2 * tan2) - 1.0f) * 0.5f;
}
RPX_HD Dual dggx_microfacet(Dual a2, V3 wo, V3 wi, V3 h) {
  Dual g2 = 1.0f / (1.0f + dggx_lambda(wo, a2) + dggx_lambda(wi, a2));
  return dggx_d(dv3(h), a2) * g2 / (4.0f * wo.z * wi.z);
}
RPX_HD Dual dschlick_f(Dual f0, Dual f90, float cos) {
  float m = clampf(1.0f - cos, 0.0f, 1.0f);
  float m2 = m * m;
  return f0 + (f90 - f0) * (m2 * m2 * m);
}
RPX_HD Dual dfresnel_dielectric(Dual cos_i, Dual eta) {
  cos_i = dclamp(cos_i, 0.0f, 1.0f);
  Dual sin2_t = (1.0f - cos_i * cos_i) / (eta * eta);
  if (sin2_t.v >= 1.0f) return dc(1.0f);
  Dual cos_t = dsqrt(1.0f - sin2_t);
  Dual rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
  Dual rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
  return 0.5f * (rs * rs + rp * rp);
}
RPX_HD Dual dthin_reflectance(Dual cos, Dual eta) {
  Dual r = dfresnel_dielectric(cos, eta);
  return 2.0f * r / (1.0f + r);
}

// ---------------------------------------------------------------- the BSDF of one surface side
struct DBsdf {
  DC3 diffuse, metal_f0, metal_f90;
  float metallic, dielectric;
  Dual f0, f90;
  Dual spec_alpha, spec_a2;
  float coat;
  Dual coat_alpha, coat_a2;
  float clear;
  Dual eta;
  DC3 tint;
  u32 thin;
  u32 ms_spec_on;
  DMsLobe ms_spec_lobe;
  DC3 ms_spec_colour;
  u32 ms_coat_on;
  DMsLobe ms_coat_lobe;
  Dual ms_coat_c;
  u32 ms_clear_on;
  Dual ms_clear_alpha;
  DLayer base_layer, coat_layer;
  Dual diffuse_norm;
};
RPX_HD bool dhas_specular(const DBsdf& b) { return (b.f90.v > 0.0f && b.dielectric > 0.0f) || b.metallic > 0.0f; }

// bsdf_surface with the parameters as dual numbers (rpx_bsdf.h, after bsdf.rs:384-493).
RPX_HD DBsdf dbsdf_surface(const EnergyTables& t, u32 kind, DC3 base_color, float metallic_in, Dual roughness_in,
                           Dual ior, float transmission_in, float coat_in, float coat_roughness_in, u32 flags,
                           float absorption_mm, bool entering) {
  DC3 base{dclamp(base_color.r, 0, 1), dclamp(base_color.g, 0, 1), dclamp(base_color.b, 0, 1)};
  float metallic = kind == MATERIAL_GLASS ? 0.0f : clampf(metallic_in, 0.0f, 1.0f);
  float transmission = clampf(transmission_in, 0.0f, 1.0f);
  ior = dmax(ior, 1.0f);
  Dual q = (ior - 1.0f) / (ior + 1.0f);
  Dual f0 = q * q;
  Dual roughness = dclamp(roughness_in, 0.0f, 1.0f);
  float coat_roughness = clampf(coat_roughness_in, 0.0f, 1.0f);
  float clear = (1.0f - metallic) * transmission;
  bool thin = (flags & MATERIAL_FLAG_THIN) != 0 || ior.v < 1.0001f;
  Dual eta = (thin || entering) ? ior : 1.0f / ior;
  DC3 tint;
  if (thin) tint = base;
  else if (absorption_mm > 0.0f) tint = dsplat(dc(1.0f));
  else tint = DC3{dsqrt(base.r), dsqrt(base.g), dsqrt(base.b)};
  float coat = (clear > 0.0f && !thin && !entering) ? 0.0f : clampf(coat_in, 0.0f, 1.0f);
  DBsdf b;
  b.diffuse = base * ((1.0f - metallic) * (1.0f - transmission));
  b.metal_f0 = base;
  b.metal_f90 = DC3{dmin(50.0f * base.r, 1.0f), dmin(50.0f * base.g, 1.0f), dmin(50.0f * base.b, 1.0f)};
  b.metallic = metallic;
  b.dielectric = (1.0f - metallic) * (1.0f - transmission);
  b.f0 = f0;
  b.f90 = dmin(50.0f * f0, 1.0f);
  b.spec_alpha = dmax(roughness * roughness, MIN_ALPHA);
  b.spec_a2 = b.spec_alpha * b.spec_alpha;
  b.coat = coat;
  b.coat_alpha = dc(fmax_r(coat_roughness * coat_roughness, MIN_ALPHA));
  b.coat_a2 = b.coat_alpha * b.coat_alpha;
  b.clear = clear;
  b.eta = eta;
  b.tint = tint;
  b.thin = thin ? 1u : 0u;
  b.base_layer = dlayer_clear();
  b.coat_layer = dlayer_clear();
  b.diffuse_norm = dc(1.0f);
  Dual alpha = b.spec_alpha;
  DMsLobe lobe;
  b.ms_spec_on = 0;
  b.ms_spec_colour = dsplat(dc(0.0f));
  b.ms_spec_lobe = DMsLobe{dc(0), dc(0), dc(0)};
  if (dms_lobe_make(t, alpha, lobe) && dhas_specular(b)) {
    Dual diel = b.dielectric > 0.0f ? b.dielectric * dmultiple_fresnel(dschlick_average(b.f0, b.f90), lobe.e_avg)
                                    : dc(0.0f);
    DC3 colour;
    if (b.metallic > 0.0f) {
      DC3 m{dmultiple_fresnel(dschlick_average(b.metal_f0.r, b.metal_f90.r), lobe.e_avg),
            dmultiple_fresnel(dschlick_average(b.metal_f0.g, b.metal_f90.g), lobe.e_avg),
            dmultiple_fresnel(dschlick_average(b.metal_f0.b, b.metal_f90.b), lobe.e_avg)};
      colour = m * b.metallic + dsplat(diel);
    } else {
      colour = dsplat(diel);
    }
    if (peak(dvalue(colour)) > 0.0f) {
      b.ms_spec_on = 1;
      b.ms_spec_lobe = lobe;
      b.ms_spec_colour = colour;
    }
  }
  b.ms_coat_on = 0;
  b.ms_coat_lobe = DMsLobe{dc(0), dc(0), dc(0)};
  b.ms_coat_c = dc(0.0f);
  if (coat > 0.0f) {
    DMsLobe cl;
    if (dms_lobe_make(t, b.coat_alpha, cl)) {
      Dual c = coat * dmultiple_fresnel(dschlick_average(dc(COAT_F0), dc(1.0f)), cl.e_avg);
      if (c.v > 0.0f) {
        b.ms_coat_on = 1;
        b.ms_coat_lobe = cl;
        b.ms_coat_c = c;
      }
    }
  }
  b.ms_clear_on = (clear > 0.0f && !thin && alpha.v >= MIN_COMPENSATED_ALPHA) ? 1u : 0u;
  b.ms_clear_alpha = alpha;
  if (b.diffuse.r.v > 0.0f || b.diffuse.g.v > 0.0f || b.diffuse.b.v > 0.0f) {
    Dual ms = b.ms_spec_on ? dmultiple_fresnel(dschlick_average(b.f0, b.f90), b.ms_spec_lobe.e_avg) : dc(0.0f);
    b.base_layer = dlayer_make(t, b.f0, b.f90, 1.0f, alpha, ms);
  }
  if (coat > 0.0f) {
    Dual ms = b.ms_coat_on ? dmultiple_fresnel(dschlick_average(dc(COAT_F0), dc(1.0f)), b.ms_coat_lobe.e_avg)
                           : dc(0.0f);
    b.coat_layer = dlayer_make(t, dc(COAT_F0), dc(1.0f), coat, b.coat_alpha, ms);
  }
  b.diffuse_norm = djoint_norm(t, b.base_layer, b.coat_layer);
  return b;
}

RPX_HD DC3 dspecular_fresnel(const DBsdf& b, float cos) {
  Dual d = dschlick_f(b.f0, b.f90, cos);
  if (b.metallic <= 0.0f && b.clear <= 0.0f) return dsplat(d);
  DC3 metal{dschlick_f(b.metal_f0.r, b.metal_f90.r, cos), dschlick_f(b.metal_f0.g, b.metal_f90.g, cos),
            dschlick_f(b.metal_f0.b, b.metal_f90.b, cos)};
  return dsplat(d) * b.dielectric + metal * b.metallic;
}
RPX_HD Dual dcoat_fresnel(const DBsdf& b, float cos) {
  return b.coat <= 0.0f ? dc(0.0f) : dc(b.coat * schlick_f(COAT_F0, 1.0f, cos));
}
RPX_HD Dual dclear_fresnel(const DBsdf& b, Dual cos) {
  return b.thin ? dthin_reflectance(cos, b.eta) : dfresnel_dielectric(cos, b.eta);
}
RPX_HD bool dclear_compensation(const EnergyTables& t, const DBsdf& b, float mu_o, Dual& k) {
  if (!b.ms_clear_on) return false;
  Dual e = de_clear_albedo(t, mu_o, b.ms_clear_alpha, b.eta);
  k = 1.0f / dmax(e, 0.05f);
  return true;
}

RPX_HD DC3 deval_refraction(const EnergyTables& t, const DBsdf& b, V3 wo, V3 wi) {
  // The half vector moves with eta: h = normalize(wo + wi eta), turned to the upper side.
  DV3 s{wo.x + wi.x * b.eta, wo.y + wi.y * b.eta, wo.z + wi.z * b.eta};
  Dual len = dsqrt(s.x * s.x + s.y * s.y + s.z * s.z);
  DV3 h = s;
  if (len.v > 0.0f && is_finite(len.v)) h = DV3{s.x / len, s.y / len, s.z / len};
  if (h.z.v < 0.0f) h = DV3{-h.x, -h.y, -h.z};
  Dual co = ddot(wo, h), ci = ddot(wi, h);
  Dual hl = h.x * h.x + h.y * h.y + h.z * h.z;
  if (!(co.v > 0.0f && ci.v < 0.0f && hl.v > 0.5f)) return dsplat(dc(0.0f));
  Dual f = dfresnel_dielectric(co, b.eta);
  Dual denom = co + b.eta * ci;
  Dual g2 = 1.0f / (1.0f + dggx_lambda(wo, b.spec_a2) + dggx_lambda(wi, b.spec_a2));
  Dual value = (1.0f - f) * dggx_d(h, b.spec_a2) * g2 * (-ci) * co / (-wi.z * wo.z * denom * denom);
  Dual k;
  if (dclear_compensation(t, b, wo.z, k)) value = value * k;
  Dual attenuation = dlayer_passes(t, b.coat_layer, wo.z);
  return b.tint * (b.clear * attenuation * value);
}

// bsdf_eval with the BSDF's dual numbers: the value and its derivative per channel.
RPX_HD DC3 dbsdf_eval(const EnergyTables& t, const DBsdf& b, V3 wo, V3 wi, bool diffuse_only) {
  if (wo.z <= 0.0f) return dsplat(dc(0.0f));
  if (wi.z <= 0.0f) {
    if (wi.z < 0.0f && !diffuse_only && b.clear > 0.0f && !b.thin) return deval_refraction(t, b, wo, wi);
    return dsplat(dc(0.0f));
  }
  Dual attenuation = dlayer_passes(t, b.coat_layer, wo.z) * dlayer_passes(t, b.coat_layer, wi.z);
  DC3 diff = b.diffuse * (FRAC_1_PI * dlayer_passes(t, b.base_layer, wo.z) * dlayer_passes(t, b.base_layer, wi.z) *
                          b.diffuse_norm);
  if (diffuse_only) return diff * attenuation;
  V3 h = normalize(wo + wi);
  float cos_h = fmax_r(dot(wo, h), 0.0f);
  DC3 below = diff;
  if (dhas_specular(b)) below = below + dspecular_fresnel(b, cos_h) * dggx_microfacet(b.spec_a2, wo, wi, h);
  if (b.ms_spec_on) below = below + b.ms_spec_colour * dms_shape(t, b.ms_spec_lobe, wo.z, wi.z);
  if (b.clear > 0.0f) {
    Dual c = b.clear * dclear_fresnel(b, dc(cos_h)) * dggx_microfacet(b.spec_a2, wo, wi, h);
    Dual k;
    if (dclear_compensation(t, b, wo.z, k)) c = c * k;
    below = below + dsplat(c);
  }
  DC3 f = below * attenuation;
  if (b.coat > 0.0f) f = f + dsplat(dcoat_fresnel(b, cos_h) * dggx_microfacet(b.coat_a2, wo, wi, h));
  if (b.ms_coat_on) f = f + dsplat(b.ms_coat_c * dms_shape(t, b.ms_coat_lobe, wo.z, wi.z));
  return f;
}

}  // namespace rpx
