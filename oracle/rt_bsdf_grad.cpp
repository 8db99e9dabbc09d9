// ORACLE — test infrastructure only (see rt_core.hpp and rt_bsdf_grad.hpp).
#include "rt_bsdf_grad.hpp"
#include <cmath>

namespace orc {
namespace {

// A value and its derivative with respect to the one parameter being differentiated.
struct Dv {
  float v = 0.0f, d = 0.0f;
  Dv() = default;
  Dv(float value) : v(value), d(0.0f) {}
  Dv(float value, float slope) : v(value), d(slope) {}
};
Dv operator+(Dv a, Dv b) { return Dv(a.v + b.v, a.d + b.d); }
Dv operator-(Dv a, Dv b) { return Dv(a.v - b.v, a.d - b.d); }
Dv operator-(Dv a) { return Dv(-a.v, -a.d); }
Dv operator*(Dv a, Dv b) { return Dv(a.v * b.v, a.d * b.v + a.v * b.d); }
Dv operator/(Dv a, Dv b) {
  float q = a.v / b.v;
  return Dv(q, (a.d - q * b.d) / b.v);
}
Dv sqrt_d(Dv a) {
  float s = std::sqrt(a.v);
  return Dv(s, s > 0.0f ? a.d * 0.5f / s : 0.0f);
}
// f32::max / f32::min keep the operand they return, and its slope.
Dv max_d(Dv a, Dv b) { return a.v >= b.v ? a : b; }
Dv min_d(Dv a, Dv b) { return a.v <= b.v ? a : b; }
Dv clamp_d(Dv x, float lo, float hi) { return x.v < lo ? Dv(lo) : (x.v > hi ? Dv(hi) : x); }

struct DRgb {
  Dv r, g, b;
};
DRgb splat_d(Dv v) { return DRgb{v, v, v}; }
DRgb operator+(DRgb a, DRgb b) { return DRgb{a.r + b.r, a.g + b.g, a.b + b.b}; }
DRgb operator*(DRgb a, DRgb b) { return DRgb{a.r * b.r, a.g * b.g, a.b * b.b}; }
DRgb operator*(DRgb a, Dv s) { return DRgb{a.r * s, a.g * s, a.b * s}; }

using energy::MIN_COMPENSATED_ALPHA;
using energy::NODES_IOR;
using energy::NODES_MU;
using energy::NODES_ROUGHNESS;
constexpr float UNIT = 1.0f / 65535.0f;
constexpr float COAT_F0_G = 0.04f;

// energy.rs lookups: bilinear in (sqrt mu, sqrt alpha), the slope that of the cell's interpolation.
void locate_d(Dv x, int n, int& i, Dv& w) {
  Dv t = clamp_d(x, 0.0f, 1.0f) * Dv((float)(n - 1));
  u64 ti = f_to_usize(t.v);
  i = (int)std::min<u64>(ti, (u64)(n - 2));
  w = t - Dv((float)i);
}
Dv lookup_d(const uint16_t* table, float mu, Dv alpha) {
  int i, j;
  Dv u, v;
  locate_d(Dv(std::sqrt(fmax_r(mu, 0.0f))), NODES_MU, i, u);
  locate_d(sqrt_d(max_d(alpha, Dv(0.0f))), NODES_ROUGHNESS, j, v);
  auto at = [&](int jj, int ii) { return (float)table[jj * NODES_MU + ii] * UNIT; };
  float low = at(j, i) + (at(j, i + 1) - at(j, i)) * u.v;
  float high = at(j + 1, i) + (at(j + 1, i + 1) - at(j + 1, i)) * u.v;
  return Dv(low) + Dv(high - low) * v;
}
Dv average_d(const std::vector<float>& a, Dv alpha) {
  int j;
  Dv v;
  locate_d(sqrt_d(max_d(alpha, Dv(0.0f))), NODES_ROUGHNESS, j, v);
  return Dv(a[j]) + Dv(a[j + 1] - a[j]) * v;
}
Dv albedo_d(float mu, Dv alpha) { return lookup_d(energy::tables().albedo.data(), mu, alpha); }
Dv schlick_albedo_d(float mu, Dv alpha) { return lookup_d(energy::tables().schlick.data(), mu, alpha); }
Dv clear_albedo_d(float mu, Dv alpha, Dv eta) {
  const energy::Tables& t = energy::tables();
  const std::vector<uint16_t>* table;
  Dv ior;
  if (eta.v >= 1.0f) {
    table = &t.clear_in;
    ior = eta;
  } else {
    table = &t.clear_out;
    ior = Dv(1.0f) / eta;
  }
  Dv x = (ior - Dv(1.0f)) / (ior + Dv(1.0f));
  const float max_x = (energy::MAX_IOR - 1.0f) / (energy::MAX_IOR + 1.0f);
  int k;
  Dv w;
  locate_d(x / Dv(max_x), NODES_IOR, k, w);
  const int slice = NODES_MU * NODES_ROUGHNESS;
  Dv a = lookup_d(table->data() + k * slice, mu, alpha);
  Dv b = lookup_d(table->data() + (k + 1) * slice, mu, alpha);
  return a + (b - a) * w;
}
Dv schlick_average_d(Dv f0, Dv f90) { return f0 + (f90 - f0) * Dv(1.0f / 21.0f); }
Dv multiple_fresnel_d(Dv f_avg, Dv e_avg) {
  Dv d = Dv(1.0f) - f_avg * (Dv(1.0f) - e_avg);
  return d.v > 0.0f ? f_avg * f_avg * e_avg / d : Dv(0.0f);
}

struct LayerD {
  Dv f0, f90;
  float strength = 0.0f;
  bool rough = false;
  Dv alpha, ms, norm = Dv(1.0f);
  Dv albedo_at(float mu) const {
    Dv e;
    if (rough) {
      Dv ee = albedo_d(mu, alpha);
      e = f0 * ee + (f90 - f0) * schlick_albedo_d(mu, alpha) + ms * (Dv(1.0f) - ee);
    } else {
      float m = clampf(1.0f - mu, 0.0f, 1.0f);
      float m2 = m * m;
      e = f0 + (f90 - f0) * Dv(m2 * m2 * m);
    }
    return clamp_d(e * Dv(strength), 0.0f, 1.0f);
  }
  Dv passes(float mu) const { return Dv(1.0f) - albedo_at(mu); }
};
LayerD make_layer(Dv f0, Dv f90, float strength, Dv alpha, Dv ms) {
  if (std::isnan(strength) || strength <= 0.0f || (f90.v <= 0.0f && f0.v <= 0.0f)) return LayerD();
  LayerD l;
  l.f0 = f0;
  l.f90 = f90;
  l.strength = strength;
  l.rough = alpha.v >= MIN_COMPENSATED_ALPHA;
  l.alpha = alpha;
  l.ms = ms;
  Dv avg;
  if (l.rough) {
    Dv e_avg = average_d(energy::tables().albedo_avg, alpha);
    avg = f0 * e_avg + (f90 - f0) * average_d(energy::tables().schlick_avg, alpha) + ms * (Dv(1.0f) - e_avg);
  } else {
    avg = schlick_average_d(f0, f90);
  }
  avg = avg * Dv(strength);
  l.norm = Dv(1.0f) / (Dv(1.0f) - min_d(avg, Dv(0.999f)));
  return l;
}
Dv joint_norm_d(const LayerD& below, const LayerD& above) {
  if (above.strength <= 0.0f) return below.norm;
  if (below.strength <= 0.0f) return above.norm;
  static const float nodes[8][2] = {
      {0.019855072f, 0.05061427f}, {0.10166676f, 0.11119052f}, {0.2372338f, 0.15685332f},
      {0.40828268f, 0.18134189f},  {0.5917173f, 0.18134189f},  {0.7627662f, 0.15685332f},
      {0.8983332f, 0.11119052f},   {0.9801449f, 0.05061427f},
  };
  Dv sum(-0.0f);
  for (int k = 0; k < 8; ++k) {
    float mu = nodes[k][0], w = nodes[k][1];
    sum = sum + Dv(w * 2.0f * mu) * below.passes(mu) * above.passes(mu);
  }
  return Dv(1.0f) / max_d(sum, Dv(0.001f));
}

struct LobeD {
  Dv alpha, norm, e_avg;
  Dv lost(float mu) const { return max_d(Dv(1.0f) - albedo_d(mu, alpha), Dv(0.0f)); }
  Dv shape(float mo, float mi) const { return lost(mo) * lost(mi) * norm; }
};
bool make_lobe(Dv alpha, LobeD& out) {
  if (std::isnan(alpha.v) || alpha.v < MIN_COMPENSATED_ALPHA) return false;
  Dv e_avg = average_d(energy::tables().albedo_avg, alpha);
  Dv lost = Dv(1.0f) - e_avg;
  if (!(lost.v > 1.0e-6f)) return false;
  out.alpha = alpha;
  out.norm = Dv(FRAC_1_PI) / lost;
  out.e_avg = e_avg;
  return true;
}

Dv ggx_d(Dv a2, Dv hx, Dv hy, Dv hz) {
  Dv t = a2 * hz * hz + (hx * hx + hy * hy);
  if (t.v <= 0.0f) return Dv(0.0f);
  return a2 / (Dv(PI) * t * t);
}
Dv ggx_lambda(Dv a2, Vec3 w) {
  float cos2 = w.z * w.z;
  if (cos2 <= 0.0f) return Dv(INF);
  float tan2 = (w.x * w.x + w.y * w.y) / cos2;
  return (sqrt_d(Dv(1.0f) + a2 * Dv(tan2)) - Dv(1.0f)) * Dv(0.5f);
}
Dv microfacet(Dv a2, Vec3 wo, Vec3 wi, Vec3 h) {
  Dv g2 = Dv(1.0f) / (Dv(1.0f) + ggx_lambda(a2, wo) + ggx_lambda(a2, wi));
  return ggx_d(a2, Dv(h.x), Dv(h.y), Dv(h.z)) * g2 / Dv(4.0f * wo.z * wi.z);
}
Dv schlick_d(Dv f0, Dv f90, float cos) {
  float m = clampf(1.0f - cos, 0.0f, 1.0f);
  float m2 = m * m;
  return f0 + (f90 - f0) * Dv(m2 * m2 * m);
}
Dv fresnel_dielectric_d(Dv cos_i, Dv eta) {
  cos_i = clamp_d(cos_i, 0.0f, 1.0f);
  Dv sin2_t = (Dv(1.0f) - cos_i * cos_i) / (eta * eta);
  if (sin2_t.v >= 1.0f) return Dv(1.0f);
  Dv cos_t = sqrt_d(Dv(1.0f) - sin2_t);
  Dv rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
  Dv rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
  return Dv(0.5f) * (rs * rs + rp * rp);
}

// Bsdf::surface_at with one parameter carried as a dual number (bsdf.rs:384-493).
struct BsdfD {
  DRgb diffuse, metal_f0, metal_f90, tint, ms_spec_colour;
  float metallic = 0, dielectric = 0, coat = 0, clear = 0;
  Dv f0, f90, spec_alpha, spec_a2, coat_alpha, coat_a2, eta, ms_coat_c, ms_clear_alpha, diffuse_norm = Dv(1.0f);
  bool thin = false, ms_spec_on = false, ms_coat_on = false, ms_clear_on = false;
  LobeD ms_spec_lobe, ms_coat_lobe;
  LayerD base_layer, coat_layer;

  bool has_specular() const { return (f90.v > 0.0f && dielectric > 0.0f) || metallic > 0.0f; }
  DRgb specular_fresnel(float cos) const {
    Dv d = schlick_d(f0, f90, cos);
    if (metallic <= 0.0f && clear <= 0.0f) return splat_d(d);
    DRgb metal{schlick_d(metal_f0.r, metal_f90.r, cos), schlick_d(metal_f0.g, metal_f90.g, cos),
               schlick_d(metal_f0.b, metal_f90.b, cos)};
    return splat_d(d) * Dv(dielectric) + metal * Dv(metallic);
  }
  Dv coat_fresnel(float cos) const {
    if (coat <= 0.0f) return Dv(0.0f);
    float m = clampf(1.0f - cos, 0.0f, 1.0f);
    float m2 = m * m;
    return Dv(coat * (COAT_F0_G + (1.0f - COAT_F0_G) * (m2 * m2 * m)));
  }
  Dv clear_fresnel(Dv cos) const {
    if (!thin) return fresnel_dielectric_d(cos, eta);
    Dv r = fresnel_dielectric_d(cos, eta);
    return Dv(2.0f) * r / (Dv(1.0f) + r);
  }
  bool clear_compensation(float mu_o, Dv& k) const {
    if (!ms_clear_on) return false;
    Dv e = clear_albedo_d(mu_o, ms_clear_alpha, eta);
    k = Dv(1.0f) / max_d(e, Dv(0.05f));
    return true;
  }
  DRgb eval_refraction(Vec3 wo, Vec3 wi) const {
    Dv sx = Dv(wo.x) + Dv(wi.x) * eta, sy = Dv(wo.y) + Dv(wi.y) * eta, sz = Dv(wo.z) + Dv(wi.z) * eta;
    Dv len = sqrt_d(sx * sx + sy * sy + sz * sz);
    Dv hx = sx, hy = sy, hz = sz;
    if (len.v > 0.0f && std::isfinite(len.v)) {
      hx = sx / len;
      hy = sy / len;
      hz = sz / len;
    }
    if (hz.v < 0.0f) {
      hx = -hx;
      hy = -hy;
      hz = -hz;
    }
    Dv co = Dv(wo.x) * hx + Dv(wo.y) * hy + Dv(wo.z) * hz;
    Dv ci = Dv(wi.x) * hx + Dv(wi.y) * hy + Dv(wi.z) * hz;
    Dv hl = hx * hx + hy * hy + hz * hz;
    if (!(co.v > 0.0f && ci.v < 0.0f && hl.v > 0.5f)) return splat_d(Dv(0.0f));
    Dv f = fresnel_dielectric_d(co, eta);
    Dv denom = co + eta * ci;
    Dv g2 = Dv(1.0f) / (Dv(1.0f) + ggx_lambda(spec_a2, wo) + ggx_lambda(spec_a2, wi));
    Dv value = (Dv(1.0f) - f) * ggx_d(spec_a2, hx, hy, hz) * g2 * (-ci) * co / (Dv(-wi.z * wo.z) * denom * denom);
    Dv k;
    if (clear_compensation(wo.z, k)) value = value * k;
    Dv attenuation = coat_layer.passes(wo.z);
    return tint * (Dv(clear) * attenuation * value);
  }
  DRgb eval(Vec3 wo, Vec3 wi, Lobes lobes) const {
    if (wo.z <= 0.0f) return splat_d(Dv(0.0f));
    if (wi.z <= 0.0f) {
      if (wi.z < 0.0f && lobes == Lobes::All && clear > 0.0f && !thin) return eval_refraction(wo, wi);
      return splat_d(Dv(0.0f));
    }
    Dv attenuation = coat_layer.passes(wo.z) * coat_layer.passes(wi.z);
    DRgb diff = diffuse * (Dv(FRAC_1_PI) * base_layer.passes(wo.z) * base_layer.passes(wi.z) * diffuse_norm);
    if (lobes == Lobes::Diffuse) return diff * attenuation;
    Vec3 h = (wo + wi).normalize();
    float cos_h = fmax_r(wo.dot(h), 0.0f);
    DRgb below = diff;
    if (has_specular()) below = below + specular_fresnel(cos_h) * microfacet(spec_a2, wo, wi, h);
    if (ms_spec_on) below = below + ms_spec_colour * ms_spec_lobe.shape(wo.z, wi.z);
    if (clear > 0.0f) {
      Dv c = Dv(clear) * clear_fresnel(Dv(cos_h)) * microfacet(spec_a2, wo, wi, h);
      Dv k;
      if (clear_compensation(wo.z, k)) c = c * k;
      below = below + splat_d(c);
    }
    DRgb f = below * attenuation;
    if (coat > 0.0f) f = f + splat_d(coat_fresnel(cos_h) * microfacet(coat_a2, wo, wi, h));
    if (ms_coat_on) f = f + splat_d(ms_coat_c * ms_coat_lobe.shape(wo.z, wi.z));
    return f;
  }
};

BsdfD surface_d(const Material& m, bool entering, int param) {
  auto var = [&](float v, int k) { return Dv(v, k == param ? 1.0f : 0.0f); };
  DRgb base{clamp_d(var(m.base_color.r, 1), 0, 1), clamp_d(var(m.base_color.g, 2), 0, 1),
            clamp_d(var(m.base_color.b, 3), 0, 1)};
  float metallic = m.kind == MATERIAL_GLASS ? 0.0f : clampf(m.metallic, 0.0f, 1.0f);
  float transmission = clampf(m.transmission, 0.0f, 1.0f);
  Dv ior = max_d(var(m.ior, 4), Dv(1.0f));
  Dv q = (ior - Dv(1.0f)) / (ior + Dv(1.0f));
  Dv roughness = clamp_d(var(m.roughness, 0), 0.0f, 1.0f);
  float coat_roughness = clampf(m.coat_roughness, 0.0f, 1.0f);
  float clear = (1.0f - metallic) * transmission;
  bool thin = (m.flags & MATERIAL_FLAG_THIN) != 0 || ior.v < 1.0001f;
  BsdfD b;
  b.eta = (thin || entering) ? ior : Dv(1.0f) / ior;
  if (thin) b.tint = base;
  else if (m.absorption_mm > 0.0f) b.tint = splat_d(Dv(1.0f));
  else b.tint = DRgb{sqrt_d(base.r), sqrt_d(base.g), sqrt_d(base.b)};
  b.coat = (clear > 0.0f && !thin && !entering) ? 0.0f : clampf(m.coat, 0.0f, 1.0f);
  b.diffuse = base * Dv((1.0f - metallic) * (1.0f - transmission));
  b.metal_f0 = base;
  b.metal_f90 = DRgb{min_d(Dv(50.0f) * base.r, Dv(1.0f)), min_d(Dv(50.0f) * base.g, Dv(1.0f)),
                     min_d(Dv(50.0f) * base.b, Dv(1.0f))};
  b.metallic = metallic;
  b.dielectric = (1.0f - metallic) * (1.0f - transmission);
  b.f0 = q * q;
  b.f90 = min_d(Dv(50.0f) * b.f0, Dv(1.0f));
  b.spec_alpha = max_d(roughness * roughness, Dv(MIN_ALPHA));
  b.spec_a2 = b.spec_alpha * b.spec_alpha;
  b.coat_alpha = Dv(fmax_r(coat_roughness * coat_roughness, MIN_ALPHA));
  b.coat_a2 = b.coat_alpha * b.coat_alpha;
  b.clear = clear;
  b.thin = thin;
  // multiple scattering (Bsdf::multiple_scattering)
  Dv alpha = b.spec_alpha;
  LobeD lobe;
  if (make_lobe(alpha, lobe) && b.has_specular()) {
    auto ms = [&](Dv a0, Dv a90) { return multiple_fresnel_d(schlick_average_d(a0, a90), lobe.e_avg); };
    Dv diel = b.dielectric > 0.0f ? Dv(b.dielectric) * ms(b.f0, b.f90) : Dv(0.0f);
    DRgb colour = b.metallic > 0.0f
                      ? DRgb{ms(b.metal_f0.r, b.metal_f90.r), ms(b.metal_f0.g, b.metal_f90.g),
                             ms(b.metal_f0.b, b.metal_f90.b)} * Dv(b.metallic) + splat_d(diel)
                      : splat_d(diel);
    if (fmax_r(fmax_r(colour.r.v, colour.g.v), colour.b.v) > 0.0f) {
      b.ms_spec_on = true;
      b.ms_spec_lobe = lobe;
      b.ms_spec_colour = colour;
    }
  }
  if (b.coat > 0.0f) {
    LobeD cl;
    if (make_lobe(b.coat_alpha, cl)) {
      Dv c = Dv(b.coat) * multiple_fresnel_d(schlick_average_d(Dv(COAT_F0_G), Dv(1.0f)), cl.e_avg);
      if (c.v > 0.0f) {
        b.ms_coat_on = true;
        b.ms_coat_lobe = cl;
        b.ms_coat_c = c;
      }
    }
  }
  b.ms_clear_on = clear > 0.0f && !thin && alpha.v >= MIN_COMPENSATED_ALPHA;
  b.ms_clear_alpha = alpha;
  // layers (Bsdf::layers)
  if (b.diffuse.r.v > 0.0f || b.diffuse.g.v > 0.0f || b.diffuse.b.v > 0.0f) {
    Dv ms = b.ms_spec_on ? multiple_fresnel_d(schlick_average_d(b.f0, b.f90), b.ms_spec_lobe.e_avg) : Dv(0.0f);
    b.base_layer = make_layer(b.f0, b.f90, 1.0f, alpha, ms);
  }
  if (b.coat > 0.0f) {
    Dv ms = b.ms_coat_on ? multiple_fresnel_d(schlick_average_d(Dv(COAT_F0_G), Dv(1.0f)), b.ms_coat_lobe.e_avg)
                         : Dv(0.0f);
    b.coat_layer = make_layer(Dv(COAT_F0_G), Dv(1.0f), b.coat, b.coat_alpha, ms);
  }
  b.diffuse_norm = joint_norm_d(b.base_layer, b.coat_layer);
  return b;
}

}  // namespace

void bsdf_param_derivs_exact(const Material& m, bool entering, Vec3 wo, Vec3 wi, Lobes lobes, Rgb out[5]) {
  for (int k = 0; k < 5; ++k) {
    DRgb f = surface_d(m, entering, k).eval(wo, wi, lobes);
    out[k] = Rgb(f.r.d, f.g.d, f.b.d);
  }
}

}  // namespace orc
