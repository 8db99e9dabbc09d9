// ORACLE — test infrastructure only (see rt_core.hpp). The denoiser of
// engine/crates/rt-post/src/denoise.rs; reference lines per function.
#include "rt_denoise.hpp"

#include <algorithm>

namespace orc {
namespace denoise {

// denoise.rs:100-119
Preset Preset::clamped() const {
  Preset p;
  p.levels = levels < 1 ? 1 : (levels > LEVELS ? LEVELS : levels);
  p.variance_radius = variance_radius < 1 ? 1 : (variance_radius > 3 ? 3 : variance_radius);
  return p;
}

namespace {
const float KERNEL[5] = {1.0f / 16.0f, 1.0f / 4.0f, 3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};
const float GAUSS3[3] = {0.25f, 0.5f, 0.25f};

// denoise.rs:410-413
inline float luminance(float r, float g, float b) { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }

// denoise.rs:416-423
inline float pow128(float x) {
  float v = x;
  for (int i = 0; i < 7; ++i) v *= v;
  return v;
}

// denoise.rs:426-447
struct Guide {
  u32 part, material;
  float n[3];
  float z;
  float dz[2];
  float a[3];
  bool foreground() const { return part != 0; }
  bool same_surface(const Guide& o) const { return part == o.part && material == o.material; }
};

inline float dot3(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// denoise.rs:750-758
inline float depth_term(const Guide& g, const Guide& q, float dx, float dy) {
  if (!(is_finite(g.z) && is_finite(q.z))) return g.z == q.z ? 0.0f : INF;
  float expected = std::fabs(g.dz[0] * dx + g.dz[1] * dy);
  float eps = 1.0e-3f * std::fabs(g.z) + 1.0e-6f;
  return std::fabs(g.z - q.z) / (SIGMA_DEPTH * expected + eps);
}

// denoise.rs:761-764
inline float albedo_term(const Guide& g, const Guide& q) {
  return (std::fabs(g.a[0] - q.a[0]) + std::fabs(g.a[1] - q.a[1]) + std::fabs(g.a[2] - q.a[2])) / SIGMA_ALBEDO;
}

// denoise.rs:769-779
inline float falloff(float x) {
  float t = 1.0f - x * (1.0f / 16.0f);
  if (is_nan(t) || t <= 0.0f) return 0.0f;
  t = t * t;
  t = t * t;
  t = t * t;
  return t * t;
}

// denoise.rs:782-785
inline float surface_weight(const Guide& g, const Guide& q, float dx, float dy) {
  return pow128(fmax_r(dot3(g.n, q.n), 0.0f)) * falloff(depth_term(g, q, dx, dy));
}

inline float finite_or_zero(float v) { return is_finite(v) ? v : 0.0f; }

// denoise.rs:264-287
bool shape_ok(const Input& in) {
  if (in.width == 0 || (u64)in.origin_y + in.rows > in.height) return false;
  return in.sums && in.lum2 && in.albedo && in.normal && in.depth && in.part_id && in.material_id;
}

// denoise.rs:338-356
Status check_rows(const Input& in, u32 y0, u32 y1, Preset preset) {
  if (!shape_ok(in)) return BAD_SHAPE;
  u32 halo = preset.halo();
  u32 band_end = in.origin_y + in.rows;
  u32 lo = y0 > halo ? y0 - halo : 0;
  u64 hi = std::min<u64>((u64)y1 + halo, in.height);
  if (y0 > y1 || y0 < in.origin_y || y1 > band_end || in.origin_y > lo || band_end < hi) return BAD_ROWS;
  return OK;
}

// denoise.rs:359-365
float strength_of(float s) {
  if (is_nan(s)) return 1.0f;
  return clampf(s, 0.0f, 1.0f);
}
}  // namespace

// denoise.rs:450-742
std::vector<float> filter_band(const Input& in, Preset preset) {
  const size_t w = in.width, h = in.rows, n = w * h;
  const u32 samples = std::max<u32>(in.samples, 1);
  const float inv = 1.0f / (float)samples;
  std::vector<float> out(in.sums, in.sums + n * 4);

  // Per pixel: the edge-stopping guides, the divisor that takes the albedo out, the demodulated
  // light and its variance.
  std::vector<Guide> guides(n);
  for (size_t p = 0; p < n; ++p) {
    Guide& g = guides[p];
    g.part = in.part_id[p];
    g.material = in.material_id[p];
    for (int k = 0; k < 3; ++k) {
      g.n[k] = finite_or_zero(in.normal[p * 3 + k]);
      g.a[k] = finite_or_zero(in.albedo[p * 3 + k]);
    }
    g.z = in.depth[p];
    g.dz[0] = g.dz[1] = 0.0f;
  }
  for (size_t y = 0; y < h; ++y)
    for (size_t x = 0; x < w; ++x) {
      size_t p = y * w + x;
      Guide g = guides[p];
      if (!g.foreground() || !is_finite(g.z)) continue;
      auto zof = [&](bool has, size_t q, float& z) {
        if (has && guides[q].same_surface(g) && is_finite(guides[q].z)) {
          z = guides[q].z;
          return true;
        }
        return false;
      };
      auto grad = [&](bool ha, size_t a, bool hb, size_t b) {
        float za = 0, zb = 0;
        bool sa = zof(ha, a, za), sb = zof(hb, b, zb);
        if (sa && sb) return 0.5f * (zb - za);
        if (sa) return g.z - za;
        if (sb) return zb - g.z;
        return 0.0f;
      };
      guides[p].dz[0] = grad(x > 0, p - 1, x + 1 < w, p + 1);
      guides[p].dz[1] = grad(y > 0, p - w, y + 1 < h, p + w);
    }
  std::vector<float> divisor(n * 3, 1.0f), light(n * 3, 0.0f), var(n, 0.0f);
  for (size_t p = 0; p < n; ++p) {
    const Guide& g = guides[p];
    if (!g.foreground()) continue;
    float c[3] = {in.sums[p * 4] * inv, in.sums[p * 4 + 1] * inv, in.sums[p * 4 + 2] * inv};
    for (int k = 0; k < 3; ++k) {
      float d = fmax_r(g.a[k], MIN_ALBEDO);
      divisor[p * 3 + k] = d;
      light[p * 3 + k] = c[k] / d;
    }
    bool enough = in.counts ? in.counts[p] >= (float)MIN_SAMPLES_FOR_VARIANCE : samples >= MIN_SAMPLES_FOR_VARIANCE;
    if (enough) {
      float l = luminance(c[0], c[1], c[2]);
      float m2 = in.lum2[p] * inv;
      float dof = in.counts ? in.counts[p] - 1.0f : (float)(samples - 1);
      float v = fmax_r(m2 - l * l, 0.0f) / dof;
      float ld = luminance(divisor[p * 3], divisor[p * 3 + 1], divisor[p * 3 + 2]);
      var[p] = v / (ld * ld);
    }
  }
  auto lum_of = [](const std::vector<float>& lt, size_t p) { return luminance(lt[p * 3], lt[p * 3 + 1], lt[p * 3 + 2]); };

  // Fireflies (denoise.rs:541-575), in place in row-major order as the reference writes it.
  std::vector<float> before(n);
  for (size_t p = 0; p < n; ++p) before[p] = lum_of(light, p);
  for (size_t y = 0; y < h; ++y)
    for (size_t x = 0; x < w; ++x) {
      size_t p = y * w + x;
      const Guide g = guides[p];
      if (!g.foreground()) continue;
      float brightest = 0.0f, vmax = 0.0f;
      u32 count = 0;
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          long qx = (long)x + dx, qy = (long)y + dy;
          if ((dx == 0 && dy == 0) || qx < 0 || qy < 0 || qx >= (long)w || qy >= (long)h) continue;
          size_t q = (size_t)qy * w + (size_t)qx;
          if (!g.same_surface(guides[q])) continue;
          brightest = fmax_r(brightest, before[q]);
          vmax = fmax_r(vmax, var[q]);
          ++count;
        }
      float limit = FIREFLY_RATIO * brightest + FIREFLY_FLOOR;
      if (count >= 3 && before[p] > limit) {
        float k = limit / before[p];
        for (int c = 0; c < 3; ++c) light[p * 3 + c] *= k;
        var[p] = vmax;
      }
    }

  // Too few samples for a pixel's own variance (denoise.rs:576-624).
  auto few = [&](size_t p) {
    return in.counts ? in.counts[p] < (float)MIN_SAMPLES_FOR_VARIANCE : samples < MIN_SAMPLES_FOR_VARIANCE;
  };
  bool any_few = false;
  for (size_t p = 0; p < n && !any_few; ++p) any_few = few(p);
  if (any_few) {
    int r = (int)preset.variance_radius;
    for (size_t y = 0; y < h; ++y)
      for (size_t x = 0; x < w; ++x) {
        size_t p = y * w + x;
        const Guide g = guides[p];
        if (!g.foreground() || !few(p)) continue;
        float sw = 0.0f, m1 = 0.0f, m2 = 0.0f;
        for (int dy = -r; dy <= r; ++dy) {
          long qy = (long)y + dy;
          if (qy < 0 || qy >= (long)h) continue;
          for (int dx = -r; dx <= r; ++dx) {
            long qx = (long)x + dx;
            if (qx < 0 || qx >= (long)w) continue;
            size_t q = (size_t)qy * w + (size_t)qx;
            const Guide& gq = guides[q];
            if (!g.same_surface(gq)) continue;
            float wgt = surface_weight(g, gq, (float)dx, (float)dy);
            float l = lum_of(light, q);
            sw += wgt;
            m1 += wgt * l;
            m2 += wgt * l * l;
          }
        }
        if (sw > 0.0f) {
          m1 = m1 / sw;
          m2 = m2 / sw;
          var[p] = fmax_r(m2 - m1 * m1, 0.0f);
        }
      }
  }

  // The passes of the a-trous transform (denoise.rs:626-724).
  std::vector<float> next_light = light, next_var = var, gvar(n, 0.0f), lum(n, 0.0f);
  for (u32 level = 0; level < preset.levels; ++level) {
    int step = 1 << level;
    for (size_t p = 0; p < n; ++p) lum[p] = lum_of(light, p);
    for (size_t y = 0; y < h; ++y)
      for (size_t x = 0; x < w; ++x) {
        size_t p = y * w + x;
        const Guide& g = guides[p];
        if (!g.foreground()) continue;
        float sw = 0.0f, sv = 0.0f;
        for (int j = 0; j < 3; ++j) {
          long qy = (long)y + j - 1;
          if (qy < 0 || qy >= (long)h) continue;
          for (int i = 0; i < 3; ++i) {
            long qx = (long)x + i - 1;
            if (qx < 0 || qx >= (long)w) continue;
            size_t q = (size_t)qy * w + (size_t)qx;
            if (!g.same_surface(guides[q])) continue;
            float k = GAUSS3[j] * GAUSS3[i];
            sw += k;
            sv += k * var[q];
          }
        }
        gvar[p] = sw > 0.0f ? sv / sw : var[p];
      }
    for (size_t y = 0; y < h; ++y)
      for (size_t x = 0; x < w; ++x) {
        size_t p = y * w + x;
        const Guide& g = guides[p];
        if (!g.foreground()) continue;
        float lp = lum[p];
        float inv_sigma_l = 1.0f / (SIGMA_LUMINANCE * std::sqrt(fmax_r(gvar[p], 0.0f) + 1.0e-10f));
        float sw = 0.0f, sv = 0.0f, sc[3] = {0.0f, 0.0f, 0.0f};
        for (int j = 0; j < 5; ++j) {
          int dy = (j - 2) * step;
          long qy = (long)y + dy;
          if (qy < 0 || qy >= (long)h) continue;
          for (int i = 0; i < 5; ++i) {
            int dx = (i - 2) * step;
            long qx = (long)x + dx;
            if (qx < 0 || qx >= (long)w) continue;
            size_t q = (size_t)qy * w + (size_t)qx;
            const Guide& gq = guides[q];
            if (!g.same_surface(gq)) continue;
            float wn = q == p ? 1.0f : pow128(fmax_r(dot3(g.n, gq.n), 0.0f));
            if (wn <= 0.0f) continue;
            float exponent = std::fabs(lp - lum[q]) * inv_sigma_l + depth_term(g, gq, (float)dx, (float)dy) +
                             albedo_term(g, gq);
            float wgt = KERNEL[j] * KERNEL[i] * wn * falloff(exponent);
            sw += wgt;
            sv += wgt * wgt * var[q];
            for (int k = 0; k < 3; ++k) sc[k] += wgt * light[q * 3 + k];
          }
        }
        float inv_w = 1.0f / sw;
        for (int k = 0; k < 3; ++k) next_light[p * 3 + k] = sc[k] * inv_w;
        next_var[p] = sv * inv_w * inv_w;
      }
    std::swap(light, next_light);
    std::swap(var, next_var);
  }

  // The albedo multiplied back, as sums of `samples` samples (denoise.rs:726-741).
  float count = (float)samples;
  for (size_t p = 0; p < n; ++p) {
    if (!guides[p].foreground()) continue;
    for (int k = 0; k < 3; ++k) {
      float c = light[p * 3 + k] * divisor[p * 3 + k];
      out[p * 4 + k] = is_finite(c) ? c * count : in.sums[p * 4 + k];
    }
  }
  return out;
}

// denoise.rs:322-335, 370-407
Status denoise_rows_at(const Input& in, u32 y0, u32 y1, Preset preset, float strength, std::vector<float>& out) {
  Preset p = preset.clamped();
  float s = strength_of(strength);
  Status st = check_rows(in, y0, y1, p);
  if (st != OK) return st;
  size_t w4 = (size_t)in.width * 4;
  size_t from = (size_t)(y0 - in.origin_y) * w4, to = (size_t)(y1 - in.origin_y) * w4;
  if (s <= 0.0f) {
    out.assign(in.sums + from, in.sums + to);
    return OK;
  }
  std::vector<float> filtered = filter_band(in, p);
  out.assign(filtered.begin() + from, filtered.begin() + to);
  if (s >= 1.0f) return OK;
  float k = 1.0f - s;
  for (size_t i = 0; i < out.size(); ++i)
    if (i % 4 != 3) out[i] = in.sums[from + i] * k + out[i] * s;
  return OK;
}

}  // namespace denoise
}  // namespace orc
