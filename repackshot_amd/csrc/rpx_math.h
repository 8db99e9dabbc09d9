// Numerics shared by the host scene preparation and the GPU kernels: the reference renderer's
// deterministic elementary functions, RNG, vectors, frames, ulp offsets, warps and the per-pixel
// counter-based Owen-scrambled Sobol sampler.
//
// Bit-exactness contract: every function is a fixed sequence of IEEE-754 single-precision
// operations (no contraction into FMA: the library is built with -ffp-contract=off, correctly
// rounded division and square root, f32 denormals kept), so the host and every GPU lane compute
// the same bits as the reference engine.
//
// Reference (paths inside engine/crates/rt-core/src): fmath.rs:40-270, rng.rs:17-55,
// math.rs:16-311, warp.rs:12-62, color.rs:45-47, sampler.rs:88-395.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RPX_HD __host__ __device__ __forceinline__

namespace rpx {

typedef uint32_t u32;
typedef uint64_t u64;
typedef int32_t i32;
typedef int64_t i64;

#define RPX_INF __builtin_huge_valf()
constexpr float F32_MAX = 3.40282347e+38f;
constexpr float F32_MIN_POSITIVE = 1.17549435e-38f;
constexpr float F32_EPSILON = 1.1920929e-07f;

RPX_HD u32 f2u(float f) { return __builtin_bit_cast(u32, f); }
RPX_HD float u2f(u32 u) { return __builtin_bit_cast(float, u); }
RPX_HD bool is_nan(float x) { return x != x; }
RPX_HD bool is_finite(float x) { return __builtin_isfinite(x); }
RPX_HD bool sign_negative(float x) { return (f2u(x) >> 31) != 0; }
RPX_HD float fabs_(float x) { return __builtin_fabsf(x); }
RPX_HD float sqrt_(float x) { return __builtin_sqrtf(x); }
RPX_HD float floor_(float x) { return __builtin_floorf(x); }
// f32::min / f32::max of the reference: the operand that is not NaN.
RPX_HD float fmin_r(float a, float b) { return __builtin_fminf(a, b); }
RPX_HD float fmax_r(float a, float b) { return __builtin_fmaxf(a, b); }
// The SIMD lanes' compare-select forms (simd/mod.rs:156-172): NaN in the first operand yields the
// second.
RPX_HD float max_sel(float a, float b) { return a > b ? a : b; }
RPX_HD float min_sel(float a, float b) { return a < b ? a : b; }
RPX_HD float clampf(float x, float lo, float hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
// Saturating float-to-integer conversions (Rust `as`): NaN gives zero.
RPX_HD i64 f_to_i64(float x) {
  if (x != x) return 0;
  if (x >= 9223372036854775807.0f) return INT64_MAX;
  if (x <= -9223372036854775808.0f) return INT64_MIN;
  return (i64)x;
}
RPX_HD i32 f_to_i32(float x) {
  if (x != x) return 0;
  if (x >= 2147483647.0f) return INT32_MAX;
  if (x <= -2147483648.0f) return INT32_MIN;
  return (i32)x;
}
RPX_HD u32 f_to_u32(float x) {
  if (!(x > 0.0f)) return 0;
  if (x >= 4294967295.0f) return UINT32_MAX;
  return (u32)x;
}

// ---------------------------------------------------------------- fmath.rs
constexpr float PI = 3.14159265358979323846264338327950288f;
constexpr float TAU = 6.28318530717958647692528676655900577f;
constexpr float FRAC_PI_2 = 1.57079632679489661923132169163975144f;
constexpr float FRAC_PI_4 = 0.785398163397448309615660845819875721f;
constexpr float FRAC_1_PI = 0.318309886183790671537767526745028724f;
constexpr float LN_2 = 0.693147180559945309417232121458176568f;
constexpr float LOG2_E = 1.44269504088896340735992468100189214f;
constexpr float SQRT_2 = 1.41421356237309504880168872420969808f;

RPX_HD float deg_to_rad() { return PI / 180.0f; }

RPX_HD void sin_cos(float x, float& s_out, float& c_out) {
  if (!is_finite(x)) {
    s_out = 0.0f;
    c_out = 1.0f;
    return;
  }
  const float PIO2_1 = 1.5703125f, PIO2_2 = 4.837513e-4f, PIO2_3 = 7.54979e-8f;
  float k = floor_(x * (2.0f / PI) + 0.5f);
  float r = ((x - k * PIO2_1) - k * PIO2_2) - k * PIO2_3;
  float r2 = r * r;
  float p = 1.0f / 362880.0f;
  p = p * r2 - 1.0f / 5040.0f;
  p = p * r2 + 1.0f / 120.0f;
  p = p * r2 - 1.0f / 6.0f;
  float s = r + r * r2 * p;
  float q = -1.0f / 3628800.0f;
  q = q * r2 + 1.0f / 40320.0f;
  q = q * r2 - 1.0f / 720.0f;
  q = q * r2 + 1.0f / 24.0f;
  q = q * r2 - 0.5f;
  float c = 1.0f + r2 * q;
  switch ((int)(f_to_i64(k) & 3)) {
    case 0: s_out = s; c_out = c; break;
    case 1: s_out = c; c_out = -s; break;
    case 2: s_out = -s; c_out = -c; break;
    default: s_out = -c; c_out = s; break;
  }
}
RPX_HD float fsin(float x) {
  float s, c;
  sin_cos(x, s, c);
  return s;
}
RPX_HD float fcos(float x) {
  float s, c;
  sin_cos(x, s, c);
  return c;
}
RPX_HD float pow2i(i32 n) { return u2f(((u32)(n + 127)) << 23); }
RPX_HD float fexp2(float x) {
  if (is_nan(x)) return x;
  if (x >= 128.0f) return RPX_INF;
  if (x < -150.0f) return 0.0f;
  float k = floor_(x + 0.5f);
  float t = (x - k) * LN_2;
  float p = 1.0f / 5040.0f;
  p = p * t + 1.0f / 720.0f;
  p = p * t + 1.0f / 120.0f;
  p = p * t + 1.0f / 24.0f;
  p = p * t + 1.0f / 6.0f;
  p = p * t + 0.5f;
  p = p * t + 1.0f;
  p = p * t + 1.0f;
  i32 ki = f_to_i32(k);
  i32 k1 = ki / 2;
  return p * pow2i(k1) * pow2i(ki - k1);
}
RPX_HD float flog2(float x) {
  if (is_nan(x)) return x;
  if (x <= 0.0f) return -RPX_INF;
  if (x == RPX_INF) return x;
  i32 e = 0;
  if (x < F32_MIN_POSITIVE) {
    x *= 8388608.0f;
    e = -23;
  }
  u32 bits = f2u(x);
  e += (i32)((bits >> 23) & 0xff) - 127;
  float m = u2f((bits & 0x007fffffu) | 0x3f800000u);
  if (m > SQRT_2) {
    m *= 0.5f;
    e += 1;
  }
  float s = (m - 1.0f) / (m + 1.0f);
  float s2 = s * s;
  float p = 1.0f / 11.0f;
  p = p * s2 + 1.0f / 9.0f;
  p = p * s2 + 1.0f / 7.0f;
  p = p * s2 + 1.0f / 5.0f;
  p = p * s2 + 1.0f / 3.0f;
  float ln_m = 2.0f * (s + s * s2 * p);
  return (float)e + ln_m * LOG2_E;
}
RPX_HD float fexp(float x) { return fexp2(x * LOG2_E); }
RPX_HD float fln(float x) { return flog2(x) * LN_2; }
RPX_HD float fpow(float x, float y) {
  if (y == 0.0f) return 1.0f;
  if (x <= 0.0f) return y > 0.0f ? 0.0f : RPX_INF;
  if (x == 1.0f) return 1.0f;
  return fexp2(y * flog2(x));
}

// ---------------------------------------------------------------- rng.rs
struct Pcg32 {
  u64 state, inc;
  RPX_HD static Pcg32 seeded(u64 seed, u64 stream) {
    Pcg32 r{0, (stream << 1) | 1};
    r.next_u32();
    r.state += seed;
    r.next_u32();
    return r;
  }
  RPX_HD u32 next_u32() {
    u64 old = state;
    state = old * 6364136223846793005ull + inc;
    u32 xs = (u32)(((old >> 18) ^ old) >> 27);
    u32 rot = (u32)(old >> 59);
    return (xs >> rot) | (xs << ((32 - rot) & 31));
  }
  RPX_HD float next_f32() { return (float)(next_u32() >> 8) * (1.0f / 16777216.0f); }
};
RPX_HD u64 mix64(u64 z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// ---------------------------------------------------------------- vectors
struct V3 {
  float x, y, z;
};
RPX_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
RPX_HD V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
RPX_HD V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
RPX_HD V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
RPX_HD V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
RPX_HD V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
RPX_HD V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
RPX_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RPX_HD V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
RPX_HD float length_sq(V3 a) { return dot(a, a); }
RPX_HD float length(V3 a) { return sqrt_(length_sq(a)); }
RPX_HD V3 normalize(V3 a) {
  float len = length(a);
  if (len > 0.0f && is_finite(len)) return v3(a.x / len, a.y / len, a.z / len);
  return a;
}
RPX_HD bool finite3(V3 a) { return is_finite(a.x) && is_finite(a.y) && is_finite(a.z); }
RPX_HD float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
RPX_HD V3 vmin(V3 a, V3 b) { return v3(fmin_r(a.x, b.x), fmin_r(a.y, b.y), fmin_r(a.z, b.z)); }
RPX_HD V3 vmax(V3 a, V3 b) { return v3(fmax_r(a.x, b.x), fmax_r(a.y, b.y), fmax_r(a.z, b.z)); }

struct C3 {
  float r, g, b;
};
RPX_HD C3 c3(float r, float g, float b) { return C3{r, g, b}; }
RPX_HD C3 splat(float v) { return C3{v, v, v}; }
RPX_HD C3 operator+(C3 a, C3 b) { return c3(a.r + b.r, a.g + b.g, a.b + b.b); }
RPX_HD C3 operator*(C3 a, C3 b) { return c3(a.r * b.r, a.g * b.g, a.b * b.b); }
RPX_HD C3 operator*(C3 a, float s) { return c3(a.r * s, a.g * s, a.b * s); }
RPX_HD C3 operator/(C3 a, float s) { return c3(a.r / s, a.g / s, a.b / s); }
RPX_HD float luminance(C3 c) { return 0.2126f * c.r + 0.7152f * c.g + 0.0722f * c.b; }
RPX_HD float peak(C3 c) { return fmax_r(fmax_r(c.r, c.g), c.b); }
RPX_HD bool finite_c(C3 c) { return is_finite(c.r) && is_finite(c.g) && is_finite(c.b); }

constexpr float RAY_EPSILON = 1.0e-3f;

struct Frame {
  V3 x, y, z;
};
// math.rs:264-285 (Duff et al. 2017)
RPX_HD Frame frame_from_normal(V3 n) {
  float sign = __builtin_copysignf(1.0f, n.z);
  float a = -1.0f / (sign + n.z);
  float b = n.x * n.y * a;
  Frame f;
  f.x = v3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
  f.y = v3(b, sign + n.y * n.y * a, -n.y);
  f.z = n;
  return f;
}
RPX_HD V3 to_local(const Frame& f, V3 v) { return v3(dot(v, f.x), dot(v, f.y), dot(v, f.z)); }
RPX_HD V3 to_world(const Frame& f, V3 v) { return f.x * v.x + f.y * v.y + f.z * v.z; }

// math.rs:293-311 (Wächter and Binder 2019)
RPX_HD float offset_axis(float p, float n) {
  if (fabs_(p) < 1.0f / 32.0f) return p + (1.0f / 65536.0f) * n;
  i32 of = f_to_i32(256.0f * n);
  u32 moved = f2u(p) + (u32)(p < 0.0f ? -of : of);
  float q = u2f(moved);
  return is_finite(q) ? q : p;
}
RPX_HD V3 offset_origin(V3 p, V3 n) { return v3(offset_axis(p.x, n.x), offset_axis(p.y, n.y), offset_axis(p.z, n.z)); }

// ---------------------------------------------------------------- warp.rs
RPX_HD void concentric_disk(float u0, float u1, float& ox, float& oy) {
  float a = 2.0f * u0 - 1.0f, b = 2.0f * u1 - 1.0f;
  if (a == 0.0f && b == 0.0f) {
    ox = 0.0f;
    oy = 0.0f;
    return;
  }
  float r, phi;
  if (fabs_(a) > fabs_(b)) {
    r = a;
    phi = FRAC_PI_4 * (b / a);
  } else {
    r = b;
    phi = FRAC_PI_2 - FRAC_PI_4 * (a / b);
  }
  float s, c;
  sin_cos(phi, s, c);
  ox = r * c;
  oy = r * s;
}
RPX_HD V3 cosine_hemisphere(float u0, float u1) {
  float x, y;
  concentric_disk(u0, u1, x, y);
  float z = sqrt_(fmax_r(1.0f - x * x - y * y, 0.0f));
  return v3(x, y, z);
}

// warp.rs:38-45: 1 - cos uniform in [0, 1 - cos_max], the sine without cancellation.
RPX_HD V3 uniform_cone(float u0, float u1, float one_minus_cos_max) {
  float one_minus_cos = u0 * one_minus_cos_max;
  float cos = 1.0f - one_minus_cos;
  float sin = sqrt_(fmax_r(one_minus_cos * (2.0f - one_minus_cos), 0.0f));
  float sp, cp;
  sin_cos(TAU * u1, sp, cp);
  return v3(sin * cp, sin * sp, cos);
}

// ---------------------------------------------------------------- sampler.rs
constexpr u32 DIM_CAMERA = 0, DIM_LENS = 2, DIM_FIRST_BOUNCE = 4, DIMS_PER_BOUNCE = 6;
constexpr u32 OFFSET_LIGHT_SELECT = 0, OFFSET_LIGHT = 1, OFFSET_BSDF = 3, OFFSET_ROULETTE = 5;
RPX_HD u32 bounce_dim(u32 b) { return DIM_FIRST_BOUNCE + b * DIMS_PER_BOUNCE; }
// sampler.rs:137-164: the hashed dimensions of a scattering medium's walk, 8 per event of bounce
// b, beyond the Sobol table.
constexpr u32 DIM_MEDIUM = 1u << 21, MEDIUM_DIMS_PER_EVENT = 8, MEDIUM_EVENTS = 1024, MEDIUM_BOUNCES = 128;
constexpr u32 MEDIUM_OFFSET_DISTANCE = 0, MEDIUM_OFFSET_PHASE = 1, MEDIUM_OFFSET_ROULETTE = 3,
              MEDIUM_OFFSET_FRESNEL = 4, MEDIUM_OFFSET_HERO = 5;
RPX_HD u32 medium_dim(u32 bounce, u32 event) {
  u32 b = bounce < MEDIUM_BOUNCES ? bounce : MEDIUM_BOUNCES - 1;
  u32 e = event < MEDIUM_EVENTS ? event : MEDIUM_EVENTS - 1;
  return DIM_MEDIUM + (b * MEDIUM_EVENTS + e) * MEDIUM_DIMS_PER_EVENT;
}

RPX_HD u32 reverse_bits(u32 v) { return __builtin_bitreverse32(v); }
RPX_HD u32 owen_scramble(u32 x, u32 seed) {
  u32 v = reverse_bits(x);
  v += seed;
  v ^= v * 0x6c50b47cu;
  v ^= v * 0xb82f1e52u;
  v ^= v * 0xc7afe638u;
  v ^= v * 0x8d22f6e6u;
  return reverse_bits(v);
}
RPX_HD float to_unit(u32 x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }
// Pixel key of (scene seed, x, y): sampler.rs:339-343.
RPX_HD u64 pixel_key(u32 seed, u32 x, u32 y) {
  u64 pixel = ((u64)y << 32) | (u64)x;
  return mix64(pixel ^ mix64((u64)seed ^ 0x5851f42d4c957f2dull));
}
RPX_HD u64 dim_seed(u64 key, u32 dim) { return mix64(key ^ ((u64)(dim + 1) * 0x9e3779b97f4a7c15ull)); }
// One number of logical dimension `dim` whose unscrambled Sobol point is `sobol_x` (the point of
// the dimension's table column at this sample index; only read for dim < 64).
RPX_HD float sample1(u64 key, u32 index, u32 dim, u32 sobol_x) {
  u64 seed = dim_seed(key, dim);
  if (dim < 64) return to_unit(owen_scramble(sobol_x, (u32)(seed >> 32)));
  return to_unit((u32)(mix64(seed ^ (u64)index) >> 32));
}

}  // namespace rpx
