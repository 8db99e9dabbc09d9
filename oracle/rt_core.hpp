// ORACLE — test infrastructure only. A CPU restatement of the reference path tracer's numeric
// core, used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
// Nothing in the product (repackshot_amd/) includes, links or calls this code.
//
// Compiled with -ffp-contract=off -fno-fast-math on x86-64 (SSE, no x87): every f32 operation
// is one IEEE operation, rounded, in the order the reference writes it.
//
// Follows (paths inside the reference's engine/crates/rt-core/src):
//   fmath.rs:40-270 (sin_cos, exp2, log2, powf, atan, atan2, acos)
//   rng.rs:17-55 (Pcg32, mix64)
//   math.rs:16-311 (Vec3, Ray, Frame::from_normal, offset_origin), 388-417 (Aabb::intersect)
//   warp.rs:12-62 (concentric_disk, cosine_hemisphere, uniform_cone, uniform_triangle)
//   color.rs:45-47 (luminance)
//   sampler.rs:88-395 (LAYOUT, JOE_KUO, build_matrices, sobol, owen_scramble, PixelSampler)
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

namespace orc {

using u32 = uint32_t;
using u64 = uint64_t;
using i32 = int32_t;
using i64 = int64_t;

constexpr float INF = std::numeric_limits<float>::infinity();
constexpr float F32_MAX = std::numeric_limits<float>::max();
constexpr float F32_MIN_POSITIVE = std::numeric_limits<float>::min();
constexpr float F32_EPSILON = std::numeric_limits<float>::epsilon();

inline u32 f2u(float f) { u32 u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(u32 u) { float f; std::memcpy(&f, &u, 4); return f; }
inline bool is_finite(float x) { return std::isfinite(x); }
inline bool is_nan(float x) { return std::isnan(x); }
inline bool sign_negative(float x) { return (f2u(x) >> 31) != 0; }

// Rust f32::min / f32::max: the non-NaN operand when one is NaN.
inline float fmin_r(float a, float b) { return std::fmin(a, b); }
inline float fmax_r(float a, float b) { return std::fmax(a, b); }
// Rust f32::clamp: NaN passes through.
inline float clampf(float x, float lo, float hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
// Rust `as` conversions saturate and send NaN to zero.
inline i64 f_to_i64(float x) {
  if (x != x) return 0;
  if (x >= 9223372036854775807.0f) return INT64_MAX;
  if (x <= -9223372036854775808.0f) return INT64_MIN;
  return (i64)x;
}
inline i32 f_to_i32(float x) {
  if (x != x) return 0;
  if (x >= 2147483647.0f) return INT32_MAX;
  if (x <= -2147483648.0f) return INT32_MIN;
  return (i32)x;
}
inline u64 f_to_usize(float x) {
  if (!(x > 0.0f)) return 0;
  if (x >= 18446744073709551615.0f) return UINT64_MAX;
  return (u64)x;
}
inline u32 f_to_u32(float x) {
  if (!(x > 0.0f)) return 0;
  if (x >= 4294967295.0f) return UINT32_MAX;
  return (u32)x;
}

// ---------------------------------------------------------------- fmath (fmath.rs)
constexpr float PI = 3.14159265358979323846264338327950288f;
constexpr float TAU = 6.28318530717958647692528676655900577f;
constexpr float FRAC_PI_2 = 1.57079632679489661923132169163975144f;
constexpr float FRAC_PI_4 = 0.785398163397448309615660845819875721f;
constexpr float FRAC_1_PI = 0.318309886183790671537767526745028724f;
constexpr float LN_2 = 0.693147180559945309417232121458176568f;
constexpr float LOG2_E = 1.44269504088896340735992468100189214f;
constexpr float SQRT_2 = 1.41421356237309504880168872420969808f;
const float DEG_TO_RAD = PI / 180.0f;

constexpr float PIO2_1 = 1.5703125f;
constexpr float PIO2_2 = 4.837513e-4f;
constexpr float PIO2_3 = 7.54979e-8f;

inline float sin_poly(float r) {
  float r2 = r * r;
  float p = 1.0f / 362880.0f;
  p = p * r2 - 1.0f / 5040.0f;
  p = p * r2 + 1.0f / 120.0f;
  p = p * r2 - 1.0f / 6.0f;
  return r + r * r2 * p;
}
inline float cos_poly(float r) {
  float r2 = r * r;
  float p = -1.0f / 3628800.0f;
  p = p * r2 + 1.0f / 40320.0f;
  p = p * r2 - 1.0f / 720.0f;
  p = p * r2 + 1.0f / 24.0f;
  p = p * r2 - 0.5f;
  return 1.0f + r2 * p;
}
// fmath.rs:62-78
inline void sin_cos(float x, float& s_out, float& c_out) {
  if (!is_finite(x)) { s_out = 0.0f; c_out = 1.0f; return; }
  float k = std::floor(x * (2.0f / PI) + 0.5f);
  float r = ((x - k * PIO2_1) - k * PIO2_2) - k * PIO2_3;
  float s = sin_poly(r), c = cos_poly(r);
  switch (f_to_i64(k) & 3) {
    case 0: s_out = s; c_out = c; break;
    case 1: s_out = c; c_out = -s; break;
    case 2: s_out = -s; c_out = -c; break;
    default: s_out = -c; c_out = s; break;
  }
}
inline float fsin(float x) { float s, c; sin_cos(x, s, c); return s; }
inline float fcos(float x) { float s, c; sin_cos(x, s, c); return c; }
inline float pow2i(i32 n) { return u2f(((u32)(n + 127)) << 23); }
// fmath.rs:99-125
inline float exp2f_(float x) {
  if (is_nan(x)) return x;
  if (x >= 128.0f) return INF;
  if (x < -150.0f) return 0.0f;
  float k = std::floor(x + 0.5f);
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
// fmath.rs:129-161
inline float log2f_(float x) {
  if (is_nan(x)) return x;
  if (x <= 0.0f) return -INF;
  if (x == INF) return x;
  i32 e = 0;
  if (x < F32_MIN_POSITIVE) { x *= 8388608.0f; e = -23; }
  u32 bits = f2u(x);
  e += (i32)((bits >> 23) & 0xff) - 127;
  float m = u2f((bits & 0x007fffffu) | 0x3f800000u);
  if (m > SQRT_2) { m *= 0.5f; e += 1; }
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
inline float expf_(float x) { return exp2f_(x * LOG2_E); }
inline float lnf_(float x) { return log2f_(x) * LN_2; }
// fmath.rs:177-188
inline float powf_(float x, float y) {
  if (y == 0.0f) return 1.0f;
  if (x <= 0.0f) return y > 0.0f ? 0.0f : INF;
  if (x == 1.0f) return 1.0f;
  return exp2f_(y * log2f_(x));
}
inline float atan_poly(float u) {
  float u2 = u * u;
  float p = 1.0f / 17.0f;
  p = -p * u2 + 1.0f / 15.0f;
  p = -p * u2 + 1.0f / 13.0f;
  p = -p * u2 + 1.0f / 11.0f;
  p = -p * u2 + 1.0f / 9.0f;
  p = -p * u2 + 1.0f / 7.0f;
  p = -p * u2 + 1.0f / 5.0f;
  p = -p * u2 + 1.0f / 3.0f;
  return u - u * u2 * p;
}
inline float atanf_(float x) {
  if (is_nan(x)) return x;
  float a = std::fabs(x);
  bool flip = a > 1.0f;
  float t = flip ? 1.0f / a : a;
  float r = t > 0.41421357f ? FRAC_PI_4 + atan_poly((t - 1.0f) / (t + 1.0f)) : atan_poly(t);
  r = flip ? FRAC_PI_2 - r : r;
  return x < 0.0f ? -r : r;
}
inline float atan2f_(float y, float x) {
  if (is_nan(x) || is_nan(y)) return std::numeric_limits<float>::quiet_NaN();
  if (x == 0.0f) return y > 0.0f ? FRAC_PI_2 : (y < 0.0f ? -FRAC_PI_2 : 0.0f);
  if (std::fabs(y) <= std::fabs(x)) {
    float r = atanf_(y / x);
    if (!(x < 0.0f)) return r;
    return (y < 0.0f) ? r - PI : r + PI;
  } else if (y > 0.0f) {
    return FRAC_PI_2 - atanf_(x / y);
  } else {
    return -FRAC_PI_2 - atanf_(x / y);
  }
}
inline float acosf_(float x) {
  x = clampf(x, -1.0f, 1.0f);
  return atan2f_(std::sqrt((1.0f - x) * (1.0f + x)), x);
}

// ---------------------------------------------------------------- rng (rng.rs)
struct Pcg32 {
  u64 state, inc;
  static Pcg32 seeded(u64 seed, u64 stream) {
    Pcg32 r{0, (stream << 1) | 1};
    r.next_u32();
    r.state += seed;
    r.next_u32();
    return r;
  }
  u32 next_u32() {
    u64 old = state;
    state = old * 6364136223846793005ull + inc;
    u32 xorshifted = (u32)(((old >> 18) ^ old) >> 27);
    u32 rot = (u32)(old >> 59);
    return (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
  }
  float next_f32() { return (float)(next_u32() >> 8) * (1.0f / 16777216.0f); }
};
inline u64 mix64(u64 z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// ---------------------------------------------------------------- vectors (math.rs)
struct Vec3 {
  float x = 0, y = 0, z = 0;
  Vec3() = default;
  constexpr Vec3(float a, float b, float c) : x(a), y(b), z(c) {}
  static Vec3 splat(float v) { return Vec3(v, v, v); }
  float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  float& at(int i) { return i == 0 ? x : (i == 1 ? y : z); }
  float dot(Vec3 o) const { return x * o.x + y * o.y + z * o.z; }
  Vec3 cross(Vec3 o) const { return Vec3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x); }
  float length_squared() const { return dot(*this); }
  float length() const { return std::sqrt(length_squared()); }
  Vec3 normalize() const {
    float len = length();
    if (len > 0.0f && is_finite(len)) return Vec3(x / len, y / len, z / len);
    return *this;
  }
  Vec3 min(Vec3 o) const { return Vec3(fmin_r(x, o.x), fmin_r(y, o.y), fmin_r(z, o.z)); }
  Vec3 max(Vec3 o) const { return Vec3(fmax_r(x, o.x), fmax_r(y, o.y), fmax_r(z, o.z)); }
  float max_component() const { return fmax_r(fmax_r(x, y), z); }
  Vec3 abs() const { return Vec3(std::fabs(x), std::fabs(y), std::fabs(z)); }
  bool finite() const { return is_finite(x) && is_finite(y) && is_finite(z); }
  bool operator==(Vec3 o) const { return x == o.x && y == o.y && z == o.z; }
  bool operator!=(Vec3 o) const { return !(*this == o); }
};
inline Vec3 operator+(Vec3 a, Vec3 b) { return Vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline Vec3 operator-(Vec3 a, Vec3 b) { return Vec3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline Vec3 operator*(Vec3 a, Vec3 b) { return Vec3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline Vec3 operator*(Vec3 a, float s) { return Vec3(a.x * s, a.y * s, a.z * s); }
inline Vec3 operator*(float s, Vec3 a) { return a * s; }
inline Vec3 operator/(Vec3 a, float s) { return Vec3(a.x / s, a.y / s, a.z / s); }
inline Vec3 operator-(Vec3 a) { return Vec3(-a.x, -a.y, -a.z); }

struct Rgb {
  float r = 0, g = 0, b = 0;
  Rgb() = default;
  constexpr Rgb(float a, float b_, float c) : r(a), g(b_), b(c) {}
  static Rgb splat(float v) { return Rgb(v, v, v); }
  float luminance() const { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }
  bool operator==(Rgb o) const { return r == o.r && g == o.g && b == o.b; }
  bool operator!=(Rgb o) const { return !(*this == o); }
};
inline Rgb operator+(Rgb a, Rgb b) { return Rgb(a.r + b.r, a.g + b.g, a.b + b.b); }
inline Rgb operator*(Rgb a, Rgb b) { return Rgb(a.r * b.r, a.g * b.g, a.b * b.b); }
inline Rgb operator*(Rgb a, float s) { return Rgb(a.r * s, a.g * s, a.b * s); }
inline Rgb operator/(Rgb a, float s) { return Rgb(a.r / s, a.g / s, a.b / s); }
inline Rgb& operator+=(Rgb& a, Rgb b) { a = a + b; return a; }
const Rgb WHITE(1.0f, 1.0f, 1.0f);
const Rgb BLACK(0.0f, 0.0f, 0.0f);
inline float peak(Rgb c) { return fmax_r(fmax_r(c.r, c.g), c.b); }

constexpr float RAY_EPSILON = 1.0e-3f;
struct Ray {
  Vec3 origin, dir;
  float t_min = RAY_EPSILON, t_max = INF;
  Ray() = default;
  Ray(Vec3 o, Vec3 d) : origin(o), dir(d), t_min(RAY_EPSILON), t_max(INF) {}
  Vec3 at(float t) const { return origin + dir * t; }
};

struct Frame {
  Vec3 x, y, z;
  // math.rs:264-275 (Duff et al. 2017)
  static Frame from_normal(Vec3 n) {
    float sign = std::copysign(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    Frame f;
    f.x = Vec3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    f.y = Vec3(b, sign + n.y * n.y * a, -n.y);
    f.z = n;
    return f;
  }
  Vec3 to_local(Vec3 v) const { return Vec3(v.dot(x), v.dot(y), v.dot(z)); }
  Vec3 to_world(Vec3 v) const { return x * v.x + y * v.y + z * v.z; }
};

// math.rs:293-311 (Wächter and Binder 2019)
inline float offset_axis(float p, float n) {
  const float ORIGIN = 1.0f / 32.0f, FLOAT_SCALE = 1.0f / 65536.0f, INT_SCALE = 256.0f;
  if (std::fabs(p) < ORIGIN) return p + FLOAT_SCALE * n;
  i32 of = f_to_i32(INT_SCALE * n);
  i32 bits = (i32)f2u(p);
  u32 moved = (u32)bits + (u32)(p < 0.0f ? -of : of);
  float q = u2f(moved);
  return is_finite(q) ? q : p;
}
inline Vec3 offset_origin(Vec3 p, Vec3 n) {
  return Vec3(offset_axis(p.x, n.x), offset_axis(p.y, n.y), offset_axis(p.z, n.z));
}

struct Aabb {
  Vec3 min{INF, INF, INF}, max{-INF, -INF, -INF};
  static Aabb empty() { return Aabb(); }
  bool is_empty() const { return min.x > max.x || min.y > max.y || min.z > max.z; }
  Aabb unite(const Aabb& o) const { Aabb r; r.min = min.min(o.min); r.max = max.max(o.max); return r; }
  Aabb extend(Vec3 p) const { Aabb r; r.min = min.min(p); r.max = max.max(p); return r; }
  Vec3 centroid() const { return (min + max) * 0.5f; }
  float surface_area() const {
    if (is_empty()) return 0.0f;
    Vec3 d = max - min;
    return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x);
  }
};

// ---------------------------------------------------------------- warps (warp.rs)
inline void concentric_disk(float u0, float u1, float& ox, float& oy) {
  float a = 2.0f * u0 - 1.0f, b = 2.0f * u1 - 1.0f;
  if (a == 0.0f && b == 0.0f) { ox = 0.0f; oy = 0.0f; return; }
  float r, phi;
  if (std::fabs(a) > std::fabs(b)) { r = a; phi = FRAC_PI_4 * (b / a); }
  else { r = b; phi = FRAC_PI_2 - FRAC_PI_4 * (a / b); }
  float s, c;
  sin_cos(phi, s, c);
  ox = r * c;
  oy = r * s;
}
inline Vec3 cosine_hemisphere(float u0, float u1) {
  float x, y;
  concentric_disk(u0, u1, x, y);
  float z = std::sqrt(fmax_r(1.0f - x * x - y * y, 0.0f));
  return Vec3(x, y, z);
}

// warp.rs:38-45: 1 - cos uniform in [0, 1 - cos_max], the sine without cancellation
inline Vec3 uniform_cone(float u0, float u1, float one_minus_cos_max) {
  float one_minus_cos = u0 * one_minus_cos_max;
  float cos = 1.0f - one_minus_cos;
  float sin = std::sqrt(fmax_r(one_minus_cos * (2.0f - one_minus_cos), 0.0f));
  float sp, cp;
  sin_cos(TAU * u1, sp, cp);
  return Vec3(sin * cp, sin * sp, cos);
}

// warp.rs:57-62: barycentrics of a point uniform on a triangle
inline void uniform_triangle(float u0, float u1, float w[3]) {
  float su = std::sqrt(u0);
  float b0 = 1.0f - su;
  float b1 = u1 * su;
  w[0] = b0;
  w[1] = b1;
  w[2] = 1.0f - b0 - b1;
}

// ---------------------------------------------------------------- sampler (sampler.rs)
constexpr u32 SOBOL_DIMENSIONS = 64;
constexpr unsigned char LAYOUT[64] = {
    0,  1,  11, 50, 19, 2,  15, 10, 13, 20, 21, 5,  27, 3,  59, 25, 26, 6,  57, 18, 47, 28,
    31, 24, 44, 22, 58, 34, 37, 49, 52, 4,  9,  39, 41, 7,  12, 16, 29, 42, 45, 8,  38, 14,
    46, 48, 51, 33, 36, 32, 56, 53, 54, 35, 63, 43, 62, 55, 60, 23, 30, 17, 40, 61};
constexpr u32 DIM_CAMERA = 0, DIM_LENS = 2, DIM_FIRST_BOUNCE = 4, DIMS_PER_BOUNCE = 6;
constexpr u32 OFFSET_LIGHT_SELECT = 0, OFFSET_LIGHT = 1, OFFSET_BSDF = 3, OFFSET_ROULETTE = 5;
inline u32 bounce_dim(u32 b) { return DIM_FIRST_BOUNCE + b * DIMS_PER_BOUNCE; }
// sampler.rs:137-164: the hashed dimensions of the scattering media's walks, beyond the Sobol
// table and the cut-outs: 8 per event, 1024 events per walk, walks after vertex 127 share a block.
constexpr u32 DIM_MEDIUM = 1u << 21, MEDIUM_DIMS_PER_EVENT = 8, MEDIUM_EVENTS = 1024, MEDIUM_BOUNCES = 128;
inline u32 medium_dim(u32 bounce, u32 event) {
  u32 b = bounce < MEDIUM_BOUNCES ? bounce : MEDIUM_BOUNCES - 1;
  u32 e = event < MEDIUM_EVENTS ? event : MEDIUM_EVENTS - 1;
  return DIM_MEDIUM + (b * MEDIUM_EVENTS + e) * MEDIUM_DIMS_PER_EVENT;
}

// The Joe–Kuo (new-joe-kuo-6.21201) degree s, coefficient a and initial direction numbers of
// dimensions 2..64 (published data; sampler.rs:200-264 lists the same rows).
struct JoeKuoRow { u32 s, a; u32 m[9]; };
extern const JoeKuoRow JOE_KUO[63];

struct SobolTable { u32 v[64][32]; SobolTable(); };
const SobolTable& sobol_table();

inline u32 sobol(u32 dim, u32 index) {
  const u32* cols = sobol_table().v[dim % SOBOL_DIMENSIONS];
  u32 x = 0;
  while (index) {
    x ^= cols[__builtin_ctz(index)];
    index &= index - 1;
  }
  return x;
}
inline u32 reverse_bits(u32 v) {
  v = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
  v = ((v >> 2) & 0x33333333u) | ((v & 0x33333333u) << 2);
  v = ((v >> 4) & 0x0F0F0F0Fu) | ((v & 0x0F0F0F0Fu) << 4);
  v = ((v >> 8) & 0x00FF00FFu) | ((v & 0x00FF00FFu) << 8);
  return (v >> 16) | (v << 16);
}
inline u32 owen_scramble(u32 x, u32 seed) {
  u32 v = reverse_bits(x);
  v += seed;
  v ^= v * 0x6c50b47cu;
  v ^= v * 0xb82f1e52u;
  v ^= v * 0xc7afe638u;
  v ^= v * 0x8d22f6e6u;
  return reverse_bits(v);
}
inline float to_unit(u32 x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

struct PixelSampler {
  u64 key;
  u32 index;
  static PixelSampler make(u32 seed, u32 x, u32 y, u32 index) {
    u64 pixel = ((u64)y << 32) | (u64)x;
    return PixelSampler{mix64(pixel ^ mix64((u64)seed ^ 0x5851f42d4c957f2dull)), index};
  }
  u64 dim_seed(u32 dim) const { return mix64(key ^ ((u64)(dim + 1) * 0x9e3779b97f4a7c15ull)); }
  float get1(u32 dim) const {
    u64 seed = dim_seed(dim);
    if (dim < 64) {
      u32 x = sobol(LAYOUT[dim], index);
      return to_unit(owen_scramble(x, (u32)(seed >> 32)));
    }
    return to_unit((u32)(mix64(seed ^ (u64)index) >> 32));
  }
  void get2(u32 dim, float& a, float& b) const { a = get1(dim); b = get1(dim + 1); }
};

}  // namespace orc
