// The denoiser's kernels (engine/crates/rt-post/src/denoise.rs; line numbers per stage). One lane
// per pixel of the band; a block is a 16 x 16 tile and consecutive tiles of the image go to the same
// XCD, so the taps of a tile and of its neighbours meet in one L2.
//
//   setup     guides (ids, finite normal, depth, finite albedo), the depth gradient, the albedo
//             divided out of the mean (the illumination) and the pixel's own variance
//   firefly   a pixel brighter than FIREFLY_RATIO x every neighbour on its surface is brought down
//   variance  below MIN_SAMPLES_FOR_VARIANCE samples: the variance of the neighbours on the surface
//   atrous    one pass per level: the 3 x 3 Gaussian of the variance, then 5 x 5 taps 2^level
//             apart weighted by the luminance, normal, depth and albedo edge-stopping functions
//   final     the albedo multiplied back, as sums, blended with the noisy sums at the strength
//
// The reference scans the firefly clamp in place; here it reads the variances before the clamp.
// That is the same: a firefly p exceeds 4 x before[q] + 0.05 > 0 for every neighbour q on its
// surface, so no such q is itself a firefly (it would need before[q] > 4 x before[p] + 0.05),
// and the variances p reads are never ones the scan had already replaced.
#include <algorithm>
#include "rpx_denoise.h"

namespace rpx {
namespace denoise {
namespace {

constexpr float SIGMA_LUMINANCE = 4.0f;
constexpr float SIGMA_DEPTH = 1.0f;
constexpr float SIGMA_ALBEDO = 0.1f;
constexpr float MIN_ALBEDO = 0.01f;
constexpr float FIREFLY_RATIO = 4.0f;
constexpr float FIREFLY_FLOOR = 0.05f;
constexpr u32 TILE = 16;

__device__ __forceinline__ float kernel5(int j) {
  return j == 2 ? 3.0f / 8.0f : ((j == 1 || j == 3) ? 1.0f / 4.0f : 1.0f / 16.0f);
}
__device__ __forceinline__ float gauss3(int j) { return j == 1 ? 0.5f : 0.25f; }

// Rec. 709 luminance (denoise.rs:410-413)
__device__ __forceinline__ float luminance(float r, float g, float b) { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }

// x^128: seven squarings (denoise.rs:416-423)
__device__ __forceinline__ float pow128(float x) {
  float v = x;
#pragma unroll
  for (int i = 0; i < 7; ++i) v *= v;
  return v;
}

// e^-x as (1 - x/16)^16 (denoise.rs:769-779)
__device__ __forceinline__ float falloff(float x) {
  float t = 1.0f - x * (1.0f / 16.0f);
  if (is_nan(t) || t <= 0.0f) return 0.0f;
  t = t * t;
  t = t * t;
  t = t * t;
  return t * t;
}

// |z_p - z_q| / (sigma_z |grad z_p . delta| + eps) (denoise.rs:750-758)
__device__ __forceinline__ float depth_term(float gz, float2 gdz, float qz, float dx, float dy) {
  if (!(is_finite(gz) && is_finite(qz))) return gz == qz ? 0.0f : RPX_INF;
  float expected = fabs_(gdz.x * dx + gdz.y * dy);
  float eps = 1.0e-3f * fabs_(gz) + 1.0e-6f;
  return fabs_(gz - qz) / (SIGMA_DEPTH * expected + eps);
}

// denoise.rs:761-764
__device__ __forceinline__ float albedo_term(float4 a, float4 b) {
  return (fabs_(a.x - b.x) + fabs_(a.y - b.y) + fabs_(a.z - b.z)) / SIGMA_ALBEDO;
}

__device__ __forceinline__ float dot3(float4 a, float4 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float finite_or_zero(float v) { return is_finite(v) ? v : 0.0f; }
__device__ __forceinline__ bool same(uint2 a, uint2 b) { return a.x == b.x && a.y == b.y; }

// The pixel of this lane: the block's tile (tiles in contiguous runs per XCD; blocks are dealt to
// the 8 XCDs round-robin), then the lane's place in it.
__device__ __forceinline__ bool pixel_of(u32 width, u32 rows, u32& x, u32& y) {
  u32 tiles_x = (width + TILE - 1) / TILE;
  u32 n = gridDim.x, b = blockIdx.x;
  u32 per = n / 8, rem = n % 8, xcd = b % 8, k = b / 8;
  u32 t = xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
  x = (t % tiles_x) * TILE + threadIdx.x % TILE;
  y = (t / tiles_x) * TILE + threadIdx.x / TILE;
  return x < width && y < rows;
}

// denoise.rs:456-534: guides, the depth gradient, the illumination and the pixel's own variance.
__global__ void __launch_bounds__(256) k_dn_setup(Planes in, u32 W, u32 R, u32 samples, uint2* ids, float4* nz,
                                                  float4* alb, float2* dz, float4* lv) {
  u32 x, y;
  if (!pixel_of(W, R, x, y)) return;
  size_t p = (size_t)y * W + x;
  uint2 id = make_uint2(in.part[p], in.material[p]);
  float z = in.depth[p];
  float4 a = make_float4(finite_or_zero(in.albedo[3 * p]), finite_or_zero(in.albedo[3 * p + 1]),
                         finite_or_zero(in.albedo[3 * p + 2]), 0.0f);
  ids[p] = id;
  nz[p] = make_float4(finite_or_zero(in.normal[3 * p]), finite_or_zero(in.normal[3 * p + 1]),
                      finite_or_zero(in.normal[3 * p + 2]), z);
  alb[p] = a;
  float2 g = make_float2(0.0f, 0.0f);
  if (id.x != 0 && is_finite(z)) {
    auto neighbour = [&](bool has, size_t q, float& zq) {
      if (!has || in.part[q] != id.x || in.material[q] != id.y) return false;
      zq = in.depth[q];
      return is_finite(zq);
    };
    auto grad = [&](bool ha, size_t qa, bool hb, size_t qb) {
      float za = 0.0f, zb = 0.0f;
      bool sa = neighbour(ha, qa, za), sb = neighbour(hb, qb, zb);
      if (sa && sb) return 0.5f * (zb - za);
      if (sa) return z - za;
      if (sb) return zb - z;
      return 0.0f;
    };
    g.x = grad(x > 0, p - 1, x + 1 < W, p + 1);
    g.y = grad(y > 0, p - W, y + 1 < R, p + W);
  }
  dz[p] = g;
  float4 l = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (id.x != 0) {
    float inv = 1.0f / (float)samples;
    const float* s = in.rgb + p * in.rgb_stride;
    float c0 = s[0] * inv, c1 = s[1] * inv, c2 = s[2] * inv;
    float d0 = fmax_r(a.x, MIN_ALBEDO), d1 = fmax_r(a.y, MIN_ALBEDO), d2 = fmax_r(a.z, MIN_ALBEDO);
    l.x = c0 / d0;
    l.y = c1 / d1;
    l.z = c2 / d2;
    bool enough = in.counts ? in.counts[p] >= (float)MIN_SAMPLES_FOR_VARIANCE : samples >= MIN_SAMPLES_FOR_VARIANCE;
    if (enough) {
      float lm = luminance(c0, c1, c2);
      float m2 = in.lum2[p] * inv;
      float dof = in.counts ? in.counts[p] - 1.0f : (float)(samples - 1);
      float v = fmax_r(m2 - lm * lm, 0.0f) / dof;
      float ld = luminance(d0, d1, d2);
      l.w = v / (ld * ld);
    }
  }
  lv[p] = l;
}

// denoise.rs:541-575
__global__ void __launch_bounds__(256) k_dn_firefly(u32 W, u32 R, const uint2* __restrict__ ids,
                                                    const float4* __restrict__ lv_in, float4* lv_out) {
  u32 x, y;
  if (!pixel_of(W, R, x, y)) return;
  size_t p = (size_t)y * W + x;
  uint2 id = ids[p];
  float4 l = lv_in[p];
  if (id.x != 0) {
    float before = luminance(l.x, l.y, l.z);
    float brightest = 0.0f, vmax = 0.0f;
    u32 count = 0;
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        int qx = (int)x + dx, qy = (int)y + dy;
        if ((dx == 0 && dy == 0) || qx < 0 || qy < 0 || qx >= (int)W || qy >= (int)R) continue;
        size_t q = (size_t)qy * W + qx;
        if (!same(id, ids[q])) continue;
        float4 o = lv_in[q];
        brightest = fmax_r(brightest, luminance(o.x, o.y, o.z));
        vmax = fmax_r(vmax, o.w);
        ++count;
      }
    float limit = FIREFLY_RATIO * brightest + FIREFLY_FLOOR;
    if (count >= 3 && before > limit) {
      float k = limit / before;
      l.x *= k;
      l.y *= k;
      l.z *= k;
      l.w = vmax;
    }
  }
  lv_out[p] = l;
}

// denoise.rs:576-624: the spread of the neighbours on the same surface (normal and depth
// weights) where a pixel has too few samples for its own.
__global__ void __launch_bounds__(256) k_dn_variance(u32 W, u32 R, int radius, u32 samples, const float* counts,
                                                     const uint2* __restrict__ ids, const float4* __restrict__ nz,
                                                     const float2* __restrict__ dz, const float4* __restrict__ lv_in,
                                                     float4* lv_out) {
  u32 x, y;
  if (!pixel_of(W, R, x, y)) return;
  size_t p = (size_t)y * W + x;
  uint2 id = ids[p];
  float4 l = lv_in[p];
  bool few = counts ? counts[p] < (float)MIN_SAMPLES_FOR_VARIANCE : samples < MIN_SAMPLES_FOR_VARIANCE;
  if (id.x != 0 && few) {
    float4 g = nz[p];
    float2 gd = dz[p];
    float sw = 0.0f, m1 = 0.0f, m2 = 0.0f;
    for (int dy = -radius; dy <= radius; ++dy) {
      int qy = (int)y + dy;
      if (qy < 0 || qy >= (int)R) continue;
      for (int dx = -radius; dx <= radius; ++dx) {
        int qx = (int)x + dx;
        if (qx < 0 || qx >= (int)W) continue;
        size_t q = (size_t)qy * W + qx;
        if (!same(id, ids[q])) continue;
        float4 gq = nz[q];
        float wgt = pow128(fmax_r(dot3(g, gq), 0.0f)) * falloff(depth_term(g.w, gd, gq.w, (float)dx, (float)dy));
        float4 o = lv_in[q];
        float lq = luminance(o.x, o.y, o.z);
        sw += wgt;
        m1 += wgt * lq;
        m2 += wgt * lq * lq;
      }
    }
    if (sw > 0.0f) {
      m1 = m1 / sw;
      m2 = m2 / sw;
      l.w = fmax_r(m2 - m1 * m1, 0.0f);
    }
  }
  lv_out[p] = l;
}

// denoise.rs:631-724: one pass of the transform, taps `step` pixels apart.
__global__ void __launch_bounds__(256) k_dn_atrous(u32 W, u32 R, int step, const uint2* __restrict__ ids,
                                                   const float4* __restrict__ nz, const float4* __restrict__ alb,
                                                   const float2* __restrict__ dz, const float4* __restrict__ lv_in,
                                                   float4* lv_out) {
  u32 x, y;
  if (!pixel_of(W, R, x, y)) return;
  size_t p = (size_t)y * W + x;
  uint2 id = ids[p];
  float4 l = lv_in[p];
  if (id.x != 0) {
    // The variance, blurred by a 3 x 3 Gaussian on the same surface.
    float sw = 0.0f, sv = 0.0f;
    for (int j = 0; j < 3; ++j) {
      int qy = (int)y + j - 1;
      if (qy < 0 || qy >= (int)R) continue;
      for (int i = 0; i < 3; ++i) {
        int qx = (int)x + i - 1;
        if (qx < 0 || qx >= (int)W) continue;
        size_t q = (size_t)qy * W + qx;
        if (!same(id, ids[q])) continue;
        float k = gauss3(j) * gauss3(i);
        sw += k;
        sv += k * lv_in[q].w;
      }
    }
    float gvar = sw > 0.0f ? sv / sw : l.w;
    float lp = luminance(l.x, l.y, l.z);
    float inv_sigma_l = 1.0f / (SIGMA_LUMINANCE * sqrt_(fmax_r(gvar, 0.0f) + 1.0e-10f));
    float4 g = nz[p];
    float4 ga = alb[p];
    float2 gd = dz[p];
    sw = 0.0f;
    sv = 0.0f;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
    for (int j = 0; j < 5; ++j) {
      int ddy = (j - 2) * step;
      int qy = (int)y + ddy;
      if (qy < 0 || qy >= (int)R) continue;
      for (int i = 0; i < 5; ++i) {
        int ddx = (i - 2) * step;
        int qx = (int)x + ddx;
        if (qx < 0 || qx >= (int)W) continue;
        size_t q = (size_t)qy * W + qx;
        if (!same(id, ids[q])) continue;
        float4 gq = nz[q];
        float wn = q == p ? 1.0f : pow128(fmax_r(dot3(g, gq), 0.0f));
        if (wn <= 0.0f) continue;
        float4 o = lv_in[q];
        float exponent = fabs_(lp - luminance(o.x, o.y, o.z)) * inv_sigma_l +
                         depth_term(g.w, gd, gq.w, (float)ddx, (float)ddy) + albedo_term(ga, alb[q]);
        float wgt = kernel5(j) * kernel5(i) * wn * falloff(exponent);
        sw += wgt;
        sv += wgt * wgt * o.w;
        s0 += wgt * o.x;
        s1 += wgt * o.y;
        s2 += wgt * o.z;
      }
    }
    // `sw` cannot be zero: the pixel's own tap carries 9/64 of the kernel's weight whatever the
    // guides say.
    float inv_w = 1.0f / sw;
    l = make_float4(s0 * inv_w, s1 * inv_w, s2 * inv_w, sv * inv_w * inv_w);
  }
  lv_out[p] = l;
}

// denoise.rs:726-741 and 370-381: the albedo multiplied back as sums of `samples` samples (the noisy
// sums where that is not finite; the background and the coverage as they were), then blended with
// the noisy sums at strength `s` in (0, 1].
__global__ void __launch_bounds__(256) k_dn_final(Planes in, u32 W, u32 R, u32 samples, float s,
                                                  const uint2* __restrict__ ids, const float4* __restrict__ alb,
                                                  const float4* __restrict__ lv, float4* out) {
  u32 x, y;
  if (!pixel_of(W, R, x, y)) return;
  size_t p = (size_t)y * W + x;
  const float* rgb = in.rgb + p * in.rgb_stride;
  float4 noisy = make_float4(rgb[0], rgb[1], rgb[2], in.alpha[p * in.alpha_stride]);
  float4 clean = noisy;
  if (ids[p].x != 0) {
    float count = (float)samples;
    float4 a = alb[p], l = lv[p];
    float c0 = l.x * fmax_r(a.x, MIN_ALBEDO), c1 = l.y * fmax_r(a.y, MIN_ALBEDO), c2 = l.z * fmax_r(a.z, MIN_ALBEDO);
    clean.x = is_finite(c0) ? c0 * count : noisy.x;
    clean.y = is_finite(c1) ? c1 * count : noisy.y;
    clean.z = is_finite(c2) ? c2 * count : noisy.z;
  }
  if (s < 1.0f) {
    float k = 1.0f - s;
    clean.x = noisy.x * k + clean.x * s;
    clean.y = noisy.y * k + clean.y * s;
    clean.z = noisy.z * k + clean.z * s;
  }
  out[p] = clean;
}

// Strength 0 (denoise.rs:393-402): no filter runs; the noisy sums as RGBA.
__global__ void __launch_bounds__(256) k_dn_noisy(Planes in, u32 W, u32 R, float4* out) {
  u32 x, y;
  if (!pixel_of(W, R, x, y)) return;
  size_t p = (size_t)y * W + x;
  const float* rgb = in.rgb + p * in.rgb_stride;
  out[p] = make_float4(rgb[0], rgb[1], rgb[2], in.alpha[p * in.alpha_stride]);
}

}  // namespace

Denoiser::~Denoiser() {
  if (mem_) hipFree(mem_);
  for (hipEvent_t& e : ev_)
    if (e) hipEventDestroy(e);
}

bool Denoiser::ensure(size_t n) {
  if (n <= cap_) return true;
  if (mem_) hipFree(mem_);
  mem_ = nullptr;
  cap_ = 0;
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t b_ids = up(n * 8), b_nz = up(n * 16), b_alb = up(n * 16), b_dz = up(n * 8), b_lv = up(n * 16);
  if (hipMalloc(&mem_, b_ids + b_nz + b_alb + b_dz + 2 * b_lv) != hipSuccess) {
    mem_ = nullptr;
    return false;
  }
  char* m = static_cast<char*>(mem_);
  ids_ = reinterpret_cast<uint2*>(m);
  nz_ = reinterpret_cast<float4*>(m + b_ids);
  alb_ = reinterpret_cast<float4*>(m + b_ids + b_nz);
  dz_ = reinterpret_cast<float2*>(m + b_ids + b_nz + b_alb);
  lv_[0] = reinterpret_cast<float4*>(m + b_ids + b_nz + b_alb + b_dz);
  lv_[1] = reinterpret_cast<float4*>(m + b_ids + b_nz + b_alb + b_dz + b_lv);
  cap_ = n;
  return true;
}

hipError_t Denoiser::run(const Planes& in, u32 W, u32 R, u32 samples, Preset preset, float strength, float4* out,
                         hipStream_t st) {
  preset = preset.clamped();
  samples = std::max<u32>(samples, 1);
  float s = is_nan(strength) ? 1.0f : clampf(strength, 0.0f, 1.0f);
  size_t n = (size_t)W * R;
  timing = DenoiseTiming{};
  if (n == 0) return hipSuccess;
  u32 tiles = ((W + TILE - 1) / TILE) * ((R + TILE - 1) / TILE);
  dim3 grid(tiles), block(TILE * TILE);
  if (s <= 0.0f) {
    hipLaunchKernelGGL(k_dn_noisy, grid, block, 0, st, in, W, R, out);
    return hipGetLastError();
  }
  if (!ensure(n)) return hipErrorOutOfMemory;
  if (profiling && !ev_[0])
    for (hipEvent_t& e : ev_)
      if (hipEventCreate(&e) != hipSuccess) return hipErrorOutOfMemory;
  auto mark = [&](int i) {
    if (profiling) hipEventRecord(ev_[i], st);
  };
  mark(0);
  hipLaunchKernelGGL(k_dn_setup, grid, block, 0, st, in, W, R, samples, ids_, nz_, alb_, dz_, lv_[0]);
  mark(1);
  hipLaunchKernelGGL(k_dn_firefly, grid, block, 0, st, W, R, ids_, lv_[0], lv_[1]);
  mark(2);
  int cur = 1;
  // The spatial estimate runs where a pixel has too few samples for its own variance (with
  // per-pixel counts the kernel decides pixel by pixel).
  if (in.counts || samples < MIN_SAMPLES_FOR_VARIANCE) {
    hipLaunchKernelGGL(k_dn_variance, grid, block, 0, st, W, R, (int)preset.radius, samples, in.counts, ids_, nz_, dz_,
                       lv_[1], lv_[0]);
    cur = 0;
  }
  mark(3);
  for (u32 level = 0; level < preset.levels; ++level) {
    hipLaunchKernelGGL(k_dn_atrous, grid, block, 0, st, W, R, 1 << level, ids_, nz_, alb_, dz_, lv_[cur], lv_[1 - cur]);
    cur = 1 - cur;
  }
  mark(4);
  hipLaunchKernelGGL(k_dn_final, grid, block, 0, st, in, W, R, samples, s, ids_, alb_, lv_[cur], out);
  mark(5);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !profiling) return e;
  e = hipEventSynchronize(ev_[5]);
  if (e != hipSuccess) return e;
  float ms[5];
  for (int i = 0; i < 5; ++i) hipEventElapsedTime(&ms[i], ev_[i], ev_[i + 1]);
  timing.setup_ms = ms[0];
  timing.firefly_ms = ms[1];
  timing.variance_ms = ms[2];
  timing.atrous_ms = ms[3];
  timing.final_ms = ms[4];
  hipEventElapsedTime(&timing.total_ms, ev_[0], ev_[5]);
  timing.levels = preset.levels;
  return hipSuccess;
}

}  // namespace denoise
}  // namespace rpx
