// The denoiser on the GPU (engine/crates/rt-post/src/denoise.rs, the reference's "E5b-1"): an
// edge-avoiding a-trous wavelet filter (B3-spline taps 2^k apart) guided by the exact passes
// (normal, depth, part and material ids, albedo) and by each pixel's sample variance, with the
// reference's firefly clamp and spatial variance estimate. Each stage is one kernel over the
// pixels of a band (one lane per pixel, 16 x 16 tiles handed to the XCDs in contiguous runs); the
// arithmetic is the reference's f32 in its order, so the result is its bits.
#pragma once
#include <vector>
#include "rpx_kernels.h"

namespace rpx {
namespace denoise {

constexpr u32 LEVELS = 5;  // passes of the full filter
constexpr u32 MIN_SAMPLES_FOR_VARIANCE = 4;

// A preset (denoise.rs:67-120): passes of the transform and the radius of the spatial variance
// estimate; `clamped` brings both into their ranges.
struct Preset {
  u32 levels, radius;
  u32 halo() const { return (2u << levels) + 8; }
  Preset clamped() const {
    return Preset{levels < 1 ? 1u : (levels > LEVELS ? LEVELS : levels), radius < 1 ? 1u : (radius > 3 ? 3u : radius)};
  }
};
constexpr Preset FULL{5, 3};
constexpr Preset VIEWPORT{3, 1};

// The input planes of a band of `rows` rows (`width` pixels each) in device memory. The radiance
// sums of pixel p are rgb[p * rgb_stride + 0..3] and its coverage sum alpha[p * alpha_stride]
// (a film's separate planes, or interleaved RGBA sums). `counts` (one per pixel) may be null:
// every pixel then has `samples` samples.
struct Planes {
  const float* rgb;
  u32 rgb_stride;
  const float* alpha;
  u32 alpha_stride;
  const float* lum2;
  const float* albedo;  // 3 per pixel
  const float* normal;  // 3 per pixel
  const float* depth;
  const u32* part;
  const u32* material;
  const float* counts;
};

// GPU time of the last run's stages (ms; events on the run's stream).
struct DenoiseTiming {
  float setup_ms = 0, firefly_ms = 0, variance_ms = 0, atrous_ms = 0, final_ms = 0, total_ms = 0;
  u32 levels = 0;
};

// Work buffers of one band size, reused across runs.
class Denoiser {
 public:
  Denoiser() = default;
  Denoiser(const Denoiser&) = delete;
  Denoiser& operator=(const Denoiser&) = delete;
  ~Denoiser();
  // Filters the whole band at `strength` (clamped to 0..1; NaN is 1; at 0 no filter runs and the
  // noisy sums come out) into `out`: RGBA sums of width x rows pixels (device memory).
  hipError_t run(const Planes& in, u32 width, u32 rows, u32 samples, Preset preset, float strength, float4* out,
                 hipStream_t st);
  bool profiling = false;
  DenoiseTiming timing;

 private:
  bool ensure(size_t n);
  size_t cap_ = 0;
  void* mem_ = nullptr;
  uint2* ids_ = nullptr;    // part, material
  float4* nz_ = nullptr;    // normal, depth
  float4* alb_ = nullptr;   // albedo (finite), unused
  float2* dz_ = nullptr;    // depth gradient per pixel along x and y
  float4* lv_[2] = {nullptr, nullptr};  // illumination (albedo divided out) and its variance
  hipEvent_t ev_[8] = {};
};

}  // namespace denoise
}  // namespace rpx
