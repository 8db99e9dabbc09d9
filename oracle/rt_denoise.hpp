// ORACLE — test infrastructure only (see rt_core.hpp).
// Restates the reference's denoiser, an edge-avoiding a-trous wavelet filter guided by the exact
// passes and each pixel's sample variance (engine/crates/rt-post/src/denoise.rs):
//   :44-130    constants, Preset (FULL, VIEWPORT, halo, clamped)
//   :186-288   DenoiseInput shape checks and bands (with_halo_of)
//   :322-407   denoise_rows_with, check_rows, strength_of, blend, denoise_rows_at
//   :409-785   luminance, pow128, Guide, filter_band, depth_term, albedo_term, falloff,
//              surface_weight
// Every operation is one f32 IEEE operation in the reference's order (the build has no FMA
// contraction), so a band gives the reference's bits.
#pragma once
#include <vector>
#include "rt_core.hpp"

namespace orc {
namespace denoise {

constexpr u32 LEVELS = 5;
constexpr float SIGMA_LUMINANCE = 4.0f;
constexpr float SIGMA_DEPTH = 1.0f;
constexpr float SIGMA_ALBEDO = 0.1f;
constexpr u32 MIN_SAMPLES_FOR_VARIANCE = 4;
constexpr float MIN_ALBEDO = 0.01f;
constexpr float FIREFLY_RATIO = 4.0f;
constexpr float FIREFLY_FLOOR = 0.05f;

struct Preset {
  u32 levels, variance_radius;
  u32 halo() const { return (2u << levels) + 8; }
  Preset clamped() const;
};
constexpr Preset FULL{5, 3};
constexpr Preset VIEWPORT{3, 1};

// A band of rows of the image (all planes `width x rows` pixels from row `origin_y`); `counts`
// (one per pixel) may be null.
struct Input {
  u32 width = 0, height = 0, origin_y = 0, rows = 0, samples = 0;
  const float* counts = nullptr;
  const float* sums = nullptr;  // RGBA, 4 per pixel
  const float* lum2 = nullptr;
  const float* albedo = nullptr;  // 3 per pixel
  const float* normal = nullptr;  // 3 per pixel
  const float* depth = nullptr;
  const u32* part_id = nullptr;
  const u32* material_id = nullptr;
};

enum Status { OK = 0, BAD_SHAPE = 1, BAD_ROWS = 2 };

// The filter over the whole band: RGBA sums of `width x rows` pixels.
std::vector<float> filter_band(const Input& in, Preset preset);
// denoise_rows_at: rows y0..y1 of the image at `strength` (RGBA sums, width x (y1 - y0)).
Status denoise_rows_at(const Input& in, u32 y0, u32 y1, Preset preset, float strength, std::vector<float>& out);

}  // namespace denoise
}  // namespace orc
