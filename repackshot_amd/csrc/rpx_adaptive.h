// Adaptive sampling (engine/crates/rt-core/src/adaptive.rs, the reference's "E10"): rounds of
// samples go to the 8 x 8 blocks whose relative error of the mean is highest. The per-pixel error,
// its 3 x 3 smoothing and the per-block maximum run on the GPU over the resident film; the choice of
// blocks (candidates, the 50 % quantile, the round's size) runs on the host over one float per
// block. Every pixel stays the plain mean of its own samples, sample indices 0..n in order.
#pragma once
#include <string>
#include <vector>
#include "rpx_kernels.h"

namespace rpx {
namespace adaptive {

constexpr u32 BLOCK = 8;
constexpr u32 WARMUP = 16;
constexpr u32 ROUND = 16;
constexpr u32 MAX_ROUND_FACTOR = 16;
constexpr float EPSILON = 0.01f;
constexpr float UNMEASURED = 1.0e30f;

struct Decision {
  u32 round, active, blocks;
  float threshold, max_error, remaining;
};

// The floor of the mean in the relative error for an exposure of `ev` stops (adaptive.rs:95-97).
float floor_for(float ev);

// Block errors of a width x height image on the GPU (adaptive.rs:149-186): `sums` device radiance
// sums with `stride` floats per pixel, `lum2` device squared-luminance sums, `counts` device
// samples per block, `error` device scratch (one float per pixel), `block_error` device output
// (one float per block).
hipError_t block_errors(u32 width, u32 height, const float* sums, u32 stride, const float* lum2, const u32* counts,
                        float floor, float* error, float* block_error, hipStream_t st);

// The rest of the decision on the host (adaptive.rs:187-262) from the block errors.
Decision choose(u32 width, u32 height, const std::vector<u32>& counts, const std::vector<float>& block_error,
                float target, u32 cap, std::vector<uint8_t>& active);

// An adaptive render's state (adaptive.rs:319-554).
struct State {
  u32 width = 0, height = 0, cols = 0, rows = 0;
  std::vector<u32> counts;
  std::vector<uint8_t> active;
  u32 round = 0, done = 0, iteration = 0, cap = 0;
  float target = 0, floor = 0;
  bool decided = false;
  void init(u32 w, u32 h, float target, float ev, u32 cap);
  u32 quota(u32 b) const;
  u32 left() const { return round - done; }
};

}  // namespace adaptive
}  // namespace rpx
