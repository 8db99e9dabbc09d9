// ORACLE — test infrastructure only (see rt_core.hpp).
// Restates the reference's adaptive sampling (engine/crates/rt-core/src/adaptive.rs): the per-pixel
// relative error of the mean (:102-127), its 3 x 3 smoothing and per-block maximum, the candidates,
// the 50 % quantile threshold and the size of the next round (decide, :149-262), and the state of an
// adaptive render in rounds (Adaptive, :319-554).
#pragma once
#include <cstdint>
#include <vector>
#include "rt_core.hpp"

namespace orc {
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

float floor_for(float ev);  // adaptive.rs:95-97 (EPSILON / 2^ev, fmath exp2)
float pixel_error(float r, float g, float b, float lum2, u32 n, float floor);
void block_grid(u32 width, u32 height, u32& cols, u32& rows);
Decision decide(u32 width, u32 height, const std::vector<u32>& counts, const float* sums, size_t stride,
                const float* lum2, float floor, float target, u32 cap, std::vector<uint8_t>& active);

class Adaptive {
 public:
  Adaptive(u32 width, u32 height, float target, float ev, u32 cap);
  bool warming_up() const { return !has_last_; }
  u32 left() const { return round_ - done_; }
  bool converged() const { return round_ == 0; }
  u32 iteration() const { return iteration_ + done_; }
  u32 blocks() const { return (u32)counts_.size(); }
  u32 block_of(u32 x, u32 y) const { return (y / BLOCK) * cols_ + x / BLOCK; }
  u32 quota(u32 b) const;
  u32 samples_of(u32 b) const;
  void min_max(u32& lo, u32& hi) const;
  void advance(u32 count) { done_ = std::min(done_ + count, round_); }
  Decision decide(const float* sums, size_t stride, const float* lum2);
  const std::vector<u32>& counts() const { return counts_; }
  const std::vector<uint8_t>& active() const { return active_; }
  u32 cols() const { return cols_; }

 private:
  u32 width_, height_, cols_, rows_;
  std::vector<u32> counts_;
  std::vector<uint8_t> active_;
  u32 round_, done_ = 0, iteration_ = 0, cap_;
  float target_, floor_;
  bool has_last_ = false;
};

}  // namespace adaptive
}  // namespace orc
