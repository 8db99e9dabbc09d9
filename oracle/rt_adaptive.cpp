// ORACLE — test infrastructure only (see rt_core.hpp). Adaptive sampling, restated from
// engine/crates/rt-core/src/adaptive.rs (line numbers per function below).
#include "rt_adaptive.hpp"
#include <algorithm>
#include <cmath>
#include <limits>

namespace orc {
namespace adaptive {

// adaptive.rs:95-97 with color.rs:146-148 (exposure = fmath exp2).
float floor_for(float ev) { return EPSILON / exp2f_(ev); }

// adaptive.rs:102-127
float pixel_error(float r, float g, float b, float lum2, u32 n, float floor) {
  if (n < 2) return UNMEASURED;
  float nf = (float)n;
  float inv = 1.0f / nf;
  float l = 0.2126f * (r * inv) + 0.7152f * (g * inv) + 0.0722f * (b * inv);
  float m2 = lum2 * inv;
  if (!(std::isfinite(l) && std::isfinite(m2))) return UNMEASURED;
  float d = m2 - l * l;
  float spread = d > 0.0f ? d : 0.0f;
  float var = spread / (nf - 1.0f);
  float mean = l > 0.0f ? l : 0.0f;
  float e = std::sqrt(var) / (mean + floor);
  return (std::isfinite(e) && e < UNMEASURED) ? e : UNMEASURED;
}

// adaptive.rs:130-132
void block_grid(u32 width, u32 height, u32& cols, u32& rows) {
  cols = (width + BLOCK - 1) / BLOCK;
  rows = (height + BLOCK - 1) / BLOCK;
}

// adaptive.rs:149-262
Decision decide(u32 width, u32 height, const std::vector<u32>& counts, const float* sums, size_t stride,
                const float* lum2, float floor, float target, u32 cap, std::vector<uint8_t>& active) {
  size_t w = width, h = height;
  u32 cols_u, rows_u;
  block_grid(width, height, cols_u, rows_u);
  size_t cols = cols_u, rows = rows_u, blocks = cols * rows;
  auto block_of = [&](size_t x, size_t y) { return (y / BLOCK) * cols + x / BLOCK; };
  std::vector<float> error(w * h, 0.0f);
  for (size_t y = 0; y < h; ++y)
    for (size_t x = 0; x < w; ++x) {
      size_t p = y * w + x, i = p * stride;
      error[p] = pixel_error(sums[i], sums[i + 1], sums[i + 2], lum2[p], counts[block_of(x, y)], floor);
    }
  std::vector<float> block_error(blocks, 0.0f);
  for (size_t y = 0; y < h; ++y) {
    size_t y0 = y > 0 ? y - 1 : 0, y1 = std::min(y + 2, h);
    for (size_t x = 0; x < w; ++x) {
      size_t x0 = x > 0 ? x - 1 : 0, x1 = std::min(x + 2, w);
      float sum = 0.0f;
      for (size_t ny = y0; ny < y1; ++ny)
        for (size_t nx = x0; nx < x1; ++nx) sum += error[ny * w + nx];
      float s = sum / (float)((y1 - y0) * (x1 - x0));
      size_t b = block_of(x, y);
      if (s > block_error[b]) block_error[b] = s;
    }
  }
  float floor_target = target > 0.0f ? target : 0.0f;
  auto candidate = [&](size_t b) { return block_error[b] > floor_target && (cap == 0 || counts[b] < cap); };
  std::vector<float> errors;
  for (size_t b = 0; b < blocks; ++b)
    if (candidate(b)) errors.push_back(block_error[b]);
  float max_error = 0.0f;
  for (float e : block_error)
    if (e > max_error) max_error = e;
  float remaining = 0.0f;
  if (floor_target > 0.0f) {
    for (size_t b = 0; b < blocks; ++b) {
      if (!candidate(b)) continue;
      float ratio = block_error[b] / floor_target;
      float n = (float)counts[b];
      float need = n * (ratio * ratio - 1.0f);
      if (cap > 0) {
        float room = (float)(cap - counts[b]);
        if (need > room) need = room;
      }
      size_t bx = b % cols, by = b / cols;
      size_t bw = std::min(w - bx * BLOCK, (size_t)BLOCK), bh = std::min(h - by * BLOCK, (size_t)BLOCK);
      remaining += need * (float)(bw * bh);
    }
  }
  active.assign(blocks, 0);
  if (errors.empty()) return Decision{0, 0, (u32)blocks, 0.0f, max_error, remaining};
  // f32::total_cmp: the errors are finite and non-negative, so the usual order.
  std::sort(errors.begin(), errors.end());
  float threshold = errors[(errors.size() - 1) / 2];
  u32 count = 0, least = std::numeric_limits<u32>::max();
  for (size_t b = 0; b < blocks; ++b)
    if (candidate(b) && block_error[b] >= threshold) {
      active[b] = 1;
      ++count;
      least = std::min(least, counts[b]);
    }
  u32 factor = std::min(std::max((u32)blocks / (4 * count), 1u), MAX_ROUND_FACTOR);
  u32 round = std::min(ROUND * factor, std::max(least, ROUND));
  return Decision{round, count, (u32)blocks, threshold, max_error, remaining};
}

// adaptive.rs:358-378
Adaptive::Adaptive(u32 width, u32 height, float target, float ev, u32 cap)
    : width_(width), height_(height), cap_(cap) {
  block_grid(width, height, cols_, rows_);
  counts_.assign((size_t)cols_ * rows_, 0);
  active_.assign((size_t)cols_ * rows_, 1);
  round_ = cap > 0 ? std::min(WARMUP, cap) : WARMUP;
  target_ = std::max(target, 0.0f);
  floor_ = floor_for(ev);
}

// adaptive.rs:421-430
u32 Adaptive::quota(u32 b) const {
  if (!active_[b]) return 0;
  if (cap_ > 0) return std::min(round_, cap_ > counts_[b] ? cap_ - counts_[b] : 0u);
  return round_;
}

u32 Adaptive::samples_of(u32 b) const { return counts_[b] + std::min(quota(b), done_); }

void Adaptive::min_max(u32& lo, u32& hi) const {
  lo = std::numeric_limits<u32>::max();
  hi = 0;
  for (u32 b = 0; b < blocks(); ++b) {
    u32 n = samples_of(b);
    lo = std::min(lo, n);
    hi = std::max(hi, n);
  }
  if (lo == std::numeric_limits<u32>::max()) lo = hi = 0;
}

// adaptive.rs:517-539
Decision Adaptive::decide(const float* sums, size_t stride, const float* lum2) {
  for (u32 b = 0; b < blocks(); ++b) counts_[b] += quota(b);
  iteration_ += round_;
  done_ = 0;
  Decision d = adaptive::decide(width_, height_, counts_, sums, stride, lum2, floor_, target_, cap_, active_);
  round_ = d.round;
  has_last_ = true;
  return d;
}

}  // namespace adaptive
}  // namespace orc
