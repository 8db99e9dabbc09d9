// Adaptive sampling: the GPU's per-pixel and per-block errors and the host's choice of blocks
// (engine/crates/rt-core/src/adaptive.rs; line numbers per function). The arithmetic is the
// reference's f32, in its order, so the decisions are the reference's bit for bit.
#include <algorithm>
#include <limits>
#include "rpx_adaptive.h"

namespace rpx {
namespace adaptive {

float floor_for(float ev) { return EPSILON / fexp2(ev); }

namespace {
// adaptive.rs:102-127
__device__ __forceinline__ float pixel_error(float r, float g, float b, float lum2, u32 n, float floor) {
  if (n < 2) return UNMEASURED;
  float nf = (float)n;
  float inv = 1.0f / nf;
  float l = 0.2126f * (r * inv) + 0.7152f * (g * inv) + 0.0722f * (b * inv);
  float m2 = lum2 * inv;
  if (!(is_finite(l) && is_finite(m2))) return UNMEASURED;
  float d = m2 - l * l;
  float spread = d > 0.0f ? d : 0.0f;
  float var = spread / (nf - 1.0f);
  float mean = l > 0.0f ? l : 0.0f;
  float e = sqrt_(var) / (mean + floor);
  return (is_finite(e) && e < UNMEASURED) ? e : UNMEASURED;
}
}  // namespace

__global__ void __launch_bounds__(256) k_pixel_error(u32 width, u32 height, u32 cols, const float* sums, u32 stride,
                                                     const float* lum2, const u32* counts, float floor, float* error) {
  u32 n = width * height;
  for (u32 p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    u32 x = p % width, y = p / width;
    const float* s = sums + (size_t)p * stride;
    error[p] = pixel_error(s[0], s[1], s[2], lum2[p], counts[(y / BLOCK) * cols + x / BLOCK], floor);
  }
}

// adaptive.rs:173-186: the 3 x 3 mean (inside the image, rows then columns) and its largest value
// per block. One wave per 8 x 8 block, lane t its pixel (t % 8, t / 8); the smoothed errors are
// finite and non-negative, so the largest float is the largest bit pattern, taken across the wave
// by shuffles (lanes outside the image hold 0).
__global__ void __launch_bounds__(64) k_block_error(u32 width, u32 height, u32 cols, u32 rows, const float* error,
                                                    u32* block_error_bits) {
  for (u32 blk = blockIdx.x; blk < cols * rows; blk += gridDim.x) {
    u32 x = (blk % cols) * BLOCK + threadIdx.x % BLOCK, y = (blk / cols) * BLOCK + threadIdx.x / BLOCK;
    u32 bits = 0;
    if (x < width && y < height) {
      u32 y0 = y > 0 ? y - 1 : 0, y1 = y + 2 < height ? y + 2 : height;
      u32 x0 = x > 0 ? x - 1 : 0, x1 = x + 2 < width ? x + 2 : width;
      float sum = 0.0f;
      for (u32 ny = y0; ny < y1; ++ny)
        for (u32 nx = x0; nx < x1; ++nx) sum += error[(size_t)ny * width + nx];
      bits = f2u(sum / (float)((y1 - y0) * (x1 - x0)));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      u32 o = (u32)__shfl_xor((int)bits, off);
      bits = o > bits ? o : bits;
    }
    if (threadIdx.x == 0) block_error_bits[blk] = bits;
  }
}

hipError_t block_errors(u32 width, u32 height, const float* sums, u32 stride, const float* lum2, const u32* counts,
                        float floor, float* error, float* block_error, hipStream_t st) {
  u32 cols = (width + BLOCK - 1) / BLOCK, rows = (height + BLOCK - 1) / BLOCK;
  u32 n = width * height;
  u32 grid = std::min<u32>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_pixel_error, dim3(grid), dim3(256), 0, st, width, height, cols, sums, stride, lum2, counts,
                     floor, error);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_block_error, dim3(std::min<u32>(cols * rows, 1u << 16)), dim3(64), 0, st, width, height, cols,
                     rows, error, (u32*)block_error);
  return hipGetLastError();
}

// adaptive.rs:187-262
Decision choose(u32 width, u32 height, const std::vector<u32>& counts, const std::vector<float>& block_error,
                float target, u32 cap, std::vector<uint8_t>& active) {
  size_t w = width, h = height;
  size_t cols = (width + BLOCK - 1) / BLOCK, rows = (height + BLOCK - 1) / BLOCK, blocks = cols * rows;
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
  // The errors are finite and non-negative: the plain order is the reference's total order.
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

// adaptive.rs:358-378, 421-430
void State::init(u32 w, u32 h, float t, float ev, u32 c) {
  width = w;
  height = h;
  cols = (w + BLOCK - 1) / BLOCK;
  rows = (h + BLOCK - 1) / BLOCK;
  counts.assign((size_t)cols * rows, 0);
  active.assign((size_t)cols * rows, 1);
  cap = c;
  round = c > 0 ? std::min(WARMUP, c) : WARMUP;
  done = 0;
  iteration = 0;
  target = t > 0.0f ? t : 0.0f;
  floor = floor_for(ev);
  decided = false;
}

u32 State::quota(u32 b) const {
  if (!active[b]) return 0;
  if (cap > 0) return std::min(round, cap > counts[b] ? cap - counts[b] : 0u);
  return round;
}

}  // namespace adaptive
}  // namespace rpx
