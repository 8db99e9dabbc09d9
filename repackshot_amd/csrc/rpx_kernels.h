// Wavefront path state (structure of arrays in HBM) and the kernel launchers of the GPU tracer.
#pragma once
#include "rpx_scene.h"

namespace rpx {

constexpr u32 FLAG_PRIMARY = 1, FLAG_HAS_PDF = 2, FLAG_MEDIUM = 4, FLAG_BY_FLUX = 8;

// One entry per path slot (a pixel of the current sample pass).
struct PathState {
  float4* ray_o;   // origin
  float4* ray_d;   // direction
  uint4* hit;      // instance, mesh triangle, t bits (written by extend)
  float4* thr;     // throughput rgb, density of the BSDF sample that made the ray
  float4* sum;     // radiance rgb, alpha
  float4* med;     // absorption coefficients of the solid the ray is in, flags (bits)
  float4* sel_p;   // light selection of the vertex the ray left: point, total
  float4* sel_w0;  // its kept weights 0..3
  float4* sel_w1;  // its kept weights 4..7
  float4* sh_o;    // shadow ray origin, t_max
  float4* sh_d;    // shadow ray direction
  float4* sh_c;    // contribution the shadow ray adds when unblocked
  u64* key;        // pixel key of the sampler
  u32* queue[2];   // extension queues (ping-pong by round)
  u32* shadow_queue;
  u32* counts;     // [0], [1] queue lengths, [2] shadow queue length
  u64* rays;       // [0] closest-hit rays traced, [1] shadow rays counted without a trace, [2] shadow rays traced
};

struct Film {
  float* rgb;
  float* weight;
  float* alpha;
  float* lum2;
};

struct PassArgs {
  u32 n_paths, x0, y0, w, sample_index;
  u32 sobol_cam[4];  // unscrambled Sobol points of logical dimensions 0..3 at this sample index
};

struct RoundArgs {
  u32 round, sample_index;
  u32 sobol[6];  // unscrambled Sobol points of the round's six vertex dimensions
};

hipError_t launch_camera(const DevScene& S, const PathState& P, const PassArgs& A, hipStream_t st);
hipError_t launch_extend(const DevScene& S, const PathState& P, u32 round, u32 max_paths, hipStream_t st);
hipError_t launch_shade(const DevScene& S, const PathState& P, const RoundArgs& R, u32 max_paths, hipStream_t st);
hipError_t launch_shadow(const DevScene& S, const PathState& P, u32 max_paths, hipStream_t st);
hipError_t launch_splat(const PathState& P, const Film& F, u32 n, hipStream_t st);
hipError_t launch_tile_out(const PathState& P, float* out, u32 n, u32 channels, hipStream_t st);
hipError_t launch_trace_rays(const DevScene& S, const float* rays, u32 n, u32* inst, u32* tri, float* t, uint8_t* any,
                             hipStream_t st);
hipError_t launch_energy(uint16_t* albedo, uint16_t* schlick, uint16_t* clear_in, uint16_t* clear_out, float* albedo_avg,
                         float* schlick_avg, hipStream_t st);

// The Sobol generator matrices (Joe–Kuo new-joe-kuo-6.21201, 64 dimensions) and the logical-to-
// Sobol dimension layout of the reference sampler.
u32 sobol_point(u32 logical_dim, u32 index);

}  // namespace rpx
