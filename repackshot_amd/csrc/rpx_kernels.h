// Wavefront path state (structure of arrays in HBM) and the kernel launchers of the GPU tracer.
//
// The tracer keeps a pool of path slots resident on the device. Work is a list of segments
// (samples [first, first + count) of a pixel rectangle of one scene); every (segment, sample,
// pixel) is one work id, numbered segment by segment, then pixel by pixel (8 x 8 tiles), the
// samples of a pixel consecutive. Each iteration runs
//   regen   the slots finished two iterations before write their radiance to `results[work id]`
//           and take the next work ids from a device cursor (camera ray of the new path),
//           appending to the live queue
//   extend  closest hit of every live path; beside it, on a second stream, the shadow kernel
//           (any hit of the previous iteration's shadow rays; unblocked ones add their light)
//   shade   surface, emission, NEE, BSDF sample, roulette: survivors to the other live queue,
//           finished paths to this parity's done queue, shadow rays to this parity's shadow queue
// until a regen leaves no live path and nothing pending. A path's additions keep their order: the
// shadow ray of a vertex adds its light before the shade of the next vertex runs. A splat then adds each pixel's samples to its film in
// sample-index order, so the film does not depend on the pool size or on the order paths finish.
#pragma once
#include "rpx_scene.h"

namespace rpx {

// Path flags. FLAG_SCATTER: the ray is inside a scattering medium (the material index in the
// bits of med.x), so its walk runs before the next vertex; FLAG_EXIT_PDF: the vertex the ray left
// was a medium's exit.
constexpr u32 FLAG_PRIMARY = 1, FLAG_HAS_PDF = 2, FLAG_MEDIUM = 4, FLAG_BY_FLUX = 8, FLAG_SCATTER = 16,
              FLAG_EXIT_PDF = 32, FLAG_WALKING = 64;
// Where a path of a scene with caustics is in the split between the photon map and the paths
// (integrator/mod.rs:1084-1113): FLAG_CAU_RECV after a receiver (a surface that is not a caster),
// with FLAG_CAU_CASTERS once a caster followed it; neither: the eye state.
constexpr u32 FLAG_CAU_RECV = 128, FLAG_CAU_CASTERS = 1u << 20;
// FLAG_WALKING: the ray is a step of a walk in progress (its state in walk_w / walk_r); the shade
// passes the path on untouched and the next extend traces the step.
constexpr u32 EVENTS_SHIFT = 8, EVENTS_MASK = 0xfffu;  // flags bits 8..19: medium steps of the path so far
constexpr u32 BOUNCE_SHIFT = 21;  // flags bits 21..31: the bounce of the path's next vertex (at most 1025)
// How a walk through a scattering medium ended (hit.w, written by the walk kernel; 0: no walk).
constexpr u32 WALK_CROSSED = 1, WALK_EXIT = 2, WALK_ESCAPED = 3, WALK_ABSORBED = 4;
constexpr u32 NO_WORK = 0xffffffffu;

struct Film {
  float* rgb;
  float* weight;
  float* alpha;
  float* lum2;
};

// Samples [first, first + count) of the pixels (x0 .. x0 + w, y0 .. y0 + h) of scene `scene`;
// work ids base .. base + count * w * h, sample-minor (base + p * count + k: sample first + k of
// the p-th pixel, pixels in 8 x 8 tiles), each path's result stored at its work id.
struct DevSegment {
  u32 scene;
  u32 x0, y0, w, h;
  u32 first, count;
  u32 base;
  Film film;
  u32 film_w;  // film row length in pixels
  // Adjoint runs: d loss / d image (rgb per film pixel) and the image's d image / d sample.
  const float* adjoint;
  float adjoint_scale;
};

// Counter slots (u32) of PathState::counts.
// C_TAIL: paths handed to the drain kernel by the last bookkeeping (0: none).
// The shadow and finished queues are kept per iteration parity: the shadow rays of one iteration
// are traced beside the next iteration's extension rays, and its finished slots are taken by the
// regen of the iteration after that, once their shadow rays have added their light.
// C_DRAIN_LEFT: paths of the drain not finished yet; C_PARKED: paths the drain handed on to the
// one-wave-per-path stage (when no more than Launch::park_at of them were left).
constexpr u32 C_LIVE0 = 0, C_LIVE1 = 1, C_SHADOW0 = 2, C_DONE0 = 3, C_CURSOR = 4, C_END = 5, C_IDLE = 6, C_TAIL = 7,
              C_SHADOW1 = 8, C_DONE1 = 9, C_DRAIN_LEFT = 10, C_PARKED = 11, C_CONT = 12;
constexpr u32 PARK_MAX = 1u << 16;
constexpr u32 N_COUNTS = 16;
// Stack entries a closest-hit walk stopped by its trip budget keeps (a deeper one goes on).
constexpr int CONT_STACK = 32;
__host__ __device__ inline u32 c_shadow(u32 parity) { return parity ? C_SHADOW1 : C_SHADOW0; }
__host__ __device__ inline u32 c_done(u32 parity) { return parity ? C_DONE1 : C_DONE0; }

// One entry per path slot.
struct PathState {
  float4* ray_o;   // origin
  float4* ray_d;   // direction
  uint4* hit;      // instance, mesh triangle, t bits (written by extend), how a walk ended
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
  uint4* ident;    // results slot of the path's work id (NO_WORK: empty slot), scene, x | y << 16, sample index
  u32* live[2];    // live queues (ping-pong by iteration)
  u32* shadow_queue[2];  // shadow rays made by the shade of an iteration of that parity
  u32* done_queue[2];    // slots finished by the shade of an iteration of that parity
  u32* parked;           // PARK_MAX slots the drain hands to one wave each (scenes with media)
  // Closest-hit walks stopped by their trip budget (C_CONT of them, at most cont_cap): per walk two
  // uint4 (slot, next entry, best instance, best triangle; best t bits, stack height, instance
  // walk's height, instance | cull << 30 | object space << 31) and 32 stack entries.
  uint4* cont;
  u32* cont_stack;
  u32 cont_cap;
  u32* counts;     // see C_*
  u64* rays;       // [0] closest-hit rays traced, [1] shadow rays counted without a trace, [2] shadow rays traced,
                   // [3] thin-wall probes after a medium's exit (counted, never needed: no such walls)
  float4* results; // per work id of the current chunk: radiance rgb, alpha
  Bsdf* bsdf_tmp;  // per slot: the BSDF of a textured hit (allocated for scenes with a texture)
  // Per slot, a walk in progress (allocated for scenes with media): its throughput weight and the
  // density of its last phase sample; its spectral-MIS ratio and (events | steps << 16) bits.
  float4* walk_w;
  float4* walk_r;
  const u32* sobol;  // 64 x 32 generator columns in logical-dimension order
  // Adjoint (path replay) state per slot: d loss / d radiance of the path, the forward's radiance of
  // the path, the pending d log f / d parameter (5 per slot) of the last BSDF sample, its target
  // (material, texel x0 | y0 << 16, tap fractions), and the next-event gradient waiting for the
  // shadow ray (5 scalars in 2 float4).
  float4* adj;
  float4* ltot;
  float4* dl;
  uint4* tgt;
  float4* gn;
  // Gradient sums: per material 5 parameters; per texel of the base-colour texture rgb.
  float* grad_mat;
  float* grad_tex;
  u32 tex_w, tex_h;
};

// Launch shape of the traversal kernels (chosen per device by the tuner).
struct TraversalShape {
  u32 block;       // lanes per block
  u32 lds_stack;   // stack entries per lane held in LDS (the rest of 64 in private memory)
  u32 min_waves;   // waves per SIMD the register allocation must allow (1: no constraint)
  u32 resident;    // 1: resident lanes (each lane walks ray after ray; grid = what the GPU holds)
};
constexpr int N_SHAPES = 8;
extern const TraversalShape SHAPES[N_SHAPES];

struct Launch {
  const DevScene* scenes;
  const DevSegment* segs;
  u32 n_segs;
  u32 pool;   // path slots in use (grid bound of the pool-sized kernels)
  int shape;         // closest-hit launch shape (index into SHAPES)
  int shadow_shape;  // any-hit launch shape
  u32 tail;          // once no new path is left, at most this many live paths are drained in one launch (0: never)
  u32 ext;           // a scene has sphere lights or scattering media: the extended shade and the walk kernel
  u32 cau;           // with ext: the scene has caustic photon maps (the shade gathers)
  u32 drain_lanes;   // with ext: lanes per wave of the drain (1..64)
  u32 park_at;       // with ext: paths left when the drain hands them to one wave each (0..PARK_MAX)
  u32 budget;        // closest-hit loop trips before a walk is handed to the resume launch (0: one launch)
  hipStream_t stream;
  hipStream_t side;  // the shadow rays' stream (default priority), forked from and joined to `stream`
};

hipError_t launch_pool_init(const PathState& P, const Launch& L, u32 work_end);
// `adj`: the adjoint (path replay) variants of the kernels.
hipError_t launch_regen(const PathState& P, const Launch& L, u32 parity, bool adj = false);
hipError_t launch_extend(const PathState& P, const Launch& L, u32 parity);
hipError_t launch_shade(const PathState& P, const Launch& L, u32 parity, bool adj = false);
// Traces the shadow rays the shade of an iteration of parity `parity` made, on L.side.
hipError_t launch_shadow(const PathState& P, const Launch& L, u32 parity, bool adj = false);
// One step of the walks of the live paths inside scattering media, at the closest hit extend found:
// a scattering event or a reflection back in gives the next step's ray (traced by the next
// extend); otherwise the walk ends (hit.w says how) and the shade takes the vertex.
hipError_t launch_walk(const PathState& P, const Launch& L, u32 parity);
// The drain: when the last bookkeeping handed paths over (C_TAIL), takes each to its end.
hipError_t launch_finish(const PathState& P, const Launch& L, u32 parity);
// Adds the chunk's samples of segment `seg` (host copy `s`) to its film, in sample order.
hipError_t launch_splat(const PathState& P, const DevSegment& s, hipStream_t st);
// Splats every segment of a device table in one launch; the segments' pixels must not overlap.
hipError_t launch_splat_segments(const PathState& P, const DevSegment* segs, u32 n_segs, hipStream_t st);
// Samples kept from an earlier run: `seg` names the pixels, the film, and how many samples each
// pixel gets (seg.count); pixel p's (tiled order) are res[p * stride + skip + k], k < seg.count.
struct KeptSplat {
  DevSegment seg;
  const float4* res;
  u32 stride, skip;
};
// Splats a device table of kept samples in one launch (pixels disjoint across entries).
hipError_t launch_splat_kept(const KeptSplat* kept, u32 n, hipStream_t st);
// Copies the one-sample results of a w x h tile (work ids 0 .., pixels in the regen's tiled order)
// out as row-major rgb or rgba floats.
hipError_t launch_tile_out(const PathState& P, float* out, u32 w, u32 h, u32 channels, hipStream_t st);
// Nearest and any hit of n rays; with wave_cap > 0, one wave per ray (the drain's traversal) with a
// shared stack of wave_cap entries (at most 1024) before it falls back to one lane.
hipError_t launch_trace_rays(const DevScene* scene, const float* rays, u32 n, u32* inst, u32* tri, float* t,
                             uint8_t* any, hipStream_t st, int wave_cap = 0);
// Planes of a pass film in device memory (rt-core/src/passes.rs): sums of the sampled passes (albedo
// rgb, coverage) with their sample count, and the exact passes of the pixels' centre rays. Planes
// a launch does not write may be null.
struct PassPlanes {
  float* albedo;  // 3 per pixel, sums
  float* alpha;   // coverage sums
  float* weight;  // samples summed
  float* normal;  // 3 per pixel
  float* depth;
  u32* part;
  u32* material;
  float* coc;
};
constexpr u32 PASSES_EXACT = 1, PASSES_SAMPLED = 2;
constexpr u32 FLOOR_ID = 0xffffffffu;  // part and material id of the studio floor
// The passes of the pixels of rectangle (x0, y0, w, h) inside the image: `what` PASSES_EXACT
// writes the exact planes, PASSES_SAMPLED adds samples [first, first + count) to the sums. Pixel
// (px, py) is at row py - oy, column px - ox of planes `ow` pixels wide.
hipError_t launch_passes(const DevScene* scene, const u32* sobol, const PassPlanes& Q, u32 x0, u32 y0, u32 w, u32 h,
                         u32 ox, u32 oy, u32 ow, u32 first, u32 count, u32 what, hipStream_t st);
// The resolved albedo (3 per pixel) and alpha of n pixels of the pass film (either may be null).
hipError_t launch_passes_resolve(const PassPlanes& Q, u32 n, float* albedo, float* alpha, hipStream_t st);
hipError_t launch_energy(uint16_t* albedo, uint16_t* schlick, uint16_t* clear_in, uint16_t* clear_out, float* albedo_avg,
                         float* schlick_avg, hipStream_t st);

// The photons of one pass by photon index, before the map is built: where each was stored
// (xyz, casters crossed in the bits of w), the direction it came from, its power (W), the normal
// of the surface it was stored on, and its key (PHOTON_STORED or PHOTON_NONE; the map builder
// turns a stored one's key into its bucket). `stored` counts the stored photons.
struct PhotonBuffers {
  float4 *position, *wi, *power, *normal;
  u32* key;
  u32* stored;
};
constexpr u32 PHOTON_STORED = 0, PHOTON_NONE = 1u << 24;  // above every bucket (at most 2^24)
// The photons in flight, by photon index: ray origin (w: the roulette's throughput), direction
// (w: casters crossed), power (w: inside an absorbing solid), that solid's coefficients, the
// state of the photon's PCG32 stream, its closest hit; two queues of photon indices (a step reads
// one and appends the survivors to the other) and their counts.
struct PhotonPaths {
  float4 *o, *d, *power, *sigma;
  u64* rng;
  uint4* hit;
  u32* live[2];
  u32* counts;  // 2
};
// Traces photons [0, photons) of pass `iteration` (caustics.rs:581-735) for at most `max_steps`
// closest hits each, storing them in B; `host_count` (pinned) receives the queue counts the
// launcher checks between steps.
hipError_t launch_photons(const DevScene* scene, u32 photons, u32 max_steps, const PhotonBuffers& B,
                          const PhotonPaths& Q, u32 iteration, u32* host_count, hipStream_t st);

// The Sobol generator matrices (Joe–Kuo new-joe-kuo-6.21201, 64 dimensions) in the reference
// sampler's logical-dimension layout: column k of logical dimension d at [d * 32 + k].
void sobol_columns(u32 out[64 * 32]);

}  // namespace rpx
