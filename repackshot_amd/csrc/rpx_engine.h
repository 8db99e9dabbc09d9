// Host side of the GPU tracer's path pool: device tables, the pool's buffers, and the loop that
// runs a list of segments (pixel rectangles x sample ranges of one or more scenes) to completion.
#pragma once
#include <string>
#include <vector>
#include "rpx_kernels.h"

namespace rpx {

class PhotonPasses;

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};
bool dev_alloc(std::vector<DevBuf>& bufs, size_t bytes, void** out);
void dev_free_all(std::vector<DevBuf>& bufs);

// Constant tables of one device: the BSDF energy tables (generated there by the recipe kernel on
// first use; the host copy prepares the BSDFs) and the Sobol generator columns.
struct DeviceTables {
  bool ready = false;
  uint16_t *albedo = nullptr, *schlick = nullptr, *clear_in = nullptr, *clear_out = nullptr;
  float *albedo_avg = nullptr, *schlick_avg = nullptr;
  u32* sobol = nullptr;
  std::vector<uint16_t> h_albedo, h_schlick, h_clear_in, h_clear_out;
  std::vector<float> h_albedo_avg, h_schlick_avg;
  EnergyTables host_view() const {
    return EnergyTables{h_albedo.data(), h_schlick.data(), h_clear_in.data(), h_clear_out.data(),
                        h_albedo_avg.data(), h_schlick_avg.data()};
  }
  EnergyTables device_view() const { return EnergyTables{albedo, schlick, clear_in, clear_out, albedo_avg, schlick_avg}; }
};
bool device_tables(int dev, DeviceTables*& out, std::string& err);

// How the pool runs.
struct EngineOptions {
  u32 pool = 1u << 21;  // path slots (capped by the work of a chunk)
  int shape = 0;        // closest-hit launch shape (index into SHAPES)
  int shadow_shape = 0;  // any-hit (shadow) launch shape
  u32 group = 8;        // iterations launched between checks for completion (even)
  u32 chunk = 1u << 27;  // most work ids (path samples) per chunk, at most RESULTS_MAX
  bool per_pass = false;  // one sample pass at a time, the pool drained after each (no regeneration across samples)
  bool graph = false;   // launch each group of iterations as one captured HIP graph
  // Once no new path is left, at most this many live paths finish in one drain launch (0: the
  // iteration loop runs to the end).
  u32 tail = 1u << 16;
  // Scenes with walks: lanes per wave the drain gives paths to, and how many paths may be left
  // when the drain hands them on to one wave each (at most PARK_MAX). On the bottle workload
  // (1920 x 1080 x 64 paths, tail 2^16) the drain took 72 ms with no hand-on, 50 ms from 1,024
  // paths left, 43 ms from 16,384 and 59 ms from all of them; fewer lanes per wave were slower.
  u32 drain_lanes = 64;
  u32 park_at = 16384;
  // Closest-hit loop trips before a walk is set aside for a second launch (0: one launch).
  u32 trip_budget = 48;
  // Adaptive rounds: a block's samples up to the cap are rendered in one run once it is a
  // candidate, and kept for the rounds (false: each round renders just its own samples).
  bool adaptive_ahead = true;
};

// GPU time of the last run, from events on the engine's stream (profiling on).
struct Timing {
  float total_ms = 0;
  // extend (the closest-hit kernel until the shadow kernel beside it has joined), shade, shadow,
  // regen, drain, walk, the closest-hit kernel alone, and the photon maps (tracing and building;
  // launches: maps made current)
  float stage_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  u32 launches[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  u32 iterations = 0, chunks = 0;
};

// Largest number of work ids (path samples) in one chunk: the results buffer holds one float4 each.
constexpr u32 RESULTS_MAX = 1u << 27;
constexpr int N_TUNE_POOLS = 6;
constexpr int N_TUNE_CANDIDATES = N_SHAPES * N_TUNE_POOLS + N_SHAPES;
constexpr u32 TUNE_POOLS[N_TUNE_POOLS] = {1u << 21, 1u << 22, 1u << 23, 1u << 24, 1u << 25, 1u << 26};
// Closest-hit trip budgets the tuner tries last, at the chosen launch shapes.
constexpr int N_TUNE_BUDGETS = 2;
constexpr u32 TUNE_BUDGETS[N_TUNE_BUDGETS] = {0u, 48u};

class Engine {
 public:
  Engine() = default;
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;
  ~Engine();
  bool init(int device, hipStream_t stream, std::string& err);
  // Renders every segment. With `splat`, each segment's samples are added to its film in sample
  // order; without, the results of a single chunk stay in `P.results` (work id order). `disjoint`
  // promises that no two segments share a film pixel: their splats then run in one launch.
  hipError_t run(const DevScene* d_scenes, const std::vector<DevSegment>& segs, bool splat, std::string& err,
                 bool disjoint = false);
  // Path replay backpropagation: for each chunk, the forward pass (radiance per path sample into
  // the results buffer), then the replay that accumulates d loss / d parameter into
  // grad_mat (n_mat x 5) and grad_tex (tex_w x tex_h x 3), device buffers the caller zeroes. The
  // segments carry the adjoint image and its scale.
  hipError_t adjoint(const DevScene* d_scenes, const std::vector<DevSegment>& segs, float* grad_mat, float* grad_tex,
                     u32 tex_w, u32 tex_h, std::string& err);
  // The tuner: renders `probe` (whose results are discarded; one chunk) with every closest-hit
  // launch shape at each of TUNE_POOLS pool sizes, then every shadow launch shape at the best of
  // those, then each of TUNE_BUDGETS at the best of those, keeps the fastest in `opt`, and returns
  // every candidate's GPU time (ms: shape-major closest-hit candidates, the shadow candidates,
  // then the budgets). `budget_only`: the current shapes and pool are kept and only the trip
  // budget is tuned (ms: the budgets' alone).
  hipError_t autotune(const DevScene* d_scenes, const std::vector<DevSegment>& probe, std::vector<float>& ms,
                      std::string& err, bool budget_only = false);
  // One sample of a tile with every per-path buffer (pool, results, segment table, a staging
  // tile) carved from `scratch`, `bytes` of device memory the caller owns: no allocation. The
  // tile's channels-per-pixel radiance (3: rgb, 4: rgba) is copied to `out` (host memory, rows
  // `out_row_floats` apart). Fails with hipErrorInvalidValue when the tile does not fit.
  hipError_t render_tile_in(const DevScene* d_scenes, const DevSegment& seg, u32 channels, void* scratch, size_t bytes,
                            float* out, size_t out_row_floats, std::string& err);
  // The scratch bytes render_tile_in needs for a tile of `pixels` pixels.
  static size_t tile_scratch_bytes(u32 pixels, bool textured, bool ext, u32 channels);
  // Closest-hit rays, shadow rays counted without a trace, shadow rays traced, thin-wall probes
  // counted after media exits (since init).
  bool ray_counts(u64 out[4]);

  EngineOptions opt;
  // A scene with caustics: the photon map of each sample pass is traced and built on the engine's
  // stream before the chunk that runs the pass; a chunk holds one segment's samples, at most
  // MAX_PASS_MAPS of them. `photon_pass` >= 0 names the map's pass for every sample instead of the
  // sample index.
  PhotonPasses* photons = nullptr;
  long long photon_pass = -1;
  bool profiling = false;
  bool textured = false;  // a scene of the next runs has a texture (slots get a BSDF record)
  bool ext = false;       // a scene of the next runs has sphere lights or scattering media
  Timing timing;
  PathState P{};

 private:
  bool ensure_pool(u32 n, std::string& err);
  void release_pool();  // frees the path pool's buffers (the next run allocates them again)
  bool ensure_adjoint(u32 n, std::string& err);
  bool plan(const std::vector<DevSegment>& segs, std::vector<std::vector<DevSegment>>& chunks, std::vector<u32>& totals,
            std::string& err);
  bool ensure_results(u32 n, std::string& err);
  bool ensure_cont(u32 n, std::string& err);  // continuation slots for a pool of n (trip budget on)
  bool ensure_segs(u32 n, std::string& err);
  // The photon maps of a chunk's sample passes made current (profiled as stage 7).
  hipError_t prepare_photons(const DevScene* d_scenes, const DevSegment& piece, bool may_allocate, std::string& err);
  // `seg_table`: device memory for the chunk's segment table (null: the engine's own).
  hipError_t run_chunk(const DevScene* d_scenes, const DevSegment* chunk, u32 n_chunk, u32 total, bool splat,
                       bool adj = false, bool disjoint = false, DevSegment* seg_table = nullptr);
  hipEvent_t next_event();
  void gather_marks();  // adds the profiled stages' event spans to `timing`

  int device_ = 0;
  hipStream_t stream_ = nullptr;
  u32 capacity_ = 0, results_cap_ = 0, segs_cap_ = 0, adjoint_cap_ = 0, bsdf_cap_ = 0, walk_cap_ = 0;
  std::vector<DevBuf> pool_bufs_, fixed_bufs_, adjoint_bufs_, bsdf_bufs_, walk_bufs_, cont_bufs_;
  u32 cont_cap_ = 0;
  DevBuf results_, segs_;
  u32* idle_host_ = nullptr;  // pinned
  std::vector<hipEvent_t> events_;
  size_t events_used_ = 0;
  std::vector<std::pair<int, size_t>> marks_;
  hipEvent_t ev_[2] = {nullptr, nullptr};
  // The shadow kernel's stream: forked from stream_ after each regen, joined before the shade.
  hipStream_t side_ = nullptr;
  hipEvent_t fork_ = nullptr, join_ = nullptr;

};

}  // namespace rpx
