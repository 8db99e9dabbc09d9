// The photon maps of a scene with caustics on the GPU (caustics.rs:737-960): each sample pass
// traces the scene's photons, then builds the hashed grid the shade kernel gathers from. A render
// runs up to MAX_PASS_MAPS passes at once, each path gathering from its own pass's map: the table
// of current maps (DevPhotonMaps) sits in device memory at a fixed address for the scene's
// lifetime, so the scene record handed to the kernels never changes between chunks.
#pragma once
#include <string>
#include <vector>
#include "rpx_kernels.h"

namespace rpx {

// The maps of the last passes are kept (at most KEPT_PHOTON_MAPS, the oldest dropped first,
// rt-wasm/src/lib.rs:58,103-121), so tiles of a pass rendered one by one trace its photons once.
constexpr u32 KEPT_PHOTON_MAPS = 16;
static_assert(KEPT_PHOTON_MAPS >= MAX_PASS_MAPS, "a chunk's maps are kept together");

class PhotonPasses {
 public:
  PhotonPasses() = default;
  PhotonPasses(const PhotonPasses&) = delete;
  PhotonPasses& operator=(const PhotonPasses&) = delete;
  ~PhotonPasses();
  // Device memory for tracing `photons` photons per pass; the header starts as an empty map.
  bool init(u32 photons, u32 max_bounces, float first_radius, hipStream_t st, std::string& err);
  // Makes the maps of passes iterations[0 .. n) current (n <= MAX_PASS_MAPS): sample pass
  // first + i gathers from the map of iterations[i]. Kept maps are reused, the others traced and
  // built from the scene (`d_scene`: its device record, whose pmaps is table()), all on `st`.
  // Without `may_allocate` a new map takes the place of the oldest kept one that this call does
  // not use rather than new device memory.
  hipError_t prepare(const DevScene* d_scene, u32 first, const u32* iterations, u32 n, hipStream_t st,
                     std::string& err, bool may_allocate = true);
  // The kept map of pass `iteration`, or -1.
  int find(u32 iteration) const;
  // Drops every kept map (the scene's materials or lights changed).
  void invalidate();
  const DevPhotonMaps* table() const { return table_; }
  u32 photons() const { return photons_; }
  u32 traced() const { return traced_; }  // maps traced since init
  // The radius of pass `iteration` (caustics.rs:210-226, Knaus and Zwicker): r_k^2 = r_0^2
  // prod_{j=1..k} (j + alpha) / (j + 1), alpha = 2/3.
  static float radius_of(float first, u32 iteration);

 private:
  struct Kept {
    u32 iteration = 0;
    bool valid = false;
    u64 age = 0;  // order of tracing
    DevPhotonMap* header = nullptr;
    float4 *pos = nullptr, *wi = nullptr, *power = nullptr, *normal = nullptr;  // bucket order
    u32* starts = nullptr;
    void* block = nullptr;
  };
  bool add_slot(hipStream_t st, std::string& err);
  // Traces pass `iteration` into a kept slot outside `avoid` (a bit per slot): a free one, a new
  // one when allowed, else the oldest.
  hipError_t trace(const DevScene* d_scene, u32 iteration, hipStream_t st, u32 avoid, int& slot, std::string& err,
                   bool may_allocate);

  u32 photons_ = 0, buckets_cap_ = 0, traced_ = 0;
  u64 clock_ = 0;
  float first_radius_ = 1.0f;
  DevPhotonMaps* table_ = nullptr;  // the current maps, read by the shade
  PhotonBuffers raw_{};
  PhotonPaths paths_{};
  u32 max_steps_ = 0;
  u32* host_count_ = nullptr;  // pinned
  u32 *keys_sorted_ = nullptr, *index_ = nullptr, *index_sorted_ = nullptr;
  void* sort_tmp_ = nullptr;
  size_t sort_bytes_ = 0;
  void* block_ = nullptr;  // one allocation holding the trace and sort buffers above
  std::vector<Kept> kept_;
};

}  // namespace rpx
