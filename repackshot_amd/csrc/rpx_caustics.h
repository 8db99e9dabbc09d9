// The photon maps of a scene with caustics on the GPU (caustics.rs:737-960): each pass traces
// the scene's photons, then builds the hashed grid the shade kernel gathers from. The current
// map's header (DevPhotonMap) sits in device memory at a fixed address for the scene's lifetime
// (a copy of a kept map's own header), so the scene record handed to the kernels never changes
// between passes.
#pragma once
#include <string>
#include <vector>
#include "rpx_kernels.h"

namespace rpx {

// The maps of the last passes are kept (at most KEPT_PHOTON_MAPS, the oldest dropped first,
// rt-wasm/src/lib.rs:58,103-121), so tiles of a pass rendered one by one trace its photons once.
constexpr u32 KEPT_PHOTON_MAPS = 16;

class PhotonPasses {
 public:
  PhotonPasses() = default;
  PhotonPasses(const PhotonPasses&) = delete;
  PhotonPasses& operator=(const PhotonPasses&) = delete;
  ~PhotonPasses();
  // Device memory for tracing `photons` photons per pass; the header starts as an empty map.
  bool init(u32 photons, float first_radius, hipStream_t st, std::string& err);
  // Makes the map of pass `iteration` the current one (header()): a kept map is reused, else the
  // scene's photons (`d_scene`: its device record, whose pmap is header()) are traced and the map
  // built, all on `st`. Without `may_allocate` a new map takes the place of the oldest kept one
  // rather than new device memory.
  hipError_t build(const DevScene* d_scene, u32 iteration, hipStream_t st, std::string& err,
                   bool may_allocate = true);
  // Drops every kept map (the scene's materials or lights changed).
  void invalidate();
  const DevPhotonMap* header() const { return header_; }
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

  u32 photons_ = 0, buckets_cap_ = 0, traced_ = 0;
  u64 clock_ = 0;
  float first_radius_ = 1.0f;
  DevPhotonMap* header_ = nullptr;  // the current map's header, read by the shade
  PhotonBuffers raw_{};
  u32 *keys_sorted_ = nullptr, *index_ = nullptr, *index_sorted_ = nullptr;
  void* sort_tmp_ = nullptr;
  size_t sort_bytes_ = 0;
  void* block_ = nullptr;  // one allocation holding the trace and sort buffers above
  std::vector<Kept> kept_;
};

}  // namespace rpx
