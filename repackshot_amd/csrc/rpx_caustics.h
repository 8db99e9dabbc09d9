// The photon maps of a scene with caustics on the GPU (caustics.rs:737-960): each pass traces
// the scene's photons, then builds the hashed grid the shade kernel gathers from. The map's
// header (DevPhotonMap) sits in device memory at a fixed address for the scene's lifetime, so
// the scene record handed to the kernels never changes between passes.
#pragma once
#include <string>
#include "rpx_kernels.h"

namespace rpx {

class PhotonPasses {
 public:
  PhotonPasses() = default;
  PhotonPasses(const PhotonPasses&) = delete;
  PhotonPasses& operator=(const PhotonPasses&) = delete;
  ~PhotonPasses();
  // Device memory for `photons` photons per pass; the header starts as an empty map.
  bool init(u32 photons, float first_radius, hipStream_t st, std::string& err);
  // Traces pass `iteration` of the scene (`d_scene`: its device record, whose pmap is header())
  // and builds its map, all on `st`.
  hipError_t build(const DevScene* d_scene, u32 iteration, hipStream_t st, std::string& err);
  const DevPhotonMap* header() const { return header_; }
  u32 photons() const { return photons_; }
  // The radius of pass `iteration` (caustics.rs:210-226, Knaus and Zwicker): r_k^2 = r_0^2
  // prod_{j=1..k} (j + alpha) / (j + 1), alpha = 2/3.
  static float radius_of(float first, u32 iteration);

 private:
  u32 photons_ = 0, buckets_cap_ = 0;
  float first_radius_ = 1.0f;
  DevPhotonMap* header_ = nullptr;
  PhotonBuffers raw_{};
  float4 *pos_ = nullptr, *wi_ = nullptr, *power_ = nullptr, *normal_ = nullptr;  // bucket order
  u32 *keys_sorted_ = nullptr, *index_ = nullptr, *index_sorted_ = nullptr, *starts_ = nullptr;
  void* sort_tmp_ = nullptr;
  size_t sort_bytes_ = 0;
  void* block_ = nullptr;  // one allocation holding every array above
};

}  // namespace rpx
