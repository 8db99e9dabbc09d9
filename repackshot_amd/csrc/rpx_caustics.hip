// The photon map of a pass on the GPU (caustics.rs:737-860): the traced photons get the bucket of
// their grid cell, a stable radix sort by bucket puts them in bucket order with trace (index)
// order kept inside a bucket, and each bucket's start is found by bisection of the sorted keys.
#include "rpx_caustics.h"

#include <hipcub/hipcub.hpp>
#include <cmath>

namespace rpx {
namespace {

constexpr u32 MAX_BUCKETS = 1u << 24;
constexpr u32 KEY_BITS = 25;  // buckets below 2^24 and PHOTON_NONE = 2^24
constexpr float MM_TO_M = 1.0e-3f;
constexpr float RADIUS_ALPHA = 2.0f / 3.0f;

__global__ void __launch_bounds__(256) k_map_iota(u32* index, u32 n) {
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) index[i] = i;
}

// caustics.rs:807-818: a table of a power of two of at least twice the stored photons' buckets
// (at most 2^24), and the pass's constants.
__global__ void k_map_header(DevPhotonMap* M, const u32* stored, u32 iteration, float radius, float inv_r2, float norm,
                             float inv_cell) {
  u32 n = *stored;
  u32 want = (n > 1u ? n : 1u) * 2u;
  u32 buckets = 1;
  while (buckets < want && buckets < MAX_BUCKETS) buckets <<= 1;
  M->count = n;
  M->mask = buckets - 1;
  M->iteration = iteration;
  M->radius = radius;
  M->inv_r2 = inv_r2;
  M->norm = norm;
  M->inv_cell = inv_cell;
}

// Each stored photon's bucket (caustics.rs:820-830).
__global__ void __launch_bounds__(256) k_map_keys(PhotonBuffers B, const DevPhotonMap* M, u32 n) {
  const u32 mask = M->mask;
  const float inv_cell = M->inv_cell;
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (B.key[i] != PHOTON_STORED) continue;
    float4 p = B.position[i];
    i32 c[3];
    photon_cell(v3(p.x, p.y, p.z), inv_cell, c);
    B.key[i] = photon_bucket(c, mask);
  }
}

// The stored photons' records in bucket order.
__global__ void __launch_bounds__(256) k_map_permute(const DevPhotonMap* M, const u32* order, PhotonBuffers B,
                                                    float4* pos, float4* wi, float4* power, float4* normal) {
  const u32 n = M->count;
  for (u32 k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    u32 s = order[k];
    pos[k] = B.position[s];
    wi[k] = B.wi[s];
    power[k] = B.power[s];
    normal[k] = B.normal[s];
  }
}

// starts[b] = the first sorted photon whose bucket is at least b, for b in [0, mask + 1] (the last
// is the stored count: every unstored key is above every bucket).
__global__ void __launch_bounds__(256) k_map_starts(const DevPhotonMap* M, const u32* keys, u32 n, u32* starts,
                                                   u32 cap) {
  const u32 last = M->mask + 1;
  for (u32 b = blockIdx.x * blockDim.x + threadIdx.x; b <= last && b <= cap; b += gridDim.x * blockDim.x) {
    u32 lo = 0, hi = n;
    while (lo < hi) {
      u32 mid = (lo + hi) >> 1;
      if (keys[mid] < b) lo = mid + 1;
      else hi = mid;
    }
    starts[b] = lo;
  }
}

// The table of current maps: pass first + i reads a copy of kept map heads.p[i]'s header.
struct MapHeads {
  const DevPhotonMap* p[MAX_PASS_MAPS];
};
__global__ void k_map_table(DevPhotonMaps* T, MapHeads heads, u32 first, u32 n) {
  u32 i = threadIdx.x;
  if (i == 0) {
    T->first = first;
    T->count = n;
  }
  if (i < n) T->map[i] = *heads.p[i];
}

inline u32 grid_of(u32 n, u32 block, u32 cap) {
  u32 g = (n + block - 1) / block;
  return g < 1 ? 1 : (g > cap ? cap : g);
}

inline size_t round_up(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

float PhotonPasses::radius_of(float first, u32 iteration) {
  float r2 = first * first;
  for (u32 k = 1; k <= iteration; ++k) r2 *= ((float)k + RADIUS_ALPHA) / ((float)k + 1.0f);
  return std::sqrt(r2);
}

PhotonPasses::~PhotonPasses() {
  for (Kept& k : kept_)
    if (k.block) (void)hipFree(k.block);
  if (block_) (void)hipFree(block_);
  if (host_count_) (void)hipHostFree(host_count_);
}

bool PhotonPasses::init(u32 photons, u32 max_bounces, float first_radius, hipStream_t st, std::string& err) {
  if (photons == 0) {
    err = "a scene with caustics traces at least one photon per pass";
    return false;
  }
  photons_ = photons;
  first_radius_ = first_radius;
  max_steps_ = max_bounces + 128;  // caustics.rs:609: the most closest hits a photon takes
  if (hipHostMalloc((void**)&host_count_, 64, hipHostMallocDefault) != hipSuccess) {
    host_count_ = nullptr;
    err = "photon map: cannot allocate the pinned step counter";
    return false;
  }
  u32 buckets = 1;
  while (buckets < 2u * photons && buckets < MAX_BUCKETS) buckets <<= 1;
  buckets_cap_ = buckets;
  size_t n = photons;
  sort_bytes_ = 0;
  if (hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes_, (u32*)nullptr, (u32*)nullptr, (u32*)nullptr,
                                         (u32*)nullptr, (int)photons, 0, (int)KEY_BITS, st) != hipSuccess) {
    err = "photon map: the sort's workspace query failed";
    return false;
  }
  // One block: the table of current maps, the traced records (4 x float4) and keys, the sorted
  // keys, the index, the sorted index, the sort's workspace, the stored counter, then the photons
  // in flight (4 x float4, the PCG32 states, the hits, two queues, their counts).
  size_t sizes[] = {round_up(sizeof(DevPhotonMaps)), round_up(16 * n), round_up(16 * n), round_up(16 * n),
                    round_up(16 * n), round_up(4 * n), round_up(4 * n), round_up(4 * n), round_up(4 * n),
                    round_up(sort_bytes_ ? sort_bytes_ : 1), round_up(4),
                    round_up(16 * n), round_up(16 * n), round_up(16 * n), round_up(16 * n), round_up(8 * n),
                    round_up(16 * n), round_up(4 * n), round_up(4 * n), round_up(8)};
  size_t total = 0;
  for (size_t b : sizes) total += b;
  if (hipMalloc(&block_, total) != hipSuccess) {
    block_ = nullptr;
    err = "photon map: out of device memory";
    return false;
  }
  char* p = (char*)block_;
  auto take = [&](int k) {
    char* q = p;
    p += sizes[k];
    return (void*)q;
  };
  table_ = (DevPhotonMaps*)take(0);
  raw_.position = (float4*)take(1);
  raw_.wi = (float4*)take(2);
  raw_.power = (float4*)take(3);
  raw_.normal = (float4*)take(4);
  raw_.key = (u32*)take(5);
  keys_sorted_ = (u32*)take(6);
  index_ = (u32*)take(7);
  index_sorted_ = (u32*)take(8);
  sort_tmp_ = take(9);
  raw_.stored = (u32*)take(10);
  paths_.o = (float4*)take(11);
  paths_.d = (float4*)take(12);
  paths_.power = (float4*)take(13);
  paths_.sigma = (float4*)take(14);
  paths_.rng = (u64*)take(15);
  paths_.hit = (uint4*)take(16);
  paths_.live[0] = (u32*)take(17);
  paths_.live[1] = (u32*)take(18);
  paths_.counts = (u32*)take(19);
  if (!add_slot(st, err)) return false;
  // Until the first render, no pass has a map.
  DevPhotonMaps none{};
  if (hipMemcpyAsync(table_, &none, sizeof(none), hipMemcpyHostToDevice, st) != hipSuccess) {
    err = "photon map: upload failed";
    return false;
  }
  hipLaunchKernelGGL(k_map_iota, dim3(grid_of(photons, 256, 4096)), dim3(256), 0, st, index_, photons);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
    err = "photon map: initialisation failed";
    return false;
  }
  return true;
}

// A kept map's own memory: its header, its records in bucket order and its bucket starts, the
// header holding an empty map (two empty buckets) until the slot is built.
bool PhotonPasses::add_slot(hipStream_t st, std::string& err) {
  size_t n = photons_;
  size_t sizes[] = {round_up(sizeof(DevPhotonMap)), round_up(16 * n), round_up(16 * n), round_up(16 * n),
                    round_up(16 * n), round_up(4 * ((size_t)buckets_cap_ + 1))};
  size_t total = 0;
  for (size_t b : sizes) total += b;
  Kept k;
  if (hipMalloc(&k.block, total) != hipSuccess) {
    err = "photon map: out of device memory";
    return false;
  }
  char* p = (char*)k.block;
  auto take = [&](int i) {
    char* q = p;
    p += sizes[i];
    return (void*)q;
  };
  k.header = (DevPhotonMap*)take(0);
  k.pos = (float4*)take(1);
  k.wi = (float4*)take(2);
  k.power = (float4*)take(3);
  k.normal = (float4*)take(4);
  k.starts = (u32*)take(5);
  DevPhotonMap h{};
  h.count = 0;
  h.mask = 1;
  h.radius = first_radius_;
  h.inv_r2 = 1.0f / (first_radius_ * first_radius_);
  h.inv_cell = 1.0f / (2.0f * first_radius_);
  float r_m = first_radius_ * MM_TO_M;
  h.norm = 2.0f / (PI * r_m * r_m);
  h.position = k.pos;
  h.wi = k.wi;
  h.power = k.power;
  h.normal = k.normal;
  h.starts = k.starts;
  u32 zeros[3] = {0, 0, 0};
  if (hipMemcpyAsync(k.header, &h, sizeof(h), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(k.starts, zeros, sizeof(zeros), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    (void)hipFree(k.block);
    err = "photon map: upload failed";
    return false;
  }
  kept_.push_back(k);
  return true;
}

void PhotonPasses::invalidate() {
  for (Kept& k : kept_) k.valid = false;
}

int PhotonPasses::find(u32 iteration) const {
  for (size_t k = 0; k < kept_.size(); ++k)
    if (kept_[k].valid && kept_[k].iteration == iteration) return (int)k;
  return -1;
}

hipError_t PhotonPasses::trace(const DevScene* d_scene, u32 iteration, hipStream_t st, u32 avoid, int& slot_out,
                               std::string& err, bool may_allocate) {
  int slot = -1;
  for (size_t k = 0; k < kept_.size() && slot < 0; ++k)
    if (!kept_[k].valid && !((avoid >> k) & 1u)) slot = (int)k;
  if (slot < 0 && may_allocate && kept_.size() < KEPT_PHOTON_MAPS) {
    std::string unused;
    if (add_slot(st, unused)) slot = (int)kept_.size() - 1;
  }
  if (slot < 0)
    for (size_t k = 0; k < kept_.size(); ++k)
      if (!((avoid >> k) & 1u) && (slot < 0 || kept_[k].age < kept_[slot].age)) slot = (int)k;
  if (slot < 0) {
    err = "photon map: no kept map to trace into";
    return hipErrorInvalidValue;
  }
  Kept& K = kept_[slot];
  K.valid = false;
  float radius = radius_of(first_radius_, iteration);
  float r_m = radius * MM_TO_M;
  float inv_cell = 1.0f / (2.0f * radius);
  float inv_r2 = 1.0f / (radius * radius);
  float norm = 2.0f / (PI * r_m * r_m);
  hipError_t e = launch_photons(d_scene, photons_, max_steps_, raw_, paths_, iteration, host_count_, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_map_header, dim3(1), dim3(1), 0, st, K.header, raw_.stored, iteration, radius, inv_r2, norm,
                       inv_cell);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_map_keys, dim3(grid_of(photons_, 256, 8192)), dim3(256), 0, st, raw_, K.header, photons_);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    size_t bytes = sort_bytes_;
    e = hipcub::DeviceRadixSort::SortPairs(sort_tmp_, bytes, raw_.key, keys_sorted_, index_, index_sorted_,
                                           (int)photons_, 0, (int)KEY_BITS, st);
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_map_permute, dim3(grid_of(photons_, 256, 8192)), dim3(256), 0, st, K.header, index_sorted_,
                       raw_, K.pos, K.wi, K.power, K.normal);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_map_starts, dim3(grid_of(buckets_cap_ + 1, 256, 8192)), dim3(256), 0, st, K.header,
                       keys_sorted_, photons_, K.starts, buckets_cap_);
    e = hipGetLastError();
  }
  if (e != hipSuccess) {
    err = std::string("photon map: ") + hipGetErrorString(e);
    return e;
  }
  K.valid = true;
  K.iteration = iteration;
  K.age = ++clock_;
  ++traced_;
  slot_out = slot;
  return hipSuccess;
}

hipError_t PhotonPasses::prepare(const DevScene* d_scene, u32 first, const u32* iterations, u32 n, hipStream_t st,
                                  std::string& err, bool may_allocate) {
  if (n == 0 || n > MAX_PASS_MAPS) {
    err = "photon map: a render runs 1 to 16 sample passes with their maps";
    return hipErrorInvalidValue;
  }
  MapHeads heads{};
  u32 used = 0;  // kept maps this call reads
  for (u32 i = 0; i < n; ++i) {
    int slot = find(iterations[i]);
    if (slot < 0) {
      hipError_t e = trace(d_scene, iterations[i], st, used, slot, err, may_allocate);
      if (e != hipSuccess) return e;
    }
    used |= 1u << slot;
    heads.p[i] = kept_[slot].header;
  }
  hipLaunchKernelGGL(k_map_table, dim3(1), dim3(64), 0, st, table_, heads, first, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) err = std::string("photon map: ") + hipGetErrorString(e);
  return e;
}

}  // namespace rpx
