// The C ABI of the GPU tracer (include/rpx.h): a registry of scene handles, each owning its
// device scene, wavefront path buffers, resident film and HIP stream.
#include <atomic>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include "../../include/rpx.h"
#include "rpx_kernels.h"

namespace rpx {
namespace {

thread_local std::string g_error;
void set_error(const std::string& e) { g_error = e; }

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

// Energy tables of one device: generated there by the recipe kernel on first use; the host copy
// prepares the BSDFs.
struct DeviceTables {
  bool ready = false;
  uint16_t *albedo = nullptr, *schlick = nullptr, *clear_in = nullptr, *clear_out = nullptr;
  float *albedo_avg = nullptr, *schlick_avg = nullptr;
  std::vector<uint16_t> h_albedo, h_schlick, h_clear_in, h_clear_out;
  std::vector<float> h_albedo_avg, h_schlick_avg;
  EnergyTables host_view() const {
    return EnergyTables{h_albedo.data(), h_schlick.data(), h_clear_in.data(), h_clear_out.data(),
                        h_albedo_avg.data(), h_schlick_avg.data()};
  }
  EnergyTables device_view() const { return EnergyTables{albedo, schlick, clear_in, clear_out, albedo_avg, schlick_avg}; }
};
std::mutex g_tables_mutex;
std::map<int, DeviceTables> g_tables;

bool device_tables(int dev, DeviceTables*& out, std::string& err) {
  std::lock_guard<std::mutex> lock(g_tables_mutex);
  DeviceTables& t = g_tables[dev];
  if (!t.ready) {
    const size_t SL = 1024, CL = 17408;
    if (hipMalloc(&t.albedo, SL * 2) != hipSuccess || hipMalloc(&t.schlick, SL * 2) != hipSuccess ||
        hipMalloc(&t.clear_in, CL * 2) != hipSuccess || hipMalloc(&t.clear_out, CL * 2) != hipSuccess ||
        hipMalloc(&t.albedo_avg, 32 * 4) != hipSuccess || hipMalloc(&t.schlick_avg, 32 * 4) != hipSuccess) {
      err = "device memory for the energy tables";
      return false;
    }
    hipError_t e = launch_energy(t.albedo, t.schlick, t.clear_in, t.clear_out, t.albedo_avg, t.schlick_avg, 0);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      err = std::string("energy table kernel: ") + hipGetErrorString(e);
      return false;
    }
    t.h_albedo.resize(SL);
    t.h_schlick.resize(SL);
    t.h_clear_in.resize(CL);
    t.h_clear_out.resize(CL);
    t.h_albedo_avg.resize(32);
    t.h_schlick_avg.resize(32);
    hipMemcpy(t.h_albedo.data(), t.albedo, SL * 2, hipMemcpyDeviceToHost);
    hipMemcpy(t.h_schlick.data(), t.schlick, SL * 2, hipMemcpyDeviceToHost);
    hipMemcpy(t.h_clear_in.data(), t.clear_in, CL * 2, hipMemcpyDeviceToHost);
    hipMemcpy(t.h_clear_out.data(), t.clear_out, CL * 2, hipMemcpyDeviceToHost);
    hipMemcpy(t.h_albedo_avg.data(), t.albedo_avg, 32 * 4, hipMemcpyDeviceToHost);
    e = hipMemcpy(t.h_schlick_avg.data(), t.schlick_avg, 32 * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      err = std::string("energy table copy: ") + hipGetErrorString(e);
      return false;
    }
    t.ready = true;
  }
  out = &t;
  return true;
}

struct Handle {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mutex;
  HostScene host;
  DevScene dev{};
  std::vector<DevBuf> scene_bufs;
  u64 triangles = 0, nodes = 0;
  // wavefront buffers
  u32 capacity = 0;
  PathState paths{};
  std::vector<DevBuf> path_bufs;
  // resident film
  Film film{};
  std::vector<DevBuf> film_bufs;
  u64 samples = 0;
  float last_ms = 0;
  // Per-stage timing of the last render call (profiling on): extend, shade, shadow ms and the
  // number of extend and shadow launches.
  bool profiling = false;
  float stage_ms[3] = {0, 0, 0};
  u32 stage_launches[2] = {0, 0};
  std::vector<hipEvent_t> pool;
  size_t pool_used = 0;
  std::vector<std::pair<int, size_t>> marks;  // (stage, first event index)
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipEvent_t next_event() {
    if (pool_used == pool.size()) {
      hipEvent_t e;
      hipEventCreate(&e);
      pool.push_back(e);
    }
    return pool[pool_used++];
  }

  ~Handle() {
    if (stream) hipStreamSynchronize(stream);
    for (auto* v : {&scene_bufs, &path_bufs, &film_bufs})
      for (DevBuf& b : *v)
        if (b.p) hipFree(b.p);
    for (hipEvent_t e : ev)
      if (e) hipEventDestroy(e);
    for (hipEvent_t e : pool) hipEventDestroy(e);
    if (stream) hipStreamDestroy(stream);
  }
};

std::mutex g_registry_mutex;
std::map<u32, std::shared_ptr<Handle>> g_registry;
u32 g_next_handle = 1;

std::shared_ptr<Handle> find(u32 h) {
  std::lock_guard<std::mutex> lock(g_registry_mutex);
  auto it = g_registry.find(h);
  return it == g_registry.end() ? nullptr : it->second;
}

template <class T>
bool upload(std::vector<DevBuf>& bufs, const std::vector<T>& v, const T*& out, hipStream_t st, std::string& err) {
  size_t bytes = std::max<size_t>(v.size() * sizeof(T), 16);
  DevBuf b;
  if (hipMalloc(&b.p, bytes) != hipSuccess) {
    err = "device memory exhausted while uploading the scene";
    return false;
  }
  b.bytes = bytes;
  bufs.push_back(b);
  if (!v.empty() && hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st) != hipSuccess) {
    err = "scene upload failed";
    return false;
  }
  out = static_cast<const T*>(b.p);
  return true;
}

bool alloc(std::vector<DevBuf>& bufs, size_t bytes, void** out) {
  DevBuf b;
  if (hipMalloc(&b.p, std::max<size_t>(bytes, 16)) != hipSuccess) return false;
  b.bytes = bytes;
  bufs.push_back(b);
  *out = b.p;
  return true;
}

// Builds and uploads the device scene of `h->host`.
bool build_device_scene(Handle* h, std::string& err) {
  DeviceTables* tables;
  if (!device_tables(h->device, tables, err)) return false;
  FlatScene flat;
  unsigned hw = std::thread::hardware_concurrency();
  if (!flatten_scene(h->host, tables->host_view(), flat, err, hw ? (int)std::min(hw, 16u) : 4)) return false;
  for (DevBuf& b : h->scene_bufs)
    if (b.p) hipFree(b.p);
  h->scene_bufs.clear();
  DevScene d = flat.consts;
  auto& B = h->scene_bufs;
  hipStream_t st = h->stream;
  if (!upload(B, flat.nodes, d.nodes, st, err) || !upload(B, flat.tris, d.tris, st, err) ||
      !upload(B, flat.instances, d.instances, st, err) || !upload(B, flat.meshes, d.meshes, st, err) ||
      !upload(B, flat.positions, d.positions, st, err) || !upload(B, flat.normals, d.normals, st, err) ||
      !upload(B, flat.indices, d.indices, st, err) || !upload(B, flat.bsdfs, d.bsdfs, st, err) ||
      !upload(B, flat.materials, d.materials, st, err) || !upload(B, flat.areas, d.areas, st, err) ||
      !upload(B, flat.cdf, d.cdf, st, err) || !upload(B, flat.flux_pmf, d.flux_pmf, st, err) ||
      !upload(B, flat.entries, d.entries, st, err))
    return false;
  d.tables = tables->device_view();
  if (hipStreamSynchronize(st) != hipSuccess) {
    err = "scene upload failed";
    return false;
  }
  h->dev = d;
  h->triangles = flat.triangles;
  h->nodes = flat.nodes.size();
  // Traversal stack bound: at most three entries per 4-wide level (LDS 32 + private 32).
  u32 bound = 3 * (flat.max_tlas_depth + flat.max_blas_depth) + 2;
  if (bound > 64) {
    err = "the BVH is too deep for the traversal stack";
    return false;
  }
  return true;
}

bool ensure_paths(Handle* h, u32 n, std::string& err) {
  if (n <= h->capacity) return true;
  for (DevBuf& b : h->path_bufs)
    if (b.p) hipFree(b.p);
  h->path_bufs.clear();
  h->capacity = 0;
  PathState& P = h->paths;
  auto& B = h->path_bufs;
  size_t f4 = (size_t)n * sizeof(float4);
  bool ok = alloc(B, f4, (void**)&P.ray_o) && alloc(B, f4, (void**)&P.ray_d) && alloc(B, f4, (void**)&P.hit) &&
            alloc(B, f4, (void**)&P.thr) && alloc(B, f4, (void**)&P.sum) && alloc(B, f4, (void**)&P.med) &&
            alloc(B, f4, (void**)&P.sel_p) && alloc(B, f4, (void**)&P.sel_w0) && alloc(B, f4, (void**)&P.sel_w1) &&
            alloc(B, f4, (void**)&P.sh_o) && alloc(B, f4, (void**)&P.sh_d) && alloc(B, f4, (void**)&P.sh_c) &&
            alloc(B, (size_t)n * 8, (void**)&P.key) && alloc(B, (size_t)n * 4, (void**)&P.queue[0]) &&
            alloc(B, (size_t)n * 4, (void**)&P.queue[1]) && alloc(B, (size_t)n * 4, (void**)&P.shadow_queue) &&
            alloc(B, 64, (void**)&P.counts) && alloc(B, 64, (void**)&P.rays);
  if (!ok) {
    err = "device memory exhausted for the path buffers";
    return false;
  }
  hipMemsetAsync(P.counts, 0, 64, h->stream);
  hipMemsetAsync(P.rays, 0, 64, h->stream);
  h->capacity = n;
  return true;
}

bool ensure_film(Handle* h, std::string& err) {
  if (h->film.rgb) return true;
  size_t px = (size_t)h->dev.width * h->dev.height;
  auto& B = h->film_bufs;
  Film& F = h->film;
  if (!alloc(B, px * 12, (void**)&F.rgb) || !alloc(B, px * 4, (void**)&F.weight) || !alloc(B, px * 4, (void**)&F.alpha) ||
      !alloc(B, px * 4, (void**)&F.lum2)) {
    err = "device memory exhausted for the film";
    return false;
  }
  hipMemsetAsync(F.rgb, 0, px * 12, h->stream);
  hipMemsetAsync(F.weight, 0, px * 4, h->stream);
  hipMemsetAsync(F.alpha, 0, px * 4, h->stream);
  hipMemsetAsync(F.lum2, 0, px * 4, h->stream);
  return true;
}

// One sample pass over the pixel rectangle (x0, y0, w, h): camera, then max_depth + 1 rounds of
// extend / shade / shadow. Path i is pixel (x0 + i % w, y0 + i / w).
hipError_t run_pass(Handle* h, u32 x0, u32 y0, u32 w, u32 hh, u32 sample_index) {
  const DevScene& S = h->dev;
  u32 n = w * hh;
  PassArgs A;
  A.n_paths = n;
  A.x0 = x0;
  A.y0 = y0;
  A.w = w;
  A.sample_index = sample_index;
  for (u32 k = 0; k < 4; ++k) A.sobol_cam[k] = sobol_point(k, sample_index);
  hipError_t e = launch_camera(S, h->paths, A, h->stream);
  auto mark = [&](int stage) {
    if (!h->profiling) return;
    h->marks.push_back({stage, h->pool_used});
    hipEventRecord(h->next_event(), h->stream);
  };
  auto done = [&]() {
    if (h->profiling) hipEventRecord(h->next_event(), h->stream);
  };
  for (u32 round = 0; round <= S.max_depth && e == hipSuccess; ++round) {
    mark(0);
    e = launch_extend(S, h->paths, round, n, h->stream);
    done();
    if (e != hipSuccess) break;
    RoundArgs R;
    R.round = round;
    R.sample_index = sample_index;
    u32 dims = bounce_dim(round);
    for (u32 k = 0; k < 6; ++k) R.sobol[k] = sobol_point(dims + k, sample_index);
    mark(1);
    e = launch_shade(S, h->paths, R, n, h->stream);
    done();
    if (e != hipSuccess) break;
    mark(2);
    e = launch_shadow(S, h->paths, n, h->stream);
    done();
  }
  return e;
}

u32 status_of(hipError_t e) { return e == hipSuccess ? RPX_OK : RPX_GPU_LAUNCH; }

u32 load_into(Handle* h, const uint8_t* ptr, u32 len, bool keep_meshes, std::string& err) {
  HostScene s;
  if (!parse_rsb(ptr, len, s, err, keep_meshes)) return RPX_BAD_SCENE;
  if (keep_meshes && s.meshes.empty()) {
    // rt_scene_update: a file without MESH sections keeps the loaded meshes (its INST sections,
    // or the implicit instances of the kept meshes, refer to them).
    bool implicit = s.instances.empty();
    s.meshes = h->host.meshes;
    if (implicit)
      for (size_t i = 0; i < s.meshes.size(); ++i) {
        HostInstance in;
        in.mesh = (u32)i;
        in.material = s.meshes[i].material;
        s.instances.push_back(in);
      }
    for (const HostMesh& m : s.meshes)
      if (m.material >= s.materials.size()) {
        err = "a kept mesh refers to a material that does not exist";
        return RPX_BAD_SCENE;
      }
    for (const HostInstance& in : s.instances)
      if (in.mesh >= s.meshes.size() || in.material >= s.materials.size()) {
        err = "an instance refers to a mesh or material that does not exist";
        return RPX_BAD_SCENE;
      }
  }
  HostScene old = std::move(h->host);
  h->host = std::move(s);
  if (!build_device_scene(h, err)) {
    h->host = std::move(old);
    return RPX_GPU_MEMORY;
  }
  return RPX_OK;
}
}  // namespace
}  // namespace rpx

using namespace rpx;

extern "C" {

uint32_t rt_version(void) { return 1; }
uint32_t rt_features(void) { return 0; }

uint32_t rpx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return (uint32_t)n;
}

uint32_t rt_scene_load(const uint8_t* ptr, uint32_t len) {
  if (!ptr) {
    set_error("rt_scene_load: null pointer");
    return 0;
  }
  if (rpx_device_count() == 0) {
    set_error("rt_scene_load: no HIP device is available");
    return 0;
  }
  auto h = std::make_shared<Handle>();
  hipGetDevice(&h->device);
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&h->ev[0]) != hipSuccess || hipEventCreate(&h->ev[1]) != hipSuccess) {
    set_error("rt_scene_load: cannot create a HIP stream");
    return 0;
  }
  std::string err;
  if (load_into(h.get(), ptr, len, false, err) != RPX_OK) {
    set_error(err);
    return 0;
  }
  std::lock_guard<std::mutex> lock(g_registry_mutex);
  u32 id = g_next_handle++;
  if (g_next_handle == 0) g_next_handle = 1;
  g_registry[id] = h;
  return id;
}

uint32_t rt_scene_update(uint32_t handle, const uint8_t* ptr, uint32_t len) {
  if (!ptr) {
    set_error("rt_scene_update: null pointer");
    return RPX_NULL_POINTER;
  }
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  std::string err;
  u32 s = load_into(h.get(), ptr, len, true, err);
  if (s != RPX_OK) {
    set_error(err);
    return s == RPX_BAD_SCENE ? RPX_BAD_SCENE : s;
  }
  // A new scene starts a new film.
  for (DevBuf& b : h->film_bufs)
    if (b.p) hipFree(b.p);
  h->film_bufs.clear();
  h->film = Film{};
  h->samples = 0;
  return RPX_OK;
}

uint32_t rt_scene_error_len(void) { return (uint32_t)g_error.size(); }
void rt_scene_error_copy(uint8_t* ptr) {
  if (ptr) std::memcpy(ptr, g_error.data(), g_error.size());
}

uint32_t rt_scene_free(uint32_t handle) {
  std::shared_ptr<Handle> h;
  {
    std::lock_guard<std::mutex> lock(g_registry_mutex);
    auto it = g_registry.find(handle);
    if (it == g_registry.end()) return RPX_BAD_HANDLE;
    h = it->second;
    g_registry.erase(it);
  }
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  return RPX_OK;
}

static uint32_t render_tile_impl(uint32_t handle, uint32_t x, uint32_t y, uint32_t w, uint32_t hh,
                                 uint32_t sample_index, float* out, u32 channels) {
  if (!out) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  if (w == 0 || hh == 0 || (u64)w * hh > (1ull << 22)) return RPX_BAD_TILE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  std::string err;
  // Pixels outside the image are written as zero (they are not rendered).
  u32 W = h->dev.width, H = h->dev.height;
  std::memset(out, 0, (size_t)w * hh * channels * 4);
  if (x >= W || y >= H) return RPX_OK;
  u32 cw = std::min(w, W - x), ch = std::min(hh, H - y);
  if (!ensure_paths(h.get(), cw * ch, err)) {
    set_error(err);
    return RPX_GPU_MEMORY;
  }
  float* d_out = nullptr;
  if (hipMalloc(&d_out, (size_t)cw * ch * channels * 4) != hipSuccess) return RPX_GPU_MEMORY;
  hipError_t e = run_pass(h.get(), x, y, cw, ch, sample_index);
  if (e == hipSuccess) e = launch_tile_out(h->paths, d_out, cw * ch, channels, h->stream);
  std::vector<float> tmp((size_t)cw * ch * channels);
  if (e == hipSuccess) e = hipMemcpyAsync(tmp.data(), d_out, tmp.size() * 4, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  hipFree(d_out);
  if (e != hipSuccess) {
    set_error(std::string("render: ") + hipGetErrorString(e));
    return RPX_GPU_LAUNCH;
  }
  for (u32 j = 0; j < ch; ++j)
    std::memcpy(out + (size_t)j * w * channels, tmp.data() + (size_t)j * cw * channels, (size_t)cw * channels * 4);
  return RPX_OK;
}

uint32_t rt_render_tile(uint32_t handle, uint32_t x, uint32_t y, uint32_t w, uint32_t h, uint32_t sample_index,
                        float* out) {
  return render_tile_impl(handle, x, y, w, h, sample_index, out, 3);
}
uint32_t rt_render_tile_rgba(uint32_t handle, uint32_t x, uint32_t y, uint32_t w, uint32_t h, uint32_t sample_index,
                             float* out) {
  return render_tile_impl(handle, x, y, w, h, sample_index, out, 4);
}

uint32_t rpx_render_samples(uint32_t handle, uint32_t first, uint32_t count) {
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  std::string err;
  u32 n = h->dev.width * h->dev.height;
  if (!ensure_paths(h.get(), n, err) || !ensure_film(h.get(), err)) {
    set_error(err);
    return RPX_GPU_MEMORY;
  }
  h->pool_used = 0;
  h->marks.clear();
  hipEventRecord(h->ev[0], h->stream);
  hipError_t e = hipSuccess;
  for (u32 k = 0; k < count && e == hipSuccess; ++k) {
    e = run_pass(h.get(), 0, 0, h->dev.width, h->dev.height, first + k);
    if (e == hipSuccess) e = launch_splat(h->paths, h->film, n, h->stream);
  }
  hipEventRecord(h->ev[1], h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    set_error(std::string("render: ") + hipGetErrorString(e));
    return RPX_GPU_LAUNCH;
  }
  hipEventElapsedTime(&h->last_ms, h->ev[0], h->ev[1]);
  h->stage_ms[0] = h->stage_ms[1] = h->stage_ms[2] = 0;
  h->stage_launches[0] = h->stage_launches[1] = 0;
  for (auto& m : h->marks) {
    float ms = 0;
    hipEventElapsedTime(&ms, h->pool[m.second], h->pool[m.second + 1]);
    h->stage_ms[m.first] += ms;
    if (m.first == 0) h->stage_launches[0]++;
    if (m.first == 2) h->stage_launches[1]++;
  }
  h->samples += count;
  return RPX_OK;
}

uint32_t rpx_film_reset(uint32_t handle) {
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  std::string err;
  if (!ensure_film(h.get(), err)) {
    set_error(err);
    return RPX_GPU_MEMORY;
  }
  size_t px = (size_t)h->dev.width * h->dev.height;
  hipMemsetAsync(h->film.rgb, 0, px * 12, h->stream);
  hipMemsetAsync(h->film.weight, 0, px * 4, h->stream);
  hipMemsetAsync(h->film.alpha, 0, px * 4, h->stream);
  hipMemsetAsync(h->film.lum2, 0, px * 4, h->stream);
  return hipStreamSynchronize(h->stream) == hipSuccess ? RPX_OK : RPX_GPU_LAUNCH;
}

uint32_t rpx_film_read(uint32_t handle, float* rgb, float* weight, float* alpha, float* lum2) {
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  std::string err;
  if (!ensure_film(h.get(), err)) {
    set_error(err);
    return RPX_GPU_MEMORY;
  }
  size_t px = (size_t)h->dev.width * h->dev.height;
  hipError_t e = hipSuccess;
  if (rgb && e == hipSuccess) e = hipMemcpyAsync(rgb, h->film.rgb, px * 12, hipMemcpyDeviceToHost, h->stream);
  if (weight && e == hipSuccess) e = hipMemcpyAsync(weight, h->film.weight, px * 4, hipMemcpyDeviceToHost, h->stream);
  if (alpha && e == hipSuccess) e = hipMemcpyAsync(alpha, h->film.alpha, px * 4, hipMemcpyDeviceToHost, h->stream);
  if (lum2 && e == hipSuccess) e = hipMemcpyAsync(lum2, h->film.lum2, px * 4, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  return e == hipSuccess ? RPX_OK : RPX_GPU_LAUNCH;
}

uint32_t rpx_film_device(uint32_t handle, void** ptrs4) {
  if (!ptrs4) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  std::string err;
  if (!ensure_film(h.get(), err)) {
    set_error(err);
    return RPX_GPU_MEMORY;
  }
  hipStreamSynchronize(h->stream);
  ptrs4[0] = h->film.rgb;
  ptrs4[1] = h->film.weight;
  ptrs4[2] = h->film.alpha;
  ptrs4[3] = h->film.lum2;
  return RPX_OK;
}

uint32_t rpx_trace_rays(uint32_t handle, const float* rays, uint32_t n, uint32_t* inst, uint32_t* tri, float* t,
                        uint8_t* any) {
  if (!rays || !inst || !tri || !t || !any) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  if (n == 0) return RPX_OK;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  void *d_rays = nullptr, *d_inst = nullptr, *d_tri = nullptr, *d_t = nullptr, *d_any = nullptr;
  std::vector<DevBuf> tmp;
  if (!alloc(tmp, (size_t)n * 32, &d_rays) || !alloc(tmp, (size_t)n * 4, &d_inst) || !alloc(tmp, (size_t)n * 4, &d_tri) ||
      !alloc(tmp, (size_t)n * 4, &d_t) || !alloc(tmp, (size_t)n, &d_any)) {
    for (DevBuf& b : tmp) hipFree(b.p);
    return RPX_GPU_MEMORY;
  }
  hipError_t e = hipMemcpyAsync(d_rays, rays, (size_t)n * 32, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess)
    e = launch_trace_rays(h->dev, (const float*)d_rays, n, (u32*)d_inst, (u32*)d_tri, (float*)d_t, (uint8_t*)d_any,
                          h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(inst, d_inst, (size_t)n * 4, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(tri, d_tri, (size_t)n * 4, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(t, d_t, (size_t)n * 4, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(any, d_any, (size_t)n, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  for (DevBuf& b : tmp) hipFree(b.p);
  return status_of(e);
}

uint32_t rpx_scene_info(uint32_t handle, uint64_t* out8) {
  if (!out8) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  out8[0] = h->dev.width;
  out8[1] = h->dev.height;
  out8[2] = h->host.max_samples;
  out8[3] = h->dev.seed;
  out8[4] = h->dev.max_depth;
  out8[5] = h->triangles;
  out8[6] = h->nodes;
  out8[7] = h->dev.n_instances;
  return RPX_OK;
}

uint32_t rpx_scene_stats(uint32_t handle, uint64_t* out4) {
  if (!out4) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  u64 r[3] = {0, 0, 0};
  if (h->paths.rays) {
    hipMemcpyAsync(r, h->paths.rays, 24, hipMemcpyDeviceToHost, h->stream);
    hipStreamSynchronize(h->stream);
  }
  out4[0] = r[0] + r[1] + r[2];
  out4[1] = r[0];
  out4[2] = r[1] + r[2];
  out4[3] = h->samples;
  return RPX_OK;
}

uint32_t rpx_set_profiling(uint32_t handle, uint32_t on) {
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  h->profiling = on != 0;
  return RPX_OK;
}

uint32_t rpx_last_timing(uint32_t handle, float* out6) {
  if (!out6) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  out6[0] = h->last_ms;
  out6[1] = h->stage_ms[0];
  out6[2] = h->stage_ms[1];
  out6[3] = h->stage_ms[2];
  out6[4] = (float)h->stage_launches[0];
  out6[5] = (float)h->stage_launches[1];
  return RPX_OK;
}

uint32_t rpx_energy_tables(uint16_t* albedo, uint16_t* schlick, uint16_t* clear_in, uint16_t* clear_out,
                           float* albedo_avg, float* schlick_avg) {
  if (rpx_device_count() == 0) return RPX_GPU_MEMORY;
  int dev = 0;
  hipGetDevice(&dev);
  DeviceTables* t;
  std::string err;
  if (!device_tables(dev, t, err)) {
    set_error(err);
    return RPX_GPU_LAUNCH;
  }
  if (albedo) std::memcpy(albedo, t->h_albedo.data(), 1024 * 2);
  if (schlick) std::memcpy(schlick, t->h_schlick.data(), 1024 * 2);
  if (clear_in) std::memcpy(clear_in, t->h_clear_in.data(), 17408 * 2);
  if (clear_out) std::memcpy(clear_out, t->h_clear_out.data(), 17408 * 2);
  if (albedo_avg) std::memcpy(albedo_avg, t->h_albedo_avg.data(), 32 * 4);
  if (schlick_avg) std::memcpy(schlick_avg, t->h_schlick_avg.data(), 32 * 4);
  return RPX_OK;
}

}  // extern "C"
