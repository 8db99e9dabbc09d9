// The C ABI of the GPU tracer (include/rpx.h): a registry of scene handles, each owning its
// device scene, path pool, resident film and HIP stream, and of render queues (several scenes'
// jobs rendered by one pool).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include "../../include/rpx.h"
#include "rpx_adaptive.h"
#include "rpx_caustics.h"
#include "rpx_denoise.h"
#include "rpx_engine.h"

namespace rpx {
namespace {

thread_local std::string g_error;
void set_error(const std::string& e) { g_error = e; }

struct Handle {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mutex;
  HostScene host;
  DevScene dev{};
  DevScene* d_scene = nullptr;  // device copy of `dev` (a one-entry scene table)
  std::vector<DevBuf> scene_bufs;
  u64 triangles = 0, nodes = 0;
  std::unique_ptr<Engine> engine;
  std::unique_ptr<PhotonPasses> photons;  // a scene with caustics: its photon maps' memory
  // resident film
  Film film{};
  std::vector<DevBuf> film_bufs;
  u64 samples = 0;
  u64 film_samples = 0;  // samples per pixel in the film since it was last cleared
  // base-colour texture of one material
  std::vector<DevBuf> tex_bufs;
  // adaptive sampling: samples per block of the last adaptive render, device scratch
  std::vector<u32> block_counts;
  std::vector<DevBuf> adaptive_bufs;
  size_t adaptive_px = 0, adaptive_blocks = 0;
  u32* d_block_counts = nullptr;
  float *d_pixel_error = nullptr, *d_block_error = nullptr;
  KeptSplat* d_kept_table = nullptr;  // one entry per block
  // pinned staging of a decision's transfers: block counts up, block errors down, kept-splat table up
  void* adaptive_pinned = nullptr;
  // samples rendered ahead of the adaptive rounds (float4 radiance + alpha each), and the last
  // adaptive render's work: path samples rendered, of them rendered ahead, runs ahead, unused
  DevBuf kept{};
  size_t kept_cap = 0;
  u64 adaptive_work[4] = {0, 0, 0, 0};
  // resident pass film (sums of the sampled passes, exact planes), the resolved albedo the
  // denoiser reads, the denoiser's work buffers and its output (RGBA sums)
  PassPlanes passes{};
  float* pass_albedo = nullptr;
  float4* denoised = nullptr;
  std::vector<DevBuf> pass_bufs;
  std::unique_ptr<denoise::Denoiser> denoiser;

  ~Handle() {
    if (stream) hipStreamSynchronize(stream);
    denoiser.reset();
    dev_free_all(pass_bufs);
    engine.reset();
    photons.reset();
    dev_free_all(scene_bufs);
    dev_free_all(film_bufs);
    dev_free_all(tex_bufs);
    dev_free_all(adaptive_bufs);
    if (adaptive_pinned) hipHostFree(adaptive_pinned);
    if (kept.p) hipFree(kept.p);
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
  void* p = nullptr;
  if (!dev_alloc(bufs, std::max<size_t>(v.size() * sizeof(T), 16), &p)) {
    err = "device memory exhausted while uploading the scene";
    return false;
  }
  if (!v.empty() && hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st) != hipSuccess) {
    err = "scene upload failed";
    return false;
  }
  out = static_cast<const T*>(p);
  return true;
}

// Builds and uploads the device scene of `h->host`.
bool build_device_scene(Handle* h, std::string& err) {
  DeviceTables* tables;
  if (!device_tables(h->device, tables, err)) return false;
  FlatScene flat;
  unsigned hw = std::thread::hardware_concurrency();
  if (!flatten_scene(h->host, tables->host_view(), flat, err, hw ? (int)std::min(hw, 16u) : 4)) return false;
  // Traversal stack bound: at most three entries per 4-wide level.
  u32 bound = 3 * (flat.max_tlas_depth + flat.max_blas_depth) + 2;
  if (bound > 64) {
    err = "the BVH is too deep for the traversal stack";
    return false;
  }
  if (flat.consts.width > 65535 || flat.consts.height > 65535) {
    err = "the image is wider or taller than 65535 pixels";
    return false;
  }
  hipStreamSynchronize(h->stream);
  if (h->engine) h->engine->photons = nullptr;
  h->photons.reset();
  dev_free_all(h->scene_bufs);
  DevScene d = flat.consts;
  auto& B = h->scene_bufs;
  hipStream_t st = h->stream;
  if (!upload(B, flat.nodes, d.nodes, st, err) || !upload(B, flat.tris, d.tris, st, err) ||
      !upload(B, flat.instances, d.instances, st, err) || !upload(B, flat.meshes, d.meshes, st, err) ||
      !upload(B, flat.positions, d.positions, st, err) || !upload(B, flat.normals, d.normals, st, err) ||
      !upload(B, flat.indices, d.indices, st, err) || !upload(B, flat.corners, d.corners, st, err) ||
      !upload(B, flat.bsdfs, d.bsdfs, st, err) ||
      !upload(B, flat.materials, d.materials, st, err) || !upload(B, flat.areas, d.areas, st, err) ||
      !upload(B, flat.cdf, d.cdf, st, err) || !upload(B, flat.flux_pmf, d.flux_pmf, st, err) ||
      !upload(B, flat.entries, d.entries, st, err) || !upload(B, flat.spheres, d.spheres, st, err) ||
      !upload(B, flat.media, d.media, st, err) || !upload(B, flat.suns, d.suns, st, err) ||
      !upload(B, flat.emitters, d.emitters, st, err) || !upload(B, flat.emit_tris, d.emit_tris, st, err) ||
      !upload(B, flat.emit_cdf, d.emit_cdf, st, err) || !upload(B, flat.instance_emitter, d.instance_emitter, st, err))
    return false;
  d.pmaps = nullptr;
  if (d.caustics_on) {
    if (!upload(B, flat.caustic_sources, d.caustics.sources, st, err) ||
        !upload(B, flat.caustic_cdf, d.caustics.source_cdf, st, err))
      return false;
    auto maps = std::make_unique<PhotonPasses>();
    if (!maps->init(d.caustics.photons, d.caustics.max_bounces, d.caustics.radius, st, err)) return false;
    d.pmaps = maps->table();
    h->photons = std::move(maps);
  }
  d.tables = tables->device_view();
  std::vector<DevScene> one(1, d);
  const DevScene* ds = nullptr;
  if (!upload(B, one, ds, st, err)) return false;
  if (hipStreamSynchronize(st) != hipSuccess) {
    err = "scene upload failed";
    return false;
  }
  h->dev = d;
  h->d_scene = const_cast<DevScene*>(ds);
  dev_free_all(h->tex_bufs);
  h->triangles = flat.triangles;
  h->nodes = flat.nodes.size();
  return true;
}

// Points the handle's engine at what its scene needs (BSDF records per slot, the extended
// shade, the photon maps of a scene with caustics, by sample index).
void prepare_engine(Handle* h) {
  h->engine->textured = h->dev.tex_material != NO_TEXTURE;
  h->engine->ext = h->dev.ext != 0;
  h->engine->photons = h->photons.get();
  h->engine->photon_pass = -1;
}

// Copies the handle's scene constants to its one-entry device scene table.
bool upload_scene_table(Handle* h) {
  if (h->photons) h->photons->invalidate();
  return hipMemcpyAsync(h->d_scene, &h->dev, sizeof(DevScene), hipMemcpyHostToDevice, h->stream) == hipSuccess &&
         hipStreamSynchronize(h->stream) == hipSuccess;
}

bool ensure_film(Handle* h, std::string& err) {
  if (h->film.rgb) return true;
  size_t px = (size_t)h->dev.width * h->dev.height;
  auto& B = h->film_bufs;
  Film& F = h->film;
  if (!dev_alloc(B, px * 12, (void**)&F.rgb) || !dev_alloc(B, px * 4, (void**)&F.weight) ||
      !dev_alloc(B, px * 4, (void**)&F.alpha) || !dev_alloc(B, px * 4, (void**)&F.lum2)) {
    err = "device memory exhausted for the film";
    return false;
  }
  hipMemsetAsync(F.rgb, 0, px * 12, h->stream);
  hipMemsetAsync(F.weight, 0, px * 4, h->stream);
  hipMemsetAsync(F.alpha, 0, px * 4, h->stream);
  hipMemsetAsync(F.lum2, 0, px * 4, h->stream);
  return true;
}

DevSegment whole_image(const Handle* h, u32 scene, u32 first, u32 count) {
  DevSegment s{};
  s.scene = scene;
  s.x0 = 0;
  s.y0 = 0;
  s.w = h->dev.width;
  s.h = h->dev.height;
  s.first = first;
  s.count = count;
  s.film = h->film;
  s.film_w = h->dev.width;
  return s;
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

// A render queue: jobs (scene handle, sample range) rendered by one path pool, each into its own
// resident film.
struct Job {
  std::shared_ptr<Handle> scene;
  u32 first, count;
  u32 width, height;
  Film film;
};

struct Queue {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mutex;
  std::vector<Job> jobs;
  std::vector<DevBuf> film_bufs;
  DevBuf scenes;  // device scene table of the last run
  std::unique_ptr<Engine> engine;
  ~Queue() {
    if (stream) hipStreamSynchronize(stream);
    engine.reset();
    dev_free_all(film_bufs);
    if (scenes.p) hipFree(scenes.p);
    if (stream) hipStreamDestroy(stream);
  }
};

std::map<u32, std::shared_ptr<Queue>> g_queues;
u32 g_next_queue = 1;

std::shared_ptr<Queue> find_queue(u32 q) {
  std::lock_guard<std::mutex> lock(g_registry_mutex);
  auto it = g_queues.find(q);
  return it == g_queues.end() ? nullptr : it->second;
}

void apply_options(Engine& e, u32 pool, u32 shape, u32 shadow_shape, u32 group, u32 chunk, u32 flags, u32 tail) {
  if (tail) e.opt.tail = tail;
  if (flags & 4u) e.opt.tail = 0;
  if (pool) e.opt.pool = std::max<u32>(pool, 64);
  if (shadow_shape < (u32)N_SHAPES) e.opt.shadow_shape = (int)shadow_shape;
  if (chunk) e.opt.chunk = std::min(chunk, RESULTS_MAX);
  if (shape < (u32)N_SHAPES) e.opt.shape = (int)shape;
  if (group) e.opt.group = std::max<u32>(2, group & ~1u);
  e.opt.per_pass = (flags & 1u) != 0;
  e.opt.graph = (flags & 2u) != 0;
  e.opt.adaptive_ahead = (flags & 8u) == 0;
}

void fill_timing(const Engine& e, float* out12) {
  const Timing& t = e.timing;
  out12[0] = t.total_ms;
  for (int k = 0; k < 4; ++k) {
    out12[1 + k] = t.stage_ms[k];
    out12[5 + k] = (float)t.launches[k];
  }
  out12[1] += t.stage_ms[5];  // the media walks are closest-hit work
  out12[9] = (float)t.iterations;
  out12[10] = (float)t.chunks;
  out12[11] = t.stage_ms[4];
}
}  // namespace
}  // namespace rpx

using namespace rpx;

extern "C" {

uint32_t rt_version(void) { return 1; }
uint32_t rt_features(void) { return 0; }

// Host buffers for callers that write scene bytes or read results through the ABI's own memory
// (the wasm hosts' pattern): zeroed, never null (a zero length gets one byte), freed with their
// length.
uint8_t* rt_alloc(uint32_t len) { return static_cast<uint8_t*>(std::calloc(len ? len : 1, 1)); }
void rt_free(uint8_t* ptr, uint32_t len) {
  (void)len;
  std::free(ptr);
}

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
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("rt_scene_load: cannot create a HIP stream");
    return 0;
  }
  std::string err;
  h->engine = std::make_unique<Engine>();
  if (!h->engine->init(h->device, h->stream, err)) {
    set_error("rt_scene_load: " + err);
    return 0;
  }
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
    return s;
  }
  // A new scene starts a new film and new passes.
  hipStreamSynchronize(h->stream);
  dev_free_all(h->film_bufs);
  h->film = Film{};
  h->samples = 0;
  h->film_samples = 0;
  dev_free_all(h->pass_bufs);
  h->passes = PassPlanes{};
  h->pass_albedo = nullptr;
  h->denoised = nullptr;
  return RPX_OK;
}

uint32_t rt_scene_error_len(void) { return (uint32_t)g_error.size(); }

uint32_t rt_scene_caustics(uint32_t handle) {
  auto h = find(handle);
  if (!h) return 0xffffffffu;
  std::lock_guard<std::mutex> lock(h->mutex);
  return h->photons ? h->photons->photons() : 0u;
}
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

// `photon_pass` >= 0: the photon map's pass (else the sample index's).
static uint32_t render_tile_impl(uint32_t handle, uint32_t x, uint32_t y, uint32_t w, uint32_t hh,
                                 uint32_t sample_index, float* out, u32 channels, long long photon_pass = -1) {
  if (!out) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  if (w == 0 || hh == 0 || (u64)w * hh > (1ull << 22)) return RPX_BAD_TILE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  // Pixels outside the image are written as zero (they are not rendered).
  u32 W = h->dev.width, H = h->dev.height;
  std::memset(out, 0, (size_t)w * hh * channels * 4);
  if (x >= W || y >= H) return RPX_OK;
  u32 cw = std::min(w, W - x), ch = std::min(hh, H - y);
  DevSegment s{};
  s.scene = 0;
  s.x0 = x;
  s.y0 = y;
  s.w = cw;
  s.h = ch;
  s.first = sample_index;
  s.count = 1;
  s.film_w = W;
  std::string err;
  float* d_out = nullptr;
  if (hipMalloc(&d_out, (size_t)cw * ch * channels * 4) != hipSuccess) {
    set_error("device memory exhausted for the tile");
    return RPX_GPU_MEMORY;
  }
  prepare_engine(h.get());
  h->engine->photon_pass = photon_pass;
  hipError_t e = h->engine->run(h->d_scene, std::vector<DevSegment>(1, s), false, err);
  if (e == hipSuccess) e = launch_tile_out(h->engine->P, d_out, cw, ch, channels, h->stream);
  std::vector<float> tmp((size_t)cw * ch * channels);
  if (e == hipSuccess) e = hipMemcpyAsync(tmp.data(), d_out, tmp.size() * 4, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  hipFree(d_out);
  if (e != hipSuccess) {
    set_error(err.empty() ? std::string("render: ") + hipGetErrorString(e) : err);
    return e == hipErrorOutOfMemory ? RPX_GPU_MEMORY : RPX_GPU_LAUNCH;
  }
  for (u32 j = 0; j < ch; ++j)
    std::memcpy(out + (size_t)j * w * channels, tmp.data() + (size_t)j * cw * channels, (size_t)cw * channels * 4);
  return RPX_OK;
}

// ---------------------------------------------------------------- allocation-free tile render
// The device memory of rpx_render_tile_noalloc: one region per GPU, held until the process ends.
// Its host-side name is not a device address; the per-device address comes from
// hipGetSymbolAddress on the entry point's first call on each device.
constexpr size_t NOALLOC_SCRATCH_BYTES = 16u << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on device `dev` (current), looked up once per device.
static void* scratch_of(int dev, std::string& err) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    err = "device index beyond the allocation-free entry point's table";
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      err = std::string("hipGetSymbolAddress of the scratch region: ") + hipGetErrorString(e);
      return nullptr;
    }
    g_scratch_ptr[dev] = ptr;
  }
  return g_scratch_ptr[dev];
}

uint32_t rpx_render_tile_noalloc(uint32_t handle, uint32_t x, uint32_t y, uint32_t w, uint32_t hh,
                                 uint32_t sample_index, float* out) {
  if (!out) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  if (w == 0 || hh == 0 || (u64)w * hh > RPX_NOALLOC_MAX_PIXELS) {
    set_error("rpx_render_tile_noalloc renders tiles of 1 to 16384 pixels");
    return RPX_BAD_TILE;
  }
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  std::string err;
  void* region = scratch_of(h->device, err);
  if (!region) {
    set_error(err);
    return RPX_GPU_LAUNCH;
  }
  u32 W = h->dev.width, H = h->dev.height;
  std::memset(out, 0, (size_t)w * hh * 3 * 4);
  if (x >= W || y >= H) return RPX_OK;
  DevSegment s{};
  s.scene = 0;
  s.x0 = x;
  s.y0 = y;
  s.w = std::min(w, W - x);
  s.h = std::min(hh, H - y);
  s.first = sample_index;
  s.count = 1;
  s.film_w = W;
  prepare_engine(h.get());
  hipError_t e = h->engine->render_tile_in(h->d_scene, s, 3, region, NOALLOC_SCRATCH_BYTES, out, (size_t)w * 3, err);
  if (e != hipSuccess) {
    set_error(err);
    return e == hipErrorInvalidValue ? RPX_BAD_TILE : RPX_GPU_LAUNCH;
  }
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
uint32_t rt_render_block_rgba(uint32_t handle, uint32_t x, uint32_t y, uint32_t w, uint32_t h, uint32_t sample_index,
                              uint32_t iteration, float* out) {
  // Scenes with edge supersampling are refused at load, so a block of a round is a tile;
  // `iteration` selects the photon map of a scene with caustics.
  return render_tile_impl(handle, x, y, w, h, sample_index, out, 4, (long long)iteration);
}

// ---------------------------------------------------------------- adaptive sampling
static bool ensure_adaptive(Handle* h, size_t px, size_t blocks, std::string& err) {
  if (px <= h->adaptive_px && blocks <= h->adaptive_blocks) return true;
  hipStreamSynchronize(h->stream);
  dev_free_all(h->adaptive_bufs);
  if (h->adaptive_pinned) hipHostFree(h->adaptive_pinned);
  h->adaptive_pinned = nullptr;
  h->adaptive_px = h->adaptive_blocks = 0;
  auto& B = h->adaptive_bufs;
  if (!dev_alloc(B, blocks * 4, (void**)&h->d_block_counts) || !dev_alloc(B, px * 4, (void**)&h->d_pixel_error) ||
      !dev_alloc(B, blocks * 4, (void**)&h->d_block_error) ||
      !dev_alloc(B, blocks * sizeof(KeptSplat), (void**)&h->d_kept_table) ||
      hipHostMalloc(&h->adaptive_pinned, blocks * (8 + sizeof(KeptSplat)), hipHostMallocDefault) != hipSuccess) {
    dev_free_all(h->adaptive_bufs);
    h->adaptive_pinned = nullptr;
    err = "device memory exhausted for the adaptive sampling's error planes";
    return false;
  }
  h->adaptive_px = px;
  h->adaptive_blocks = blocks;
  return true;
}

uint32_t rpx_render_adaptive(uint32_t handle, uint32_t cap, uint64_t* out8) {
  if (!out8) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  if (cap == 0) cap = h->host.max_samples;
  if (cap == 0) {
    set_error("rpx_render_adaptive: neither a cap nor the scene's max_samples bounds the rounds");
    return RPX_BAD_TILE;
  }
  if (h->photons) {
    set_error("rpx_render_adaptive: a scene with caustics renders its samples pass by pass (rpx_render_samples)");
    return RPX_BAD_SCENE;
  }
  std::string err;
  u32 W = h->dev.width, H = h->dev.height;
  adaptive::State ad;
  ad.init(W, H, h->host.noise_threshold, h->host.exposure_ev, cap);
  size_t px = (size_t)W * H, blocks = ad.counts.size();
  if (!ensure_film(h.get(), err) || !ensure_adaptive(h.get(), px, blocks, err)) {
    set_error(err);
    return RPX_GPU_MEMORY;
  }
  hipStream_t st = h->stream;
  const Film& F = h->film;
  hipMemsetAsync(F.rgb, 0, px * 12, st);
  hipMemsetAsync(F.weight, 0, px * 4, st);
  hipMemsetAsync(F.alpha, 0, px * 4, st);
  hipMemsetAsync(F.lum2, 0, px * 4, st);
  prepare_engine(h.get());
  hipEvent_t ev[2];
  hipEventCreate(&ev[0]);
  hipEventCreate(&ev[1]);
  hipEventRecord(ev[0], st);
  u64 decisions = 0, rounds = 0, paths = 0;
  adaptive::Decision last{};
  std::vector<float> block_error(blocks);
  std::vector<DevSegment> segs;
  hipError_t e = hipSuccess;
  // Rendering ahead: a block's k-th sample is the same whichever round renders it, so once a block
  // may take part in the rounds (a candidate of a decision), all its samples up to the cap are
  // rendered in one run and kept; each round then adds the block's next samples from what was
  // kept, in sample order, and only blocks chosen without kept samples need a run of their own.
  // The decisions and the film are those of rendering each round's samples alone.
  const bool ahead = h->engine->opt.adaptive_ahead;
  u32* pinned_counts = (u32*)h->adaptive_pinned;
  float* pinned_errors = (float*)(pinned_counts + blocks);
  KeptSplat* pinned_table = (KeptSplat*)(pinned_errors + blocks);
  struct Kept {
    size_t base = 0;  // pixel p's kept samples start at kept + base + p * count
    u32 first = 0, count = 0;
  };
  std::vector<Kept> kept(blocks);
  size_t kept_used = 0;
  u64 rendered = 0, rendered_ahead = 0, runs_ahead = 0;
  std::vector<KeptSplat> table;
  auto rect_of = [&](u32 b, DevSegment& s) {
    s = DevSegment{};
    s.scene = 0;
    s.x0 = (b % ad.cols) * adaptive::BLOCK;
    s.y0 = (b / ad.cols) * adaptive::BLOCK;
    s.w = std::min(adaptive::BLOCK, W - s.x0);
    s.h = std::min(adaptive::BLOCK, H - s.y0);
    s.film = F;
    s.film_w = W;
  };
  // Renders blocks `need` from their current count to the cap and keeps the results; one run per
  // group of blocks whose results fit one chunk.
  auto render_ahead = [&](const std::vector<u32>& need) -> hipError_t {
    const u32 limit = std::min(std::max<u32>(h->engine->opt.chunk, 1), RESULTS_MAX);
    size_t total = 0;
    for (u32 b : need) {
      DevSegment s;
      rect_of(b, s);
      total += (size_t)s.w * s.h * (ad.cap - ad.counts[b]);
    }
    if (kept_used + total > h->kept_cap) {
      size_t cap_new = std::max(kept_used + total + total / 4, 2 * h->kept_cap);
      void* p = nullptr;
      if (hipMalloc(&p, cap_new * sizeof(float4)) != hipSuccess) {
        err = "device memory exhausted for the samples rendered ahead of the adaptive rounds";
        return hipErrorOutOfMemory;
      }
      hipError_t r = hipSuccess;
      if (kept_used) r = hipMemcpyAsync(p, h->kept.p, kept_used * sizeof(float4), hipMemcpyDeviceToDevice, st);
      if (r == hipSuccess) r = hipStreamSynchronize(st);
      if (r != hipSuccess) {
        hipFree(p);
        return r;
      }
      if (h->kept.p) hipFree(h->kept.p);
      h->kept.p = p;
      h->kept.bytes = cap_new * sizeof(float4);
      h->kept_cap = cap_new;
    }
    const bool per_pass = h->engine->opt.per_pass;
    h->engine->opt.per_pass = false;  // the run's results must stay in one chunk
    hipError_t r = hipSuccess;
    size_t k = 0;
    while (k < need.size() && r == hipSuccess) {
      std::vector<DevSegment> run_segs;
      std::vector<u32> run_blocks;
      u32 base = 0;
      for (; k < need.size(); ++k) {
        u32 b = need[k];
        DevSegment s;
        rect_of(b, s);
        s.first = ad.counts[b];
        s.count = ad.cap - ad.counts[b];
        u32 n = s.w * s.h * s.count;
        if (!run_segs.empty() && (u64)base + n > limit) break;
        s.base = base;
        base += n;
        run_segs.push_back(s);
        run_blocks.push_back(b);
      }
      r = h->engine->run(h->d_scene, run_segs, false, err);
      if (r == hipSuccess)
        r = hipMemcpyAsync((float4*)h->kept.p + kept_used, h->engine->P.results, (size_t)base * sizeof(float4),
                           hipMemcpyDeviceToDevice, st);
      if (r != hipSuccess) break;
      for (size_t i = 0; i < run_blocks.size(); ++i)
        kept[run_blocks[i]] = Kept{kept_used + run_segs[i].base, run_segs[i].first, run_segs[i].count};
      kept_used += base;
      rendered += base;
      rendered_ahead += base;
      ++runs_ahead;
    }
    h->engine->opt.per_pass = per_pass;
    return r;
  };
  for (;;) {
    if (ad.left() == 0) {
      // adaptive.rs:517-539: the round's samples are counted, then the next round is chosen.
      for (u32 b = 0; b < blocks; ++b) ad.counts[b] += ad.quota(b);
      ad.iteration += ad.round;
      ad.done = 0;
      // (The pinned staging is free again: the previous decision's transfers were waited for.)
      std::memcpy(pinned_counts, ad.counts.data(), blocks * 4);
      e = hipMemcpyAsync(h->d_block_counts, pinned_counts, blocks * 4, hipMemcpyHostToDevice, st);
      if (e == hipSuccess)
        e = adaptive::block_errors(W, H, F.rgb, 3, F.lum2, h->d_block_counts, ad.floor, h->d_pixel_error,
                                   h->d_block_error, st);
      if (e == hipSuccess) e = hipMemcpyAsync(pinned_errors, h->d_block_error, blocks * 4, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) break;
      std::memcpy(block_error.data(), pinned_errors, blocks * 4);
      last = adaptive::choose(W, H, ad.counts, block_error, ad.target, ad.cap, ad.active);
      ad.round = last.round;
      ++decisions;
      if (ad.round == 0) break;
    }
    u32 count = ad.left();
    if (!ahead || decisions == 0) {
      // Every block with a quota renders its next samples; the blocks' pixels are disjoint.
      segs.clear();
      for (u32 b = 0; b < blocks; ++b) {
        u32 q = ad.quota(b);
        if (q == 0) continue;
        DevSegment s;
        rect_of(b, s);
        s.first = ad.counts[b];
        s.count = std::min(q, count);
        segs.push_back(s);
        paths += (u64)s.w * s.h * s.count;
      }
      e = h->engine->run(h->d_scene, segs, true, err, true);
      if (e != hipSuccess) break;
      for (const DevSegment& s : segs) rendered += (u64)s.w * s.h * s.count;
    } else {
      // Chosen blocks without their next samples kept, and (with them, in the same run) every
      // candidate of the decision not yet rendered ahead; at the first decision, every candidate.
      const float floor_target = ad.target > 0.0f ? ad.target : 0.0f;
      std::vector<u32> need;
      bool chosen_missing = false;
      for (u32 b = 0; b < blocks; ++b) {
        u32 q = ad.quota(b);
        if (q == 0) continue;
        const Kept& k = kept[b];
        u32 n = std::min(q, count);
        if (!(k.count && k.first <= ad.counts[b] && ad.counts[b] + n <= k.first + k.count)) chosen_missing = true;
      }
      if (chosen_missing || decisions == 1)
        for (u32 b = 0; b < blocks; ++b) {
          bool candidate = block_error[b] > floor_target && ad.counts[b] < ad.cap;
          bool chosen = ad.quota(b) > 0;
          const Kept& k = kept[b];
          bool has = k.count && k.first <= ad.counts[b] && ad.counts[b] < k.first + k.count;
          if ((candidate || chosen) && !has) need.push_back(b);
        }
      if (!need.empty()) {
        e = render_ahead(need);
        if (e != hipSuccess) break;
      }
      table.clear();
      for (u32 b = 0; b < blocks; ++b) {
        u32 q = ad.quota(b);
        if (q == 0) continue;
        const Kept& k = kept[b];
        KeptSplat t{};
        rect_of(b, t.seg);
        t.seg.first = ad.counts[b];
        t.seg.count = std::min(q, count);
        t.res = (const float4*)h->kept.p + k.base;
        t.stride = k.count;
        t.skip = ad.counts[b] - k.first;
        table.push_back(t);
        paths += (u64)t.seg.w * t.seg.h * t.seg.count;
      }
      std::memcpy(pinned_table, table.data(), table.size() * sizeof(KeptSplat));
      e = hipMemcpyAsync(h->d_kept_table, pinned_table, table.size() * sizeof(KeptSplat), hipMemcpyHostToDevice, st);
      if (e == hipSuccess) e = launch_splat_kept(h->d_kept_table, (u32)table.size(), st);
      if (e != hipSuccess) break;
    }
    ad.done = std::min(ad.done + count, ad.round);
    ++rounds;
  }
  // Samples rendered ahead that no round used.
  u64 unused = 0;
  if (e == hipSuccess)
    for (u32 b = 0; b < blocks; ++b) {
      const Kept& k = kept[b];
      if (k.count && ad.counts[b] < k.first + k.count) {
        DevSegment s;
        rect_of(b, s);
        unused += (u64)s.w * s.h * (k.first + k.count - std::max(ad.counts[b], k.first));
      }
    }
  h->adaptive_work[0] = rendered;
  h->adaptive_work[1] = rendered_ahead;
  h->adaptive_work[2] = runs_ahead;
  h->adaptive_work[3] = unused;
  hipEventRecord(ev[1], st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  float ms = 0;
  if (e == hipSuccess) hipEventElapsedTime(&ms, ev[0], ev[1]);
  hipEventDestroy(ev[0]);
  hipEventDestroy(ev[1]);
  if (e != hipSuccess) {
    set_error(err.empty() ? std::string("rpx_render_adaptive: ") + hipGetErrorString(e) : err);
    return e == hipErrorOutOfMemory ? RPX_GPU_MEMORY : RPX_GPU_LAUNCH;
  }
  h->block_counts = ad.counts;
  h->samples = 0;
  h->film_samples = 0;
  u32 lo = ~0u, hi = 0;
  for (u32 c : ad.counts) {
    lo = std::min(lo, c);
    hi = std::max(hi, c);
  }
  u32 bits[2];
  std::memcpy(&bits[0], &last.max_error, 4);
  std::memcpy(&bits[1], &last.remaining, 4);
  out8[0] = decisions;
  out8[1] = rounds;
  out8[2] = paths;
  out8[3] = blocks ? lo : 0;
  out8[4] = hi;
  out8[5] = bits[0];
  out8[6] = bits[1];
  out8[7] = (u64)((double)ms * 1.0e6);
  return RPX_OK;
}

uint32_t rpx_adaptive_counts(uint32_t handle, uint32_t* counts, uint32_t cap) {
  if (!counts) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  if (cap < h->block_counts.size()) return RPX_BAD_TILE;
  std::memcpy(counts, h->block_counts.data(), h->block_counts.size() * 4);
  return RPX_OK;
}

uint32_t rpx_adaptive_work(uint32_t handle, uint64_t* out4) {
  if (!out4) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  for (int k = 0; k < 4; ++k) out4[k] = h->adaptive_work[k];
  return RPX_OK;
}

uint32_t rt_adaptive_decide(uint32_t width, uint32_t height, float ev, float target, uint32_t cap,
                            const uint8_t* counts_ptr, const uint8_t* sums_ptr, const uint8_t* lum2_ptr,
                            uint8_t* active_ptr, uint8_t* decision_ptr) {
  if (!counts_ptr || !sums_ptr || !lum2_ptr || !active_ptr || !decision_ptr) return RPX_NULL_POINTER;
  if (width == 0 || height == 0 || (u64)width * height > (1ull << 28)) return RPX_BAD_TILE;
  if (rpx_device_count() == 0) {
    set_error("rt_adaptive_decide: no HIP device is available");
    return RPX_GPU_MEMORY;
  }
  u32 cols = (width + adaptive::BLOCK - 1) / adaptive::BLOCK, rows = (height + adaptive::BLOCK - 1) / adaptive::BLOCK;
  size_t blocks = (size_t)cols * rows, px = (size_t)width * height;
  // The planes are little-endian words with no alignment promised.
  std::vector<u32> counts(blocks);
  std::memcpy(counts.data(), counts_ptr, blocks * 4);
  std::vector<DevBuf> bufs;
  float *d_sums = nullptr, *d_lum2 = nullptr, *d_error = nullptr, *d_block_error = nullptr;
  u32* d_counts = nullptr;
  if (!dev_alloc(bufs, px * 16, (void**)&d_sums) || !dev_alloc(bufs, px * 4, (void**)&d_lum2) ||
      !dev_alloc(bufs, px * 4, (void**)&d_error) || !dev_alloc(bufs, blocks * 4, (void**)&d_block_error) ||
      !dev_alloc(bufs, blocks * 4, (void**)&d_counts)) {
    dev_free_all(bufs);
    set_error("rt_adaptive_decide: device memory exhausted");
    return RPX_GPU_MEMORY;
  }
  hipStream_t st = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMemcpyAsync(d_sums, sums_ptr, px * 16, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(d_lum2, lum2_ptr, px * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(d_counts, counts.data(), blocks * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = adaptive::block_errors(width, height, d_sums, 4, d_lum2, d_counts, adaptive::floor_for(ev), d_error,
                               d_block_error, st);
  std::vector<float> block_error(blocks);
  if (e == hipSuccess) e = hipMemcpyAsync(block_error.data(), d_block_error, blocks * 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (st) hipStreamDestroy(st);
  dev_free_all(bufs);
  if (e != hipSuccess) {
    set_error(std::string("rt_adaptive_decide: ") + hipGetErrorString(e));
    return RPX_GPU_LAUNCH;
  }
  std::vector<uint8_t> active;
  adaptive::Decision d = adaptive::choose(width, height, counts, block_error, target, cap, active);
  u32 words[6] = {d.round, d.active, d.blocks, 0, 0, 0};
  std::memcpy(&words[3], &d.threshold, 4);
  std::memcpy(&words[4], &d.max_error, 4);
  std::memcpy(&words[5], &d.remaining, 4);
  std::memcpy(active_ptr, active.data(), blocks);
  std::memcpy(decision_ptr, words, sizeof(words));
  return RPX_OK;
}

uint32_t rpx_render_samples(uint32_t handle, uint32_t first, uint32_t count) {
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  if (count == 0) return RPX_OK;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  std::string err;
  if (!ensure_film(h.get(), err)) {
    set_error(err);
    return RPX_GPU_MEMORY;
  }
  prepare_engine(h.get());
  hipError_t e = h->engine->run(h->d_scene, std::vector<DevSegment>(1, whole_image(h.get(), 0, first, count)), true, err);
  if (e != hipSuccess) {
    set_error(err);
    return e == hipErrorOutOfMemory ? RPX_GPU_MEMORY : RPX_GPU_LAUNCH;
  }
  h->samples += count;
  h->film_samples += count;
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
  h->film_samples = 0;
  return hipStreamSynchronize(h->stream) == hipSuccess ? RPX_OK : RPX_GPU_LAUNCH;
}

static uint32_t read_film(const Film& F, size_t px, hipStream_t st, float* rgb, float* weight, float* alpha,
                          float* lum2) {
  hipError_t e = hipSuccess;
  if (rgb && e == hipSuccess) e = hipMemcpyAsync(rgb, F.rgb, px * 12, hipMemcpyDeviceToHost, st);
  if (weight && e == hipSuccess) e = hipMemcpyAsync(weight, F.weight, px * 4, hipMemcpyDeviceToHost, st);
  if (alpha && e == hipSuccess) e = hipMemcpyAsync(alpha, F.alpha, px * 4, hipMemcpyDeviceToHost, st);
  if (lum2 && e == hipSuccess) e = hipMemcpyAsync(lum2, F.lum2, px * 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e == hipSuccess ? RPX_OK : RPX_GPU_LAUNCH;
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
  return read_film(h->film, (size_t)h->dev.width * h->dev.height, h->stream, rgb, weight, alpha, lum2);
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

static uint32_t trace_rays_impl(uint32_t handle, const float* rays, uint32_t n, uint32_t* inst, uint32_t* tri, float* t,
                                uint8_t* any, int wave_cap) {
  if (!rays || !inst || !tri || !t || !any) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  if (n == 0) return RPX_OK;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  void *d_rays = nullptr, *d_inst = nullptr, *d_tri = nullptr, *d_t = nullptr, *d_any = nullptr;
  std::vector<DevBuf> tmp;
  if (!dev_alloc(tmp, (size_t)n * 32, &d_rays) || !dev_alloc(tmp, (size_t)n * 4, &d_inst) ||
      !dev_alloc(tmp, (size_t)n * 4, &d_tri) || !dev_alloc(tmp, (size_t)n * 4, &d_t) || !dev_alloc(tmp, (size_t)n, &d_any)) {
    dev_free_all(tmp);
    return RPX_GPU_MEMORY;
  }
  hipError_t e = hipMemcpyAsync(d_rays, rays, (size_t)n * 32, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess)
    e = launch_trace_rays(h->d_scene, (const float*)d_rays, n, (u32*)d_inst, (u32*)d_tri, (float*)d_t, (uint8_t*)d_any,
                          h->stream, wave_cap);
  if (e == hipSuccess) e = hipMemcpyAsync(inst, d_inst, (size_t)n * 4, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(tri, d_tri, (size_t)n * 4, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(t, d_t, (size_t)n * 4, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(any, d_any, (size_t)n, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  dev_free_all(tmp);
  return status_of(e);
}

uint32_t rpx_trace_rays(uint32_t handle, const float* rays, uint32_t n, uint32_t* inst, uint32_t* tri, float* t,
                        uint8_t* any) {
  return trace_rays_impl(handle, rays, n, inst, tri, t, any, 0);
}

uint32_t rpx_trace_rays_wave(uint32_t handle, const float* rays, uint32_t n, uint32_t stack_entries, uint32_t* inst,
                             uint32_t* tri, float* t, uint8_t* any) {
  if (stack_entries == 0 || stack_entries > 1024) {
    set_error("rpx_trace_rays_wave: stack_entries must be 1 .. 1024");
    return RPX_BAD_TILE;
  }
  return trace_rays_impl(handle, rays, n, inst, tri, t, any, (int)stack_entries);
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

static void stats_of(Engine& e, u64 samples, uint64_t* out4) {
  u64 r[4] = {0, 0, 0, 0};
  e.ray_counts(r);
  out4[0] = r[0] + r[1] + r[2] + r[3];
  out4[1] = r[0];
  out4[2] = r[1] + r[2];
  out4[3] = samples;
}

uint32_t rpx_scene_stats(uint32_t handle, uint64_t* out4) {
  if (!out4) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  stats_of(*h->engine, h->samples, out4);
  return RPX_OK;
}

uint32_t rpx_set_trip_budget(uint32_t handle, uint32_t trips) {
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  h->engine->opt.trip_budget = trips;
  return RPX_OK;
}

uint32_t rpx_trip_budget(uint32_t handle, uint32_t* out) {
  if (!out) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  *out = h->engine->opt.trip_budget;
  return RPX_OK;
}

uint32_t rpx_tune_trip_budget(uint32_t handle, uint32_t samples, float* out3) {
  if (!out3) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  u32 npix = h->dev.width * h->dev.height;
  u32 count = std::max<u32>(1, std::min<u32>(samples ? samples : 32, RESULTS_MAX / std::max<u32>(npix, 1)));
  DevSegment s = whole_image(h.get(), 0, 0, count);
  std::vector<float> ms;
  std::string err;
  prepare_engine(h.get());
  hipError_t r = h->engine->autotune(h->d_scene, std::vector<DevSegment>(1, s), ms, err, true);
  if (r != hipSuccess) {
    set_error(err);
    return r == hipErrorOutOfMemory ? RPX_GPU_MEMORY : RPX_GPU_LAUNCH;
  }
  out3[0] = (float)h->engine->opt.trip_budget;
  for (int b = 0; b < N_TUNE_BUDGETS; ++b) out3[1 + b] = ms[b];
  return RPX_OK;
}

uint32_t rpx_set_options(uint32_t handle, uint32_t pool, uint32_t shape, uint32_t shadow_shape, uint32_t group,
                         uint32_t chunk, uint32_t flags, uint32_t tail) {
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  apply_options(*h->engine, pool, shape, shadow_shape, group, chunk, flags, tail);
  return RPX_OK;
}

static uint32_t tune_report(Engine& e, hipError_t r, const std::vector<float>& ms, const std::string& err, float* out) {
  if (r != hipSuccess) {
    set_error(err);
    return r == hipErrorOutOfMemory ? RPX_GPU_MEMORY : RPX_GPU_LAUNCH;
  }
  out[0] = (float)e.opt.pool;
  out[1] = (float)e.opt.shape;
  out[2] = (float)e.opt.shadow_shape;
  out[3] = 0.0f;
  for (size_t k = 0; k < ms.size() && k < (size_t)N_TUNE_CANDIDATES; ++k) {
    out[4 + k] = ms[k];
    if (out[3] == 0.0f || (ms[k] > 0.0f && ms[k] < out[3])) out[3] = ms[k];
  }
  return RPX_OK;
}

uint32_t rpx_autotune(uint32_t handle, uint32_t samples, float* out60) {
  if (!out60) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  u32 npix = h->dev.width * h->dev.height;
  u32 count = std::max<u32>(1, std::min<u32>(samples ? samples : 32, RESULTS_MAX / std::max<u32>(npix, 1)));
  DevSegment s = whole_image(h.get(), 0, 0, count);
  std::vector<float> ms;
  std::string err;
  prepare_engine(h.get());
  hipError_t r = h->engine->autotune(h->d_scene, std::vector<DevSegment>(1, s), ms, err);
  return tune_report(*h->engine, r, ms, err, out60);
}

uint32_t rpx_set_profiling(uint32_t handle, uint32_t on) {
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  h->engine->profiling = on != 0;
  return RPX_OK;
}

uint32_t rpx_last_timing(uint32_t handle, float* out12) {
  if (!out12) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  fill_timing(*h->engine, out12);
  return RPX_OK;
}

uint32_t rpx_last_kernel_ms(uint32_t handle, float* out10) {
  if (!out10) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  const Timing& t = h->engine->timing;
  float* out8 = out10;
  out8[0] = t.stage_ms[6];
  out8[1] = (float)t.launches[6];
  out8[2] = t.stage_ms[2];
  out8[3] = (float)t.launches[2];
  out8[4] = t.stage_ms[0];
  out8[5] = t.stage_ms[5];
  out8[6] = t.stage_ms[4];
  out8[7] = t.stage_ms[1];
  out10[8] = t.stage_ms[7];
  out10[9] = (float)(h->photons ? h->photons->traced() : 0u);
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

// ------------------------------------------------------------------ differentiable parameters
uint32_t rpx_material_params(uint32_t handle, float* out, uint32_t cap) {
  auto h = find(handle);
  if (!h) return 0;
  std::lock_guard<std::mutex> lock(h->mutex);
  u32 n = (u32)h->host.materials.size();
  for (u32 m = 0; m < n && out && (m + 1) * N_MAT_PARAMS <= cap; ++m) {
    const HostMaterial& hm = h->host.materials[m];
    float* o = out + N_MAT_PARAMS * m;
    o[0] = hm.roughness;
    o[1] = hm.base_color.r;
    o[2] = hm.base_color.g;
    o[3] = hm.base_color.b;
    o[4] = hm.ior;
  }
  return n;
}

uint32_t rpx_set_material_params(uint32_t handle, const float* params, uint32_t n) {
  if (!params) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  if (n != h->host.materials.size()) {
    set_error("rpx_set_material_params: expected one parameter set per material");
    return RPX_BAD_SCENE;
  }
  DeviceTables* tables;
  std::string err;
  if (!device_tables(h->device, tables, err)) {
    set_error(err);
    return RPX_GPU_LAUNCH;
  }
  std::vector<Bsdf> bsdfs;
  std::vector<DevMaterial> mats;
  std::vector<DevMedium> media;
  std::vector<HostMaterial> updated = h->host.materials;
  for (u32 m = 0; m < n; ++m) {
    HostMaterial& hm = updated[m];
    const float* p = params + N_MAT_PARAMS * m;
    hm.roughness = p[0];
    hm.base_color = c3(p[1], p[2], p[3]);
    hm.ior = p[4];
    Bsdf sides[2];
    material_sides(tables->host_view(), hm, sides);
    bsdfs.push_back(sides[0]);
    bsdfs.push_back(sides[1]);
    mats.push_back(dev_material(hm));
    DevMedium md{};
    if (scatter_medium_of(hm, md) && !h->dev.ext) {
      // The scene's kernels were chosen without media (an index moved into a medium's range).
      set_error("rpx_set_material_params: the parameters give a material a scattering medium it did not have");
      return RPX_BAD_SCENE;
    }
    media.push_back(md);
  }
  h->host.materials = updated;
  if (h->host.caustic_photons > 0) {
    // Which instances cast caustics, and the photons' paths, follow the materials: the scene is
    // built again (its kept photon maps dropped with it).
    if (!build_device_scene(h.get(), err) || !upload_scene_table(h.get())) {
      set_error("rpx_set_material_params: " + err);
      return RPX_GPU_LAUNCH;
    }
    return RPX_OK;
  }
  hipStreamSynchronize(h->stream);
  hipError_t e = hipMemcpy(const_cast<Bsdf*>(h->dev.bsdfs), bsdfs.data(), bsdfs.size() * sizeof(Bsdf),
                           hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(const_cast<DevMaterial*>(h->dev.materials), mats.data(), mats.size() * sizeof(DevMaterial),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess && n)
    e = hipMemcpy(const_cast<DevMedium*>(h->dev.media), media.data(), media.size() * sizeof(DevMedium),
                  hipMemcpyHostToDevice);
  return e == hipSuccess ? RPX_OK : RPX_GPU_LAUNCH;
}

uint32_t rpx_set_texture(uint32_t handle, uint32_t material, uint32_t w, uint32_t hh, const float* rgb, float origin_x,
                         float origin_y, float scale_x, float scale_y) {
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  hipStreamSynchronize(h->stream);
  dev_free_all(h->tex_bufs);
  h->dev.texture = nullptr;
  h->dev.tex_material = NO_TEXTURE;
  h->dev.tex_w = h->dev.tex_h = 0;
  if (material != NO_TEXTURE) {
    if (!rgb) return RPX_NULL_POINTER;
    if (material >= h->host.materials.size() || w == 0 || hh == 0 || w > 65535 || hh > 65535) {
      set_error("rpx_set_texture: no such material, or an empty or oversized texture");
      return RPX_BAD_SCENE;
    }
    void* p = nullptr;
    size_t bytes = (size_t)w * hh * 3 * sizeof(float);
    if (!dev_alloc(h->tex_bufs, bytes, &p)) return RPX_GPU_MEMORY;
    if (hipMemcpy(p, rgb, bytes, hipMemcpyHostToDevice) != hipSuccess) return RPX_GPU_LAUNCH;
    h->dev.texture = (const float*)p;
    h->dev.tex_material = material;
    h->dev.tex_w = w;
    h->dev.tex_h = hh;
    h->dev.tex_origin_x = origin_x;
    h->dev.tex_origin_y = origin_y;
    h->dev.tex_scale_x = scale_x;
    h->dev.tex_scale_y = scale_y;
  }
  return upload_scene_table(h.get()) ? RPX_OK : RPX_GPU_LAUNCH;
}

uint32_t rpx_adjoint(uint32_t handle, uint32_t first, uint32_t count, const float* d_image, float* grad_materials,
                     float* grad_texture) {
  if (!d_image || !grad_materials) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  if (count == 0) return RPX_OK;
  std::lock_guard<std::mutex> lock(h->mutex);
  if (h->dev.ext) {
    set_error("rpx_adjoint: the replay follows area lights and the environment only (no sphere lights, suns, "
              "emissive meshes sampled as lights, caustics or scattering media)");
    return RPX_BAD_SCENE;
  }
  hipSetDevice(h->device);
  size_t px = (size_t)h->dev.width * h->dev.height;
  size_t n_mat = h->host.materials.size() * N_MAT_PARAMS;
  size_t n_tex = (size_t)h->dev.tex_w * h->dev.tex_h * 3;
  std::vector<DevBuf> tmp;
  void *d_adj = nullptr, *d_gm = nullptr, *d_gt = nullptr;
  if (!dev_alloc(tmp, px * 12, &d_adj) || !dev_alloc(tmp, std::max<size_t>(n_mat, 1) * 4, &d_gm) ||
      !dev_alloc(tmp, std::max<size_t>(n_tex, 1) * 4, &d_gt)) {
    dev_free_all(tmp);
    set_error("device memory exhausted for the adjoint");
    return RPX_GPU_MEMORY;
  }
  hipError_t e = hipMemcpyAsync(d_adj, d_image, px * 12, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipMemsetAsync(d_gm, 0, std::max<size_t>(n_mat, 1) * 4, h->stream);
  if (e == hipSuccess) e = hipMemsetAsync(d_gt, 0, std::max<size_t>(n_tex, 1) * 4, h->stream);
  std::string err;
  if (e == hipSuccess) {
    DevSegment s = whole_image(h.get(), 0, first, count);
    s.adjoint = (const float*)d_adj;
    // The image is the mean of the samples: d image / d sample = f32(1 / count) (image.rs:158-169).
    s.adjoint_scale = 1.0f / (float)count;
    h->engine->textured = h->dev.tex_material != NO_TEXTURE;
    h->engine->ext = false;
    e = h->engine->adjoint(h->d_scene, std::vector<DevSegment>(1, s), (float*)d_gm, (float*)d_gt, h->dev.tex_w,
                           h->dev.tex_h, err);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(grad_materials, d_gm, n_mat * 4, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess && grad_texture && n_tex)
    e = hipMemcpyAsync(grad_texture, d_gt, n_tex * 4, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  dev_free_all(tmp);
  if (e != hipSuccess) {
    set_error(err.empty() ? std::string("adjoint: ") + hipGetErrorString(e) : err);
    return e == hipErrorOutOfMemory ? RPX_GPU_MEMORY : RPX_GPU_LAUNCH;
  }
  return RPX_OK;
}

// ------------------------------------------------------------------ render queues
uint32_t rpx_queue_create(void) {
  if (rpx_device_count() == 0) {
    set_error("rpx_queue_create: no HIP device is available");
    return 0;
  }
  auto q = std::make_shared<Queue>();
  hipGetDevice(&q->device);
  if (hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("rpx_queue_create: cannot create a HIP stream");
    return 0;
  }
  std::string err;
  q->engine = std::make_unique<Engine>();
  if (!q->engine->init(q->device, q->stream, err)) {
    set_error("rpx_queue_create: " + err);
    return 0;
  }
  std::lock_guard<std::mutex> lock(g_registry_mutex);
  u32 id = g_next_queue++;
  if (g_next_queue == 0) g_next_queue = 1;
  g_queues[id] = q;
  return id;
}

uint32_t rpx_queue_add(uint32_t queue, uint32_t scene, uint32_t first, uint32_t count, uint32_t* job) {
  if (!job) return RPX_NULL_POINTER;
  auto q = find_queue(queue);
  auto h = find(scene);
  if (!q || !h) return RPX_BAD_HANDLE;
  if (h->device != q->device) {
    set_error("rpx_queue_add: the scene lives on another device");
    return RPX_BAD_SCENE;
  }
  std::lock_guard<std::mutex> lock(q->mutex);
  hipSetDevice(q->device);
  Job j;
  j.scene = h;
  j.first = first;
  j.count = count;
  {
    std::lock_guard<std::mutex> hl(h->mutex);
    if (h->photons) {
      set_error("rpx_queue_add: a scene with caustics renders its samples pass by pass, not in a queue");
      return RPX_BAD_SCENE;
    }
    j.width = h->dev.width;
    j.height = h->dev.height;
  }
  size_t px = (size_t)j.width * j.height;
  auto& B = q->film_bufs;
  Film& F = j.film;
  if (!dev_alloc(B, px * 12, (void**)&F.rgb) || !dev_alloc(B, px * 4, (void**)&F.weight) ||
      !dev_alloc(B, px * 4, (void**)&F.alpha) || !dev_alloc(B, px * 4, (void**)&F.lum2)) {
    set_error("device memory exhausted for a job's film");
    return RPX_GPU_MEMORY;
  }
  hipMemsetAsync(F.rgb, 0, px * 12, q->stream);
  hipMemsetAsync(F.weight, 0, px * 4, q->stream);
  hipMemsetAsync(F.alpha, 0, px * 4, q->stream);
  hipMemsetAsync(F.lum2, 0, px * 4, q->stream);
  if (hipStreamSynchronize(q->stream) != hipSuccess) return RPX_GPU_LAUNCH;
  *job = (u32)q->jobs.size();
  q->jobs.push_back(j);
  return RPX_OK;
}

uint32_t rpx_queue_run(uint32_t queue) {
  auto q = find_queue(queue);
  if (!q) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(q->mutex);
  hipSetDevice(q->device);
  if (q->jobs.empty()) return RPX_OK;
  // The scenes stay locked (in handle order) while the pool reads them.
  std::vector<Handle*> scenes;
  for (const Job& j : q->jobs) scenes.push_back(j.scene.get());
  std::sort(scenes.begin(), scenes.end());
  scenes.erase(std::unique(scenes.begin(), scenes.end()), scenes.end());
  std::vector<std::unique_lock<std::mutex>> locks;
  for (Handle* h : scenes) locks.emplace_back(h->mutex);
  std::vector<DevScene> table;
  std::map<Handle*, u32> index;
  for (Handle* h : scenes) {
    hipStreamSynchronize(h->stream);
    index[h] = (u32)table.size();
    table.push_back(h->dev);
  }
  std::vector<DevSegment> segs;
  for (const Job& j : q->jobs) {
    if (j.scene->dev.width != j.width || j.scene->dev.height != j.height) {
      set_error("rpx_queue_run: a job's scene changed its image size since it was queued");
      return RPX_BAD_SCENE;
    }
    DevSegment s{};
    s.scene = index[j.scene.get()];
    s.w = j.width;
    s.h = j.height;
    s.first = j.first;
    s.count = j.count;
    s.film = j.film;
    s.film_w = j.width;
    if (s.count) segs.push_back(s);
  }
  size_t bytes = table.size() * sizeof(DevScene);
  if (!q->scenes.p || q->scenes.bytes < bytes) {
    if (q->scenes.p) hipFree(q->scenes.p);
    q->scenes = DevBuf{};
    if (hipMalloc(&q->scenes.p, bytes) != hipSuccess) {
      q->scenes.p = nullptr;
      set_error("device memory exhausted for the scene table");
      return RPX_GPU_MEMORY;
    }
    q->scenes.bytes = bytes;
  }
  if (hipMemcpyAsync(q->scenes.p, table.data(), bytes, hipMemcpyHostToDevice, q->stream) != hipSuccess)
    return RPX_GPU_LAUNCH;
  std::string err;
  q->engine->textured = false;
  q->engine->ext = false;
  for (const DevScene& d : table) {
    q->engine->textured = q->engine->textured || d.tex_material != NO_TEXTURE;
    q->engine->ext = q->engine->ext || d.ext != 0;
  }
  hipError_t e = q->engine->run((const DevScene*)q->scenes.p, segs, true, err);
  if (e != hipSuccess) {
    set_error(err);
    return e == hipErrorOutOfMemory ? RPX_GPU_MEMORY : RPX_GPU_LAUNCH;
  }
  return RPX_OK;
}

uint32_t rpx_queue_reset(uint32_t queue) {
  auto q = find_queue(queue);
  if (!q) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(q->mutex);
  hipSetDevice(q->device);
  for (const Job& j : q->jobs) {
    size_t px = (size_t)j.width * j.height;
    hipMemsetAsync(j.film.rgb, 0, px * 12, q->stream);
    hipMemsetAsync(j.film.weight, 0, px * 4, q->stream);
    hipMemsetAsync(j.film.alpha, 0, px * 4, q->stream);
    hipMemsetAsync(j.film.lum2, 0, px * 4, q->stream);
  }
  return hipStreamSynchronize(q->stream) == hipSuccess ? RPX_OK : RPX_GPU_LAUNCH;
}

uint32_t rpx_queue_film_read(uint32_t queue, uint32_t job, float* rgb, float* weight, float* alpha, float* lum2) {
  auto q = find_queue(queue);
  if (!q) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(q->mutex);
  if (job >= q->jobs.size()) return RPX_BAD_HANDLE;
  hipSetDevice(q->device);
  const Job& j = q->jobs[job];
  return read_film(j.film, (size_t)j.width * j.height, q->stream, rgb, weight, alpha, lum2);
}

uint32_t rpx_queue_set_options(uint32_t queue, uint32_t pool, uint32_t shape, uint32_t shadow_shape,
                               uint32_t group, uint32_t chunk, uint32_t flags, uint32_t tail) {
  auto q = find_queue(queue);
  if (!q) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(q->mutex);
  apply_options(*q->engine, pool, shape, shadow_shape, group, chunk, flags, tail);
  return RPX_OK;
}

uint32_t rpx_queue_autotune(uint32_t queue, uint32_t paths, float* out60) {
  if (!out60) return RPX_NULL_POINTER;
  auto q = find_queue(queue);
  if (!q) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(q->mutex);
  hipSetDevice(q->device);
  if (q->jobs.empty()) return RPX_OK;
  std::vector<Handle*> scenes;
  for (const Job& j : q->jobs) scenes.push_back(j.scene.get());
  std::sort(scenes.begin(), scenes.end());
  scenes.erase(std::unique(scenes.begin(), scenes.end()), scenes.end());
  std::vector<std::unique_lock<std::mutex>> locks;
  for (Handle* h : scenes) locks.emplace_back(h->mutex);
  std::vector<DevScene> table;
  std::map<Handle*, u32> index;
  for (Handle* h : scenes) {
    hipStreamSynchronize(h->stream);
    index[h] = (u32)table.size();
    table.push_back(h->dev);
  }
  // The probe: the jobs' first samples, in queue order, up to `paths` path samples.
  u32 budget = std::min<u32>(paths ? paths : (1u << 25), RESULTS_MAX);
  std::vector<DevSegment> probe;
  u32 used = 0;
  for (const Job& j : q->jobs) {
    u32 npix = j.width * j.height;
    if (j.count == 0 || used + npix > budget) continue;
    DevSegment s{};
    s.scene = index[j.scene.get()];
    s.w = j.width;
    s.h = j.height;
    s.first = j.first;
    s.count = std::min(j.count, (budget - used) / npix);
    s.film = j.film;
    s.film_w = j.width;
    used += s.count * npix;
    probe.push_back(s);
  }
  if (probe.empty()) return RPX_OK;
  void* d = nullptr;
  if (hipMalloc(&d, table.size() * sizeof(DevScene)) != hipSuccess) return RPX_GPU_MEMORY;
  hipMemcpy(d, table.data(), table.size() * sizeof(DevScene), hipMemcpyHostToDevice);
  std::vector<float> ms;
  std::string err;
  q->engine->textured = false;
  q->engine->ext = false;
  for (const DevScene& ds : table) {
    q->engine->textured = q->engine->textured || ds.tex_material != NO_TEXTURE;
    q->engine->ext = q->engine->ext || ds.ext != 0;
  }
  hipError_t r = q->engine->autotune((const DevScene*)d, probe, ms, err);
  hipFree(d);
  return tune_report(*q->engine, r, ms, err, out60);
}

uint32_t rpx_queue_trip_budget(uint32_t queue, uint32_t* out) {
  if (!out) return RPX_NULL_POINTER;
  auto q = find_queue(queue);
  if (!q) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(q->mutex);
  *out = q->engine->opt.trip_budget;
  return RPX_OK;
}

uint32_t rpx_queue_set_profiling(uint32_t queue, uint32_t on) {
  auto q = find_queue(queue);
  if (!q) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(q->mutex);
  q->engine->profiling = on != 0;
  return RPX_OK;
}

uint32_t rpx_queue_last_timing(uint32_t queue, float* out12) {
  if (!out12) return RPX_NULL_POINTER;
  auto q = find_queue(queue);
  if (!q) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(q->mutex);
  fill_timing(*q->engine, out12);
  return RPX_OK;
}

uint32_t rpx_queue_stats(uint32_t queue, uint64_t* out4) {
  if (!out4) return RPX_NULL_POINTER;
  auto q = find_queue(queue);
  if (!q) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(q->mutex);
  hipSetDevice(q->device);
  u64 samples = 0;
  for (const Job& j : q->jobs) samples += j.count;
  stats_of(*q->engine, samples, out4);
  return RPX_OK;
}

uint32_t rpx_queue_free(uint32_t queue) {
  std::shared_ptr<Queue> q;
  {
    std::lock_guard<std::mutex> lock(g_registry_mutex);
    auto it = g_queues.find(queue);
    if (it == g_queues.end()) return RPX_BAD_HANDLE;
    q = it->second;
    g_queues.erase(it);
  }
  std::lock_guard<std::mutex> lock(q->mutex);
  hipSetDevice(q->device);
  return RPX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- passes and the denoiser
namespace rpx {
namespace {

constexpr u32 PASS_NORMAL = 2, PASS_DEPTH = 4, PASS_PART_ID = 8, PASS_MATERIAL_ID = 16, PASS_COC = 32,
              PASS_ALBEDO = 1, PASS_ALPHA = 64, PASS_ALL = 127;
constexpr u32 PASS_EXACT = PASS_ALL & ~(PASS_ALBEDO | PASS_ALPHA);

const u32* sobol_of(int device) {
  DeviceTables* t = nullptr;
  std::string err;
  return device_tables(device, t, err) ? t->sobol : nullptr;
}

// Device planes of n pixels: sums (albedo, alpha, weight) zero, the exact planes background
// (normal 0, depth +inf, ids 0, coc 0).
bool alloc_passes(std::vector<DevBuf>& B, size_t n, PassPlanes& Q, hipStream_t st) {
  if (!dev_alloc(B, n * 12, (void**)&Q.albedo) || !dev_alloc(B, n * 4, (void**)&Q.alpha) ||
      !dev_alloc(B, n * 4, (void**)&Q.weight) || !dev_alloc(B, n * 12, (void**)&Q.normal) ||
      !dev_alloc(B, n * 4, (void**)&Q.depth) || !dev_alloc(B, n * 4, (void**)&Q.part) ||
      !dev_alloc(B, n * 4, (void**)&Q.material) || !dev_alloc(B, n * 4, (void**)&Q.coc))
    return false;
  hipMemsetAsync(Q.albedo, 0, n * 12, st);
  hipMemsetAsync(Q.alpha, 0, n * 4, st);
  hipMemsetAsync(Q.weight, 0, n * 4, st);
  hipMemsetAsync(Q.normal, 0, n * 12, st);
  hipMemsetD32Async((hipDeviceptr_t)Q.depth, 0x7f800000, n, st);
  hipMemsetAsync(Q.part, 0, n * 4, st);
  hipMemsetAsync(Q.material, 0, n * 4, st);
  hipMemsetAsync(Q.coc, 0, n * 4, st);
  return true;
}

bool ensure_passes(Handle* h, std::string& err) {
  if (h->passes.albedo) return true;
  size_t n = (size_t)h->dev.width * h->dev.height;
  auto& B = h->pass_bufs;
  if (!alloc_passes(B, n, h->passes, h->stream) || !dev_alloc(B, n * 12, (void**)&h->pass_albedo) ||
      !dev_alloc(B, n * 16, (void**)&h->denoised)) {
    dev_free_all(B);
    h->passes = PassPlanes{};
    err = "device memory exhausted for the passes";
    return false;
  }
  return true;
}

// The resolved planes of `mask` (known bits, in the order of their bits) of an n-pixel pass film,
// one after another as little-endian words, to host memory.
hipError_t read_planes(const PassPlanes& Q, size_t n, u32 mask, uint8_t* out, hipStream_t st) {
  std::vector<DevBuf> tmp;
  float *albedo = nullptr, *alpha = nullptr;
  if ((mask & (PASS_ALBEDO | PASS_ALPHA)) &&
      (!dev_alloc(tmp, n * 12, (void**)&albedo) || !dev_alloc(tmp, n * 4, (void**)&alpha))) {
    dev_free_all(tmp);
    return hipErrorOutOfMemory;
  }
  hipError_t e = hipSuccess;
  if (albedo) e = launch_passes_resolve(Q, (u32)n, albedo, alpha, st);
  const void* planes[7] = {albedo, Q.normal, Q.depth, Q.part, Q.material, Q.coc, alpha};
  const size_t channels[7] = {3, 3, 1, 1, 1, 1, 1};
  size_t at = 0;
  for (int k = 0; k < 7 && e == hipSuccess; ++k) {
    if (!(mask & (1u << k))) continue;
    e = hipMemcpyAsync(out + at, planes[k], n * channels[k] * 4, hipMemcpyDeviceToHost, st);
    at += n * channels[k] * 4;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  dev_free_all(tmp);
  return e;
}

// The one-sample albedo (3 per pixel) or coverage of a tile (render.rs:558-640 for one sample).
uint32_t tile_sampled(uint32_t handle, uint32_t x, uint32_t y, uint32_t w, uint32_t hh, uint32_t sample_index, float* out,
                      bool albedo) {
  if (!out) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  if (w == 0 || hh == 0 || (u64)w * hh > (1ull << 22)) return RPX_BAD_TILE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  size_t n = (size_t)w * hh;
  std::vector<DevBuf> tmp;
  PassPlanes Q{};
  const u32* sobol = sobol_of(h->device);
  if (!sobol || !dev_alloc(tmp, n * 12, (void**)&Q.albedo) || !dev_alloc(tmp, n * 4, (void**)&Q.alpha) ||
      !dev_alloc(tmp, n * 4, (void**)&Q.weight)) {
    dev_free_all(tmp);
    set_error("device memory exhausted for the tile's passes");
    return RPX_GPU_MEMORY;
  }
  hipStream_t st = h->stream;
  hipMemsetAsync(Q.albedo, 0, n * 12, st);
  hipMemsetAsync(Q.alpha, 0, n * 4, st);
  hipMemsetAsync(Q.weight, 0, n * 4, st);
  hipError_t e = launch_passes(h->d_scene, sobol, Q, x, y, w, hh, x, y, w, sample_index, 1, PASSES_SAMPLED, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, albedo ? (void*)Q.albedo : (void*)Q.alpha, n * (albedo ? 12 : 4), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  dev_free_all(tmp);
  if (e != hipSuccess) {
    set_error(std::string("passes: ") + hipGetErrorString(e));
    return RPX_GPU_LAUNCH;
  }
  return RPX_OK;
}

// The device state of the handle-less denoiser calls: one per device, calls serialised.
struct DenoiseContext {
  std::mutex mutex;
  hipStream_t stream = nullptr;
  denoise::Denoiser dn;
  DevBuf in, out;
};
std::mutex g_dn_mutex;
std::map<int, std::unique_ptr<DenoiseContext>> g_dn;

DenoiseContext* denoise_context(int dev) {
  std::lock_guard<std::mutex> lock(g_dn_mutex);
  auto& c = g_dn[dev];
  if (!c) {
    auto fresh = std::make_unique<DenoiseContext>();
    if (hipStreamCreateWithFlags(&fresh->stream, hipStreamNonBlocking) != hipSuccess) return nullptr;
    c = std::move(fresh);
  }
  return c.get();
}

bool grow(DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes) return true;
  if (b.p) hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
  if (hipMalloc(&b.p, std::max<size_t>(bytes, 256)) != hipSuccess) return false;
  b.bytes = std::max<size_t>(bytes, 256);
  return true;
}

bool preset_of(u32 id, denoise::Preset& p) {
  if (id == 0) p = denoise::FULL;
  else if (id == 1) p = denoise::VIEWPORT;
  else return false;
  return true;
}

// rt-wasm/src/lib.rs:420-475 and rt-post/src/denoise.rs:322-407: the band's planes (`words` per
// pixel: 14, or 15 with each pixel's own sample count) in host memory, rows y0..y1 of RGBA sums out.
uint32_t denoise_band(u32 preset_id, float strength, u32 width, u32 height, u32 origin_y, u32 rows, u32 samples, u32 y0,
                      u32 y1, const uint8_t* in_ptr, float* out_ptr, u32 words) {
  if (!in_ptr || !out_ptr) return RPX_NULL_POINTER;
  denoise::Preset preset;
  if (!preset_of(preset_id, preset)) return RPX_BAD_DENOISE;
  preset = preset.clamped();
  // The band (denoise.rs:264-287) and the rows with their halo (denoise.rs:338-356).
  if (width == 0 || (u64)origin_y + rows > height) return RPX_BAD_DENOISE;
  u32 halo = preset.halo(), band_end = origin_y + rows;
  u32 lo = y0 > halo ? y0 - halo : 0;
  u64 hi = std::min<u64>((u64)y1 + halo, height);
  if (y0 > y1 || y0 < origin_y || y1 > band_end || origin_y > lo || band_end < hi) return RPX_BAD_DENOISE;
  size_t n = (size_t)width * rows, w4 = (size_t)width * 4;
  if (y1 == y0) return RPX_OK;
  if (rpx_device_count() == 0) {
    set_error("rt_denoise: no HIP device is available");
    return RPX_GPU_MEMORY;
  }
  int dev = 0;
  hipGetDevice(&dev);
  DenoiseContext* c = denoise_context(dev);
  if (!c) {
    set_error("rt_denoise: cannot create a HIP stream");
    return RPX_GPU_MEMORY;
  }
  std::lock_guard<std::mutex> lock(c->mutex);
  if (!grow(c->in, n * words * 4) || !grow(c->out, n * 16)) {
    set_error("rt_denoise: device memory exhausted for the band");
    return RPX_GPU_MEMORY;
  }
  hipStream_t st = c->stream;
  hipError_t e = hipMemcpyAsync(c->in.p, in_ptr, n * words * 4, hipMemcpyHostToDevice, st);
  const float* f = static_cast<const float*>(c->in.p);
  denoise::Planes P{};
  P.rgb = f;
  P.rgb_stride = 4;
  P.alpha = f + 3;
  P.alpha_stride = 4;
  P.lum2 = f + n * 4;
  P.albedo = f + n * 5;
  P.normal = f + n * 8;
  P.depth = f + n * 11;
  P.part = reinterpret_cast<const u32*>(f + n * 12);
  P.material = reinterpret_cast<const u32*>(f + n * 13);
  P.counts = words > 14 ? f + n * 14 : nullptr;
  float4* d_out = static_cast<float4*>(c->out.p);
  if (e == hipSuccess) e = c->dn.run(P, width, rows, samples, preset, strength, d_out, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out_ptr, d_out + (size_t)(y0 - origin_y) * width, (size_t)(y1 - y0) * w4 * 4,
                       hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    set_error(std::string("rt_denoise: ") + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? RPX_GPU_MEMORY : RPX_GPU_LAUNCH;
  }
  return RPX_OK;
}

}  // namespace
}  // namespace rpx

extern "C" {

uint32_t rt_render_tile_passes(uint32_t handle, uint32_t x, uint32_t y, uint32_t w, uint32_t hh, uint32_t mask,
                               uint8_t* out_ptr) {
  if (!out_ptr) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  if (w == 0 || hh == 0 || (u64)w * hh > (1ull << 22)) return RPX_BAD_TILE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  u32 m = mask & PASS_EXACT;
  if (m == 0) return RPX_OK;
  size_t n = (size_t)w * hh;
  std::vector<DevBuf> tmp;
  PassPlanes Q{};
  const u32* sobol = sobol_of(h->device);
  if (!sobol || !alloc_passes(tmp, n, Q, h->stream)) {
    dev_free_all(tmp);
    set_error("device memory exhausted for the tile's passes");
    return RPX_GPU_MEMORY;
  }
  hipError_t e = launch_passes(h->d_scene, sobol, Q, x, y, w, hh, x, y, w, 0, 0, PASSES_EXACT, h->stream);
  if (e == hipSuccess) e = read_planes(Q, n, m, out_ptr, h->stream);
  dev_free_all(tmp);
  if (e != hipSuccess) {
    set_error(std::string("rt_render_tile_passes: ") + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? RPX_GPU_MEMORY : RPX_GPU_LAUNCH;
  }
  return RPX_OK;
}

uint32_t rt_render_tile_albedo(uint32_t handle, uint32_t x, uint32_t y, uint32_t w, uint32_t h, uint32_t sample_index,
                               float* out_ptr) {
  return tile_sampled(handle, x, y, w, h, sample_index, out_ptr, true);
}

uint32_t rt_render_tile_alpha(uint32_t handle, uint32_t x, uint32_t y, uint32_t w, uint32_t h, uint32_t sample_index,
                              float* out_ptr) {
  return tile_sampled(handle, x, y, w, h, sample_index, out_ptr, false);
}

uint32_t rt_denoise(uint32_t width, uint32_t height, uint32_t origin_y, uint32_t rows, uint32_t samples, uint32_t y0,
                    uint32_t y1, const uint8_t* in_ptr, float* out_ptr) {
  return denoise_band(0, 1.0f, width, height, origin_y, rows, samples, y0, y1, in_ptr, out_ptr, 14);
}

uint32_t rt_denoise_with(uint32_t preset, float strength, uint32_t width, uint32_t height, uint32_t origin_y,
                         uint32_t rows, uint32_t samples, uint32_t y0, uint32_t y1, const uint8_t* in_ptr,
                         float* out_ptr) {
  return denoise_band(preset, strength, width, height, origin_y, rows, samples, y0, y1, in_ptr, out_ptr, 14);
}

uint32_t rt_denoise_halo(uint32_t preset) {
  denoise::Preset p;
  return preset_of(preset, p) ? p.halo() : 0;
}

uint32_t rpx_denoise_counted(uint32_t preset, float strength, uint32_t width, uint32_t height, uint32_t origin_y,
                             uint32_t rows, uint32_t samples, uint32_t y0, uint32_t y1, const uint8_t* in_ptr,
                             float* out_ptr) {
  return denoise_band(preset, strength, width, height, origin_y, rows, samples, y0, y1, in_ptr, out_ptr, 15);
}

uint32_t rpx_render_passes(uint32_t handle, uint32_t first, uint32_t count) {
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  std::string err;
  const u32* sobol = sobol_of(h->device);
  if (!sobol || !ensure_passes(h.get(), err)) {
    set_error(err);
    return RPX_GPU_MEMORY;
  }
  u32 W = h->dev.width, H = h->dev.height;
  hipError_t e = launch_passes(h->d_scene, sobol, h->passes, 0, 0, W, H, 0, 0, W, first, count,
                               PASSES_EXACT | (count ? PASSES_SAMPLED : 0u), h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    set_error(std::string("rpx_render_passes: ") + hipGetErrorString(e));
    return RPX_GPU_LAUNCH;
  }
  return RPX_OK;
}

uint32_t rpx_passes_reset(uint32_t handle) {
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  std::string err;
  if (!ensure_passes(h.get(), err)) {
    set_error(err);
    return RPX_GPU_MEMORY;
  }
  size_t n = (size_t)h->dev.width * h->dev.height;
  hipMemsetAsync(h->passes.albedo, 0, n * 12, h->stream);
  hipMemsetAsync(h->passes.alpha, 0, n * 4, h->stream);
  hipMemsetAsync(h->passes.weight, 0, n * 4, h->stream);
  return hipStreamSynchronize(h->stream) == hipSuccess ? RPX_OK : RPX_GPU_LAUNCH;
}

uint32_t rpx_passes_read(uint32_t handle, uint32_t mask, uint8_t* out) {
  if (!out) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  std::string err;
  if (!ensure_passes(h.get(), err)) {
    set_error(err);
    return RPX_GPU_MEMORY;
  }
  hipError_t e = read_planes(h->passes, (size_t)h->dev.width * h->dev.height, mask & PASS_ALL, out, h->stream);
  return e == hipSuccess ? RPX_OK : (e == hipErrorOutOfMemory ? RPX_GPU_MEMORY : RPX_GPU_LAUNCH);
}

uint32_t rpx_denoise_film(uint32_t handle, uint32_t preset, float strength, uint32_t samples, float* out) {
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  denoise::Preset p;
  if (!preset_of(preset, p)) return RPX_BAD_DENOISE;
  std::lock_guard<std::mutex> lock(h->mutex);
  hipSetDevice(h->device);
  std::string err;
  if (!ensure_film(h.get(), err) || !ensure_passes(h.get(), err)) {
    set_error(err);
    return RPX_GPU_MEMORY;
  }
  if (!h->denoiser) {
    h->denoiser = std::make_unique<denoise::Denoiser>();
    h->denoiser->profiling = true;
  }
  u32 W = h->dev.width, H = h->dev.height;
  size_t n = (size_t)W * H;
  if (samples == 0) samples = (u32)std::min<u64>(h->film_samples, 0xffffffffu);
  hipStream_t st = h->stream;
  hipError_t e = launch_passes_resolve(h->passes, (u32)n, h->pass_albedo, nullptr, st);
  denoise::Planes P{};
  P.rgb = h->film.rgb;
  P.rgb_stride = 3;
  P.alpha = h->film.alpha;
  P.alpha_stride = 1;
  P.lum2 = h->film.lum2;
  P.albedo = h->pass_albedo;
  P.normal = h->passes.normal;
  P.depth = h->passes.depth;
  P.part = h->passes.part;
  P.material = h->passes.material;
  P.counts = nullptr;
  if (e == hipSuccess) e = h->denoiser->run(P, W, H, samples, p, strength, h->denoised, st);
  if (e == hipSuccess && out) e = hipMemcpyAsync(out, h->denoised, n * 16, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    set_error(std::string("rpx_denoise_film: ") + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? RPX_GPU_MEMORY : RPX_GPU_LAUNCH;
  }
  return RPX_OK;
}

uint32_t rpx_denoise_timing(uint32_t handle, float* out8) {
  if (!out8) return RPX_NULL_POINTER;
  auto h = find(handle);
  if (!h) return RPX_BAD_HANDLE;
  std::lock_guard<std::mutex> lock(h->mutex);
  denoise::DenoiseTiming t = h->denoiser ? h->denoiser->timing : denoise::DenoiseTiming{};
  out8[0] = t.total_ms;
  out8[1] = t.setup_ms;
  out8[2] = t.firefly_ms;
  out8[3] = t.variance_ms;
  out8[4] = t.atrous_ms;
  out8[5] = t.final_ms;
  out8[6] = (float)t.levels;
  out8[7] = 0.0f;
  return RPX_OK;
}

}  // extern "C"
