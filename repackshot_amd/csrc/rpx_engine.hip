// Host side of the path pool: per-device tables, buffers, and the iteration loop.
//
// A run splits its segments into chunks of at most RESULTS_MAX work ids (whole sample passes of a
// segment), and for each chunk resets the pool, then launches groups of iterations
// (regen, extend, shade, shadow) until the extend kernel finds no live path (a flag read back
// after each group), then splats the chunk's samples into the films. Each chunk's films receive
// whole samples in increasing sample order, so the films do not depend on the chunking, the pool
// size or the launch shape.
#include <algorithm>
#include <map>
#include <mutex>
#include "rpx_engine.h"
#include "rpx_caustics.h"

namespace rpx {

bool dev_alloc(std::vector<DevBuf>& bufs, size_t bytes, void** out) {
  DevBuf b;
  if (hipMalloc(&b.p, std::max<size_t>(bytes, 16)) != hipSuccess) return false;
  b.bytes = bytes;
  bufs.push_back(b);
  *out = b.p;
  return true;
}

void dev_free_all(std::vector<DevBuf>& bufs) {
  for (DevBuf& b : bufs)
    if (b.p) hipFree(b.p);
  bufs.clear();
}

namespace {
std::mutex g_tables_mutex;
std::map<int, DeviceTables> g_tables;
}  // namespace

bool device_tables(int dev, DeviceTables*& out, std::string& err) {
  std::lock_guard<std::mutex> lock(g_tables_mutex);
  DeviceTables& t = g_tables[dev];
  if (!t.ready) {
    const size_t SL = 1024, CL = 17408;
    if (hipMalloc(&t.albedo, SL * 2) != hipSuccess || hipMalloc(&t.schlick, SL * 2) != hipSuccess ||
        hipMalloc(&t.clear_in, CL * 2) != hipSuccess || hipMalloc(&t.clear_out, CL * 2) != hipSuccess ||
        hipMalloc(&t.albedo_avg, 32 * 4) != hipSuccess || hipMalloc(&t.schlick_avg, 32 * 4) != hipSuccess ||
        hipMalloc(&t.sobol, 64 * 32 * 4) != hipSuccess) {
      err = "device memory for the energy and Sobol tables";
      return false;
    }
    std::vector<u32> cols(64 * 32);
    sobol_columns(cols.data());
    hipError_t e = hipMemcpy(t.sobol, cols.data(), cols.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_energy(t.albedo, t.schlick, t.clear_in, t.clear_out, t.albedo_avg, t.schlick_avg, 0);
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

Engine::~Engine() {
  if (stream_) hipStreamSynchronize(stream_);
  if (side_) {
    hipStreamSynchronize(side_);
    hipStreamDestroy(side_);
  }
  if (fork_) hipEventDestroy(fork_);
  if (join_) hipEventDestroy(join_);

  dev_free_all(pool_bufs_);
  dev_free_all(adjoint_bufs_);
  dev_free_all(bsdf_bufs_);
  dev_free_all(walk_bufs_);
  dev_free_all(cont_bufs_);
  dev_free_all(fixed_bufs_);
  if (results_.p) hipFree(results_.p);
  if (segs_.p) hipFree(segs_.p);
  if (idle_host_) hipHostFree(idle_host_);
  for (hipEvent_t e : ev_)
    if (e) hipEventDestroy(e);
  for (hipEvent_t e : events_) hipEventDestroy(e);
}

bool Engine::init(int device, hipStream_t stream, std::string& err) {
  device_ = device;
  stream_ = stream;
  DeviceTables* tables;
  if (!device_tables(device, tables, err)) return false;
  P.sobol = tables->sobol;
  if (!dev_alloc(fixed_bufs_, N_COUNTS * 4, (void**)&P.counts) || !dev_alloc(fixed_bufs_, 64, (void**)&P.rays)) {
    err = "device memory exhausted for the pool counters";
    return false;
  }
  if (hipHostMalloc((void**)&idle_host_, 64, hipHostMallocDefault) != hipSuccess || hipEventCreate(&ev_[0]) != hipSuccess ||
      hipEventCreate(&ev_[1]) != hipSuccess) {
    err = "cannot create the pool's host flag or events";
    return false;
  }
  if (hipStreamCreateWithFlags(&side_, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&fork_, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&join_, hipEventDisableTiming) != hipSuccess) {
    err = "cannot create the shadow rays' stream or its events";
    return false;
  }

  hipMemsetAsync(P.counts, 0, N_COUNTS * 4, stream_);
  hipMemsetAsync(P.rays, 0, 64, stream_);
  return hipStreamSynchronize(stream_) == hipSuccess;
}

bool Engine::ensure_pool(u32 n, std::string& err) {
  if (textured && n > bsdf_cap_) {
    hipStreamSynchronize(stream_);
    dev_free_all(bsdf_bufs_);
    bsdf_cap_ = 0;
    if (!dev_alloc(bsdf_bufs_, (size_t)n * sizeof(Bsdf), (void**)&P.bsdf_tmp)) {
      err = "device memory exhausted for the textured BSDF records";
      return false;
    }
    bsdf_cap_ = n;
  }
  if (ext && n > walk_cap_) {
    hipStreamSynchronize(stream_);
    dev_free_all(walk_bufs_);
    walk_cap_ = 0;
    size_t f4 = (size_t)n * sizeof(float4);
    if (!dev_alloc(walk_bufs_, f4, (void**)&P.walk_w) || !dev_alloc(walk_bufs_, f4, (void**)&P.walk_r) ||
        !dev_alloc(walk_bufs_, (size_t)PARK_MAX * 4, (void**)&P.parked)) {
      dev_free_all(walk_bufs_);
      err = "device memory exhausted for the media walks' state";
      return false;
    }
    walk_cap_ = n;
  }
  if (n <= capacity_) return true;
  hipStreamSynchronize(stream_);
  dev_free_all(pool_bufs_);
  capacity_ = 0;
  auto& B = pool_bufs_;
  size_t f4 = (size_t)n * sizeof(float4);
  bool ok = dev_alloc(B, f4, (void**)&P.ray_o) && dev_alloc(B, f4, (void**)&P.ray_d) && dev_alloc(B, f4, (void**)&P.hit) &&
            dev_alloc(B, f4, (void**)&P.thr) && dev_alloc(B, f4, (void**)&P.sum) && dev_alloc(B, f4, (void**)&P.med) &&
            dev_alloc(B, f4, (void**)&P.sel_p) && dev_alloc(B, f4, (void**)&P.sel_w0) &&
            dev_alloc(B, f4, (void**)&P.sel_w1) && dev_alloc(B, f4, (void**)&P.sh_o) &&
            dev_alloc(B, f4, (void**)&P.sh_d) && dev_alloc(B, f4, (void**)&P.sh_c) &&
            dev_alloc(B, f4, (void**)&P.ident) && dev_alloc(B, (size_t)n * 8, (void**)&P.key) &&
            dev_alloc(B, (size_t)n * 4, (void**)&P.live[0]) && dev_alloc(B, (size_t)n * 4, (void**)&P.live[1]) &&
            dev_alloc(B, (size_t)n * 4, (void**)&P.shadow_queue[0]) &&
            dev_alloc(B, (size_t)n * 4, (void**)&P.shadow_queue[1]) &&
            dev_alloc(B, (size_t)n * 4, (void**)&P.done_queue[0]) && dev_alloc(B, (size_t)n * 4, (void**)&P.done_queue[1]);
  if (!ok) {
    dev_free_all(pool_bufs_);
    err = "device memory exhausted for the path pool";
    return false;
  }
  capacity_ = n;
  return true;
}

bool Engine::ensure_cont(u32 n, std::string& err) {
  u32 cap = opt.trip_budget ? std::max<u32>(n / 4, 1024u) : 0u;
  if (cap <= cont_cap_) return true;
  hipStreamSynchronize(stream_);
  dev_free_all(cont_bufs_);
  P.cont = nullptr;
  P.cont_stack = nullptr;
  P.cont_cap = cont_cap_ = 0;
  if (!dev_alloc(cont_bufs_, (size_t)cap * 2 * sizeof(uint4), (void**)&P.cont) ||
      !dev_alloc(cont_bufs_, (size_t)cap * CONT_STACK * 4, (void**)&P.cont_stack)) {
    dev_free_all(cont_bufs_);
    P.cont = nullptr;
    P.cont_stack = nullptr;
    err = "device memory exhausted for the closest-hit continuations";
    return false;
  }
  P.cont_cap = cont_cap_ = cap;
  return true;
}

bool Engine::ensure_results(u32 n, std::string& err) {
  if (n <= results_cap_) return true;
  hipStreamSynchronize(stream_);
  if (results_.p) hipFree(results_.p);
  results_ = DevBuf{};
  results_cap_ = 0;
  if (hipMalloc(&results_.p, (size_t)n * sizeof(float4)) != hipSuccess) {
    results_.p = nullptr;
    err = "device memory exhausted for the sample results";
    return false;
  }
  results_cap_ = n;
  P.results = (float4*)results_.p;
  return true;
}

bool Engine::ensure_segs(u32 n, std::string& err) {
  if (n <= segs_cap_) return true;
  hipStreamSynchronize(stream_);
  if (segs_.p) hipFree(segs_.p);
  segs_ = DevBuf{};
  segs_cap_ = 0;
  if (hipMalloc(&segs_.p, (size_t)n * sizeof(DevSegment)) != hipSuccess) {
    segs_.p = nullptr;
    err = "device memory exhausted for the segment table";
    return false;
  }
  segs_cap_ = n;
  return true;
}

hipEvent_t Engine::next_event() {
  if (events_used_ == events_.size()) {
    hipEvent_t e;
    hipEventCreate(&e);
    events_.push_back(e);
  }
  return events_[events_used_++];
}

bool Engine::ray_counts(u64 out[4]) {
  return hipMemcpyAsync(out, P.rays, 32, hipMemcpyDeviceToHost, stream_) == hipSuccess &&
         hipStreamSynchronize(stream_) == hipSuccess;
}

hipError_t Engine::run_chunk(const DevScene* d_scenes, const DevSegment* chunk, u32 n_chunk, u32 total, bool splat,
                             bool adj, bool disjoint, DevSegment* seg_table) {
  DevSegment* table = seg_table ? seg_table : (DevSegment*)segs_.p;
  hipError_t e = hipMemcpyAsync(table, chunk, n_chunk * sizeof(DevSegment), hipMemcpyHostToDevice, stream_);
  if (e != hipSuccess) return e;
  Launch L;
  L.scenes = d_scenes;
  L.segs = table;
  L.n_segs = n_chunk;
  L.pool = std::min(opt.pool, total);
  L.shape = opt.shape;
  L.shadow_shape = opt.shadow_shape;
  // The drain runs the forward vertices only; the replay keeps the loop.
  L.ext = ext ? 1u : 0u;
  L.cau = photons ? 1u : 0u;
  L.tail = adj ? 0u : opt.tail;
  L.drain_lanes = std::min(std::max(opt.drain_lanes, 1u), 64u);
  L.park_at = std::min(opt.park_at, PARK_MAX);
  L.budget = (opt.trip_budget && P.cont) ? opt.trip_budget : 0u;
  L.stream = stream_;
  L.side = side_;
  e = launch_pool_init(P, L, total);
  if (e != hipSuccess) return e;
  u32 group = std::max<u32>(2, opt.group & ~1u);
  auto mark = [&](int stage) {
    if (profiling && !opt.graph) {
      marks_.push_back({stage, events_used_});
      hipEventRecord(next_event(), stream_);
    }
  };
  auto done = [&]() {
    if (profiling && !opt.graph) hipEventRecord(next_event(), stream_);
  };
  auto mark_side = [&](int stage) {
    if (profiling && !opt.graph) {
      marks_.push_back({stage, events_used_});
      hipEventRecord(next_event(), side_);
    }
  };
  auto done_side = [&]() {
    if (profiling && !opt.graph) hipEventRecord(next_event(), side_);
  };
  auto iterations = [&](u32 count) -> hipError_t {
    hipError_t r = hipSuccess;
    for (u32 k = 0; k < count && r == hipSuccess; ++k) {
      u32 parity = k & 1u;
      mark(3);
      r = launch_regen(P, L, parity, adj);
      done();
      if (r != hipSuccess) break;
      // The last iteration's shadow rays, beside this iteration's extension rays.
      r = hipEventRecord(fork_, stream_);
      if (r == hipSuccess) r = hipStreamWaitEvent(side_, fork_, 0);
      if (r != hipSuccess) break;
      mark_side(2);
      r = launch_shadow(P, L, parity ^ 1u, adj);
      done_side();
      if (r == hipSuccess) r = hipEventRecord(join_, side_);
      if (r != hipSuccess) break;
      // The extend stage's time spans both traversals (closest hit here, any hit beside it).
      mark(0);
      r = launch_extend(P, L, parity);
      done();  // the closest-hit kernel's own end (stage 0 spans three events: start, kernel, join)
      // Joined even after an error, so that no side work outlives the chunk unawaited.
      hipError_t w = hipStreamWaitEvent(stream_, join_, 0);
      if (r == hipSuccess) r = w;
      done();
      if (r != hipSuccess) break;
      if (L.ext) {
        mark(5);
        r = launch_walk(P, L, parity);
        done();
        if (r != hipSuccess) break;
      }
      if (L.tail) {
        mark(4);
        r = launch_finish(P, L, parity);
        done();
        if (r != hipSuccess) break;
      }
      mark(1);
      r = launch_shade(P, L, parity, adj);
      done();
    }
    return r;
  };
  hipGraphExec_t exec = nullptr;
  if (opt.graph && !adj) {
    hipGraph_t graph = nullptr;
    e = hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipError_t ce = iterations(group);
    e = hipStreamEndCapture(stream_, &graph);
    if (ce != hipSuccess) e = ce;
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (graph) hipGraphDestroy(graph);
    if (e != hipSuccess) return e;
  }
  for (;;) {
    e = exec ? hipGraphLaunch(exec, stream_) : iterations(group);
    if (e != hipSuccess) break;
    timing.iterations += group;
    e = hipMemcpyAsync(idle_host_, P.counts + C_IDLE, 4, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    if (e != hipSuccess || *idle_host_) break;
  }
  if (exec) hipGraphExecDestroy(exec);
  if (e != hipSuccess) return e;
  if (splat && disjoint && n_chunk > 1) {
    e = launch_splat_segments(P, L.segs, L.n_segs, stream_);
    if (e != hipSuccess) return e;
  } else if (splat)
    for (u32 k = 0; k < n_chunk; ++k) {
      e = launch_splat(P, chunk[k], stream_);
      if (e != hipSuccess) return e;
    }
  timing.chunks++;
  return hipSuccess;
}

namespace {
// Hands out 256-byte aligned pieces of one device region (nullptr base: only counts the bytes).
struct Carver {
  char* base;
  size_t cap, used = 0;
  bool ok = true;
  template <class T>
  T* take(size_t n) {
    size_t bytes = (n * sizeof(T) + 255) & ~(size_t)255;
    if (used + bytes > cap) {
      ok = false;
      return nullptr;
    }
    T* p = base ? reinterpret_cast<T*>(base + used) : nullptr;
    used += bytes;
    return p;
  }
};

// The per-path buffers of an n-slot pool that runs n paths at once, its results, its segment
// table and an n x channels staging tile, carved from `c` into Q.
void carve_pool(Carver& c, PathState& Q, u32 n, bool textured, bool ext, u32 channels, DevSegment*& table,
                float*& staging) {
  Q.ray_o = c.take<float4>(n);
  Q.ray_d = c.take<float4>(n);
  Q.hit = c.take<uint4>(n);
  Q.thr = c.take<float4>(n);
  Q.sum = c.take<float4>(n);
  Q.med = c.take<float4>(n);
  Q.sel_p = c.take<float4>(n);
  Q.sel_w0 = c.take<float4>(n);
  Q.sel_w1 = c.take<float4>(n);
  Q.sh_o = c.take<float4>(n);
  Q.sh_d = c.take<float4>(n);
  Q.sh_c = c.take<float4>(n);
  Q.ident = c.take<uint4>(n);
  Q.key = c.take<u64>(n);
  for (int k = 0; k < 2; ++k) {
    Q.live[k] = c.take<u32>(n);
    Q.shadow_queue[k] = c.take<u32>(n);
    Q.done_queue[k] = c.take<u32>(n);
  }
  Q.results = c.take<float4>(n);
  Q.bsdf_tmp = textured ? c.take<Bsdf>(n) : nullptr;
  Q.walk_w = ext ? c.take<float4>(n) : nullptr;
  Q.walk_r = ext ? c.take<float4>(n) : nullptr;
  Q.parked = ext ? c.take<u32>(PARK_MAX) : nullptr;
  table = c.take<DevSegment>(1);
  staging = c.take<float>((size_t)n * channels);
}
}  // namespace

size_t Engine::tile_scratch_bytes(u32 pixels, bool textured, bool ext, u32 channels) {
  Carver c{nullptr, ~(size_t)0};
  PathState Q{};
  DevSegment* table;
  float* staging;
  carve_pool(c, Q, pixels, textured, ext, channels, table, staging);
  return c.used;
}

hipError_t Engine::render_tile_in(const DevScene* d_scenes, const DevSegment& seg, u32 channels, void* scratch,
                                  size_t bytes, float* out, size_t out_row_floats, std::string& err) {
  u32 n = seg.w * seg.h;
  if (n == 0 || seg.count != 1) {
    err = "one sample of a non-empty tile";
    return hipErrorInvalidValue;
  }
  Carver c{static_cast<char*>(scratch), bytes};
  PathState keep = P;
  DevSegment* table = nullptr;
  float* staging = nullptr;
  carve_pool(c, P, n, textured, ext, channels, table, staging);
  if (!c.ok) {
    P = keep;
    err = "the tile's path state does not fit the scratch region";
    return hipErrorInvalidValue;
  }
  EngineOptions keep_opt = opt;
  bool keep_profiling = profiling;
  // Every path in flight at once, launched eagerly (a captured graph would allocate).
  opt.pool = n;
  opt.graph = false;
  opt.per_pass = false;
  profiling = false;
  timing = Timing{};
  DevSegment s = seg;
  s.base = 0;
  hipError_t e = hipSuccess;
  // The photon maps' memory is the scene's, set aside when it was loaded.
  if (photons) e = prepare_photons(d_scenes, s, false, err);
  if (e == hipSuccess) e = run_chunk(d_scenes, &s, 1, n, false, false, false, table);
  if (e == hipSuccess) e = launch_tile_out(P, staging, seg.w, seg.h, channels, stream_);
  if (e == hipSuccess)
    e = hipMemcpy2DAsync(out, out_row_floats * 4, staging, (size_t)seg.w * channels * 4, (size_t)seg.w * channels * 4,
                         seg.h, hipMemcpyDeviceToHost, stream_);
  hipError_t w = hipStreamSynchronize(stream_);
  if (e == hipSuccess) e = w;
  P = keep;
  opt = keep_opt;
  profiling = keep_profiling;
  if (e != hipSuccess) err = std::string("render: ") + hipGetErrorString(e);
  return e;
}

bool Engine::plan(const std::vector<DevSegment>& segs, std::vector<std::vector<DevSegment>>& chunks,
                  std::vector<u32>& totals, std::string& err) {
  // Chunks of whole sample passes, at most `limit` work ids each.
  const u32 limit = std::min(std::max<u32>(opt.chunk, 1), RESULTS_MAX);
  chunks.clear();
  totals.clear();
  std::vector<DevSegment> cur;
  u32 base = 0;
  auto flush = [&]() {
    if (!cur.empty()) {
      chunks.push_back(cur);
      totals.push_back(base);
      cur.clear();
      base = 0;
    }
  };
  for (const DevSegment& s : segs) {
    u32 npix = s.w * s.h;
    if (npix == 0 || npix > limit) {
      err = "a segment's pixel rectangle is empty or larger than the results buffer";
      return false;
    }
    u32 k0 = 0;
    while (k0 < s.count) {
      u32 room = (limit - base) / npix;
      if (opt.per_pass) room = std::min<u32>(room, 1);
      if (photons) room = std::min<u32>(room, MAX_PASS_MAPS);  // the passes' maps are kept together
      if (room == 0) {
        flush();
        continue;
      }
      u32 take = std::min(room, s.count - k0);
      DevSegment piece = s;
      piece.first = s.first + k0;
      piece.count = take;
      piece.base = base;
      cur.push_back(piece);
      base += take * npix;
      k0 += take;
      if (opt.per_pass || photons) flush();  // a chunk of a scene with caustics: one segment's passes
    }
  }
  flush();
  return true;
}

hipError_t Engine::prepare_photons(const DevScene* d_scenes, const DevSegment& piece, bool may_allocate,
                                   std::string& err) {
  u32 passes[MAX_PASS_MAPS];
  for (u32 i = 0; i < piece.count && i < MAX_PASS_MAPS; ++i)
    passes[i] = photon_pass >= 0 ? (u32)photon_pass : piece.first + i;
  if (profiling) {
    marks_.push_back({7, events_used_});
    hipEventRecord(next_event(), stream_);
  }
  hipError_t e = photons->prepare(d_scenes + piece.scene, piece.first, passes, piece.count, stream_, err, may_allocate);
  if (profiling) hipEventRecord(next_event(), stream_);
  return e;
}

hipError_t Engine::run(const DevScene* d_scenes, const std::vector<DevSegment>& segs, bool splat, std::string& err,
                       bool disjoint) {
  timing = Timing{};
  events_used_ = 0;
  marks_.clear();
  std::vector<std::vector<DevSegment>> chunks;
  std::vector<u32> totals;
  if (!plan(segs, chunks, totals, err)) return hipErrorInvalidValue;
  if (!splat && chunks.size() > 1) {
    err = "unsplatted results must fit one chunk";
    return hipErrorInvalidValue;
  }
  u32 max_total = 0, max_segs = 0;
  for (size_t c = 0; c < chunks.size(); ++c) {
    max_total = std::max(max_total, totals[c]);
    max_segs = std::max<u32>(max_segs, (u32)chunks[c].size());
  }
  if (!ensure_pool(std::min(opt.pool, max_total), err) || !ensure_results(max_total, err) ||
      !ensure_segs(max_segs, err) || !ensure_cont(std::min(opt.pool, max_total), err))
    return hipErrorOutOfMemory;
  hipEventRecord(ev_[0], stream_);
  hipError_t e = hipSuccess;
  for (size_t c = 0; c < chunks.size() && e == hipSuccess; ++c) {
    if (photons) e = prepare_photons(d_scenes, chunks[c][0], true, err);
    if (e == hipSuccess)
      e = run_chunk(d_scenes, chunks[c].data(), (u32)chunks[c].size(), totals[c], splat, false, disjoint);
  }
  hipEventRecord(ev_[1], stream_);
  if (e == hipSuccess) e = hipStreamSynchronize(stream_);
  if (e != hipSuccess) {
    err = std::string("render: ") + hipGetErrorString(e);
    return e;
  }
  hipEventElapsedTime(&timing.total_ms, ev_[0], ev_[1]);
  gather_marks();
  return hipSuccess;
}

void Engine::gather_marks() {
  for (auto& m : marks_) {
    float ms = 0;
    if (m.first == 0) {
      // The closest-hit stage: its kernel alone, and the span until the shadow kernel beside it
      // has joined.
      hipEventElapsedTime(&ms, events_[m.second], events_[m.second + 1]);
      timing.stage_ms[6] += ms;
      timing.launches[6]++;
      hipEventElapsedTime(&ms, events_[m.second], events_[m.second + 2]);
    } else {
      hipEventElapsedTime(&ms, events_[m.second], events_[m.second + 1]);
    }
    timing.stage_ms[m.first] += ms;
    timing.launches[m.first]++;
  }
}

bool Engine::ensure_adjoint(u32 n, std::string& err) {
  if (n <= adjoint_cap_) return true;
  hipStreamSynchronize(stream_);
  dev_free_all(adjoint_bufs_);
  adjoint_cap_ = 0;
  auto& B = adjoint_bufs_;
  size_t f4 = (size_t)n * sizeof(float4);
  bool ok = dev_alloc(B, f4, (void**)&P.adj) && dev_alloc(B, f4, (void**)&P.ltot) &&
            dev_alloc(B, f4 * N_MAT_PARAMS, (void**)&P.dl) && dev_alloc(B, f4, (void**)&P.tgt) &&
            dev_alloc(B, 2 * f4, (void**)&P.gn);
  if (!ok) {
    dev_free_all(adjoint_bufs_);
    err = "device memory exhausted for the adjoint's path state";
    return false;
  }
  adjoint_cap_ = n;
  return true;
}

hipError_t Engine::adjoint(const DevScene* d_scenes, const std::vector<DevSegment>& segs, float* grad_mat,
                           float* grad_tex, u32 tex_w, u32 tex_h, std::string& err) {
  timing = Timing{};
  events_used_ = 0;
  marks_.clear();
  EngineOptions keep = opt;
  opt.per_pass = false;
  opt.graph = false;
  std::vector<std::vector<DevSegment>> chunks;
  std::vector<u32> totals;
  bool planned = plan(segs, chunks, totals, err);
  opt = keep;
  if (!planned) return hipErrorInvalidValue;
  u32 max_total = 0, max_segs = 0;
  for (size_t c = 0; c < chunks.size(); ++c) {
    max_total = std::max(max_total, totals[c]);
    max_segs = std::max<u32>(max_segs, (u32)chunks[c].size());
  }
  u32 pool = std::min(opt.pool, max_total);
  if (!ensure_pool(pool, err) || !ensure_adjoint(pool, err) || !ensure_results(max_total, err) ||
      !ensure_segs(max_segs, err))
    return hipErrorOutOfMemory;
  P.grad_mat = grad_mat;
  P.grad_tex = grad_tex;
  P.tex_w = tex_w;
  P.tex_h = tex_h;
  bool keep_graph = opt.graph;
  opt.graph = false;
  hipEventRecord(ev_[0], stream_);
  hipError_t e = hipSuccess;
  for (size_t c = 0; c < chunks.size() && e == hipSuccess; ++c) {
    // The forward (radiance per path sample), then the replay.
    e = run_chunk(d_scenes, chunks[c].data(), (u32)chunks[c].size(), totals[c], false, false);
    if (e == hipSuccess)
      e = run_chunk(d_scenes, chunks[c].data(), (u32)chunks[c].size(), totals[c], false, true);
  }
  hipEventRecord(ev_[1], stream_);
  opt.graph = keep_graph;
  if (e == hipSuccess) e = hipStreamSynchronize(stream_);
  P.grad_mat = nullptr;
  P.grad_tex = nullptr;
  if (e != hipSuccess) {
    err = std::string("adjoint: ") + hipGetErrorString(e);
    return e;
  }
  hipEventElapsedTime(&timing.total_ms, ev_[0], ev_[1]);
  gather_marks();
  return hipSuccess;
}

hipError_t Engine::autotune(const DevScene* d_scenes, const std::vector<DevSegment>& probe, std::vector<float>& ms,
                            std::string& err, bool budget_only) {
  EngineOptions keep = opt;
  bool keep_profiling = profiling;
  profiling = false;
  opt.per_pass = false;
  opt.graph = false;
  opt.chunk = RESULTS_MAX;  // the probe's results stay unsplatted: one chunk
  // A scene with caustics: the candidates render the probe's first MAX_PASS_MAPS passes against
  // their photon maps, made first.
  PhotonPasses* keep_photons = photons;
  std::vector<DevSegment> probe_run = probe;
  if (photons && !probe_run.empty()) {
    probe_run.resize(1);
    probe_run[0].count = std::min<u32>(probe_run[0].count, MAX_PASS_MAPS);
    hipError_t b = prepare_photons(d_scenes, probe_run[0], true, err);
    if (b == hipSuccess) b = hipStreamSynchronize(stream_);
    if (b != hipSuccess) {
      opt = keep;
      profiling = keep_profiling;
      return b;
    }
  }
  photons = nullptr;
  // 0: the candidate did not fit in device memory
  ms.assign(budget_only ? N_TUNE_BUDGETS : N_TUNE_CANDIDATES + N_TUNE_BUDGETS, 0.0f);
  float* budget_ms = ms.data() + (budget_only ? 0 : N_TUNE_CANDIDATES);
  float best = 0.0f;
  EngineOptions chosen = keep;
  hipError_t e = hipSuccess;
  // A candidate whose pool does not fit in device memory is skipped, not fatal.
  auto time_it = [&](float& t) -> bool {
    e = run(d_scenes, probe_run, false, err);  // warm (allocations, caches)
    if (e == hipSuccess) e = run(d_scenes, probe_run, false, err);
    if (e == hipErrorOutOfMemory) {
      e = hipSuccess;
      err.clear();
      release_pool();
      return false;
    }
    t = timing.total_ms;
    return e == hipSuccess;
  };
  opt.shadow_shape = budget_only ? keep.shadow_shape : 0;
  for (int sh = 0; sh < N_SHAPES && e == hipSuccess && !budget_only; ++sh)
    for (int p = 0; p < N_TUNE_POOLS && e == hipSuccess; ++p) {
      opt.shape = sh;
      opt.pool = TUNE_POOLS[p];
      float t = 0.0f;
      if (!time_it(t)) continue;
      ms[sh * N_TUNE_POOLS + p] = t;
      if (best == 0.0f || t < best) {
        best = t;
        chosen.shape = sh;
        chosen.pool = TUNE_POOLS[p];
        chosen.shadow_shape = 0;
      }
    }
  if (e == hipSuccess && best == 0.0f && !budget_only) {
    err = "no tuner candidate fits in device memory";
    e = hipErrorOutOfMemory;
  }
  opt.shape = chosen.shape;
  opt.pool = chosen.pool;
  for (int sh = 0; sh < N_SHAPES && e == hipSuccess && !budget_only; ++sh) {
    opt.shadow_shape = sh;
    float t = 0.0f;
    if (!time_it(t)) continue;
    ms[N_SHAPES * N_TUNE_POOLS + sh] = t;
    if (t < best) {
      best = t;
      chosen.shadow_shape = sh;
    }
  }
  // The trip budgets, each timed afresh (the stages above ran at the handle's own budget).
  opt.shadow_shape = chosen.shadow_shape;
  float best_budget = 0.0f;
  for (int b = 0; b < N_TUNE_BUDGETS && e == hipSuccess; ++b) {
    opt.trip_budget = TUNE_BUDGETS[b];
    float t = 0.0f;
    if (!time_it(t)) continue;
    budget_ms[b] = t;
    if (best_budget == 0.0f || t < best_budget) {
      best_budget = t;
      chosen.trip_budget = TUNE_BUDGETS[b];
    }
  }
  if (e == hipSuccess && best_budget == 0.0f) {
    err = "no tuner candidate fits in device memory";
    e = hipErrorOutOfMemory;
  }
  profiling = keep_profiling;
  opt = keep;
  photons = keep_photons;
  // The probes grew the pool to the largest candidate; keep no more than the chosen size.
  if (capacity_ > chosen.pool) release_pool();
  if (e != hipSuccess) return e;
  opt.shape = chosen.shape;
  opt.shadow_shape = chosen.shadow_shape;
  opt.pool = chosen.pool;
  opt.trip_budget = chosen.trip_budget;
  return hipSuccess;
}

void Engine::release_pool() {
  hipStreamSynchronize(stream_);
  dev_free_all(pool_bufs_);
  capacity_ = 0;
}

}  // namespace rpx
