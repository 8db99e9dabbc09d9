// The GPU tracer's kernels (gfx950, wave64), in wavefront form over a resident pool of path
// slots: one kernel per stage, each grid-striding over a compacted queue whose length lives in
// device memory, so the host never waits between stages.
//
//   regen    finished slots hand their radiance to the results buffer and take new work ids
//            (camera ray: sampler key, filter offset, pinhole / thin lens / orthographic)
//   extend   closest hit over the 4-wide BVHs (instance BVH, then the mesh's), LDS stack
//   shade    surface details, emission, MIS-weighted light crossings, next-event estimation
//            (light selection + sample + BSDF eval, a shadow ray queued), BSDF sample, Russian
//            roulette; survivors go to the next live queue, finished paths to the done queue
//   shadow   any hit of the queued shadow rays; unblocked ones add their contribution
//   splat    per-pixel film sums, samples added in sample-index order
//
// Queues are compacted per block: a ballot per wave, each lane's slot from the popcount of the lanes
// below it (mbcnt), a prefix over the block's waves in LDS and one atomic per block. New paths take
// consecutive work ids and queue entries without atomics (k_regen, then k_bookkeep moves the
// counters). Paths of different samples, bounces
// and scenes share a launch: every lane computes its own Sobol points from the generator columns
// and its sample index, and scene constants are read per distinct scene of a wave (a waterfall
// over the wave's scene ids, one pass when the wave holds one scene).
//
// Reference (paths inside engine/crates/rt-core/src): integrator/mod.rs:1140-1891 (the path
// loop), light.rs:1026-1551 (select, pick_at, sample_selected, crossed_from, escaped),
// environment.rs:293-331 (ramp sampling), geometry/traverse.rs (slab test with the robust far
// scale, tie-break), geometry/triangle.rs:113-154 (watertight test), geometry/mod.rs:682-808
// (hit details, terminator lift), bevel.rs:232-252 (shading normal), render.rs:79-97 and
// scene/mod.rs:294-357 (camera), image.rs:81-94 (film).
#include <algorithm>
#include "rpx_kernels.h"
#include "rpx_bsdf_grad.h"

namespace rpx {

const TraversalShape SHAPES[N_SHAPES] = {{128, 24, 1, 0}, {128, 32, 1, 0}, {128, 16, 1, 0}, {256, 16, 1, 0},
                                         {128, 16, 8, 0}, {128, 24, 1, 1}, {128, 16, 1, 1}, {256, 16, 1, 1}};

// ================================================================== traversal
namespace {
constexpr int STACK_TOTAL = 64;  // entries per lane; the host refuses deeper trees

// The geometry a lane traverses (its scene's arrays and instance-BVH root).
struct Geo {
  const Node4* nodes;
  const float4* tris;
  const DevInstance* instances;
  u32 root;
};
__device__ __forceinline__ Geo geo_of(const DevScene* scenes, u32 sid) {
  const DevScene& S = scenes[sid];
  return Geo{S.nodes, S.tris, S.instances, S.tlas_root};
}

// PreparedRay (triangle.rs:41-91): origin, reciprocal direction, octant, and the shear of the
// watertight test, whose axis permutation is held as lane masks (k*0: axis is x, k*1: axis is y).
// Scalars only, so the whole ray stays in registers.
struct PRay {
  float ox, oy, oz;
  float ix, iy, iz;
  float sx, sy, sz;
  float t_min, t_max;
  bool nx, ny, nz;
  bool kx0, kx1, ky0, ky1, kz0, kz1;
};

__device__ __forceinline__ float sel3(float a, float b, float c, int k) { return k == 0 ? a : (k == 1 ? b : c); }
__device__ __forceinline__ float pick(float a, float b, float c, bool k0, bool k1) { return k0 ? a : (k1 ? b : c); }

__device__ __forceinline__ PRay prepare(V3 o, V3 d, float t_min, float t_max) {
  PRay r;
  r.ix = 1.0f / d.x;
  r.iy = 1.0f / d.y;
  r.iz = 1.0f / d.z;
  float a0 = fabs_(d.x), a1 = fabs_(d.y), a2 = fabs_(d.z);
  int kz = (a0 >= a1 && a0 >= a2) ? 0 : (a1 >= a2 ? 1 : 2);
  // (kz + 1) % 3 and (kz + 2) % 3, swapped for a negative d[kz] (keeps the winding's sign)
  int kx = kz == 2 ? 0 : kz + 1, ky = kz == 0 ? 2 : kz - 1;
  if (sel3(d.x, d.y, d.z, kz) < 0.0f) {
    int t = kx;
    kx = ky;
    ky = t;
  }
  float szz = sel3(r.ix, r.iy, r.iz, kz);
  r.ox = o.x;
  r.oy = o.y;
  r.oz = o.z;
  r.nx = sign_negative(r.ix);
  r.ny = sign_negative(r.iy);
  r.nz = sign_negative(r.iz);
  r.kx0 = kx == 0;
  r.kx1 = kx == 1;
  r.ky0 = ky == 0;
  r.ky1 = ky == 1;
  r.kz0 = kz == 0;
  r.kz1 = kz == 1;
  r.sx = sel3(d.x, d.y, d.z, kx) * szz;
  r.sy = sel3(d.x, d.y, d.z, ky) * szz;
  r.sz = szz;
  r.t_min = t_min;
  r.t_max = fmin_r(t_max, F32_MAX);
  return r;
}

// Edge functions of the watertight test (triangle.rs:113-154) for vertices given relative to the
// ray origin (the same differences as permuting first, then subtracting): u, v, w and the sheared
// z of each vertex.
__device__ __forceinline__ void shear(const PRay& r, V3 pa, V3 pb, V3 pc, float& u, float& v, float& w, float& az,
                                      float& bz, float& cz) {
  float a_x = pick(pa.x, pa.y, pa.z, r.kx0, r.kx1), a_y = pick(pa.x, pa.y, pa.z, r.ky0, r.ky1),
        a_z = pick(pa.x, pa.y, pa.z, r.kz0, r.kz1);
  float b_x = pick(pb.x, pb.y, pb.z, r.kx0, r.kx1), b_y = pick(pb.x, pb.y, pb.z, r.ky0, r.ky1),
        b_z = pick(pb.x, pb.y, pb.z, r.kz0, r.kz1);
  float c_x = pick(pc.x, pc.y, pc.z, r.kx0, r.kx1), c_y = pick(pc.x, pc.y, pc.z, r.ky0, r.ky1),
        c_z = pick(pc.x, pc.y, pc.z, r.kz0, r.kz1);
  float ax = a_x - r.sx * a_z, ay = a_y - r.sy * a_z;
  float bx = b_x - r.sx * b_z, by = b_y - r.sy * b_z;
  float cx = c_x - r.sx * c_z, cy = c_y - r.sy * c_z;
  u = cx * by - cy * bx;
  v = ax * cy - ay * cx;
  w = bx * ay - by * ax;
  az = r.sz * a_z;
  bz = r.sz * b_z;
  cz = r.sz * c_z;
}

// The watertight test: true with t when the triangle is hit in [t_min, t_max].
__device__ __forceinline__ bool watertight(const PRay& r, float4 va, float4 vb, float4 vc, float t_max, bool cull_back,
                                           float& t_out) {
  V3 pa = v3(va.x - r.ox, va.y - r.oy, va.z - r.oz);
  V3 pb = v3(vb.x - r.ox, vb.y - r.oy, vb.z - r.oz);
  V3 pc = v3(vc.x - r.ox, vc.y - r.oy, vc.z - r.oz);
  float u, v, w, az, bz, cz;
  shear(r, pa, pb, pc, u, v, w, az, bz, cz);
  bool negative = u < 0.0f || v < 0.0f || w < 0.0f;
  bool positive = u > 0.0f || v > 0.0f || w > 0.0f;
  bool rejected = cull_back ? negative : (negative && positive);
  float det = u + v + w;
  if (rejected || det == 0.0f) return false;
  float t = (u * az + v * bz + w * cz) / det;
  t_out = t;
  return t >= r.t_min && t <= t_max;
}

// Slab test of one child (traverse.rs:184-222): the near distance, and whether near <= far.
__device__ __forceinline__ bool slab(const PRay& r, float lx, float hx, float ly, float hy, float lz, float hz,
                                     float t_max, float& near) {
  const float ROBUST = 1.0f + 4.0f * F32_EPSILON;
  float x0 = r.nx ? hx : lx, x1 = r.nx ? lx : hx;
  float y0 = r.ny ? hy : ly, y1 = r.ny ? ly : hy;
  float z0 = r.nz ? hz : lz, z1 = r.nz ? lz : hz;
  float tx0 = (x0 - r.ox) * r.ix, ty0 = (y0 - r.oy) * r.iy, tz0 = (z0 - r.oz) * r.iz;
  float tx1 = (x1 - r.ox) * r.ix, ty1 = (y1 - r.oy) * r.iy, tz1 = (z1 - r.oz) * r.iz;
  float nr = max_sel(tz0, max_sel(ty0, max_sel(tx0, r.t_min)));
  float fr = min_sel(tz1, min_sel(ty1, min_sel(tx1, t_max))) * ROBUST;
  near = nr;
  return nr <= fr;
}

// The four children of a node: hit mask, near distances, child references.
__device__ __forceinline__ u32 node_test(const Node4* __restrict__ node, const PRay& r, float t_max, float4& near,
                                         uint4& child) {
  const float4* n = reinterpret_cast<const float4*>(node);
  float4 lx = n[0], hx = n[1], ly = n[2], hy = n[3], lz = n[4], hz = n[5];
  child = reinterpret_cast<const uint4*>(node)[6];
  u32 bits = 0;
  bits |= slab(r, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, t_max, near.x) ? 1u : 0u;
  bits |= slab(r, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, t_max, near.y) ? 2u : 0u;
  bits |= slab(r, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, t_max, near.z) ? 4u : 0u;
  bits |= slab(r, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, t_max, near.w) ? 8u : 0u;
  return bits;
}

// Per-lane traversal stack: the first LDSN entries in this lane's LDS column (entry k at
// lds[k * BLK]), the rest in private memory (reached only by deep trees).
template <int BLK, int LDSN>
struct Stack {
  u32* lds;
  u32 tail[STACK_TOTAL - LDSN];
  int sp;
  __device__ __forceinline__ void push(u32 v) {
    if (sp < LDSN) lds[sp * BLK] = v;
    else tail[sp - LDSN] = v;
    ++sp;
  }
  __device__ __forceinline__ u32 pop() {
    --sp;
    return sp < LDSN ? lds[sp * BLK] : tail[sp - LDSN];
  }
  __device__ __forceinline__ u32 at(int e) const { return e < LDSN ? lds[e * BLK] : tail[e - LDSN]; }
};

__device__ __forceinline__ void order2(u64& a, u64& b) {
  u64 lo = a < b ? a : b, hi = a < b ? b : a;
  a = lo;
  b = hi;
}

// Pushes the hit children of a node, farthest first, and returns the nearest (a five-exchange
// sorting network over (near bits, child) keys; missed children sort last).
template <class St>
__device__ __forceinline__ u32 push_ordered(u32 bits, float4 near, uint4 ch, St& st) {
  const u64 MISS = ~0ull;
  u64 k0 = (bits & 1u) ? (((u64)f2u(near.x) << 32) | ch.x) : MISS;
  u64 k1 = (bits & 2u) ? (((u64)f2u(near.y) << 32) | ch.y) : MISS;
  u64 k2 = (bits & 4u) ? (((u64)f2u(near.z) << 32) | ch.z) : MISS;
  u64 k3 = (bits & 8u) ? (((u64)f2u(near.w) << 32) | ch.w) : MISS;
  order2(k0, k1);
  order2(k2, k3);
  order2(k0, k2);
  order2(k1, k3);
  order2(k1, k2);
  u32 n = (u32)__popc(bits);
  if (n > 3) st.push((u32)k3);
  if (n > 2) st.push((u32)k2);
  if (n > 1) st.push((u32)k1);
  return (u32)k0;
}

__device__ __forceinline__ V3 xf_point(const float* m, V3 p) {
  return v3(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
            m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}
__device__ __forceinline__ V3 xf_vector(const float* m, V3 v) {
  return v3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
            m[8] * v.x + m[9] * v.y + m[10] * v.z);
}
__device__ __forceinline__ V3 xf_transpose(const float* m, V3 v) {
  return v3(m[0] * v.x + m[4] * v.y + m[8] * v.z, m[1] * v.x + m[5] * v.y + m[9] * v.z,
            m[2] * v.x + m[6] * v.y + m[10] * v.z);
}

__device__ __forceinline__ bool usable_ray(V3 o, V3 d) {
  return finite3(o) && finite3(d) && !(d.x == 0.0f && d.y == 0.0f && d.z == 0.0f);
}

// Nearest hit: instance, mesh triangle id and t (traverse.rs:665-740), the lexicographic minimum
// of (t, instance, triangle) over the triangles hit in [t_min, t_max]. `camera` skips the
// instances hidden to the camera. Returns false on a miss.
template <int BLK, int LDSN>
__device__ bool trace_nearest(const Geo& g, u32* lds_col, V3 o, V3 d, float t_min, float t_max, bool camera,
                              u32& hit_inst, u32& hit_tri, float& hit_t) {
  if (!usable_ray(o, d)) return false;
  PRay ray = prepare(o, d, t_min, t_max);
  if (!(ray.t_min <= ray.t_max)) return false;
  float best_t = ray.t_max;
  u32 best_inst = 0xffffffffu, best_tri = 0xffffffffu;
  Stack<BLK, LDSN> st;
  st.lds = lds_col;
  st.sp = 0;
  bool local = false;  // `ray` is in a transformed instance's object space
  int tlas_sp = -1;  // stack height where the current instance's walk began (-1: in the TLAS)
  u32 inst = 0;
  bool cull = false;
  u32 cur = g.root;
  for (;;) {
    if (cur != REF_EMPTY) {
      if (!(cur & REF_LEAF)) {
        float4 near;
        uint4 ch;
        u32 bits = node_test(g.nodes + cur, ray, best_t, near, ch);
        if (bits) {
          cur = push_ordered(bits, near, ch, st);
          continue;
        }
      } else if (tlas_sp < 0) {
        // An instance leaf: walk its mesh in object space.
        u32 i = (cur & ~REF_LEAF) >> LEAF_COUNT_BITS;
        const DevInstance& in = g.instances[i];
        u32 fl = in.flags;
        if ((fl & INST_ACTIVE) && !(camera && (fl & INST_CAMERA_HIDDEN))) {
          inst = i;
          cull = (fl & INST_ONE_SIDED) != 0;
          if (!(fl & INST_IDENTITY)) {
            V3 lo = xf_point(in.to_local, o), ld = xf_vector(in.to_local, d);
            ray = prepare(lo, ld, t_min, t_max);
            local = true;
          }
          tlas_sp = st.sp;
          cur = in.blas_root;
          continue;
        }
      } else {
        u32 first = (cur & ~REF_LEAF) >> LEAF_COUNT_BITS, count = (cur & 7u) + 1;
        const float4* tp = g.tris + 3 * (size_t)first;
        for (u32 k = 0; k < count; ++k, tp += 3) {
          float4 va = tp[0], vb = tp[1], vc = tp[2];
          float t;
          if (watertight(ray, va, vb, vc, best_t, cull, t)) {
            u32 tri = f2u(va.w);
            if (t < best_t || (t == best_t && (inst < best_inst || (inst == best_inst && tri < best_tri)))) {
              best_t = t;
              best_inst = inst;
              best_tri = tri;
            }
          }
        }
      }
    }
    // Next entry; leaving a mesh walk goes back to world space.
    if (tlas_sp >= 0 && st.sp == tlas_sp) {
      tlas_sp = -1;
      if (local) ray = prepare(o, d, t_min, t_max);
      local = false;
    }
    if (st.sp == 0) break;
    cur = st.pop();
  }
  if (best_inst == 0xffffffffu) return false;
  hit_inst = best_inst;
  hit_tri = best_tri;
  hit_t = best_t;
  return true;
}

// A closest-hit walk that may stop after a number of loop trips and go on later (trace_nearest's
// loop with its state in `w`): the entry about to be visited, the best hit so far, the instance
// being walked and its stack height, the object-space flag. The hit is the lexicographic minimum of
// (t, instance, triangle) over the triangles hit, which no order of visits changes, so a walk cut
// in two gives trace_nearest's result.
struct WalkState {
  u32 cur, best_inst, best_tri, inst;
  float best_t;
  int tlas_sp;
  bool local, cull;
};
// Runs the walk until it ends (true) or `budget` trips have been made (false, `w` and `st` hold
// where it stands). `ray` is the ray in the space of w (world, or the instance's when w.local).
template <int BLK, int LDSN>
__device__ bool walk_nearest(const Geo& g, Stack<BLK, LDSN>& st, V3 o, V3 d, float t_min, float t_max, bool camera,
                             PRay& ray, WalkState& w, u32 budget) {
  for (u32 trips = 0;; ++trips) {
    if (trips == budget) return false;
    u32 cur = w.cur;
    if (cur != REF_EMPTY) {
      if (!(cur & REF_LEAF)) {
        float4 near;
        uint4 ch;
        u32 bits = node_test(g.nodes + cur, ray, w.best_t, near, ch);
        if (bits) {
          w.cur = push_ordered(bits, near, ch, st);
          continue;
        }
      } else if (w.tlas_sp < 0) {
        u32 i = (cur & ~REF_LEAF) >> LEAF_COUNT_BITS;
        const DevInstance& in = g.instances[i];
        u32 fl = in.flags;
        if ((fl & INST_ACTIVE) && !(camera && (fl & INST_CAMERA_HIDDEN))) {
          w.inst = i;
          w.cull = (fl & INST_ONE_SIDED) != 0;
          if (!(fl & INST_IDENTITY)) {
            V3 lo = xf_point(in.to_local, o), ld = xf_vector(in.to_local, d);
            ray = prepare(lo, ld, t_min, t_max);
            w.local = true;
          }
          w.tlas_sp = st.sp;
          w.cur = in.blas_root;
          continue;
        }
      } else {
        u32 first = (cur & ~REF_LEAF) >> LEAF_COUNT_BITS, count = (cur & 7u) + 1;
        const float4* tp = g.tris + 3 * (size_t)first;
        for (u32 k = 0; k < count; ++k, tp += 3) {
          float4 va = tp[0], vb = tp[1], vc = tp[2];
          float t;
          if (watertight(ray, va, vb, vc, w.best_t, w.cull, t)) {
            u32 tri = f2u(va.w);
            if (t < w.best_t ||
                (t == w.best_t && (w.inst < w.best_inst || (w.inst == w.best_inst && tri < w.best_tri)))) {
              w.best_t = t;
              w.best_inst = w.inst;
              w.best_tri = tri;
            }
          }
        }
      }
    }
    if (w.tlas_sp >= 0 && st.sp == w.tlas_sp) {
      w.tlas_sp = -1;
      if (w.local) ray = prepare(o, d, t_min, t_max);
      w.local = false;
    }
    if (st.sp == 0) return true;
    w.cur = st.pop();
  }
}

// Any hit in [t_min, t_max] (traverse.rs:742-802).
template <int BLK, int LDSN>
__device__ bool trace_any(const Geo& g, u32* lds_col, V3 o, V3 d, float t_min, float t_max) {
  if (!usable_ray(o, d)) return false;
  PRay ray = prepare(o, d, t_min, t_max);
  if (!(ray.t_min <= ray.t_max)) return false;
  Stack<BLK, LDSN> st;
  st.lds = lds_col;
  st.sp = 0;
  bool local = false;  // `ray` is in a transformed instance's object space
  int tlas_sp = -1;
  bool cull = false;
  u32 cur = g.root;
  for (;;) {
    if (cur != REF_EMPTY) {
      if (!(cur & REF_LEAF)) {
        float4 near;
        uint4 ch;
        u32 bits = node_test(g.nodes + cur, ray, ray.t_max, near, ch);
        if (bits) {
          // Nearest first: an occluder tends to be found sooner (3.6 % fewer shadow ms than slot
          // order on the large render).
          cur = push_ordered(bits, near, ch, st);
          continue;
        }
      } else if (tlas_sp < 0) {
        u32 i = (cur & ~REF_LEAF) >> LEAF_COUNT_BITS;
        const DevInstance& in = g.instances[i];
        u32 fl = in.flags;
        if (fl & INST_ACTIVE) {
          cull = (fl & INST_ONE_SIDED) != 0;
          if (!(fl & INST_IDENTITY)) {
            V3 lo = xf_point(in.to_local, o), ld = xf_vector(in.to_local, d);
            ray = prepare(lo, ld, t_min, t_max);
            local = true;
          }
          tlas_sp = st.sp;
          cur = in.blas_root;
          continue;
        }
      } else {
        u32 first = (cur & ~REF_LEAF) >> LEAF_COUNT_BITS, count = (cur & 7u) + 1;
        const float4* tp = g.tris + 3 * (size_t)first;
        for (u32 k = 0; k < count; ++k, tp += 3) {
          float t;
          if (watertight(ray, tp[0], tp[1], tp[2], ray.t_max, cull, t)) return true;
        }
      }
    }
    if (tlas_sp >= 0 && st.sp == tlas_sp) {
      tlas_sp = -1;
      if (local) ray = prepare(o, d, t_min, t_max);
      local = false;
    }
    if (st.sp == 0) break;
    cur = st.pop();
  }
  return false;
}

// ------------------------------------------------------------------ one ray per wave
// For a handful of rays left alone on the GPU (the last steps of the longest walks), one lane per
// ray leaves the traversal a chain of dependent fetches: one node, then one triangle at a time.
// Here the whole wave traverses one ray: each round takes up to 16 entries off a shared LDS stack,
// four lanes per entry (a child slot of a node, or two triangles of a leaf), so a round costs one
// fetch latency for up to 16 nodes. The closest hit is the lexicographic minimum of (t, instance,
// triangle) over the triangles hit within the best distance so far, and the slab test is
// conservative, so the order of the visits does not change it: the result is that of trace_nearest,
// bit for bit. Every lane of the wave must call these with the same ray.
constexpr u32 WAVE_WORLD = 0xffffffffu;  // an entry of the instance tree
constexpr int WAVE_STACK = 1024;         // entries of the shared stack
constexpr u32 WAVE_ENTRIES = 16;         // entries per round (four lanes each)
struct WaveStack {
  u32 ref[WAVE_STACK];
  u32 space[WAVE_STACK];  // instance whose object space the entry is in, or WAVE_WORLD
};
enum class WaveEnd { done, hit, overflow };

template <bool ANY>
__device__ WaveEnd wave_traverse(const Geo& g, WaveStack& S, int cap, V3 o, V3 d, float t_min, float t_max,
                                 bool camera, float& best_t, u32& best_inst, u32& best_tri) {
  const u32 lane = __lane_id(), e = lane >> 2, c = lane & 3u;
  const PRay world = prepare(o, d, t_min, t_max);
  best_t = world.t_max;
  best_inst = best_tri = 0xffffffffu;
  if (!usable_ray(o, d) || !(world.t_min <= world.t_max)) return WaveEnd::done;
  if (lane == 0) {
    S.ref[0] = g.root;
    S.space[0] = WAVE_WORLD;
  }
  int sp = 1;
  __syncthreads();
  while (sp > 0) {
    const int n = sp < (int)WAVE_ENTRIES ? sp : (int)WAVE_ENTRIES;
    const int base = sp - n;
    const bool mine = (int)e < n;
    u32 ref = REF_EMPTY, space = WAVE_WORLD;
    if (mine) {
      ref = S.ref[base + e];
      space = S.space[base + e];
    }
    __syncthreads();
    sp = base;
    PRay ray = world;
    bool cull = false;
    if (mine && space != WAVE_WORLD) {
      const DevInstance& in = g.instances[space];
      cull = (in.flags & INST_ONE_SIDED) != 0;
      if (!(in.flags & INST_IDENTITY)) ray = prepare(xf_point(in.to_local, o), xf_vector(in.to_local, d), t_min, t_max);
    }
    bool push = false, hit = false;
    u32 push_ref = REF_EMPTY, push_space = WAVE_WORLD, h_tri = 0xffffffffu;
    float h_t = RPX_INF;
    if (mine && ref != REF_EMPTY) {
      if (!(ref & REF_LEAF)) {
        const Node4* nd = g.nodes + ref;
        float near;
        u32 ch = nd->child[c];
        if (ch != REF_EMPTY && slab(ray, nd->lo_x[c], nd->hi_x[c], nd->lo_y[c], nd->hi_y[c], nd->lo_z[c], nd->hi_z[c],
                                    ANY ? ray.t_max : best_t, near)) {
          push = true;
          push_ref = ch;
          push_space = space;
        }
      } else if (space == WAVE_WORLD) {
        // An instance leaf: its mesh's tree, in the instance's object space.
        if (c == 0) {
          u32 i = (ref & ~REF_LEAF) >> LEAF_COUNT_BITS;
          const DevInstance& in = g.instances[i];
          u32 fl = in.flags;
          if ((fl & INST_ACTIVE) && (ANY || !(camera && (fl & INST_CAMERA_HIDDEN))) && in.blas_root != REF_EMPTY) {
            push = true;
            push_ref = in.blas_root;
            push_space = i;
          }
        }
      } else {
        u32 first = (ref & ~REF_LEAF) >> LEAF_COUNT_BITS, count = (ref & 7u) + 1;
        for (u32 k = c; k < count; k += 4) {
          const float4* tp = g.tris + 3 * (size_t)(first + k);
          float4 va = tp[0], vb = tp[1], vc = tp[2];
          float t;
          if (watertight(ray, va, vb, vc, ANY ? ray.t_max : best_t, cull, t)) {
            u32 tri = f2u(va.w);
            if (!hit || t < h_t || (t == h_t && tri < h_tri)) {
              h_t = t;
              h_tri = tri;
            }
            hit = true;
          }
        }
      }
    }
    if (ANY) {
      if (__ballot(hit)) return WaveEnd::hit;
    } else if (__ballot(hit)) {
      // The round's lexicographic minimum over the wave, then against the best so far.
      float t = hit ? h_t : RPX_INF;
      u32 a = hit ? space : 0xffffffffu, b = hit ? h_tri : 0xffffffffu;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        float t2 = __shfl_xor(t, off);
        u32 a2 = (u32)__shfl_xor((int)a, off), b2 = (u32)__shfl_xor((int)b, off);
        if (t2 < t || (t2 == t && (a2 < a || (a2 == a && b2 < b)))) {
          t = t2;
          a = a2;
          b = b2;
        }
      }
      if (t < best_t || (t == best_t && (a < best_inst || (a == best_inst && b < best_tri)))) {
        best_t = t;
        best_inst = a;
        best_tri = b;
      }
    }
    const u64 m = __ballot(push);
    const int np = (int)__popcll(m);
    if (sp + np > cap) return WaveEnd::overflow;
    if (push) {
      int pos = sp + (int)__builtin_amdgcn_mbcnt_hi((u32)(m >> 32), __builtin_amdgcn_mbcnt_lo((u32)m, 0u));
      S.ref[pos] = push_ref;
      S.space[pos] = push_space;
    }
    sp += np;
    __syncthreads();
  }
  return WaveEnd::done;
}

// Nearest hit of one ray by the whole wave: trace_nearest's result. A stack deeper than `cap`
// entries hands the ray to trace_nearest in lane 0 (its stack's first 16 entries in the shared
// stack's memory, one every 64 words, as in a 64-lane block's LDS columns).
__device__ bool trace_nearest_wave(const Geo& g, WaveStack& S, int cap, V3 o, V3 d, float t_min, float t_max,
                                   bool camera, u32& hit_inst, u32& hit_tri, float& hit_t) {
  float bt;
  u32 bi, bn;
  WaveEnd end = wave_traverse<false>(g, S, cap, o, d, t_min, t_max, camera, bt, bi, bn);
  if (end == WaveEnd::overflow) {
    bool found = false;
    if (__lane_id() == 0) found = trace_nearest<64, 16>(g, S.ref, o, d, t_min, t_max, camera, bi, bn, bt);
    found = __shfl((int)found, 0) != 0;
    bi = (u32)__shfl((int)bi, 0);
    bn = (u32)__shfl((int)bn, 0);
    bt = __shfl(bt, 0);
    if (!found) return false;
  } else if (bi == 0xffffffffu) {
    return false;
  }
  hit_inst = bi;
  hit_tri = bn;
  hit_t = bt;
  return true;
}

// Any hit of one ray by the whole wave: trace_any's answer.
__device__ bool trace_any_wave(const Geo& g, WaveStack& S, int cap, V3 o, V3 d, float t_min, float t_max) {
  float bt;
  u32 bi, bn;
  WaveEnd end = wave_traverse<true>(g, S, cap, o, d, t_min, t_max, false, bt, bi, bn);
  if (end != WaveEnd::overflow) return end == WaveEnd::hit;
  bool blocked = false;
  if (__lane_id() == 0) blocked = trace_any<64, 16>(g, S.ref, o, d, t_min, t_max);
  return __shfl((int)blocked, 0) != 0;
}

// ================================================================== queue compaction

// Appends for a whole block to NC queues at once: per-wave ballots, a prefix over the block's waves
// in LDS, and one atomic per queue per block (every thread of the block must call this).
constexpr int MAX_BLOCK_WAVES = 16;
template <int NC>
__device__ __forceinline__ void block_append(u32* const counters[NC], const bool want[NC], u32 slot[NC]) {
  __shared__ u32 wave_count[NC][MAX_BLOCK_WAVES];
  __shared__ u32 base[NC];
  u32 lane = __lane_id(), wave = threadIdx.x / 64u, waves = (blockDim.x + 63u) / 64u;
  u32 below[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    u64 mask = __ballot(want[c]);
    below[c] = __builtin_amdgcn_mbcnt_hi((u32)(mask >> 32), __builtin_amdgcn_mbcnt_lo((u32)mask, 0u));
    if (lane == 0) wave_count[c][wave] = (u32)__popcll(mask);
  }
  __syncthreads();
  if (threadIdx.x < (u32)NC) {
    u32 c = threadIdx.x, total = 0;
    for (u32 w = 0; w < waves; ++w) {
      u32 t = wave_count[c][w];
      wave_count[c][w] = total;
      total += t;
    }
    base[c] = total ? atomicAdd(counters[c], total) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < NC; ++c) slot[c] = base[c] + wave_count[c][wave] + below[c];
  __syncthreads();
}

// The wave-uniform value of `v` in the lowest lane of `mask` (mask must be non-zero).
__device__ __forceinline__ u32 lowest_lane_value(u64 mask, u32 v) {
  int leader = (int)__ffsll((long long)mask) - 1;
  return (u32)__builtin_amdgcn_readlane((int)v, leader);
}

// ================================================================== sampler
// One number of logical dimension `dim` at sample `index` of the pixel with sampler key `key`
// (sampler.rs:329-395): the Sobol point from the generator columns, Owen-scrambled per pixel.
__device__ __forceinline__ float sample_at(const u32* __restrict__ cols, u64 key, u32 index, u32 dim) {
  u32 x = 0;
  if (dim < 64) {
    const u32* c = cols + dim * 32;
    u32 bits = index;
    while (bits) {
      x ^= c[__builtin_ctz(bits)];
      bits &= bits - 1;
    }
  }
  return sample1(key, index, dim, x);
}
}  // namespace

// ================================================================== adjoint records
namespace {
// The adjoint's gradient of one vertex: its target (a material's five parameters, and for a
// textured material the texels of its base colour instead of parameters 1..3) and the scalar
// d loss / d parameter contributions.
struct VertexGrad {
  u32 mat, tex;  // NO_TARGET: none; tex packs the tap's texel (x0 | y0 << 16)
  float ax, ay;  // the tap's fractions
  float g[N_MAT_PARAMS];
};
constexpr u32 NO_TARGET = 0xffffffffu;

__device__ __forceinline__ void no_grad(VertexGrad& v) {
  v.mat = NO_TARGET;
  v.tex = NO_TARGET;
  v.ax = v.ay = 0.0f;
  for (int k = 0; k < N_MAT_PARAMS; ++k) v.g[k] = 0.0f;
}

}  // namespace

// ================================================================== regen: results + camera (render.rs:79-97)
namespace {
__device__ __forceinline__ float filter_offset(u32 filter, float u) {
  if (filter == 1) return 0.5f;
  if (filter == 2) return 0.5f + (u < 0.5f ? sqrt_(2.0f * u) - 1.0f : 1.0f - sqrt_(2.0f - 2.0f * u));
  return u;
}

// The camera ray of pixel (px, py) at sample `index` (scene/mod.rs:294-357).
__device__ __forceinline__ void camera_ray(const DevScene& S, const u32* cols, u64 key, u32 px, u32 py, u32 index,
                                           V3& o, V3& d) {
  float du = sample_at(cols, key, index, DIM_CAMERA);
  float dv = sample_at(cols, key, index, DIM_CAMERA + 1);
  du = filter_offset(S.pixel_filter, du);
  dv = filter_offset(S.pixel_filter, dv);
  float u = ((float)px + du) / (float)(S.width > 1 ? S.width : 1);
  float v = ((float)py + dv) / (float)(S.height > 1 ? S.height : 1);
  if (S.ortho) {
    float wdt = S.ortho_width;
    o = S.cam_position + S.cam_right * ((u - 0.5f) * wdt) + S.cam_up * ((0.5f - v) * wdt / S.aspect);
    d = S.cam_forward;
  } else {
    float x = (u - 0.5f) * S.sensor_w;
    float y = (0.5f - v) * S.sensor_h;
    V3 dir = normalize(S.cam_forward * S.focal_mm + S.cam_right * x + S.cam_up * y);
    if (S.lens_radius == 0.0f) {
      o = S.cam_position;
      d = dir;
    } else {
      float l0 = sample_at(cols, key, index, DIM_LENS);
      float l1 = sample_at(cols, key, index, DIM_LENS + 1);
      float lx, ly;
      concentric_disk(l0, l1, lx, ly);
      V3 focus_point = S.cam_position + dir * (S.focus_dist / dot(dir, S.cam_forward));
      o = S.cam_position + (S.cam_right * lx + S.cam_up * ly) * S.lens_radius;
      d = normalize(focus_point - o);
    }
  }
}

// The segment holding work id `wid` (segments sorted by base).
__device__ __forceinline__ u32 find_segment(const DevSegment* __restrict__ segs, u32 n_segs, u32 wid) {
  u32 lo = 0, hi = n_segs - 1;
  while (lo < hi) {
    u32 mid = (lo + hi + 1) >> 1;
    if (segs[mid].base <= wid) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
}  // namespace

// n / d for a divisor that is the same across the wave: a shift when it is a power of two (the
// usual sample counts and image widths), where a u32 division costs about 35 instructions.
__device__ __forceinline__ u32 udiv(u32 n, u32 d) { return (d & (d - 1u)) == 0u ? n >> __builtin_ctz(d) : n / d; }

// Pixel p (0 .. w * h) of a w x h rectangle, in 8 x 8 tiles: bands of 8 rows, each in chunks of 8
// columns (the last band and the last chunk narrower), row-major inside a chunk. Consecutive work
// ids, and so the lanes of a wave, then cover a tile of the image rather than a strip of a row: its
// camera rays, and the first bounces from them, walk much the same nodes. A bijection, so every
// pixel's samples still reach its film in sample order.
__device__ __forceinline__ void tiled_pixel(u32 w, u32 h, u32 p, u32& x, u32& y) {
  const u32 band_px = 8u * w;
  const u32 b = udiv(p, band_px);
  const u32 rows = h - 8u * b < 8u ? h - 8u * b : 8u;
  const u32 q = p - b * band_px, chunk_px = 8u * rows, full = w / 8u;
  const u32 c = udiv(q, chunk_px);
  if (c < full) {
    const u32 j = q - c * chunk_px;
    x = 8u * c + j % 8u;
    y = 8u * b + j / 8u;
  } else {
    const u32 cw = w - 8u * full, j = q - full * chunk_px;
    const u32 r = udiv(j, cw);
    x = 8u * full + (j - r * cw);
    y = 8u * b + r;
  }
}

// Resets the pool: every slot empty and queued as finished, the work cursor at 0.
__global__ void __launch_bounds__(256) k_pool_init(PathState P, u32 pool, u32 work_end) {
  u32 stride = gridDim.x * blockDim.x;
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < pool; i += stride) {
    P.done_queue[0][i] = i;
    P.ident[i] = make_uint4(NO_WORK, 0u, 0u, 0u);
  }
  if (blockIdx.x == 0 && threadIdx.x < N_COUNTS) {
    u32 c = threadIdx.x;
    P.counts[c] = c == C_DONE0 ? pool : (c == C_END ? work_end : 0u);
  }
}

// With ADJ, finished paths keep the forward's results and new paths load their adjoint and the
// forward's radiance of the same work id.
//
// The j-th finished slot takes work id cursor + j and live-queue entry survivors + j (no atomics);
// k_bookkeep then advances the counters.
__device__ __forceinline__ u32 new_paths(u32 done, u32 cursor, u32 end) {
  return cursor < end ? (done < end - cursor ? done : end - cursor) : 0u;
}

template <bool ADJ>
__global__ void __launch_bounds__(256) k_regen(const DevScene* __restrict__ scenes, const DevSegment* __restrict__ segs,
                                               u32 n_segs, PathState P, u32 parity) {
  u32 n = P.counts[c_done(parity)];
  u32 end = P.counts[C_END];
  u32 cursor = P.counts[C_CURSOR];
  u32 survivors = P.counts[C_LIVE0 + parity];
  u32 fresh = new_paths(n, cursor, end);
  u32* out_q = P.live[parity];
  u32 stride = gridDim.x * blockDim.x;
  // Whole waves iterate together so that the ballots below see every lane.
  u32 n_iter = (n + stride - 1) / stride;
  u32 gid = blockIdx.x * blockDim.x + threadIdx.x;
  for (u32 it = 0; it < n_iter; ++it) {
    u32 j = it * stride + gid;
    bool active = j < n;
    u32 slot = 0;
    if (active) {
      slot = P.done_queue[parity][j];
      u32 done_work = P.ident[slot].x;
      if (!ADJ && done_work != NO_WORK) P.results[done_work] = P.sum[slot];
    }
    u32 wid = cursor + j;
    bool got = active && j < fresh;
    u32 sid = 0, px = 0, py = 0, index = 0, rslot = 0;
    if (got) {
      const DevSegment& sg = segs[find_segment(segs, n_segs, wid)];
      // Sample-minor: consecutive work ids are the samples of one pixel (a wave's paths start as
      // nearly one ray), pixels in tiled order. The result is stored at the work id, so a pixel's
      // results are contiguous.
      u32 local = wid - sg.base;
      u32 p = udiv(local, sg.count), k = local - p * sg.count;
      rslot = wid;
      sid = sg.scene;
      u32 lx, ly;
      tiled_pixel(sg.w, sg.h, p, lx, ly);
      px = sg.x0 + lx;
      py = sg.y0 + ly;
      index = sg.first + k;
      if constexpr (ADJ) {
        const float* a = sg.adjoint + 3 * ((size_t)py * sg.film_w + px);
        float sc = sg.adjoint_scale;
        P.adj[slot] = make_float4(a[0] * sc, a[1] * sc, a[2] * sc, 0.0f);
        P.ltot[slot] = P.results[rslot];
        P.tgt[slot] = make_uint4(NO_TARGET, NO_TARGET, 0u, 0u);
      }
    } else if (active) {
      P.ident[slot] = make_uint4(NO_WORK, 0u, 0u, 0u);
    }
    // One pass per distinct scene of the wave.
    bool todo = got;
    u64 pending = __ballot(todo);
    while (pending) {
      u32 cur = lowest_lane_value(pending, sid);
      if (todo && sid == cur) {
        const DevScene& S = scenes[cur];
        u64 key = pixel_key(S.seed, px, py);
        V3 o, d;
        camera_ray(S, P.sobol, key, px, py, index, o, d);
        P.ray_o[slot] = make_float4(o.x, o.y, o.z, 0.0f);
        P.ray_d[slot] = make_float4(d.x, d.y, d.z, 0.0f);
        P.thr[slot] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        P.sum[slot] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        P.med[slot] = make_float4(0.0f, 0.0f, 0.0f, u2f(FLAG_PRIMARY));
        P.key[slot] = key;
        P.ident[slot] = make_uint4(rslot, sid, px | (py << 16), index);
        todo = false;
      }
      pending = __ballot(todo);
    }
    if (got) out_q[survivors + j] = slot;
  }
}

// The counters after a regen: the live queue holds survivors + new paths, the cursor moves past
// the new paths, and the queues this iteration's shade fills (this parity's finished and shadow
// queues, both consumed by now) start empty; the other parity's keep what the last shade left
// for this iteration's shadow kernel and the next regen. No live path and nothing left in those
// marks the pool idle (every path of the chunk done). Once no new path is left and at most `tail`
// paths are live, they go to the drain kernel instead (C_TAIL) and this iteration's extend and
// shade find empty queues.
__global__ void k_bookkeep(PathState P, u32 parity, u32 tail) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  u32 n = P.counts[c_done(parity)], end = P.counts[C_END], cursor = P.counts[C_CURSOR];
  u32 fresh = new_paths(n, cursor, end);
  u32 live = P.counts[C_LIVE0 + parity] + fresh;
  P.counts[C_LIVE0 + parity] = live;
  P.counts[C_CURSOR] = cursor + fresh;
  P.counts[c_done(parity)] = 0;
  P.counts[C_LIVE0 + (parity ^ 1)] = 0;
  P.counts[c_shadow(parity)] = 0;
  P.counts[C_TAIL] = 0;
  P.counts[C_PARKED] = 0;
  P.counts[C_CONT] = 0;
  P.rays[0] += live;
  u32 pending = P.counts[c_done(parity ^ 1)] + P.counts[c_shadow(parity ^ 1)];
  if (live == 0) {
    if (pending == 0) P.counts[C_IDLE] = 1;
  } else if (live <= tail && cursor + fresh >= end) {
    P.counts[C_TAIL] = live;
    P.counts[C_DRAIN_LEFT] = live;
    P.counts[C_LIVE0 + parity] = 0;
  }
}

// ================================================================== extend
template <int BLK, int LDSN, int WPE>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(WPE))) k_extend(const DevScene* __restrict__ scenes, PathState P, u32 parity) {
  __shared__ u32 lds[LDSN * BLK];
  u32* col = lds + threadIdx.x;
  const u32* q = P.live[parity];
  u32 n = P.counts[C_LIVE0 + parity];
  u32 stride = gridDim.x * blockDim.x;
  for (u32 j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
    u32 i = q[j];
    float4 o = P.ray_o[i], d = P.ray_d[i];
    bool primary = (f2u(P.med[i].w) & FLAG_PRIMARY) != 0;
    Geo g = geo_of(scenes, P.ident[i].y);
    u32 inst = 0xffffffffu, tri = 0xffffffffu;
    float t = RPX_INF;
    trace_nearest<BLK, LDSN>(g, col, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), RAY_EPSILON, RPX_INF, primary, inst, tri, t);
    P.hit[i] = make_uint4(inst, tri, f2u(t), 0u);
  }
}

// The closest hit in two launches: every walk runs at most `budget` loop trips; the walks still
// going then take continuation slots (their state and stack saved, one atomic per wave) and
// k_extend_resume finishes them in waves of their own, so a wave of the first launch is not held
// by its longest walks. A walk that finds no slot (or a stack deeper than CONT_STACK) goes on where
// it is. The hits are trace_nearest's.
template <int BLK, int LDSN, int WPE>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(WPE)))
k_extend_budget(const DevScene* __restrict__ scenes, PathState P, u32 parity, u32 budget) {
  __shared__ u32 lds[LDSN * BLK];
  u32* col = lds + threadIdx.x;
  const u32* q = P.live[parity];
  u32 n = P.counts[C_LIVE0 + parity];
  u32 stride = gridDim.x * blockDim.x;
  // Whole waves iterate together (the continuation slots are taken per wave).
  u32 n_iter = (n + stride - 1) / stride;
  u32 gid = blockIdx.x * blockDim.x + threadIdx.x;
  for (u32 it = 0; it < n_iter; ++it) {
    u32 j = it * stride + gid;
    bool active = j < n, stopped = false;
    u32 i = 0;
    Stack<BLK, LDSN> st;
    st.lds = col;
    st.sp = 0;
    WalkState w{REF_EMPTY, 0xffffffffu, 0xffffffffu, 0u, RPX_INF, -1, false, false};
    V3 o = v3(0.0f, 0.0f, 0.0f), d = o;
    bool primary = false;
    Geo g{nullptr, nullptr, nullptr, REF_EMPTY};
    PRay ray;
    if (active) {
      i = q[j];
      float4 o4 = P.ray_o[i], d4 = P.ray_d[i];
      o = v3(o4.x, o4.y, o4.z);
      d = v3(d4.x, d4.y, d4.z);
      primary = (f2u(P.med[i].w) & FLAG_PRIMARY) != 0;
      g = geo_of(scenes, P.ident[i].y);
      if (usable_ray(o, d)) {
        ray = prepare(o, d, RAY_EPSILON, RPX_INF);
        if (ray.t_min <= ray.t_max) {
          w.cur = g.root;
          w.best_t = ray.t_max;
          stopped = !walk_nearest<BLK, LDSN>(g, st, o, d, RAY_EPSILON, RPX_INF, primary, ray, w, budget);
        }
      }
    }
    const u64 m = __ballot(stopped && st.sp <= CONT_STACK);
    u32 base = 0;
    if (m) {
      const int leader = (int)__ffsll((long long)m) - 1;
      if ((int)__lane_id() == leader) base = atomicAdd(&P.counts[C_CONT], (u32)__popcll(m));
      base = (u32)__shfl((int)base, leader);
    }
    if (stopped) {
      const u32 k = base + __builtin_amdgcn_mbcnt_hi((u32)(m >> 32), __builtin_amdgcn_mbcnt_lo((u32)m, 0u));
      if (st.sp <= CONT_STACK && k < P.cont_cap) {
        P.cont[2 * (size_t)k] = make_uint4(i, w.cur, w.best_inst, w.best_tri);
        P.cont[2 * (size_t)k + 1] = make_uint4(f2u(w.best_t), (u32)st.sp, (u32)w.tlas_sp,
                                               w.inst | (w.cull ? 1u << 30 : 0u) | (w.local ? 1u << 31 : 0u));
        u32* sv = P.cont_stack + (size_t)CONT_STACK * k;
        for (int e = 0; e < st.sp; ++e) sv[e] = st.at(e);
      } else {
        walk_nearest<BLK, LDSN>(g, st, o, d, RAY_EPSILON, RPX_INF, primary, ray, w, 0xffffffffu);
        stopped = false;
      }
    }
    if (active && !stopped) {
      bool hit = w.best_inst != 0xffffffffu;
      P.hit[i] = make_uint4(hit ? w.best_inst : 0xffffffffu, hit ? w.best_tri : 0xffffffffu,
                            f2u(hit ? w.best_t : RPX_INF), 0u);
    }
  }
}

// The walks k_extend_budget stopped, each to its end.
template <int BLK, int LDSN, int WPE>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(WPE)))
k_extend_resume(const DevScene* __restrict__ scenes, PathState P) {
  __shared__ u32 lds[LDSN * BLK];
  u32* col = lds + threadIdx.x;
  const u32 count = P.counts[C_CONT];
  const u32 n = count < P.cont_cap ? count : P.cont_cap;
  u32 stride = gridDim.x * blockDim.x;
  for (u32 k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    uint4 a = P.cont[2 * (size_t)k], b = P.cont[2 * (size_t)k + 1];
    const u32 i = a.x;
    float4 o4 = P.ray_o[i], d4 = P.ray_d[i];
    V3 o = v3(o4.x, o4.y, o4.z), d = v3(d4.x, d4.y, d4.z);
    const bool primary = (f2u(P.med[i].w) & FLAG_PRIMARY) != 0;
    Geo g = geo_of(scenes, P.ident[i].y);
    WalkState w{a.y, a.z, a.w, b.w & 0x3fffffffu, u2f(b.x), (int)b.z, (b.w >> 31) != 0, ((b.w >> 30) & 1u) != 0};
    Stack<BLK, LDSN> st;
    st.lds = col;
    st.sp = 0;
    const u32* sv = P.cont_stack + (size_t)CONT_STACK * k;
    for (u32 e = 0; e < b.y; ++e) st.push(sv[e]);
    PRay ray;
    if (w.local) {
      const DevInstance& in = g.instances[w.inst];
      ray = prepare(xf_point(in.to_local, o), xf_vector(in.to_local, d), RAY_EPSILON, RPX_INF);
    } else {
      ray = prepare(o, d, RAY_EPSILON, RPX_INF);
    }
    walk_nearest<BLK, LDSN>(g, st, o, d, RAY_EPSILON, RPX_INF, primary, ray, w, 0xffffffffu);
    bool hit = w.best_inst != 0xffffffffu;
    P.hit[i] = make_uint4(hit ? w.best_inst : 0xffffffffu, hit ? w.best_tri : 0xffffffffu,
                          f2u(hit ? w.best_t : RPX_INF), 0u);
  }
}

// ================================================================== shadow
// The adjoint's gradient sums: texels per lane; material parameters summed over the lanes of a
// wave that share a material, then one atomic per parameter (the whole wave must call this).
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ void commit_grad(const PathState& P, bool has, const VertexGrad& v) {
  bool tex = has && v.tex != NO_TARGET;
  if (tex) {
    TexTap t;
    t.x0 = v.tex & 0xffffu;
    t.y0 = v.tex >> 16;
    t.x1 = t.x0 + 1 == P.tex_w ? 0 : t.x0 + 1;
    t.y1 = t.y0 + 1 == P.tex_h ? 0 : t.y0 + 1;
    t.ax = v.ax;
    t.ay = v.ay;
    float w[4];
    texture_weights(t, w);
    u32 xs[4] = {t.x0, t.x1, t.x0, t.x1}, ys[4] = {t.y0, t.y0, t.y1, t.y1};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      float* dst = P.grad_tex + 3 * ((size_t)ys[q] * P.tex_w + xs[q]);
      atomicAdd(dst, v.g[1] * w[q]);
      atomicAdd(dst + 1, v.g[2] * w[q]);
      atomicAdd(dst + 2, v.g[3] * w[q]);
    }
  }
  bool todo = has && v.mat != NO_TARGET;
  u64 pending = __ballot(todo);
  while (pending) {
    u32 cur = lowest_lane_value(pending, v.mat);
    int leader = (int)__ffsll((long long)pending) - 1;
    bool mine = todo && v.mat == cur;
#pragma unroll
    for (int k = 0; k < N_MAT_PARAMS; ++k) {
      // A textured material's base colour is its texture: those terms went to the texels.
      bool to_texels = tex && k >= 1 && k <= 3;
      float x = wave_sum(mine && !to_texels ? v.g[k] : 0.0f);
      if ((int)__lane_id() == leader) atomicAdd(P.grad_mat + (size_t)N_MAT_PARAMS * cur + k, x);
    }
    todo = todo && !mine;
    pending = __ballot(todo);
  }
}

template <int BLK, int LDSN, int WPE, bool ADJ>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(WPE))) k_shadow(const DevScene* __restrict__ scenes, PathState P, u32 parity) {
  __shared__ u32 lds[LDSN * BLK];
  u32* col = lds + threadIdx.x;
  u32 n = P.counts[c_shadow(parity)];
  const u32* q = P.shadow_queue[parity];
  if (blockIdx.x == 0 && threadIdx.x == 0) P.rays[2] += n;
  u32 stride = gridDim.x * blockDim.x;
  // Whole waves iterate together (the adjoint's gradient sums are wave operations).
  u32 n_iter = (n + stride - 1) / stride;
  u32 gid = blockIdx.x * blockDim.x + threadIdx.x;
  for (u32 it = 0; it < n_iter; ++it) {
    u32 j = it * stride + gid;
    bool visible = false;
    u32 i = 0;
    if (j < n) {
      i = q[j];
      float4 o = P.sh_o[i], d = P.sh_d[i];
      Geo g = geo_of(scenes, P.ident[i].y);
      if (!trace_any<BLK, LDSN>(g, col, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), RAY_EPSILON, o.w)) {
        visible = true;
        float4 c = P.sh_c[i];
        float4 s = P.sum[i];
        P.sum[i] = make_float4(s.x + c.x, s.y + c.y, s.z + c.z, s.w);
      }
    }
    if constexpr (ADJ) {
      VertexGrad v;
      no_grad(v);
      if (visible) {
        uint4 tg = P.tgt[i];
        float4 a = P.gn[2 * (size_t)i], b = P.gn[2 * (size_t)i + 1];
        v.mat = tg.x;
        v.tex = tg.y;
        v.ax = u2f(tg.z);
        v.ay = u2f(tg.w);
        v.g[0] = a.x;
        v.g[1] = a.y;
        v.g[2] = a.z;
        v.g[3] = a.w;
        v.g[4] = b.x;
      }
      commit_grad(P, visible, v);
    }
  }
}


// ================================================================== traversal with resident lanes
// The same closest-hit and any-hit walks as above, one node or leaf per loop trip, each lane taking
// its next ray (queue entries gid, gid + stride, ...) as soon as its current one ends, so a wave is
// not held by its slowest ray. Launched with about as many lanes as the GPU keeps resident.
template <int BLK, int LDSN, int WPE>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(WPE)))
k_extend_flat(const DevScene* __restrict__ scenes, PathState P, u32 parity) {
  __shared__ u32 lds[LDSN * BLK];
  const u32* q = P.live[parity];
  const u32 n = P.counts[C_LIVE0 + parity];
  const u32 stride = gridDim.x * blockDim.x;
  u32 j = blockIdx.x * blockDim.x + threadIdx.x;
  Stack<BLK, LDSN> st;
  st.lds = lds + threadIdx.x;
  st.sp = 0;
  bool busy = false, local = false, cull = false, camera = false;
  int tlas_sp = -1;
  u32 slot = 0, inst = 0, cur = REF_EMPTY, best_inst = 0, best_tri = 0;
  float best_t = 0.0f;
  V3 o = v3(0, 0, 0), d = v3(0, 0, 1);
  Geo g{};
  PRay ray = prepare(o, d, RAY_EPSILON, RPX_INF);
  for (;;) {
    while (!busy && j < n) {
      slot = q[j];
      float4 ro = P.ray_o[slot], rd = P.ray_d[slot];
      o = v3(ro.x, ro.y, ro.z);
      d = v3(rd.x, rd.y, rd.z);
      camera = (f2u(P.med[slot].w) & FLAG_PRIMARY) != 0;
      g = geo_of(scenes, P.ident[slot].y);
      if (usable_ray(o, d)) {
        ray = prepare(o, d, RAY_EPSILON, RPX_INF);
        busy = ray.t_min <= ray.t_max;
      }
      if (busy) {
        best_t = ray.t_max;
        best_inst = best_tri = 0xffffffffu;
        st.sp = 0;
        local = false;
        cull = false;
        tlas_sp = -1;
        inst = 0;
        cur = g.root;
      } else {
        P.hit[slot] = make_uint4(0xffffffffu, 0xffffffffu, f2u(RPX_INF), 0u);
        j += stride;
      }
    }
    if (!busy) break;
    bool moved = false;
    if (cur != REF_EMPTY) {
      if (!(cur & REF_LEAF)) {
        float4 near;
        uint4 ch;
        u32 bits = node_test(g.nodes + cur, ray, best_t, near, ch);
        if (bits) {
          cur = push_ordered(bits, near, ch, st);
          moved = true;
        }
      } else if (tlas_sp < 0) {
        u32 i = (cur & ~REF_LEAF) >> LEAF_COUNT_BITS;
        const DevInstance& in = g.instances[i];
        u32 fl = in.flags;
        if ((fl & INST_ACTIVE) && !(camera && (fl & INST_CAMERA_HIDDEN))) {
          inst = i;
          cull = (fl & INST_ONE_SIDED) != 0;
          if (!(fl & INST_IDENTITY)) {
            V3 lo = xf_point(in.to_local, o), ld = xf_vector(in.to_local, d);
            ray = prepare(lo, ld, RAY_EPSILON, RPX_INF);
            local = true;
          }
          tlas_sp = st.sp;
          cur = in.blas_root;
          moved = true;
        }
      } else {
        u32 first = (cur & ~REF_LEAF) >> LEAF_COUNT_BITS, count = (cur & 7u) + 1;
        const float4* tp = g.tris + 3 * (size_t)first;
        for (u32 k = 0; k < count; ++k, tp += 3) {
          float4 va = tp[0], vb = tp[1], vc = tp[2];
          float t;
          if (watertight(ray, va, vb, vc, best_t, cull, t)) {
            u32 tri = f2u(va.w);
            if (t < best_t || (t == best_t && (inst < best_inst || (inst == best_inst && tri < best_tri)))) {
              best_t = t;
              best_inst = inst;
              best_tri = tri;
            }
          }
        }
      }
    }
    if (!moved) {
      if (tlas_sp >= 0 && st.sp == tlas_sp) {
        tlas_sp = -1;
        if (local) ray = prepare(o, d, RAY_EPSILON, RPX_INF);
        local = false;
      }
      if (st.sp == 0) {
        P.hit[slot] = best_inst == 0xffffffffu ? make_uint4(0xffffffffu, 0xffffffffu, f2u(RPX_INF), 0u)
                                               : make_uint4(best_inst, best_tri, f2u(best_t), 0u);
        busy = false;
        j += stride;
      } else {
        cur = st.pop();
      }
    }
  }
}

template <int BLK, int LDSN, int WPE>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(WPE)))
k_shadow_flat(const DevScene* __restrict__ scenes, PathState P, u32 parity) {
  __shared__ u32 lds[LDSN * BLK];
  const u32 n = P.counts[c_shadow(parity)];
  const u32* q = P.shadow_queue[parity];
  if (blockIdx.x == 0 && threadIdx.x == 0) P.rays[2] += n;
  const u32 stride = gridDim.x * blockDim.x;
  u32 j = blockIdx.x * blockDim.x + threadIdx.x;
  Stack<BLK, LDSN> st;
  st.lds = lds + threadIdx.x;
  st.sp = 0;
  bool busy = false, local = false, cull = false;
  int tlas_sp = -1;
  u32 slot = 0, cur = REF_EMPTY;
  float t_max = 0.0f;
  V3 o = v3(0, 0, 0), d = v3(0, 0, 1);
  Geo g{};
  PRay ray = prepare(o, d, RAY_EPSILON, RPX_INF);
  for (;;) {
    while (!busy && j < n) {
      slot = q[j];
      float4 so = P.sh_o[slot], sd = P.sh_d[slot];
      o = v3(so.x, so.y, so.z);
      d = v3(sd.x, sd.y, sd.z);
      t_max = so.w;
      g = geo_of(scenes, P.ident[slot].y);
      if (usable_ray(o, d)) {
        ray = prepare(o, d, RAY_EPSILON, t_max);
        busy = ray.t_min <= ray.t_max;
      }
      if (busy) {
        st.sp = 0;
        local = false;
        cull = false;
        tlas_sp = -1;
        cur = g.root;
      } else {
        // No walk: nothing can block the ray.
        float4 c = P.sh_c[slot], s = P.sum[slot];
        P.sum[slot] = make_float4(s.x + c.x, s.y + c.y, s.z + c.z, s.w);
        j += stride;
      }
    }
    if (!busy) break;
    bool moved = false, blocked = false;
    if (cur != REF_EMPTY) {
      if (!(cur & REF_LEAF)) {
        float4 near;
        uint4 ch;
        u32 bits = node_test(g.nodes + cur, ray, ray.t_max, near, ch);
        if (bits) {
          cur = push_ordered(bits, near, ch, st);
          moved = true;
        }
      } else if (tlas_sp < 0) {
        u32 i = (cur & ~REF_LEAF) >> LEAF_COUNT_BITS;
        const DevInstance& in = g.instances[i];
        u32 fl = in.flags;
        if (fl & INST_ACTIVE) {
          cull = (fl & INST_ONE_SIDED) != 0;
          if (!(fl & INST_IDENTITY)) {
            V3 lo = xf_point(in.to_local, o), ld = xf_vector(in.to_local, d);
            ray = prepare(lo, ld, RAY_EPSILON, t_max);
            local = true;
          }
          tlas_sp = st.sp;
          cur = in.blas_root;
          moved = true;
        }
      } else {
        u32 first = (cur & ~REF_LEAF) >> LEAF_COUNT_BITS, count = (cur & 7u) + 1;
        const float4* tp = g.tris + 3 * (size_t)first;
        for (u32 k = 0; k < count && !blocked; ++k, tp += 3) {
          float t;
          blocked = watertight(ray, tp[0], tp[1], tp[2], ray.t_max, cull, t);
        }
      }
    }
    if (blocked) {
      busy = false;
      j += stride;
    } else if (!moved) {
      if (tlas_sp >= 0 && st.sp == tlas_sp) {
        tlas_sp = -1;
        if (local) ray = prepare(o, d, RAY_EPSILON, t_max);
        local = false;
      }
      if (st.sp == 0) {
        float4 c = P.sh_c[slot], s = P.sum[slot];
        P.sum[slot] = make_float4(s.x + c.x, s.y + c.y, s.z + c.z, s.w);
        busy = false;
        j += stride;
      } else {
        cur = st.pop();
      }
    }
  }
}

// ================================================================== shade
namespace {
__device__ __forceinline__ float power_heuristic(float pa, float pb) {
  if (!(pa > 0.0f)) return 0.0f;
  float r = pb / pa;
  return 1.0f / (1.0f + r * r);
}
__device__ __forceinline__ float env_value(const DevScene& S, float z) {
  z = clampf(z, -1.0f, 1.0f);
  if (z < 0.0f) return S.env_floor * -z + S.env_horizon * (1.0f + z);
  return S.env_horizon * (1.0f - z) + S.env_zenith * z;
}
__device__ __forceinline__ float env_pdf(const DevScene& S, float z) {
  if (S.env_total_mass <= 0.0f) return 0.0f;
  return env_value(S, z) / (TAU * S.env_total_mass);
}
__device__ __forceinline__ float invert_linear(float a, float b, float u) {
  float c = u * 0.5f * (a + b);
  float denom = a + sqrt_(fmax_r(a * a + 2.0f * (b - a) * c, 0.0f));
  if (denom > 0.0f) return clampf(2.0f * c / denom, 0.0f, 1.0f);
  return 0.0f;
}
__device__ __forceinline__ float orientation_bound(V3 v, float d2, V3 axis, float cos_o, float sin_o, float bound) {
  if (d2 <= bound * bound) return 1.0f;
  float inv_d = 1.0f / sqrt_(d2);
  float sin_u = bound * inv_d;
  float cos_u = sqrt_(fmax_r(1.0f - sin_u * sin_u, 0.0f));
  float cos_s = cos_o * cos_u - sin_o * sin_u;
  float sin_s = sin_o * cos_u + cos_o * sin_u;
  if (sin_s < 0.0f) return 1.0f;
  float cos_t = dot(v, axis) * inv_d;
  if (cos_t >= cos_s) return 1.0f;
  float sin_t = sqrt_(fmax_r(1.0f - cos_t * cos_t, 0.0f));
  return fmax_r(cos_t * cos_s + sin_t * sin_s, 0.0f);
}
__device__ __forceinline__ float area_weight(const DevArea& a, V3 x) {
  V3 v = x - a.centre;
  float d2 = length_sq(v);
  float orientation = orientation_bound(v, d2, a.frame.z, 1.0f, 0.0f, a.bound);
  if (orientation <= 0.0f) return 0.0f;
  return a.base * orientation / (fmax_r(d2, a.area) * 1.0e-6f);
}
// light.rs:465-481: an emissive mesh's selection weight from its centroid, with the cone of its
// normals.
__device__ __forceinline__ float emitter_weight_at(const DevEmitter& m, V3 x) {
  V3 v = x - m.centroid;
  float d2 = length_sq(v);
  float orientation = m.cos_o <= -1.0f ? 1.0f : orientation_bound(v, d2, m.axis, m.cos_o, m.sin_o, m.bound);
  if (orientation <= 0.0f) return 0.0f;
  return m.base * orientation / (fmax_r(d2, m.area) * 1.0e-6f);
}
// EXT: entries may also be sphere lights, suns and emissive meshes (entry_kind).
template <bool EXT>
__device__ __forceinline__ float entry_weight(const DevScene& S, u32 e, V3 x) {
  i32 a = S.entries[e];
  if constexpr (EXT) {
    if (a < -1) {
      u32 k;
      EntryKind kind = entry_kind(a, k);
      if (kind == E_SPHERE) return sphere_weight(S.spheres[k], x);
      if (kind == E_SUN) return S.suns[k].base;
      return emitter_weight_at(S.emitters[k], x);
    }
  }
  return a < 0 ? S.env_base : area_weight(S.areas[a], x);
}
__device__ __forceinline__ C3 radiance_at(const DevArea& a, float cos) {
  return a.exponent == 0.0f ? a.radiance : a.radiance * fpow(cos, a.exponent);
}
__device__ __forceinline__ bool area_intersect(const DevArea& a, V3 o, V3 d, float t_max, float& t_out, float& cos_out) {
  V3 n = a.frame.z;
  float denom = dot(d, n);
  if (denom >= 0.0f) return false;
  float t = dot(a.centre - o, n) / denom;
  if (!(t > RAY_EPSILON && t < t_max)) return false;
  V3 q = (o + d * t) - a.centre;
  float x = dot(q, a.frame.x), y = dot(q, a.frame.y);
  bool inside = a.rect ? (fabs_(x) <= a.half_x && fabs_(y) <= a.half_y) : (x * x + y * y <= a.radius * a.radius);
  if (!inside) return false;
  t_out = t;
  cos_out = -denom;
  return true;
}

// The light selection of a vertex (light.rs:1026-1085): by flux, or per point with the first
// KEPT_WEIGHTS weights kept.
struct Sel {
  V3 point;
  float total;
  u32 by_flux;
  float w[KEPT_WEIGHTS];
};
template <bool EXT>
__device__ __forceinline__ float sel_weight(const DevScene& S, const Sel& s, u32 e) {
  return e < (u32)KEPT_WEIGHTS ? s.w[e] : entry_weight<EXT>(S, e, s.point);
}
template <bool EXT>
__device__ __forceinline__ float sel_pmf(const DevScene& S, const Sel& s, u32 e) {
  if (e >= S.n_entries) return 0.0f;
  if (s.by_flux) return S.flux_pmf[e];
  if (!(s.total > 0.0f && is_finite(s.total))) return 0.0f;
  return sel_weight<EXT>(S, s, e) / s.total;
}

// The exit lobe of a scattering medium (bsdf.rs:549-556, 697-701, 783-785, 837-847): Lambertian
// transmission of albedo `exit` through the surface, nothing reflected.
__device__ __forceinline__ C3 exit_eval(float exit, V3 wo, V3 wi) {
  if (wo.z <= 0.0f) return splat(0.0f);
  return wi.z < 0.0f ? splat(exit * FRAC_1_PI) : splat(0.0f);
}
__device__ __forceinline__ float exit_pdf(V3 wo, V3 wi) {
  if (wo.z <= 0.0f) return 0.0f;
  return fmax_r(-wi.z, 0.0f) * FRAC_1_PI;
}
__device__ __forceinline__ bool exit_sample(float exit, V3 wo, float u0, float u1, BsdfSample& out) {
  if (wo.z <= 0.0f) return false;
  V3 d = cosine_hemisphere(u0, u1);
  float pdf = d.z * FRAC_1_PI;
  if (!(pdf > 0.0f)) return false;
  out.wi = v3(d.x, d.y, -d.z);
  out.weight = splat(exit);
  out.pdf = pdf;
  return true;
}

__device__ __forceinline__ void add_to(C3& sum, C3 c, bool indirect, float clamp_ind) {
  if (!finite_c(c)) return;
  float pk = fmax_r(fmax_r(c.r, c.g), c.b);
  if (indirect && clamp_ind > 0.0f && pk > clamp_ind) sum = sum + c * (clamp_ind / pk);
  else sum = sum + c;
}

// d f / d parameter of the BSDF at (wo, wi_a) (diffuse-only flag `diffuse_a`) and at (wo, wi_b):
// forward-mode derivatives (rpx_bsdf_grad.h), one parameter at a time, of the BSDF built from the
// material with that parameter as the dual variable (the directions stay fixed: the adjoint
// differentiates the integrand in path space, its sampling densities held constant).
__device__ void bsdf_param_derivs(const EnergyTables& T, const DevMaterial& M, C3 base, bool entering, V3 wo,
                                  bool want_a, V3 wi_a, bool diffuse_a, bool want_b, V3 wi_b, C3 da[N_MAT_PARAMS],
                                  C3 db[N_MAT_PARAMS]) {
#pragma unroll 1
  for (int k = 0; k < N_MAT_PARAMS; ++k) {
    Dual r{M.roughness, k == 0 ? 1.0f : 0.0f}, ior{M.ior, k == 4 ? 1.0f : 0.0f};
    DC3 bc{Dual{base.r, k == 1 ? 1.0f : 0.0f}, Dual{base.g, k == 2 ? 1.0f : 0.0f}, Dual{base.b, k == 3 ? 1.0f : 0.0f}};
    DBsdf b = dbsdf_surface(T, M.kind, bc, M.metallic, r, ior, M.transmission, M.coat, M.coat_roughness, M.flags,
                            M.absorption_mm, entering);
    da[k] = want_a ? dderiv(dbsdf_eval(T, b, wo, wi_a, diffuse_a)) : splat(0.0f);
    db[k] = want_b ? dderiv(dbsdf_eval(T, b, wo, wi_b, false)) : splat(0.0f);
  }
}

__device__ __forceinline__ float dot3(C3 a, C3 b) { return a.r * b.r + a.g * b.g + a.b * b.b; }

// One vertex of the path in slot i (integrator/mod.rs:1140-1891, one loop iteration): `cont` when
// the path goes on with a new ray, `want_shadow` when a shadow ray was queued behind it.
// With ADJ (path replay backpropagation), the same vertex is replayed and its gradient found:
// `g_prev` the term of the previous vertex's BSDF sample (it needs the radiance the path gathered
// after it, the forward's total minus what is summed so far), `g_now` the next-event term of this
// vertex when its light casts no shadow; a shadowed next-event term waits for the shadow kernel.
// The surface a ray (o, d) met at instance hit h (instance, mesh triangle): hit details
// (geometry/mod.rs:731-808), the shading normal without a bevel (bevel.rs:232-252) and the
// terminator lift (geometry/mod.rs:682-728).
// A mesh triangle's corner positions and normals from its packed record (one dependent fetch after
// the mesh, two cache lines), the values the mesh's vertex arrays hold.
__device__ __forceinline__ void corners_of(const DevScene& S, const DevMesh& mesh, u32 tri, V3& p0, V3& p1, V3& p2,
                                           V3& n0, V3& n1, V3& n2) {
  const float4* c = S.corners + CORNER_FLOAT4S * ((size_t)(mesh.index_offset / 3) + tri);
  float4 a = c[0], b = c[1], e = c[2], f = c[3], g = c[4];
  p0 = v3(a.x, a.y, a.z);
  p1 = v3(a.w, b.x, b.y);
  p2 = v3(b.z, b.w, e.x);
  n0 = v3(e.y, e.z, e.w);
  n1 = v3(f.x, f.y, f.z);
  n2 = v3(f.w, g.x, g.y);
}

__device__ __forceinline__ void corner_positions_of(const DevScene& S, const DevMesh& mesh, u32 tri, V3& p0, V3& p1,
                                                    V3& p2) {
  const float4* c = S.corners + CORNER_FLOAT4S * ((size_t)(mesh.index_offset / 3) + tri);
  float4 a = c[0], b = c[1], e = c[2];
  p0 = v3(a.x, a.y, a.z);
  p1 = v3(a.w, b.x, b.y);
  p2 = v3(b.z, b.w, e.x);
}

struct SurfaceHit {
  V3 point, hit_point, normal, geometric;
  bool front, outside;
  u32 inst;
};
__device__ __forceinline__ SurfaceHit surface_hit(const DevScene& S, V3 o, V3 d, uint4 h) {
  V3 s_point, s_hit_point, s_normal, s_geometric;
  bool s_instance, s_front, s_outside;
  u32 inst;
  // Hit details (geometry/mod.rs:731-808) and the terminator lift (:682-728).
  inst = h.x;
  u32 tri = h.y;
  const DevInstance& in = S.instances[inst];
  const DevMesh& mesh = S.meshes[in.mesh];
  bool ident = (in.flags & INST_IDENTITY) != 0;
  V3 p0, p1, p2, n0, n1, n2;
  corners_of(S, mesh, tri, p0, p1, p2, n0, n1, n2);
  V3 lo = ident ? o : xf_point(in.to_local, o), ld = ident ? d : xf_vector(in.to_local, d);
  PRay r = prepare(lo, ld, RAY_EPSILON, RPX_INF);
  float b0, b1, b2;
  {
    float u, v, w, az, bz, cz;
    shear(r, p0 - lo, p1 - lo, p2 - lo, u, v, w, az, bz, cz);
    bool negative = u < 0.0f || v < 0.0f || w < 0.0f;
    bool positive = u > 0.0f || v > 0.0f || w > 0.0f;
    float det = u + v + w;
    bool ok = !(negative && positive) && det != 0.0f;
    if (ok) {
      float t = (u * az + v * bz + w * cz) / det;
      ok = t >= r.t_min && t <= F32_MAX;
    }
    if (ok) {
      b0 = u / det;
      b1 = v / det;
      b2 = w / det;
    } else {
      b0 = b1 = b2 = 1.0f / 3.0f;
    }
  }
  V3 point_local = p0 * b0 + p1 * b1 + p2 * b2;
  V3 wind = cross(p1 - p0, p2 - p0);
  V3 winding = normalize(ident ? wind : xf_transpose(in.to_local, wind));
  V3 incoming = normalize(d);
  V3 geometric = length_sq(winding) > 0.0f ? winding : -incoming;
  bool front_face = dot(winding, d) < 0.0f;
  V3 smooth = v3(0.0f, 0.0f, 0.0f);
  if (mesh.has_normals) {
    V3 interp = n0 * b0 + n1 * b1 + n2 * b2;
    smooth = ident ? interp : xf_transpose(in.to_local, interp);
  }
  V3 normal = length_sq(smooth) > 1.0e-12f ? normalize(smooth) : geometric;
  V3 f_normal = dot(normal, d) > 0.0f ? -normal : normal;
  V3 ng = dot(geometric, d) > 0.0f ? -geometric : geometric;
  V3 hit_point = ident ? point_local : xf_point(in.to_world, point_local);
  // bevel::shading_point without a bevel
  V3 sn = f_normal;
  if (dot(sn, ng) < 0.0f) sn = -sn;
  if (dot(sn, d) >= -1.0e-4f) sn = ng;
  // terminator_point
  V3 tp = hit_point;
  if (mesh.has_normals) {
    V3 pt = p0 * b0 + p1 * b1 + p2 * b2;
    V3 interp = n0 * b0 + n1 * b1 + n2 * b2;
    V3 world_normal = ident ? interp : xf_transpose(in.to_local, interp);
    float side = dot(world_normal, ng) < 0.0f ? -1.0f : 1.0f;
    V3 lifted = pt;
    V3 cs[3] = {p0, p1, p2};
    V3 nsv[3] = {n0, n1, n2};
    float ws[3] = {b0, b1, b2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      V3 nk = normalize(nsv[k]) * side;
      float below = dot(pt - cs[k], nk);
      if (below < 0.0f) lifted = lifted - nk * (below * ws[k]);
    }
    if (finite3(lifted)) tp = ident ? lifted : xf_point(in.to_world, lifted);
  }
  s_point = tp;
  s_hit_point = hit_point;
  s_normal = sn;
  s_geometric = ng;
  s_instance = true;
  s_front = front_face;
  s_outside = front_face != (mesh.inside_out != 0);
  (void)s_instance;
  return SurfaceHit{s_point, s_hit_point, s_normal, s_geometric, s_front, s_outside, inst};
}

// ================================================================== caustic photons (caustics.rs)
// pick (light.rs:1059-1074) on a short cdf: the first bucket whose end passes u.
__device__ __forceinline__ u32 pick_bucket(const float* cdf, u32 n, float u) {
  u32 k = 0;
  while (k < n && cdf[k] <= u) ++k;
  return k < n - 1 ? k : n - 1;
}

// caustics.rs:171-178
__device__ __forceinline__ float cone_omc(float r, float d) {
  float sin2 = (r * r) / (d * d);
  float cos = sqrt_(fmax_r(1.0f - sin2, 0.0f));
  return sin2 / (1.0f + cos);
}

// The cone from x to caster sphere i: its axis, 1 - cos of its half angle (0: x is inside the
// sphere, which then takes the hemisphere or the sphere) and its weight in the mixture.
__device__ __forceinline__ void caster_cone(const DevCaustics& C, u32 i, V3 x, bool one_sided, V3 n, V3& axis,
                                            float& omc, float& weight) {
  V3 d = C.sphere_c[i] - x;
  float dist = length(d);
  axis = dist > 0.0f ? d / dist : v3(0.0f, 0.0f, 1.0f);
  float omega;
  if (dist <= C.sphere_r[i]) {
    omc = 0.0f;
    omega = PI;
  } else {
    omc = cone_omc(C.sphere_r[i], dist);
    omega = TAU * omc;
  }
  float facing = one_sided ? 0.05f + fmax_r(dot(n, axis), 0.0f) : 1.0f;
  weight = omega * facing;
}

// caustics.rs:495-559: a direction from x into the casters' cones and its density under the
// mixture of every cone (`one_sided`: the emitting side n of a one-sided light). The cones'
// cdf (light.rs:636-654) is walked again rather than kept, so nothing lands in scratch memory.
__device__ bool photon_towards(const DevCaustics& C, V3 x, bool one_sided, V3 n, float u_pick, float u0, float u1,
                               V3& dir, float& pdf_out) {
  const u32 count = C.n_spheres;
  V3 axis;
  float omc, w;
  float total = -0.0f;
  for (u32 i = 0; i < count; ++i) {
    caster_cone(C, i, x, one_sided, n, axis, omc, w);
    total = total + w;
  }
  if (!(total > 0.0f && is_finite(total))) return false;
  // pick (light.rs:1059-1074): the first cone whose cdf end passes u_pick.
  u32 k = count - 1;
  float acc = 0.0f;
  for (u32 i = 0; i < count; ++i) {
    caster_cone(C, i, x, one_sided, n, axis, omc, w);
    acc += w;
    float end = i == count - 1 ? 1.0f : acc / total;
    if (!(end <= u_pick)) {
      k = i;
      break;
    }
  }
  caster_cone(C, k, x, one_sided, n, axis, omc, w);
  if (omc > 0.0f) {
    dir = normalize(to_world(frame_from_normal(axis), uniform_cone(u0, u1, omc)));
  } else if (one_sided) {
    dir = normalize(to_world(frame_from_normal(n), cosine_hemisphere(u0, u1)));
  } else {
    float z = 1.0f - 2.0f * u0;
    float r = sqrt_(fmax_r(1.0f - z * z, 0.0f));
    float sn, cs;
    sin_cos(TAU * u1, sn, cs);
    dir = v3(r * cs, r * sn, z);
  }
  float pdf = 0.0f;
  acc = 0.0f;
  float prev = 0.0f;
  for (u32 i = 0; i < count; ++i) {
    caster_cone(C, i, x, one_sided, n, axis, omc, w);
    acc += w;
    float end = i == count - 1 ? 1.0f : acc / total;
    float p = end - prev;
    prev = end;
    if (!(p > 0.0f)) continue;
    if (omc > 0.0f) {
      if (1.0f - dot(dir, axis) <= omc) pdf += p / (TAU * omc);
    } else {
      pdf += p * (one_sided ? fmax_r(dot(dir, n), 0.0f) * FRAC_1_PI : 0.25f * FRAC_1_PI);
    }
  }
  if (!(pdf > 0.0f && is_finite(pdf))) return false;
  pdf_out = pdf;
  return true;
}

// caustics.rs:369-491: where a photon leaves its light, its direction and its flux (W).
__device__ bool photon_emit(const DevScene& S, Pcg32& rng, V3& o, V3& d, C3& flux) {
  const DevCaustics& C = S.caustics;
  float n = (float)C.photons;
  u32 j = pick_bucket(C.source_cdf, C.n_sources, rng.next_f32());
  float p_light = C.source_cdf[j] - (j == 0 ? 0.0f : C.source_cdf[j - 1]);
  if (!(p_light > 0.0f)) return false;
  float up0 = rng.next_f32();
  float up1 = rng.next_f32();
  float u_pick = rng.next_f32();
  float ud0 = rng.next_f32();
  float ud1 = rng.next_f32();
  u32 src = C.sources[j], kind = src >> 24, idx = src & 0xffffffu;
  if (kind == 0) {
    const DevArea& L = S.areas[idx];
    float x0, y0;
    if (L.rect) {
      x0 = (2.0f * up0 - 1.0f) * L.half_x;
      y0 = (2.0f * up1 - 1.0f) * L.half_y;
    } else {
      float dx, dy;
      concentric_disk(up0, up1, dx, dy);
      x0 = dx * L.radius;
      y0 = dy * L.radius;
    }
    V3 x = L.centre + L.frame.x * x0 + L.frame.y * y0;
    float pdf_dir;
    if (!photon_towards(C, x, true, L.frame.z, u_pick, ud0, ud1, d, pdf_dir)) return false;
    float cos = dot(d, L.frame.z);
    if (!(cos > 0.0f)) return false;
    float pdf_area = 1.0f / (L.area * 1.0e-6f);
    C3 rad = L.exponent == 0.0f ? L.radiance : L.radiance * fpow(cos, L.exponent);
    flux = rad * (cos / (pdf_area * pdf_dir * p_light * n));
    o = x;
    return true;
  }
  if (kind == 2) {
    const DevEmitter& E = S.emitters[idx];
    const float* cdf = S.emit_cdf + E.first;
    u32 lo_i = 0, hi_i = E.count;  // the first triangle whose cdf end passes up0, by bisection
    while (lo_i < hi_i) {
      u32 mid = (lo_i + hi_i) >> 1;
      if (cdf[mid] <= up0) lo_i = mid + 1;
      else hi_i = mid;
    }
    u32 t = lo_i < E.count - 1 ? lo_i : E.count - 1;
    float lo = t == 0 ? 0.0f : cdf[t - 1];
    float width = cdf[t] - lo;
    float rest = width > 0.0f ? clampf((up0 - lo) / width, 0.0f, 0.99999994f) : 0.0f;
    const float4* tp = S.emit_tris + 4 * (size_t)(E.first + t);
    float su = sqrt_(rest);
    float w0 = 1.0f - su, w1 = up1 * su;
    float w2 = 1.0f - w0 - w1;
    V3 x = (v3(tp[0].x, tp[0].y, tp[0].z) * w0 + v3(tp[1].x, tp[1].y, tp[1].z) * w1) + v3(tp[2].x, tp[2].y, tp[2].z) * w2;
    V3 normal = v3(tp[3].x, tp[3].y, tp[3].z);
    float pdf_dir;
    if (!photon_towards(C, x, !E.two_sided, normal, u_pick, ud0, ud1, d, pdf_dir)) return false;
    float facing = dot(d, normal);
    if (!E.two_sided && facing <= 0.0f) return false;
    float cos = fabs_(facing);
    if (!(cos > 0.0f)) return false;
    float pdf_area = 1.0f / (E.area * 1.0e-6f);
    flux = E.radiance * (cos / (pdf_area * pdf_dir * p_light * n));
    o = offset_origin(x, facing > 0.0f ? normal : -normal);
    return true;
  }
  const DevSun& L = S.suns[idx];
  V3 to_sun = L.frame.z;
  float cos, pdf_dir;
  if (L.delta) {
    d = -to_sun;
    cos = 1.0f;
    pdf_dir = 1.0f;
  } else {
    V3 local = uniform_cone(ud0, ud1, L.one_minus_cos_max);
    V3 dd = normalize(to_world(L.frame, local));
    d = -dd;
    cos = local.z;
    pdf_dir = 1.0f / L.solid_angle;
  }
  V3 plane = C.scene_centre + to_sun * (2.0f * C.scene_radius);
  float tilt = 0.0f;
  if (!L.delta) {
    float s2 = 2.0f * L.one_minus_cos_max - L.one_minus_cos_max * L.one_minus_cos_max;
    tilt = sqrt_(fmax_r(s2, 0.0f)) * 4.0f * C.scene_radius;
  }
  // The sphere discs projected on the plane, by area (light.rs:636-654), walked again rather
  // than kept.
  const u32 count = C.n_spheres;
  float total = -0.0f;
  for (u32 i = 0; i < count; ++i) total = total + PI * (C.sphere_r[i] + tilt) * (C.sphere_r[i] + tilt);
  if (!(total > 0.0f && is_finite(total))) return false;
  u32 k = count - 1;
  float acc = 0.0f;
  for (u32 i = 0; i < count; ++i) {
    acc += PI * (C.sphere_r[i] + tilt) * (C.sphere_r[i] + tilt);
    float end = i == count - 1 ? 1.0f : acc / total;
    if (!(end <= u_pick)) {
      k = i;
      break;
    }
  }
  Frame frame = frame_from_normal(to_sun);
  float dx, dy;
  concentric_disk(up0, up1, dx, dy);
  float rk = C.sphere_r[k] + tilt;
  V3 ck = C.sphere_c[k];
  V3 x = ((ck - to_sun * dot(ck - plane, to_sun)) + frame.x * (dx * rk)) + frame.y * (dy * rk);
  float pdf_area = 0.0f;
  acc = 0.0f;
  float prev = 0.0f;
  for (u32 i = 0; i < count; ++i) {
    float ri = C.sphere_r[i] + tilt;
    acc += PI * ri * ri;
    float end = i == count - 1 ? 1.0f : acc / total;
    float p = end - prev;
    prev = end;
    V3 ci = C.sphere_c[i];
    if (length_sq(x - (ci - to_sun * dot(ci - plane, to_sun))) <= ri * ri) pdf_area += p / (PI * ri * ri * 1.0e-6f);
  }
  if (!(pdf_area > 0.0f)) return false;
  flux = L.radiance * (cos / (pdf_area * pdf_dir * p_light * n));
  o = x;
  return true;
}

// caustics.rs:902-950: the radiance the photons around `point` send towards the camera, f giving
// the BSDF of the visible point for a photon's direction (world space).
template <class F>
__device__ C3 photon_gather(const DevPhotonMap& M, V3 point, V3 normal, u32 max_bounces, F f) {
  C3 sum = splat(0.0f);
  if (M.count == 0) return sum;
  float r = M.radius;
  i32 lo[3], hi[3];
  photon_cell(point - v3(r, r, r), M.inv_cell, lo);
  photon_cell(point + v3(r, r, r), M.inv_cell, hi);
  for (i32 cz = lo[2]; cz <= min(hi[2], lo[2] + 1); ++cz)
    for (i32 cy = lo[1]; cy <= min(hi[1], lo[1] + 1); ++cy)
      for (i32 cx = lo[0]; cx <= min(hi[0], lo[0] + 1); ++cx) {
        i32 cell[3] = {cx, cy, cz};
        u32 b = photon_bucket(cell, M.mask);
        u32 end = M.starts[b + 1];
        for (u32 k = M.starts[b]; k < end; ++k) {
          float4 pp = M.position[k];
          float4 pn = M.normal[k];
          if (f2u(pp.w) > max_bounces || dot(v3(pn.x, pn.y, pn.z), normal) < 0.5f) continue;
          V3 pos = v3(pp.x, pp.y, pp.z);
          float d2 = length_sq(pos - point) * M.inv_r2;
          i32 pc[3];
          photon_cell(pos, M.inv_cell, pc);
          if (d2 >= 1.0f || pc[0] != cx || pc[1] != cy || pc[2] != cz) continue;
          float kern = M.norm * (1.0f - d2);
          float4 pw = M.wi[k];
          C3 fv = f(v3(pw.x, pw.y, pw.z));
          if (peak(fv) > 0.0f) {
            float4 po = M.power[k];
            sum = sum + (fv * c3(po.x, po.y, po.z)) * kern;
          }
        }
      }
  return sum;
}

// EXT adds sphere, point and spot lights and the vertices after a walk through a scattering medium
// (the walk kernel ran between extend and shade: hit.w says how it ended).
template <bool ADJ, bool EXT>
__device__ __forceinline__ void shade_path(const DevScene& S, const PathState& P, u32 i, bool& cont, bool& want_shadow,
                                           VertexGrad& g_prev, VertexGrad& g_now) {
  const EnergyTables& T = S.tables;
  {
  float4 ro = P.ray_o[i], rd = P.ray_d[i];
  V3 o = v3(ro.x, ro.y, ro.z), d = v3(rd.x, rd.y, rd.z);
  uint4 h = P.hit[i];
  float4 th = P.thr[i];
  float4 sm = P.sum[i];
  float4 md = P.med[i];
  u32 flags = f2u(md.w);
  bool primary = (flags & FLAG_PRIMARY) != 0;
  bool has_pdf = (flags & FLAG_HAS_PDF) != 0;
  bool in_medium = (flags & FLAG_MEDIUM) != 0;
  C3 thr = c3(th.x, th.y, th.z);
  float prev_pdf = th.w;
  C3 sum = c3(sm.x, sm.y, sm.z);
  float alpha = sm.w;
  u64 key = P.key[i];
  const u32 bounce = flags >> BOUNCE_SHIFT;
  const u32 dims = bounce_dim(bounce);
  const u32 sample_index = P.ident[i].w;
  // Caustics (integrator/mod.rs:1084-1113, 1261-1272): once a caster follows a receiver, the light
  // of photon lights reaching the path is the photon map's.
  bool caustics = false;
  if constexpr (EXT) caustics = S.caustics_on != 0;
  const bool dropped = caustics && (flags & FLAG_CAU_CASTERS) != 0;
  // A scattering walk that reached its solid's surface (integrator/mod.rs:1209-1242) leaves by
  // the exit lobe, whose albedo the walk kernel carried here; an absorbed walk ends the path.
  bool exit = false;
  float exit_eta2 = 0.0f;
  if constexpr (EXT) {
    // A walk in progress: its next step's ray goes to the next extend as it is.
    if (flags & FLAG_WALKING) {
      cont = true;
      return;
    }
    if (h.w == WALK_ABSORBED) return;
    exit = h.w == WALK_EXIT;
    exit_eta2 = rd.w;
  }
  C3 adj = splat(0.0f);
  if constexpr (ADJ) {
    float4 a = P.adj[i];
    adj = c3(a.x, a.y, a.z);
    uint4 tg = P.tgt[i];
    if (tg.x != NO_TARGET || tg.y != NO_TARGET) {
      float4 lt = P.ltot[i];
      C3 rest = c3(lt.x - sum.r, lt.y - sum.g, lt.z - sum.b);
      g_prev.mat = tg.x;
      g_prev.tex = tg.y;
      g_prev.ax = u2f(tg.z);
      g_prev.ay = u2f(tg.w);
      for (int k = 0; k < N_MAT_PARAMS; ++k) {
        float4 dl = P.dl[(size_t)N_MAT_PARAMS * i + k];
        g_prev.g[k] = dot3(adj, rest * c3(dl.x, dl.y, dl.z));
      }
    }
    P.tgt[i] = make_uint4(NO_TARGET, NO_TARGET, 0u, 0u);
  }
  // The selection of the vertex the ray left (only read once a BSDF sample made the ray).
  Sel from;
  if (has_pdf) {
    float4 sp = P.sel_p[i];
    from.point = v3(sp.x, sp.y, sp.z);
    from.total = sp.w;
    from.by_flux = (flags & FLAG_BY_FLUX) != 0;
    float4 w0 = P.sel_w0[i], w1 = P.sel_w1[i];
    from.w[0] = w0.x; from.w[1] = w0.y; from.w[2] = w0.z; from.w[3] = w0.w;
    from.w[4] = w1.x; from.w[5] = w1.y; from.w[6] = w1.z; from.w[7] = w1.w;
  }
  bool hit_ok = h.x != 0xffffffffu;
  float t_geometry = hit_ok ? u2f(h.z) : RPX_INF;
  float t_floor = RPX_INF;
  if ((S.floor_mode == 1) && d.z != 0.0f && !exit) {
    float tf = (S.floor_z - o.z) / d.z;
    if (tf > RAY_EPSILON && tf < RPX_INF) t_floor = tf;
  }
  float t_nearest = fmin_r(t_geometry, t_floor);
  if (in_medium && is_finite(t_nearest)) {
    C3 sg = c3(md.x, md.y, md.z);
    thr = thr * c3(fexp(-sg.r * t_nearest), fexp(-sg.g * t_nearest), fexp(-sg.b * t_nearest));
  }
  // Area lights crossed before the surface (light.rs:1354-1384); after a walk, the segment to its
  // exit vertex crosses none, and none count once their light is the photon map's.
  C3 crossed = splat(0.0f);
  for (u32 a = 0; a < (exit || dropped ? 0u : S.n_areas); ++a) {
    const DevArea& L = S.areas[a];
    bool visible = has_pdf ? L.glossy : L.camera;
    if (!visible) continue;
    float t, cos;
    if (!area_intersect(L, o, d, t_nearest, t, cos)) continue;
    float weight = 1.0f;
    if (has_pdf) {
      float dd = 0.0f + t;
      float light_pdf = sel_pmf<EXT>(S, from, L.entry) * dd * dd / (L.area * cos);
      weight = power_heuristic(prev_pdf, light_pdf);
    }
    crossed = crossed + radiance_at(L, cos) * weight;
  }
  if constexpr (EXT) {
    // The spheres after the areas, seen from the vertex the ray left (light.rs:1386-1425).
    for (u32 k = 0; k < (exit ? 0u : S.n_spheres); ++k) {
      const DevSphere& L = S.spheres[k];
      bool visible = has_pdf ? L.glossy : L.camera;
      if (!visible) continue;
      float t;
      if (!sphere_intersect(L, o, d, t_nearest, t)) continue;
      V3 vertex = o - d * 0.0f;
      float mask = sphere_mask(L, vertex);
      if (!(mask > 0.0f)) continue;
      float weight = 1.0f;
      if (has_pdf) {
        float omc;
        if (sphere_cone(L, length_sq(L.centre - vertex), omc) && omc > 0.0f)
          weight = power_heuristic(prev_pdf, sel_pmf<EXT>(S, from, L.entry) / (TAU * omc));
        else
          weight = 0.0f;
      }
      crossed = crossed + L.radiance * (mask * weight);
    }
  }
  add_to(sum, thr * crossed, bounce >= 2, S.clamp_indirect);
  bool done = false;
  if (t_nearest == RPX_INF) {
    if (primary) {
      if (S.background == 0) {
        C3 seen = splat(env_value(S, d.z));
        if constexpr (EXT) {
          // The sun discs flagged camera (light.rs:1502-1520).
          for (u32 k = 0; k < S.n_suns; ++k) {
            const DevSun& L = S.suns[k];
            if (L.camera && sun_contains(L, d)) seen = seen + L.radiance * 1.0f;
          }
        }
        if (S.exterior != 1.0f) seen = seen * S.exterior;
        sum = sum + thr * seen;
      } else if (S.background == 1) {
        sum = sum + thr * S.background_color;
      } else {
        alpha = peak(crossed) > 0.0f ? 1.0f : 0.0f;
      }
    } else {
      C3 sky = splat(env_value(S, d.z));
      if (has_pdf && S.has_env_entry)
        sky = sky * power_heuristic(prev_pdf, sel_pmf<EXT>(S, from, S.env_entry) * env_pdf(S, d.z));
      if constexpr (EXT) {
        // The sun discs a BSDF-sampled ray finds, weighted against next-event estimation (none
        // once their light is the photon map's: light.rs:1522-1533).
        for (u32 k = 0; k < (dropped && has_pdf ? 0u : S.n_suns); ++k) {
          const DevSun& L = S.suns[k];
          bool visible = has_pdf ? L.glossy : L.camera;
          if (visible && sun_contains(L, d)) {
            float weight = has_pdf ? power_heuristic(prev_pdf, sel_pmf<EXT>(S, from, L.entry) / L.solid_angle) : 1.0f;
            sky = sky + L.radiance * weight;
          }
        }
      }
      add_to(sum, thr * sky, bounce >= 2, S.clamp_indirect);
    }
    done = true;
  }
  V3 s_point, s_hit_point, s_normal, s_geometric;
  bool s_instance = false, s_front = true, s_outside = true;
  u32 inst = 0;
  if (!done) {
    if (t_floor < t_geometry) {
      V3 p = o + d * t_floor;
      V3 point = v3(p.x, p.y, S.floor_z);
      V3 nn = v3(0.0f, 0.0f, d.z < 0.0f ? 1.0f : -1.0f);
      s_point = point;
      s_hit_point = point;
      s_normal = nn;
      s_geometric = nn;
    } else {
      SurfaceHit sh = surface_hit(S, o, d, h);
      s_point = sh.point;
      s_hit_point = sh.hit_point;
      s_normal = sh.normal;
      s_geometric = sh.geometric;
      s_instance = true;
      s_front = sh.front;
      s_outside = sh.outside;
      inst = sh.inst;
    }
  }
  u32 mat = 0xffffffffu;
  if (!done && s_instance) {
    mat = S.instances[inst].material;
    if (mat < S.n_materials) {
      const DevMaterial& M = S.materials[mat];
      if (M.emissive) {
        bool one_sided = (S.instances[inst].flags & INST_ONE_SIDED) != 0;
        bool mapped = false;  // an emissive mesh sampled as a photon light, its light the map's
        if constexpr (EXT) mapped = dropped && S.instance_emitter[inst] >= 0;
        if ((s_front || !one_sided) && !mapped) {
          // A BSDF-sampled ray weighs an emitter sampled as a light against next-event estimation
          // (integrator/mod.rs:1548-1568, light.rs:1537-1551); others count at weight one.
          float weight = 1.0f;
          if constexpr (EXT) {
            i32 em = S.instance_emitter[inst];
            if (has_pdf && em >= 0) {
              const DevEmitter& E = S.emitters[em];
              float cos = dot(s_geometric, d);
              float light_pdf = sel_pmf<EXT>(S, from, E.entry) * t_geometry * t_geometry /
                                (E.area * fmax_r(fabs_(cos), 1.0e-7f));
              weight = power_heuristic(prev_pdf, light_pdf);
            }
          }
          add_to(sum, thr * M.radiance * weight, bounce >= 2, S.clamp_indirect);
        }
        done = true;
      }
    }
  }
  if (!done && bounce >= S.max_depth) done = true;
  if (!done) {
    const bool has_mat = s_instance && mat < S.n_materials;
    // A textured material's BSDF is built at the hit from the texture's base colour.
    const bool textured = has_mat && mat == S.tex_material;
    C3 base = has_mat ? S.materials[mat].base : splat(0.0f);
    TexTap tap{};
    if (textured) {
      // Built into this slot's BSDF record, so that B is a global-memory reference either way.
      tap = texture_tap(S.tex_w, S.tex_h, S.tex_origin_x, S.tex_origin_y, S.tex_scale_x, S.tex_scale_y, s_hit_point.x,
                        s_hit_point.y);
      base = texture_value(S.texture, S.tex_w, tap);
      const DevMaterial& M = S.materials[mat];
      P.bsdf_tmp[i] = material_bsdf(T, M, base, M.roughness, M.ior, s_outside);
    }
    const Bsdf& B = textured ? P.bsdf_tmp[i]
                             : (s_instance ? (mat < S.n_materials ? S.bsdfs[2 * mat + (s_outside ? 0 : 1)]
                                                                  : S.fallback_bsdf[s_outside ? 0 : 1])
                                           : S.floor_bsdf);
    VertexGrad target;
    no_grad(target);
    if constexpr (ADJ) {
      if (has_mat) {
        target.mat = mat;
        if (textured) {
          target.tex = tap.x0 | (tap.y0 << 16);
          target.ax = tap.ax;
          target.ay = tap.ay;
        }
      }
    }
    // The next-event term's derivative inputs (ADJ), and where its gradient goes: 1 this vertex's
    // g_now, 2 the shadow ray's record.
    bool nee_d = false, nee_diffuse = false;
    V3 nee_wi = v3(0.0f, 0.0f, 0.0f);
    C3 nee_thr = splat(0.0f), nee_rad = splat(0.0f);
    float nee_scale = 0.0f;
    u32 nee_to = 0;
    Frame frame = frame_from_normal(s_normal);
    V3 wo = to_local(frame, -d);
    V3 ng = s_geometric;
    V3 origin = offset_origin(s_point, ng);
    // Caustics (integrator/mod.rs:1672-1694): every receiver gathers the photons (the first one's
    // light direct, a later one's indirect); the state after this vertex says whether its light of
    // photon lights is the map's.
    u32 next_cflags = flags & (FLAG_CAU_RECV | FLAG_CAU_CASTERS);
    if constexpr (EXT) {
      if (caustics) {
        bool caster = s_instance && (S.instances[inst].flags & INST_CASTER) != 0;
        if (!caster) {
          u32 budget = S.max_depth > bounce + 1 ? S.max_depth - (bounce + 1) : 0u;
          V3 side = exit ? -s_geometric : s_geometric;
          // The map of this path's sample pass (the engine runs at most MAX_PASS_MAPS at once).
          const DevPhotonMaps& maps = *S.pmaps;
          u32 slot = sample_index - maps.first;
          C3 c = splat(0.0f);
          if (slot < maps.count)
            c = photon_gather(maps.map[slot], s_hit_point, side, budget, [&](V3 w) {
              V3 wl = to_local(frame, w);
              return exit ? exit_eval(exit_eta2, wo, wl) : bsdf_eval(T, B, wo, wl, false);
            });
          add_to(sum, thr * c, (flags & FLAG_CAU_RECV) != 0, S.clamp_indirect);
          next_cflags = FLAG_CAU_RECV;
        } else if (flags & FLAG_CAU_RECV) {
          next_cflags = FLAG_CAU_RECV | FLAG_CAU_CASTERS;
        }
      }
    }
    const bool drop_nee = next_cflags == (FLAG_CAU_RECV | FLAG_CAU_CASTERS);
    // Light selection at this vertex (light.rs:1026-1046).
    Sel sel;
    sel.point = s_point;
    sel.total = 0.0f;
    sel.by_flux = 1;
    for (int k = 0; k < KEPT_WEIGHTS; ++k) sel.w[k] = 0.0f;
    u32 ne = S.n_entries;
    if (S.per_point && ne > 0) {
      sel.by_flux = 0;
      for (u32 e = 0; e < ne; ++e) {
        float w = entry_weight<EXT>(S, e, s_point);
        if (e < (u32)KEPT_WEIGHTS) sel.w[e] = w;
        sel.total += w;
      }
    }
    // Next-event estimation (light.rs:1150-1315).
    float u_sel = sample_at(P.sobol, key, sample_index, dims + OFFSET_LIGHT_SELECT);
    float ul0 = sample_at(P.sobol, key, sample_index, dims + OFFSET_LIGHT);
    float ul1 = sample_at(P.sobol, key, sample_index, dims + OFFSET_LIGHT + 1);
    bool have_ls = false;
    V3 ls_wi = v3(0, 0, 0);
    float ls_dist = 0.0f, ls_pdf = 0.0f;
    C3 ls_rad = splat(0.0f);
    bool ls_glossy = true, ls_shadow = true, ls_delta = false;
    bool ls_photon = false;  // a light whose photons the map holds
    if (ne > 0) {
      u32 index = 0;
      float p = 0.0f, rest = 0.0f;
      bool ok = true;
      if (sel.by_flux) {
        // pick (light.rs:1059-1074): the first bucket whose end passes u, by bisection.
        u32 lo = 0, hi = ne;
        while (lo < hi) {
          u32 mid = (lo + hi) >> 1;
          if (S.cdf[mid] <= u_sel) lo = mid + 1;
          else hi = mid;
        }
        index = lo < ne - 1 ? lo : ne - 1;
        float lo_c = index == 0 ? 0.0f : S.cdf[index - 1];
        float width = S.cdf[index] - lo_c;
        rest = width > 0.0f ? clampf((u_sel - lo_c) / width, 0.0f, 0.99999994f) : 0.0f;
        p = sel_pmf<EXT>(S, sel, index);
      } else if (!(sel.total > 0.0f && is_finite(sel.total))) {
        ok = false;
      } else {
        float acc = 0.0f, lo_c = 0.0f;
        u32 last = ne - 1;
        for (u32 e = 0; e <= last; ++e) {
          float w = sel_weight<EXT>(S, sel, e);
          acc += w;
          float hi = e == last ? 1.0f : acc / sel.total;
          if (hi > u_sel || e == last) {
            float width = hi - lo_c;
            rest = width > 0.0f ? clampf((u_sel - lo_c) / width, 0.0f, 0.99999994f) : 0.0f;
            index = e;
            p = w / sel.total;
            break;
          }
          lo_c = hi;
        }
      }
      if (ok && p > 0.0f) {
        i32 a = S.entries[index];
        u32 ek = 0;
        EntryKind kind = a >= 0 ? E_AREA : (a == -1 ? E_ENV : E_SPHERE);
        if constexpr (EXT) kind = entry_kind(a, ek);
        if (EXT && kind == E_SUN) {
          // light.rs:1197-1223: the delta direction, or a direction in the disc's cone.
          const DevSun& L = S.suns[ek];
          if (L.delta) {
            ls_wi = L.frame.z;
            ls_pdf = p;
          } else {
            V3 local = uniform_cone(ul0, ul1, L.one_minus_cos_max);
            ls_wi = normalize(to_world(L.frame, local));
            ls_pdf = p / L.solid_angle;
          }
          ls_dist = RPX_INF;
          ls_rad = L.radiance;
          ls_delta = L.delta != 0;
          ls_glossy = L.glossy != 0;
          ls_shadow = L.shadow != 0;
          ls_photon = L.glossy != 0;
          have_ls = true;
        } else if (EXT && kind == E_MESH) {
          // light.rs:1270-1297: a triangle by area with the selection's remainder, a point
          // uniform on it.
          const DevEmitter& E = S.emitters[ek];
          const float* cdf = S.emit_cdf + E.first;
          u32 lo = 0, hi = E.count;
          while (lo < hi) {
            u32 mid = (lo + hi) >> 1;
            if (cdf[mid] <= rest) lo = mid + 1;
            else hi = mid;
          }
          u32 t = lo < E.count - 1 ? lo : E.count - 1;
          const float4* tp = S.emit_tris + 4 * (size_t)(E.first + t);
          float4 ta = tp[0], tb = tp[1], tc = tp[2], tn = tp[3];
          float su = sqrt_(ul0);
          float w0 = 1.0f - su, w1 = ul1 * su;
          float w2 = 1.0f - w0 - w1;
          V3 pt = (v3(ta.x, ta.y, ta.z) * w0 + v3(tb.x, tb.y, tb.z) * w1) + v3(tc.x, tc.y, tc.z) * w2;
          V3 dv = pt - s_point;
          float dist2 = length_sq(dv);
          float distance = sqrt_(dist2);
          V3 wi = dv / distance;
          float facing = -dot(wi, v3(tn.x, tn.y, tn.z));
          float cos = fabs_(facing);
          if (distance > 0.0f && (E.two_sided || facing > 0.0f) && cos > 0.0f) {
            ls_wi = wi;
            ls_dist = distance;
            ls_rad = E.radiance;
            ls_pdf = p * dist2 / (E.area * cos);
            ls_glossy = true;
            ls_shadow = true;
            ls_photon = true;
            have_ls = true;
          }
        } else if (EXT && kind == E_SPHERE) {
          // light.rs:1224-1268: a point (or a sphere too small to tell from one) by its
          // intensity over d^2, else a direction in the cone the sphere subtends.
          const DevSphere& L = S.spheres[ek];
          float mask = sphere_mask(L, s_point);
          V3 dv = L.centre - s_point;
          float dist2 = length_sq(dv);
          if (mask > 0.0f && dist2 > 0.0f) {
            float dist = sqrt_(dist2);
            V3 towards = dv / dist;
            float omc = 0.0f;
            if (!(L.radius > 0.0f) || sphere_cone(L, dist2, omc)) {
              if (omc == 0.0f) {
                ls_wi = towards;
                ls_dist = dist;
                ls_rad = L.intensity * (mask / (dist2 * 1.0e-6f));
                ls_pdf = p;
                ls_delta = true;
              } else {
                V3 local = uniform_cone(ul0, ul1, omc);
                V3 wi = normalize(to_world(frame_from_normal(towards), local));
                float b = dot(wi, dv);
                float r2 = L.radius * L.radius;
                float near_t = b - sqrt_(fmax_r(r2 - length_sq(dv - wi * b), 0.0f));
                ls_wi = wi;
                ls_dist = fmax_r(near_t, 0.0f);
                ls_rad = L.radiance * mask;
                ls_pdf = p / (TAU * omc);
                ls_delta = false;
              }
              ls_glossy = L.glossy != 0;
              ls_shadow = L.shadow != 0;
              have_ls = true;
            }
          }
        } else if (kind == E_AREA) {
          const DevArea& L = S.areas[a];
          float x0, y0;
          if (L.rect) {
            x0 = (2.0f * ul0 - 1.0f) * L.half_x;
            y0 = (2.0f * ul1 - 1.0f) * L.half_y;
          } else {
            float dx, dy;
            concentric_disk(ul0, ul1, dx, dy);
            x0 = dx * L.radius;
            y0 = dy * L.radius;
          }
          V3 pt = L.centre + L.frame.x * x0 + L.frame.y * y0;
          V3 dv = pt - s_point;
          float dist2 = length_sq(dv);
          float distance = sqrt_(dist2);
          V3 wi = dv / distance;
          float cos = -dot(wi, L.frame.z);
          if (cos > 0.0f && distance > 0.0f) {
            ls_wi = wi;
            ls_dist = distance;
            ls_rad = radiance_at(L, cos);
            ls_pdf = p * dist2 / (L.area * cos);
            ls_glossy = L.glossy != 0;
            ls_shadow = L.shadow != 0;
            ls_photon = L.glossy != 0;
            have_ls = true;
          }
        } else if (S.env_total_mass > 0.0f) {
          float target = ul0 * S.env_total_mass;
          float z;
          if (target < S.env_lower_mass) {
            z = -1.0f + invert_linear(S.env_floor, S.env_horizon, target / S.env_lower_mass);
          } else {
            float upper = S.env_total_mass - S.env_lower_mass;
            float fraction = clampf((target - S.env_lower_mass) / upper, 0.0f, 1.0f);
            z = invert_linear(S.env_horizon, S.env_zenith, fraction);
          }
          z = clampf(z, -1.0f, 1.0f);
          float rr = sqrt_(fmax_r((1.0f - z) * (1.0f + z), 0.0f));
          float sn, cs;
          sin_cos(TAU * ul1, sn, cs);
          V3 dir = v3(rr * cs, rr * sn, z);
          float pdf = env_pdf(S, dir.z);
          if (pdf > 0.0f) {
            ls_wi = dir;
            ls_dist = RPX_INF;
            ls_rad = splat(env_value(S, dir.z));
            ls_pdf = p * pdf;
            ls_glossy = true;
            ls_shadow = true;
            have_ls = true;
          }
        }
        if (have_ls && !(ls_pdf > 0.0f && is_finite(ls_pdf) && luminance(ls_rad) > 0.0f)) have_ls = false;
        if (drop_nee && ls_photon) have_ls = false;
      }
    }
    if (have_ls) {
      V3 wi = to_local(frame, ls_wi);
      C3 f = exit ? exit_eval(exit_eta2, wo, wi) : bsdf_eval(T, B, wo, wi, !ls_glossy);
      if (peak(f) > 0.0f) {
        V3 from_p = wi.z < 0.0f ? offset_origin(s_hit_point, -ng) : origin;
        float weight = (ls_delta || !ls_glossy)
                           ? 1.0f
                           : power_heuristic(ls_pdf, exit ? exit_pdf(wo, wi) : bsdf_pdf(T, B, wo, wi));
        C3 c = thr * f * ls_rad * (fabs_(wi.z) * weight / ls_pdf);
        // What `add` would add (finite, clamped), queued behind the shadow ray.
        C3 add_c = splat(0.0f);
        bool adds = finite_c(c);
        if (adds) {
          float pk = fmax_r(fmax_r(c.r, c.g), c.b);
          add_c = (bounce >= 1 && S.clamp_indirect > 0.0f && pk > S.clamp_indirect) ? c * (S.clamp_indirect / pk) : c;
        }
        if constexpr (ADJ) {
          // d(added contribution)/d parameter with the light sample and the MIS weight held; the
          // derivative is taken below, with the sampled direction's, from the same BSDF builds.
          if (adds && has_mat) {
            float pk = fmax_r(fmax_r(c.r, c.g), c.b);
            float scale = (bounce >= 1 && S.clamp_indirect > 0.0f && pk > S.clamp_indirect) ? S.clamp_indirect / pk : 1.0f;
            nee_d = true;
            nee_wi = wi;
            nee_diffuse = !ls_glossy;
            nee_thr = thr;
            nee_rad = ls_rad;
            nee_scale = fabs_(wi.z) * weight / ls_pdf * scale;
          }
        }
        if (!ls_shadow) {
          if (adds) sum = sum + add_c;
          if constexpr (ADJ) {
            if (adds && has_mat) nee_to = 1;
          }
        } else {
          float t_max = is_finite(ls_dist) ? ls_dist * (1.0f - 1.0e-4f) : RPX_INF;
          bool floor_blocks = false;
          if (S.floor_mode == 1 && ls_wi.z != 0.0f) {
            float tf = (S.floor_z - from_p.z) / ls_wi.z;
            floor_blocks = tf > RAY_EPSILON && tf < t_max;
          }
          if (floor_blocks) {
            atomicAdd((unsigned long long*)&P.rays[1], 1ull);
          } else if (adds) {
            want_shadow = true;
            P.sh_o[i] = make_float4(from_p.x, from_p.y, from_p.z, t_max);
            P.sh_d[i] = make_float4(ls_wi.x, ls_wi.y, ls_wi.z, 0.0f);
            P.sh_c[i] = make_float4(add_c.r, add_c.g, add_c.b, 0.0f);
            if constexpr (ADJ) nee_to = 2;
          } else {
            // A shadow ray is still cast (and counted) for a contribution that adds nothing.
            atomicAdd((unsigned long long*)&P.rays[1], 1ull);
          }
        }
      }
    }
    // The next direction (integrator/mod.rs:1772-1888).
    float ub0 = sample_at(P.sobol, key, sample_index, dims + OFFSET_BSDF);
    float ub1 = sample_at(P.sobol, key, sample_index, dims + OFFSET_BSDF + 1);
    BsdfSample bs;
    bool sampled_d = false;  // ADJ: the path goes on along bs.wi from a surface with a material
    if (exit ? exit_sample(exit_eta2, wo, ub0, ub1, bs) : bsdf_sample(T, B, wo, ub0, ub1, bs)) {
      bool transmitted = bs.wi.z < 0.0f;
      V3 wi = normalize(to_world(frame, bs.wi));
      bool ok = transmitted ? !(dot(wi, ng) >= 0.0f) : !(dot(wi, ng) <= 0.0f);
      if (ok) {
        thr = thr * bs.weight;
        if (bounce >= 2) {
          float survive = clampf(peak(thr), 0.05f, 1.0f);
          float u_rr = sample_at(P.sobol, key, sample_index, dims + OFFSET_ROULETTE);
          if (u_rr >= survive) ok = false;
          else thr = thr / survive;
        }
        if (ok && !(peak(thr) > 0.0f && is_finite(peak(thr)))) ok = false;
      }
      if (ok) {
        V3 next_o = transmitted ? offset_origin(s_hit_point, -ng) : origin;
        u32 nflags = FLAG_HAS_PDF | (sel.by_flux ? FLAG_BY_FLUX : 0u) | next_cflags | ((bounce + 1) << BOUNCE_SHIFT);
        C3 sg = c3(md.x, md.y, md.z);
        bool med = in_medium;
        bool scatter = false;
        if constexpr (EXT) {
          // The medium steps carry over; a sample transmitted into a solid right after an exit
          // vertex is probed for a thin wall (integrator/mod.rs:1822-1857: one ray counted).
          nflags |= flags & (EVENTS_MASK << EVENTS_SHIFT);
          if (exit) nflags |= FLAG_EXIT_PDF;
          else if ((flags & FLAG_EXIT_PDF) && s_instance && transmitted && s_outside)
            atomicAdd((unsigned long long*)&P.rays[3], 1ull);
          scatter = (flags & FLAG_SCATTER) != 0;
        }
        if (transmitted) {
          med = false;
          scatter = false;
          if (s_outside && s_instance && mat < S.n_materials) {
            u32 hm = S.materials[mat].has_medium;
            if (EXT ? hm == MEDIUM_ABSORBING : hm != 0) {
              med = true;
              sg = S.materials[mat].sigma;
            } else if (EXT && hm == MEDIUM_SCATTERING) {
              scatter = true;
              sg = c3(u2f(mat), 0.0f, 0.0f);
            }
          }
        }
        if (med) nflags |= FLAG_MEDIUM;
        if (scatter) nflags |= FLAG_SCATTER;
        P.ray_o[i] = make_float4(next_o.x, next_o.y, next_o.z, 0.0f);
        P.ray_d[i] = make_float4(wi.x, wi.y, wi.z, 0.0f);
        P.med[i] = make_float4(sg.r, sg.g, sg.b, u2f(nflags));
        P.thr[i] = make_float4(thr.r, thr.g, thr.b, bs.pdf);
        P.sel_p[i] = make_float4(sel.point.x, sel.point.y, sel.point.z, sel.total);
        P.sel_w0[i] = make_float4(sel.w[0], sel.w[1], sel.w[2], sel.w[3]);
        P.sel_w1[i] = make_float4(sel.w[4], sel.w[5], sel.w[6], sel.w[7]);
        cont = true;
        if constexpr (ADJ) sampled_d = has_mat;
      }
    }
    if constexpr (ADJ) {
      // The BSDF's derivatives at the next-event direction and along the sampled one, each
      // parameter's dual BSDF built once for both.
      if (nee_d || sampled_d) {
        C3 d_nee[N_MAT_PARAMS], d_s[N_MAT_PARAMS];
        bsdf_param_derivs(T, S.materials[mat], base, s_outside, wo, nee_d, nee_wi, nee_diffuse, sampled_d,
                          sampled_d ? bs.wi : nee_wi, d_nee, d_s);
        if (nee_d)
          for (int k = 0; k < N_MAT_PARAMS; ++k) target.g[k] = dot3(adj, nee_thr * d_nee[k] * nee_rad * nee_scale);
        if (sampled_d) {
          // d log f / d parameter along the sampled direction, applied once the radiance
          // gathered after this vertex is known (at the next vertex).
          C3 f_s = bsdf_eval(T, B, wo, bs.wi, false);
          for (int k = 0; k < N_MAT_PARAMS; ++k) {
            C3 q = c3(f_s.r > 0.0f ? d_s[k].r / f_s.r : 0.0f, f_s.g > 0.0f ? d_s[k].g / f_s.g : 0.0f,
                      f_s.b > 0.0f ? d_s[k].b / f_s.b : 0.0f);
            P.dl[(size_t)N_MAT_PARAMS * i + k] = make_float4(q.r, q.g, q.b, 0.0f);
          }
        }
      }
      if (nee_to == 1) g_now = target;
      if (nee_to == 2) {
        P.gn[2 * (size_t)i] = make_float4(target.g[0], target.g[1], target.g[2], target.g[3]);
        P.gn[2 * (size_t)i + 1] = make_float4(target.g[4], 0.0f, 0.0f, 0.0f);
      }
      if (nee_to == 2 || sampled_d) P.tgt[i] = make_uint4(target.mat, target.tex, f2u(target.ax), f2u(target.ay));
    }
  }
  P.sum[i] = make_float4(sum.r, sum.g, sum.b, alpha);
  }
}
}  // namespace

// Three waves per SIMD: the register budget (168) the shade fits without spilling; two waves leave
// the loads of a path's records unhidden, four spill. The extended shade spills 164 registers at
// three waves and none at two (222 registers): two waves are faster for scenes with media (the
// bottle's step 185 -> 180 ms) and slower with caustics, whose gather loops want the waves (1357 ->
// 1386 ms), so the extended shade has an instance of each (WAVES). The adjoint's spills (83 at
// three waves, 12 at two) cost it nothing measurable.
template <bool ADJ, bool EXT, int WAVES = 3>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES)))
k_shade(const DevScene* __restrict__ scenes, PathState P, u32 parity) {
  const u32* q = P.live[parity];
  u32 n = P.counts[C_LIVE0 + parity];
  u32* out_q = P.live[parity ^ 1];
  u32* out_n = &P.counts[C_LIVE0 + (parity ^ 1)];
  u32 stride = gridDim.x * blockDim.x;
  // Whole waves iterate together so that the ballots below see every lane.
  u32 n_iter = (n + stride - 1) / stride;
  u32 gid = blockIdx.x * blockDim.x + threadIdx.x;
  for (u32 it = 0; it < n_iter; ++it) {
    u32 j = it * stride + gid;
    bool active = j < n;
    bool cont = false, want_shadow = false;
    u32 i = 0, sid = 0;
    if (active) {
      i = q[j];
      sid = P.ident[i].y;
    }
    VertexGrad g_prev, g_now;
    no_grad(g_prev);
    no_grad(g_now);
    // One pass per distinct scene of the wave.
    bool todo = active;
    u64 pending = __ballot(todo);
    while (pending) {
      u32 cur = lowest_lane_value(pending, sid);
      if (todo && sid == cur) {
        shade_path<ADJ, EXT>(scenes[cur], P, i, cont, want_shadow, g_prev, g_now);
        todo = false;
      }
      pending = __ballot(todo);
    }
    if constexpr (ADJ) {
      commit_grad(P, active, g_prev);
      commit_grad(P, active, g_now);
    }
    // Compaction of the survivors, the finished paths and the shadow rays (one atomic per block
    // and queue).
    bool done = active && !cont;
    u32* const counters[3] = {out_n, &P.counts[c_done(parity)], &P.counts[c_shadow(parity)]};
    const bool want[3] = {cont, done, want_shadow};
    u32 slots[3];
    block_append<3>(counters, want, slots);
    if (cont) out_q[slots[0]] = i;
    if (done) P.done_queue[parity][slots[1]] = i;
    if (want_shadow) P.shadow_queue[parity][slots[2]] = i;
  }
}

// ================================================================== scattering media
namespace {
// integrator/medium.rs:143-150: the Henyey-Greenstein phase function per steradian.
__device__ __forceinline__ float hg_phase(float cos, float g) {
  float d = 1.0f + g * g - 2.0f * g * cos;
  return (1.0f - g * g) / (4.0f * PI * d * sqrt_(fmax_r(d, 1.0e-12f)));
}
// integrator/medium.rs:152-169: a direction about `dir` from the phase function, and its density.
__device__ __forceinline__ V3 hg_sample(V3 dir, float g, float u0, float u1, float& density) {
  float cos;
  if (fabs_(g) < 1.0e-3f) {
    cos = 1.0f - 2.0f * u0;
  } else {
    float s = (1.0f - g * g) / (1.0f - g + 2.0f * g * u0);
    cos = clampf((1.0f + g * g - s * s) / (2.0f * g), -1.0f, 1.0f);
  }
  float sin = sqrt_(fmax_r(1.0f - cos * cos, 0.0f));
  float sp, cp;
  sin_cos(TAU * u1, sp, cp);
  Frame frame = frame_from_normal(dir);
  density = hg_phase(cos, g);
  return normalize(to_world(frame, v3(sin * cp, sin * sp, cos)));
}
__device__ __forceinline__ C3 transmittance(C3 sigma, float t) {
  return c3(fexp(-sigma.r * t), fexp(-sigma.g * t), fexp(-sigma.b * t));
}
__device__ __forceinline__ float channel(C3 c, u32 k) { return k == 0 ? c.r : (k == 1 ? c.g : c.b); }
// The hero channel's estimate of the throughput (spectral MIS over the three channels).
__device__ __forceinline__ C3 effective(C3 w, C3 ratio) { return w * (3.0f / (ratio.r + ratio.g + ratio.b)); }

// The hit point and the geometric normal facing the ray of triangle `tri` of instance `inst`
// (geometry/mod.rs:731-808): what the walk needs at the solid's boundary.
__device__ void boundary_of(const DevScene& S, u32 inst, u32 tri, V3 o, V3 d, V3& point, V3& ng) {
  const DevInstance& in = S.instances[inst];
  const DevMesh& mesh = S.meshes[in.mesh];
  bool ident = (in.flags & INST_IDENTITY) != 0;
  V3 p0, p1, p2;
  corner_positions_of(S, mesh, tri, p0, p1, p2);
  V3 lo = ident ? o : xf_point(in.to_local, o), ld = ident ? d : xf_vector(in.to_local, d);
  PRay r = prepare(lo, ld, RAY_EPSILON, RPX_INF);
  float b0, b1, b2;
  {
    float u, v, w, az, bz, cz;
    shear(r, p0 - lo, p1 - lo, p2 - lo, u, v, w, az, bz, cz);
    bool negative = u < 0.0f || v < 0.0f || w < 0.0f;
    bool positive = u > 0.0f || v > 0.0f || w > 0.0f;
    float det = u + v + w;
    bool ok = !(negative && positive) && det != 0.0f;
    if (ok) {
      float t = (u * az + v * bz + w * cz) / det;
      ok = t >= r.t_min && t <= F32_MAX;
    }
    if (ok) {
      b0 = u / det;
      b1 = v / det;
      b2 = w / det;
    } else {
      b0 = b1 = b2 = 1.0f / 3.0f;
    }
  }
  V3 point_local = p0 * b0 + p1 * b1 + p2 * b2;
  V3 wind = cross(p1 - p0, p2 - p0);
  V3 winding = normalize(ident ? wind : xf_transpose(in.to_local, wind));
  V3 geometric = length_sq(winding) > 0.0f ? winding : -normalize(d);
  ng = dot(geometric, d) > 0.0f ? -geometric : geometric;
  point = ident ? point_local : xf_point(in.to_world, point_local);
}
}  // namespace

// One step of the random walk (integrator/medium.rs:226-360) of the path in slot i when it is
// inside a scattering medium, at the closest hit extend found for its ray: the loop body of the walk,
// whose trace is the extend of the iteration. A scattering event, or a reflection back in at the
// boundary, gives the next step's ray and keeps the path walking (the shade passes it on, the
// next extend traces it, one closest-hit ray of the reference's count); otherwise the walk ends:
// it crossed without an event (the surface's own BSDF follows), leaves by the exit lobe after
// scattering, escaped, or was absorbed. Distances are drawn in a hero channel picked per walk, the
// others weighted by spectral MIS; after MEDIUM_EVENTS steps of the path the walk goes straight
// to the boundary. One lane per path, no traversal: every step's ray is traced at full width.
__device__ void walk_step(const DevScene* __restrict__ scenes, const PathState& P, u32 i) {
  {
    const float4 md = P.med[i];
    const u32 flags = f2u(md.w);
    if (!(flags & FLAG_SCATTER)) return;
    const uint4 id = P.ident[i];
    const DevScene& S = scenes[id.y];
    const DevMedium M = S.media[f2u(md.x)];
    const u64 key = P.key[i];
    const u32 index = id.w;
    const u32 bounce = flags >> BOUNCE_SHIFT;
    u32 path_events = (flags >> EVENTS_SHIFT) & EVENTS_MASK;
    const float4 ro = P.ray_o[i], rd = P.ray_d[i];
    V3 o = v3(ro.x, ro.y, ro.z), d = v3(rd.x, rd.y, rd.z);
    const uint4 h = P.hit[i];
    const bool has_hit = h.x != 0xffffffffu;
    const float4 th = P.thr[i];
    const C3 entry = c3(th.x, th.y, th.z);
    C3 w = splat(1.0f), ratio = splat(1.0f), thr = entry;
    u32 events = 0, steps = 0;
    float pdf = 0.0f;
    if (flags & FLAG_WALKING) {
      float4 a = P.walk_w[i], b = P.walk_r[i];
      w = c3(a.x, a.y, a.z);
      pdf = a.w;
      ratio = c3(b.x, b.y, b.z);
      u32 es = f2u(b.w);
      events = es & 0xffffu;
      steps = es >> 16;
    }
    float uh = sample_at(P.sobol, key, index, medium_dim(bounce, 0) + MEDIUM_OFFSET_HERO) * 3.0f;
    u32 hero = f_to_u32(uh);
    hero = hero < 2u ? hero : 2u;
    const float sigma_hero = channel(M.sigma_t, hero);
    u32 end = 0;  // 0: the walk goes on with the ray (o, d)
    const float t_hit = has_hit ? u2f(h.z) : RPX_INF;
    const bool capped = path_events >= MEDIUM_EVENTS || is_nan(sigma_hero) || sigma_hero <= 0.0f;
    const u32 dims = medium_dim(bounce, steps);
    steps += 1;
    float t = RPX_INF;
    if (!capped) {
      float u = sample_at(P.sobol, key, index, dims + MEDIUM_OFFSET_DISTANCE);
      t = -fln(1.0f - u) / sigma_hero;
    }
    if (t < t_hit) {
      // A scattering event.
      C3 tr = transmittance(M.sigma_t, t);
      C3 pe = M.sigma_t * tr;
      C3 fe = M.sigma_s * tr;
      float ph = channel(pe, hero);
      if (is_nan(ph) || ph <= 0.0f) {
        end = WALK_ABSORBED;
      } else {
        w = w * (fe / ph);
        ratio = ratio * (pe / ph);
        float mx = peak(ratio);
        if (mx > 0.0f) {
          w = w / mx;
          ratio = ratio / mx;
        }
        events += 1;
        path_events += 1;
        C3 current = entry * effective(w, ratio);
        if (!(peak(current) > 0.0f && is_finite(peak(current)))) {
          end = WALK_ABSORBED;
        } else if (events > 32) {
          float survive = clampf(peak(current) / peak(entry), 0.05f, 1.0f);
          if (sample_at(P.sobol, key, index, dims + MEDIUM_OFFSET_ROULETTE) >= survive) end = WALK_ABSORBED;
          else w = w / survive;
        }
        if (end == 0) {
          V3 x = o + d * t;
          float u0 = sample_at(P.sobol, key, index, dims + MEDIUM_OFFSET_PHASE);
          float u1 = sample_at(P.sobol, key, index, dims + MEDIUM_OFFSET_PHASE + 1);
          float density;
          V3 dir = hg_sample(d, M.g, u0, u1, density);
          pdf = density;
          o = x;
          d = dir;
        }
      }
    } else {
      // No scattering before the surface (or the step cap is reached: the path then goes to the
      // surface with only the absorption applied).
      if (is_finite(t_hit)) {
        if (capped) {
          w = w * transmittance(M.sigma_a, t_hit);
        } else {
          C3 pe = transmittance(M.sigma_t, t_hit);
          float ph = channel(pe, hero);
          if (is_nan(ph) || ph <= 0.0f) {
            end = WALK_ABSORBED;
          } else {
            w = w * (pe / ph);
            ratio = ratio * (pe / ph);
          }
        }
      }
      if (end == 0) {
        thr = entry * effective(w, ratio);
        if (!(peak(thr) > 0.0f && is_finite(peak(thr)))) end = WALK_ABSORBED;
        else if (!has_hit) end = WALK_ESCAPED;
        else if (events == 0) end = WALK_CROSSED;
      }
      if (end == 0) {
        V3 point, ng;
        boundary_of(S, h.x, h.y, o, d, point, ng);
        float cos = clampf(-dot(d, ng), 0.0f, 1.0f);
        // integrator/medium.rs:218-224: leave with the probability of the Fresnel transmittance out.
        bool leave = capped || sample_at(P.sobol, key, index, dims + MEDIUM_OFFSET_FRESNEL) >=
                                   fresnel_dielectric(cos, 1.0f / fmax_r(M.ior, 1.0001f));
        if (leave) {
          end = WALK_EXIT;
        } else {
          // Fresnel sent it inside again: this counts against the walk's step cap like a scattering event.
          path_events += 1;
          V3 dir = normalize(d + ng * (2.0f * cos));
          o = offset_origin(point, ng);
          d = dir;
        }
      }
    }
    u32 nflags = (flags & ~(EVENTS_MASK << EVENTS_SHIFT)) | (path_events << EVENTS_SHIFT);
    if (end == 0) {
      P.walk_w[i] = make_float4(w.r, w.g, w.b, pdf);
      P.walk_r[i] = make_float4(ratio.r, ratio.g, ratio.b, u2f(events | (steps << 16)));
      P.ray_o[i] = make_float4(o.x, o.y, o.z, ro.w);
      P.ray_d[i] = make_float4(d.x, d.y, d.z, rd.w);
      P.med[i] = make_float4(md.x, md.y, md.z, u2f(nflags | FLAG_WALKING));
      return;
    }
    // The walk is over: the flags lose FLAG_WALKING whatever the end (an absorbed path must reach
    // the shade's end of the path, not be passed on as a step).
    nflags &= ~FLAG_WALKING;
    if (end == WALK_EXIT) nflags = (nflags & ~(FLAG_SCATTER | FLAG_PRIMARY)) | FLAG_HAS_PDF;
    P.med[i] = make_float4(md.x, md.y, md.z, u2f(nflags));
    if (end != WALK_ABSORBED) {
      // The walk's last ray and the throughput it leaves with (the exit's eta^2 for its lobe).
      P.ray_o[i] = make_float4(o.x, o.y, o.z, ro.w);
      P.ray_d[i] = make_float4(d.x, d.y, d.z, end == WALK_EXIT ? M.ior * M.ior : 0.0f);
      P.thr[i] = make_float4(thr.r, thr.g, thr.b, end == WALK_EXIT ? pdf : th.w);
    }
    P.hit[i] = make_uint4(h.x, h.y, h.z, end);
  }
}

// The walk steps of every live path (a step kernel between extend and shade).
__global__ void __launch_bounds__(256) k_walk(const DevScene* __restrict__ scenes, PathState P, u32 parity) {
  const u32* q = P.live[parity];
  const u32 n = P.counts[C_LIVE0 + parity];
  const u32 stride = gridDim.x * blockDim.x;
  for (u32 j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) walk_step(scenes, P, q[j]);
}

// ================================================================== drain
// The last paths of a chunk, once no new path is left: launched iteration by iteration, every
// bounce would wait for the slowest ray of the whole pool (a few rays in thousands walk hundreds
// of nodes). Here each lane takes its path to its end in one launch (closest hit, shade, shadow
// ray, again), so a path waits only for its own rays and those of its wave. The vertices run in
// the loop's order on the same per-slot state, so every path's radiance is the loop's, bit for
// bit; the radiance goes to the results as the regen would put it there. EXT: the extended shade,
// and each closest hit inside a scattering medium is a walk step (a walking path traces again).
template <int BLK, int LDSN, bool EXT>
__global__ void __launch_bounds__(BLK) k_finish(const DevScene* __restrict__ scenes, PathState P, u32 parity,
                                                u32 lanes, u32 park_at) {
  const u32 n = P.counts[C_TAIL];
  if (n == 0) return;
  __shared__ u32 lds[LDSN * BLK];
  u32* col = lds + threadIdx.x;
  const u32* q = P.live[parity];
  // The first `lanes` lanes of each wave take paths (fewer lanes: a wave's step waits for fewer
  // traversals).
  const u32 wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64u, lane_id = __lane_id();
  const u32 stride = (gridDim.x * blockDim.x / 64u) * lanes;
  // Whole waves iterate together (the shade's scene passes are wave operations).
  u32 n_iter = (n + stride - 1) / stride;
  u32 nearest = 0, shadow = 0;
  // With walks, once no more than park_at paths of the drain are left anywhere, the waves hand
  // theirs on to k_finish_wave (one wave per path) at the end of a vertex.
  const bool can_park = EXT && n_iter == 1 && P.parked != nullptr && park_at > 0;
  for (u32 it = 0; it < n_iter; ++it) {
    u32 j = it * stride + wave * lanes + lane_id;
    bool active = lane_id < lanes && j < n;
    u32 i = 0, sid = 0;
    if (active) {
      i = q[j];
      sid = P.ident[i].y;
    }
    // The first closest-hit ray was counted by the bookkeeping that handed the path over.
    bool counted = true;
    while (__ballot(active)) {
      if (active) {
        float4 o = P.ray_o[i], d = P.ray_d[i];
        bool primary = (f2u(P.med[i].w) & FLAG_PRIMARY) != 0;
        Geo g = geo_of(scenes, sid);
        u32 inst = 0xffffffffu, tri = 0xffffffffu;
        float t = RPX_INF;
        trace_nearest<BLK, LDSN>(g, col, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), RAY_EPSILON, RPX_INF, primary, inst,
                                 tri, t);
        P.hit[i] = make_uint4(inst, tri, f2u(t), 0u);
        if (!counted) ++nearest;
        counted = false;
        if constexpr (EXT) walk_step(scenes, P, i);
      }
      bool cont = false, want_shadow = false;
      VertexGrad g_prev, g_now;
      no_grad(g_prev);
      no_grad(g_now);
      bool todo = active;
      u64 pending = __ballot(todo);
      while (pending) {
        u32 cur = lowest_lane_value(pending, sid);
        if (todo && sid == cur) {
          shade_path<false, EXT>(scenes[cur], P, i, cont, want_shadow, g_prev, g_now);
          todo = false;
        }
        pending = __ballot(todo);
      }
      if (want_shadow) {
        ++shadow;
        float4 o = P.sh_o[i], d = P.sh_d[i];
        Geo g = geo_of(scenes, sid);
        if (!trace_any<BLK, LDSN>(g, col, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), RAY_EPSILON, o.w)) {
          float4 c = P.sh_c[i];
          float4 s = P.sum[i];
          P.sum[i] = make_float4(s.x + c.x, s.y + c.y, s.z + c.z, s.w);
        }
      }
      bool finished = active && !cont;
      if (finished) {
        P.results[P.ident[i].x] = P.sum[i];
        P.ident[i] = make_uint4(NO_WORK, 0u, 0u, 0u);
        active = false;
      }
      if (can_park) {
        const u64 fin = __ballot(finished);
        u32 left = 0;
        if (__lane_id() == 0) {
          u32 k = (u32)__popcll(fin);
          left = k ? atomicSub(&P.counts[C_DRAIN_LEFT], k) - k : atomicAdd(&P.counts[C_DRAIN_LEFT], 0u);
        }
        left = (u32)__shfl((int)left, 0);
        const u64 m = __ballot(active);
        if (m != 0 && left <= park_at) {
          u32 base = 0;
          if (__lane_id() == 0) base = atomicAdd(&P.counts[C_PARKED], (u32)__popcll(m));
          base = (u32)__shfl((int)base, 0);
          if (active) {
            u32 pos = base + __builtin_amdgcn_mbcnt_hi((u32)(m >> 32), __builtin_amdgcn_mbcnt_lo((u32)m, 0u));
            // At most park_at paths were left when any wave parked, so pos < park_at; a path that
            // found no room would stay here.
            if (pos < park_at) {
              P.parked[pos] = i;
              active = false;
            }
          }
        }
      }
    }
  }
  // One atomic per wave and counter.
  u32 wn = nearest, ws = shadow;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    wn += __shfl_xor(wn, off);
    ws += __shfl_xor(ws, off);
  }
  if (__lane_id() == 0) {
    if (wn) atomicAdd((unsigned long long*)&P.rays[0], (unsigned long long)wn);
    if (ws) atomicAdd((unsigned long long*)&P.rays[2], (unsigned long long)ws);
  }
}

// The paths the drain parked, one wave each (block b takes parked path b): the loop of k_finish for
// one path, its closest-hit and shadow rays traced by the whole wave, everything else in lane 0.
// Lane 0 alone reads and writes the path's state; the other lanes get the rays from its registers.
__global__ void __launch_bounds__(64) k_finish_wave(const DevScene* __restrict__ scenes, PathState P, u32 park_at) {
  const u32 n = P.counts[C_PARKED];
  if (blockIdx.x >= (n < park_at ? n : park_at)) return;
  __shared__ WaveStack stack;
  const u32 lane = __lane_id();
  const u32 i = P.parked[blockIdx.x];
  const u32 sid = P.ident[i].y;
  const Geo g = geo_of(scenes, sid);
  u32 nearest = 0, shadow = 0;
  for (;;) {
    float4 o = make_float4(0, 0, 0, 0), d = make_float4(0, 0, 0, 0);
    u32 flags = 0;
    if (lane == 0) {
      o = P.ray_o[i];
      d = P.ray_d[i];
      flags = f2u(P.med[i].w);
    }
    V3 wo = v3(__shfl(o.x, 0), __shfl(o.y, 0), __shfl(o.z, 0));
    V3 wd = v3(__shfl(d.x, 0), __shfl(d.y, 0), __shfl(d.z, 0));
    bool primary = ((u32)__shfl((int)flags, 0) & FLAG_PRIMARY) != 0;
    u32 inst = 0xffffffffu, tri = 0xffffffffu;
    float t = RPX_INF;
    trace_nearest_wave(g, stack, WAVE_STACK, wo, wd, RAY_EPSILON, RPX_INF, primary, inst, tri, t);
    bool cont = false, want_shadow = false;
    float4 so = make_float4(0, 0, 0, 0), sd = make_float4(0, 0, 0, 0);
    if (lane == 0) {
      P.hit[i] = make_uint4(inst, tri, f2u(t), 0u);
      ++nearest;
      walk_step(scenes, P, i);
      VertexGrad g_prev, g_now;
      no_grad(g_prev);
      no_grad(g_now);
      shade_path<false, true>(scenes[sid], P, i, cont, want_shadow, g_prev, g_now);
      if (want_shadow) {
        so = P.sh_o[i];
        sd = P.sh_d[i];
      }
    }
    cont = __shfl((int)cont, 0) != 0;
    if (__shfl((int)want_shadow, 0)) {
      ++shadow;
      V3 po = v3(__shfl(so.x, 0), __shfl(so.y, 0), __shfl(so.z, 0));
      V3 pd = v3(__shfl(sd.x, 0), __shfl(sd.y, 0), __shfl(sd.z, 0));
      bool blocked = trace_any_wave(g, stack, WAVE_STACK, po, pd, RAY_EPSILON, __shfl(so.w, 0));
      if (!blocked && lane == 0) {
        float4 c = P.sh_c[i];
        float4 s = P.sum[i];
        P.sum[i] = make_float4(s.x + c.x, s.y + c.y, s.z + c.z, s.w);
      }
    }
    if (!cont) {
      if (lane == 0) {
        P.results[P.ident[i].x] = P.sum[i];
        P.ident[i] = make_uint4(NO_WORK, 0u, 0u, 0u);
      }
      break;
    }
  }
  if (lane == 0) {
    atomicAdd((unsigned long long*)&P.rays[0], (unsigned long long)nearest);
    if (shadow) atomicAdd((unsigned long long*)&P.rays[2], (unsigned long long)shadow);
  }
}

// ================================================================== splat / tile output
// Adds samples first .. first + count of pixel p of segment s to its film, in sample order
// (image.rs:81-94): the sums are carried in registers and stored once, which performs the same f32
// additions in the same order as adding sample by sample. Results slot base + p * count + k holds
// sample first + k of the segment's p-th pixel in tiled order (sample-minor: a pixel's samples are
// contiguous).
// Adds the samples of pixels p0 .. p0 + T - 1 of segment `s` to the film, each pixel's in sample
// order (lane t: pixel p0 + t). A pixel's results are contiguous (work ids sample-minor); the block
// stages SPLAT_CH samples of each of its pixels at a time in LDS with coalesced loads, then each
// lane adds its own pixel's. Every lane of the block must call this with the same p0.
// `res`, `stride` and `skip`: pixel p's samples start at res + p * stride + skip (a segment's own
// results: stride count, skip 0; results kept from an earlier run may hold more samples per pixel
// than the segment adds).
constexpr u32 SPLAT_CH = 8;
template <u32 T>
__device__ __forceinline__ void splat_block(const float4* res, u32 stride, u32 skip, const DevSegment& s, u32 p0,
                                            float4 (*tile)[SPLAT_CH + 1]) {
  const u32 n = s.w * s.h, count = s.count, t = threadIdx.x, p = p0 + t;
  const bool mine = p < n;
  size_t f = 0;
  float r = 0.0f, g = 0.0f, b = 0.0f, wt = 0.0f, al = 0.0f, l2 = 0.0f;
  if (mine) {
    u32 lx, ly;
    tiled_pixel(s.w, s.h, p, lx, ly);
    f = (size_t)(s.y0 + ly) * s.film_w + (s.x0 + lx);
    r = s.film.rgb[3 * f];
    g = s.film.rgb[3 * f + 1];
    b = s.film.rgb[3 * f + 2];
    wt = s.film.weight[f];
    al = s.film.alpha[f];
    l2 = s.film.lum2[f];
  }
  for (u32 c0 = 0; c0 < count; c0 += SPLAT_CH) {
    const u32 m = count - c0 < SPLAT_CH ? count - c0 : SPLAT_CH;
    __syncthreads();
    for (u32 e = t; e < T * SPLAT_CH; e += T) {
      u32 q = e / SPLAT_CH, k = e - q * SPLAT_CH;
      if (k < m && p0 + q < n) tile[q][k] = res[(size_t)(p0 + q) * stride + skip + c0 + k];
    }
    __syncthreads();
    if (mine)
      for (u32 k = 0; k < m; ++k) {
        float4 v = tile[t][k];
        C3 c = c3(v.x, v.y, v.z) * splat(1.0f);
        r += c.r;
        g += c.g;
        b += c.b;
        wt += 1.0f;
        al += v.w;
        float l = luminance(c);
        l2 += l * l;
      }
  }
  if (mine) {
    s.film.rgb[3 * f] = r;
    s.film.rgb[3 * f + 1] = g;
    s.film.rgb[3 * f + 2] = b;
    s.film.weight[f] = wt;
    s.film.alpha[f] = al;
    s.film.lum2[f] = l2;
  }
}

__global__ void __launch_bounds__(256) k_splat(PathState P, DevSegment s) {
  __shared__ float4 tile[256][SPLAT_CH + 1];
  const u32 n = s.w * s.h;
  for (u32 p0 = blockIdx.x * 256; p0 < n; p0 += gridDim.x * 256)
    splat_block<256>(P.results + s.base, s.count, 0u, s, p0, tile);
}

// Many small segments whose pixels are disjoint (the blocks of an adaptive round): one wave per
// segment, in one launch.
__global__ void __launch_bounds__(64) k_splat_segments(PathState P, const DevSegment* __restrict__ segs, u32 n_segs) {
  __shared__ float4 tile[64][SPLAT_CH + 1];
  for (u32 i = blockIdx.x; i < n_segs; i += gridDim.x) {
    DevSegment s = segs[i];
    u32 n = s.w * s.h;
    for (u32 p0 = 0; p0 < n; p0 += 64) splat_block<64>(P.results + s.base, s.count, 0u, s, p0, tile);
  }
}

// Segments whose samples were rendered ahead by an earlier run (the adaptive rounds' kept
// results): entry i adds `seg.count` samples per pixel from its own results, one wave per entry,
// the entries' pixels disjoint.
__global__ void __launch_bounds__(64) k_splat_kept(const KeptSplat* __restrict__ kept, u32 n) {
  __shared__ float4 tile[64][SPLAT_CH + 1];
  for (u32 i = blockIdx.x; i < n; i += gridDim.x) {
    KeptSplat k = kept[i];
    u32 npx = k.seg.w * k.seg.h;
    for (u32 p0 = 0; p0 < npx; p0 += 64) splat_block<64>(k.res, k.stride, k.skip, k.seg, p0, tile);
  }
}

// The one-sample results of a w x h tile (work ids in tiled pixel order) as row-major rgb or rgba.
__global__ void __launch_bounds__(256) k_tile_out(PathState P, float* out, u32 w, u32 h, u32 channels) {
  u32 n = w * h;
  u32 stride = gridDim.x * blockDim.x;
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    float4 s = P.results[i];
    u32 x, y;
    tiled_pixel(w, h, i, x, y);
    size_t o = (size_t)y * w + x;
    // A one-sample film resolves to the sample times 1/1.
    float inv = 1.0f / 1.0f;
    out[channels * o] = s.x * inv;
    out[channels * o + 1] = s.y * inv;
    out[channels * o + 2] = s.z * inv;
    if (channels == 4) out[channels * o + 3] = s.w * inv;
  }
}

// ================================================================== ray queries for parity tests
__global__ void __launch_bounds__(128) k_trace_rays(const DevScene* __restrict__ scene, const float* rays, u32 n,
                                                     u32* inst, u32* tri, float* t, uint8_t* any) {
  __shared__ u32 lds[32 * 128];
  u32* col = lds + threadIdx.x;
  Geo g = geo_of(scene, 0);
  u32 stride = gridDim.x * blockDim.x;
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float* r = rays + 8 * (size_t)i;
    V3 o = v3(r[0], r[1], r[2]), d = v3(r[3], r[4], r[5]);
    u32 hi = 0xffffffffu, ht = 0xffffffffu;
    float tt = RPX_INF;
    trace_nearest<128, 32>(g, col, o, d, r[6], r[7], false, hi, ht, tt);
    inst[i] = hi;
    tri[i] = ht;
    t[i] = tt;
    any[i] = trace_any<128, 32>(g, col, o, d, r[6], r[7]) ? 1 : 0;
  }
}

// The same queries, one wave per ray (the drain's traversal of its last paths); `cap` entries of
// the wave's stack (at most WAVE_STACK) before it falls back to one lane.
__global__ void __launch_bounds__(64) k_trace_rays_wave(const DevScene* __restrict__ scene, const float* rays, u32 n,
                                                        int cap, u32* inst, u32* tri, float* t, uint8_t* any) {
  __shared__ WaveStack stack;
  Geo g = geo_of(scene, 0);
  for (u32 i = blockIdx.x; i < n; i += gridDim.x) {
    const float* r = rays + 8 * (size_t)i;
    V3 o = v3(r[0], r[1], r[2]), d = v3(r[3], r[4], r[5]);
    u32 hi = 0xffffffffu, ht = 0xffffffffu;
    float tt = RPX_INF;
    trace_nearest_wave(g, stack, cap, o, d, r[6], r[7], false, hi, ht, tt);
    bool a = trace_any_wave(g, stack, cap, o, d, r[6], r[7]);
    if (__lane_id() == 0) {
      inst[i] = hi;
      tri[i] = ht;
      t[i] = tt;
      any[i] = a ? 1 : 0;
    }
  }
}

// ================================================================== caustic photons: tracing
// caustics.rs:581-735 as a wavefront: photons leave their lights (k_photon_emit), then each step
// traces the closest hit of every photon still in flight (k_photon_trace, the lean traversal
// kernel) and lets it scatter (k_photon_scatter: stored on the first receiver unless it crossed no
// caster, through a caster by BSDF sampling, roulette from three casters); the survivors go to the
// other queue. A photon's random numbers come from its own PCG32 stream, so the order in which
// photons are processed changes nothing.
namespace {
// Appends `value` to `queue` for the lanes that want it: one atomic per wave.
__device__ __forceinline__ void wave_append(u32* queue, u32* count, bool want, u32 value) {
  u64 m = __ballot(want);
  if (m == 0) return;
  u32 lane = __lane_id();
  u32 leader = (u32)(__ffsll((long long)m) - 1);
  u32 base = 0;
  if (lane == leader) base = atomicAdd(count, (u32)__popcll(m));
  base = (u32)__shfl((int)base, (int)leader);
  if (want) queue[base + (u32)__popcll(m & ((1ull << lane) - 1ull))] = value;
}
__device__ __forceinline__ void wave_count(u32* count, bool want) {
  u64 m = __ballot(want);
  if (m != 0 && __lane_id() == (u32)(__ffsll((long long)m) - 1)) atomicAdd(count, (u32)__popcll(m));
}
__global__ void k_clear_u32(u32* p, u32 n) {
  for (u32 i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}
__device__ __forceinline__ u64 photon_seed(const DevCaustics& C, u32 iteration) {
  u64 key = ((u64)C.seed << 32) | (u64)iteration;
  return mix64(key ^ 0x0e5b0001c0decafeull);
}
}  // namespace

__global__ void __launch_bounds__(256) k_photon_emit(const DevScene* __restrict__ scene, u32 photons, PhotonBuffers B,
                                                     PhotonPaths Q, u32 iteration) {
  const DevScene& S = *scene;
  const u64 seed = photon_seed(S.caustics, iteration);
  u32 stride = gridDim.x * blockDim.x;
  for (u32 base = blockIdx.x * blockDim.x; base < photons; base += stride) {
    u32 index = base + threadIdx.x;
    bool live = false;
    if (index < photons) {
      B.key[index] = PHOTON_NONE;
      Pcg32 rng = Pcg32::seeded(seed, (u64)index);
      V3 o, d;
      C3 power;
      if (photon_emit(S, rng, o, d, power) && peak(power) > 0.0f && is_finite(peak(power))) {
        Q.o[index] = make_float4(o.x, o.y, o.z, 1.0f);  // w: the roulette's throughput
        Q.d[index] = make_float4(d.x, d.y, d.z, u2f(0u));  // w: casters crossed
        Q.power[index] = make_float4(power.r, power.g, power.b, u2f(0u));  // w: inside an absorbing solid
        Q.rng[index] = rng.state;
        live = true;
      }
    }
    wave_append(Q.live[0], Q.counts + 0, live, index);
  }
}

template <int BLK, int LDSN>
__global__ void __launch_bounds__(BLK) k_photon_trace(const DevScene* __restrict__ scene, PhotonPaths Q, u32 parity) {
  __shared__ u32 lds[LDSN * BLK];
  u32* col = lds + threadIdx.x;
  const DevScene& S = *scene;
  // The other queue was last read by the previous step's scatter: it starts empty for this one's.
  if (blockIdx.x == 0 && threadIdx.x == 0) Q.counts[parity ^ 1u] = 0;
  const u32* q = Q.live[parity];
  u32 n = Q.counts[parity];
  Geo g{S.nodes, S.tris, S.instances, S.tlas_root};
  u32 stride = gridDim.x * blockDim.x;
  for (u32 j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
    u32 i = q[j];
    float4 o = Q.o[i], d = Q.d[i];
    u32 hi = 0xffffffffu, ht = 0xffffffffu;
    float t = RPX_INF;
    trace_nearest<BLK, LDSN>(g, col, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), RAY_EPSILON, RPX_INF, false, hi, ht, t);
    Q.hit[i] = make_uint4(hi, ht, f2u(t), 0u);
  }
}

// One step of caustics.rs:600-735 after the closest hit; `last`: no step may follow this one.
__global__ void __launch_bounds__(256) k_photon_scatter(const DevScene* __restrict__ scene, PhotonBuffers B,
                                                        PhotonPaths Q, u32 parity, u32 iteration, u32 last) {
  const DevScene& S = *scene;
  const DevCaustics& C = S.caustics;
  const u32* q = Q.live[parity];
  u32 n = Q.counts[parity];
  u32 stride = gridDim.x * blockDim.x;
  for (u32 base = blockIdx.x * blockDim.x; base < n; base += stride) {
    u32 j = base + threadIdx.x;
    bool more = false, stored = false;
    u32 i = 0;
    if (j < n) {
      i = q[j];
      float4 o4 = Q.o[i], d4 = Q.d[i], p4 = Q.power[i];
      uint4 h = Q.hit[i];
      V3 o = v3(o4.x, o4.y, o4.z), d = v3(d4.x, d4.y, d4.z);
      float throughput = o4.w;
      u32 bounces = f2u(d4.w);
      C3 power = c3(p4.x, p4.y, p4.z);
      bool in_medium = f2u(p4.w) != 0u;
      bool has_hit = h.x != 0xffffffffu;
      float t_geometry = has_hit ? u2f(h.z) : RPX_INF;
      float t_floor = RPX_INF;
      if (S.floor_mode == 1 && d.z != 0.0f) {
        float tf = (S.floor_z - o.z) / d.z;
        if (tf > RAY_EPSILON && tf < RPX_INF) t_floor = tf;
      }
      float t_nearest = fmin_r(t_geometry, t_floor);
      if (in_medium && is_finite(t_nearest)) {
        float4 sg = Q.sigma[i];
        power = power * c3(fexp(-sg.x * t_nearest), fexp(-sg.y * t_nearest), fexp(-sg.z * t_nearest));
      }
      V3 at = v3(0.0f, 0.0f, 0.0f), normal_at = v3(0.0f, 0.0f, 1.0f);
      if (t_nearest == RPX_INF) {
        // escaped
      } else if (t_floor < t_geometry) {
        if (bounces > 0) {
          V3 p = o + d * t_floor;
          at = v3(p.x, p.y, S.floor_z);
          normal_at = v3(0.0f, 0.0f, d.z < 0.0f ? 1.0f : -1.0f);
          stored = true;
        }
      } else {
        SurfaceHit sh = surface_hit(S, o, d, h);
        u32 mat = S.instances[sh.inst].material;
        bool has_mat = mat < S.n_materials;
        V3 ng = sh.geometric;
        if (has_mat && S.materials[mat].emissive) {
          // absorbed by a light
        } else if (!(S.instances[sh.inst].flags & INST_CASTER)) {
          // A receiver ends the photon (no thin walls in the supported scenes).
          if (bounces > 0) {
            at = sh.hit_point;
            normal_at = ng;
            stored = true;
          }
        } else if (bounces < C.max_bounces) {
          Pcg32 rng;
          rng.state = Q.rng[i];
          rng.inc = ((u64)i << 1) | 1ull;
          Frame frame = frame_from_normal(sh.normal);
          V3 wo_world = -d;
          V3 wo = to_local(frame, wo_world);
          float u0 = rng.next_f32();
          float u1 = rng.next_f32();
          const Bsdf& Bs = has_mat ? S.bsdfs[2 * mat + (sh.outside ? 0 : 1)] : S.fallback_bsdf[sh.outside ? 0 : 1];
          BsdfSample bs;
          if (bsdf_sample(S.tables, Bs, wo, u0, u1, bs)) {
            bool transmitted = bs.wi.z < 0.0f;
            V3 wi = normalize(to_world(frame, bs.wi));
            if (!(transmitted ? dot(wi, ng) >= 0.0f : dot(wi, ng) <= 0.0f)) {
              C3 weight = bs.weight;
              if (transmitted) weight = weight * (Bs.eta * Bs.eta);
              // The adjoint BSDF's shading-normal correction (Veach), bounded.
              float num = fabs_(dot(wo_world, sh.normal)) * fabs_(dot(wi, ng));
              float den = fabs_(dot(wo_world, ng)) * fabs_(dot(wi, sh.normal));
              if (den > 0.0f) weight = weight * fmin_r(num / den, 10.0f);
              ++bounces;
              power = power * weight;
              throughput *= peak(weight);
              bool alive = true;
              if (bounces >= 3) {
                float survive = clampf(throughput, 0.05f, 1.0f);
                if (rng.next_f32() >= survive) {
                  alive = false;
                } else {
                  power = power * (1.0f / survive);
                  throughput /= survive;
                }
              }
              if (alive && peak(power) > 0.0f && is_finite(peak(power))) {
                V3 next_o = transmitted ? offset_origin(sh.hit_point, -ng) : offset_origin(sh.point, ng);
                if (transmitted) {
                  in_medium = false;
                  if (sh.outside && has_mat && S.materials[mat].has_medium == MEDIUM_ABSORBING) {
                    in_medium = true;
                    C3 sg = S.materials[mat].sigma;
                    Q.sigma[i] = make_float4(sg.r, sg.g, sg.b, 0.0f);
                  }
                }
                Q.o[i] = make_float4(next_o.x, next_o.y, next_o.z, throughput);
                Q.d[i] = make_float4(wi.x, wi.y, wi.z, u2f(bounces));
                Q.power[i] = make_float4(power.r, power.g, power.b, u2f(in_medium ? 1u : 0u));
                Q.rng[i] = rng.state;
                more = !last;
              }
            }
          }
        }
      }
      if (stored) {
        B.position[i] = make_float4(at.x, at.y, at.z, u2f(bounces));
        B.wi[i] = make_float4(-d.x, -d.y, -d.z, 0.0f);
        B.power[i] = make_float4(power.r, power.g, power.b, 0.0f);
        B.normal[i] = make_float4(normal_at.x, normal_at.y, normal_at.z, 0.0f);
        B.key[i] = PHOTON_STORED;
      }
    }
    (void)iteration;
    wave_append(Q.live[parity ^ 1u], Q.counts + (parity ^ 1u), more, i);
    wave_count(B.stored, stored);
  }
}

// ================================================================== exact and sampled passes
namespace {
constexpr u32 FIRST_BACKGROUND = 0, FIRST_FLOOR = 1, FIRST_SURFACE = 2;

// What a camera ray meets first (integrator/mod.rs:911-990): the background (nothing, or a light
// panel flagged `camera` in front of everything), the floor, or a surface with its face-forwarded
// shading normal (bevel.rs:372-377 without a bevel) and base colour.
struct First {
  u32 kind, inst;
  float t;
  V3 normal;
  C3 albedo;
};

template <int BLK, int LDSN>
__device__ First first_hit(const DevScene& S, u32* lds_col, V3 o, V3 d) {
  First f;
  f.kind = FIRST_BACKGROUND;
  f.inst = 0;
  f.t = 0.0f;
  f.normal = v3(0.0f, 0.0f, 0.0f);
  f.albedo = splat(0.0f);
  Geo g{S.nodes, S.tris, S.instances, S.tlas_root};
  u32 hi = 0, ht = 0;
  float tt = RPX_INF;
  bool hit = trace_nearest<BLK, LDSN>(g, lds_col, o, d, RAY_EPSILON, RPX_INF, true, hi, ht, tt);
  float t_geometry = hit ? tt : RPX_INF;
  float t_floor = RPX_INF;
  if (S.floor_mode == 1 && d.z != 0.0f) {
    float tf = (S.floor_z - o.z) / d.z;
    if (tf > RAY_EPSILON && tf < RPX_INF) t_floor = tf;
  }
  float t_nearest = fmin_r(t_geometry, t_floor);
  if (t_nearest == RPX_INF) return f;
  for (u32 a = 0; a < S.n_areas; ++a) {
    const DevArea& L = S.areas[a];
    float t, cos;
    if (L.camera && area_intersect(L, o, d, t_nearest, t, cos)) return f;
  }
  for (u32 k = 0; k < S.n_spheres; ++k) {
    const DevSphere& L = S.spheres[k];
    float t;
    if (L.camera && sphere_intersect(L, o, d, t_nearest, t) && sphere_mask(L, o) > 0.0f) return f;
  }
  if (t_floor < t_geometry) {
    f.kind = FIRST_FLOOR;
    f.t = t_floor;
    f.normal = v3(0.0f, 0.0f, d.z < 0.0f ? 1.0f : -1.0f);
    return f;
  }
  // Hit details (geometry/mod.rs:731-808): barycentrics from the scalar re-test, the shading
  // normal facing the ray, the hit point.
  const DevInstance& in = S.instances[hi];
  const DevMesh& mesh = S.meshes[in.mesh];
  bool ident = (in.flags & INST_IDENTITY) != 0;
  V3 p0, p1, p2, n0, n1, n2;
  corners_of(S, mesh, ht, p0, p1, p2, n0, n1, n2);
  V3 lo = ident ? o : xf_point(in.to_local, o), ld = ident ? d : xf_vector(in.to_local, d);
  PRay r = prepare(lo, ld, RAY_EPSILON, RPX_INF);
  float b0, b1, b2;
  {
    float u, v, w, az, bz, cz;
    shear(r, p0 - lo, p1 - lo, p2 - lo, u, v, w, az, bz, cz);
    bool negative = u < 0.0f || v < 0.0f || w < 0.0f;
    bool positive = u > 0.0f || v > 0.0f || w > 0.0f;
    float det = u + v + w;
    bool ok = !(negative && positive) && det != 0.0f;
    if (ok) {
      float t = (u * az + v * bz + w * cz) / det;
      ok = t >= r.t_min && t <= F32_MAX;
    }
    if (ok) {
      b0 = u / det;
      b1 = v / det;
      b2 = w / det;
    } else {
      b0 = b1 = b2 = 1.0f / 3.0f;
    }
  }
  V3 wind = cross(p1 - p0, p2 - p0);
  V3 winding = normalize(ident ? wind : xf_transpose(in.to_local, wind));
  V3 geometric = length_sq(winding) > 0.0f ? winding : -normalize(d);
  V3 smooth = v3(0.0f, 0.0f, 0.0f);
  if (mesh.has_normals) {
    V3 interp = n0 * b0 + n1 * b1 + n2 * b2;
    smooth = ident ? interp : xf_transpose(in.to_local, interp);
  }
  V3 normal = length_sq(smooth) > 1.0e-12f ? normalize(smooth) : geometric;
  f.kind = FIRST_SURFACE;
  f.inst = hi;
  f.t = tt;
  f.normal = dot(normal, d) > 0.0f ? -normal : normal;
  u32 mat = in.material;
  // An index outside the materials renders with the default material (base colour 0.8).
  C3 base = mat < S.n_materials ? S.materials[mat].base : c3(0.8f, 0.8f, 0.8f);
  if (mat == S.tex_material && mat < S.n_materials) {
    V3 point_local = p0 * b0 + p1 * b1 + p2 * b2;
    V3 hp = ident ? point_local : xf_point(in.to_world, point_local);
    TexTap tap = texture_tap(S.tex_w, S.tex_h, S.tex_origin_x, S.tex_origin_y, S.tex_scale_x, S.tex_scale_y, hp.x, hp.y);
    base = texture_value(S.texture, S.tex_w, tap);
  }
  f.albedo = base;
  return f;
}
}  // namespace

// Exact passes from the centre ray of each pixel (render.rs:500-556, passes.rs:177-206) and the
// albedo and coverage of samples [first, first + count) added to their sums in sample-index order
// (render.rs:558-640, passes.rs:209-235, 333-361), for the pixels of rectangle (x0, y0, w, h)
// inside the image. Pixel (px, py) is written at row py - oy, column px - ox of planes `ow` wide.
template <int BLK, int LDSN>
__global__ void __launch_bounds__(BLK) k_passes(const DevScene* __restrict__ scene, const u32* __restrict__ cols,
                                                PassPlanes Q, u32 x0, u32 y0, u32 w, u32 h, u32 ox, u32 oy, u32 ow,
                                                u32 first, u32 count, u32 what) {
  __shared__ u32 lds[LDSN * BLK];
  u32* col = lds + threadIdx.x;
  const DevScene& S = *scene;
  u32 n = w * h;
  u32 stride = gridDim.x * blockDim.x;
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    u32 px = x0 + i % w, py = y0 + i / w;
    if (px >= S.width || py >= S.height) continue;
    size_t o_ix = (size_t)(py - oy) * ow + (px - ox);
    if (what & PASSES_EXACT) {
      // centre_ray: the pinhole ray through the pixel's centre and the lens centre.
      float u = ((float)px + 0.5f) / (float)(S.width > 1 ? S.width : 1);
      float v = ((float)py + 0.5f) / (float)(S.height > 1 ? S.height : 1);
      V3 o, d;
      if (S.ortho) {
        float wdt = S.ortho_width;
        o = S.cam_position + S.cam_right * ((u - 0.5f) * wdt) + S.cam_up * ((0.5f - v) * wdt / S.aspect);
        d = S.cam_forward;
      } else {
        float x = (u - 0.5f) * S.sensor_w;
        float y = (0.5f - v) * S.sensor_h;
        V3 dir = normalize(S.cam_forward * S.focal_mm + S.cam_right * x + S.cam_up * y);
        if (S.lens_radius == 0.0f) {
          o = S.cam_position;
          d = dir;
        } else {
          V3 focus_point = S.cam_position + dir * (S.focus_dist / dot(dir, S.cam_forward));
          o = S.cam_position + (S.cam_right * 0.0f + S.cam_up * 0.0f) * S.lens_radius;
          d = normalize(focus_point - o);
        }
      }
      First f = first_hit<BLK, LDSN>(S, col, o, d);
      V3 nrm = v3(0.0f, 0.0f, 0.0f);
      float depth = RPX_INF, coc = 0.0f;
      u32 part = 0, material = 0;
      if (f.kind != FIRST_BACKGROUND) {
        if (f.kind == FIRST_FLOOR) {
          part = FLOOR_ID;
          material = FLOOR_ID;
        } else {
          u32 m = S.instances[f.inst].material;
          material = m < S.n_materials ? m + 1 : 0;
          part = f.inst + 1;
        }
        nrm = f.normal;
        depth = f.t * dot(d, S.cam_forward) + dot(o - S.cam_position, S.cam_forward);
        // coc_diameter (passes.rs:144-153)
        if (S.lens_radius > 0.0f && depth > 0.0f && is_finite(depth)) {
          float aperture = S.focal_mm / S.cam_fstop;
          float on_focus_plane = aperture * fabs_(depth - S.focus_dist) / depth;
          u32 long_side = S.width > S.height ? S.width : S.height;
          float pixel = S.cam_sensor_mm / (float)(long_side > 1 ? long_side : 1) * (S.focus_dist / S.focal_mm);
          coc = on_focus_plane / pixel;
        }
      }
      if (Q.normal) {
        Q.normal[3 * o_ix] = nrm.x;
        Q.normal[3 * o_ix + 1] = nrm.y;
        Q.normal[3 * o_ix + 2] = nrm.z;
      }
      if (Q.depth) Q.depth[o_ix] = depth;
      if (Q.part) Q.part[o_ix] = part;
      if (Q.material) Q.material[o_ix] = material;
      if (Q.coc) Q.coc[o_ix] = coc;
    }
    if ((what & PASSES_SAMPLED) && count > 0) {
      C3 sum = c3(Q.albedo[3 * o_ix], Q.albedo[3 * o_ix + 1], Q.albedo[3 * o_ix + 2]);
      float cover = Q.alpha[o_ix], wsum = Q.weight[o_ix];
      u64 key = pixel_key(S.seed, px, py);
      for (u32 k = 0; k < count; ++k) {
        V3 o, d;
        camera_ray(S, cols, key, px, py, first + k, o, d);
        First f = first_hit<BLK, LDSN>(S, col, o, d);
        C3 a = f.kind == FIRST_BACKGROUND ? splat(0.0f) : (f.kind == FIRST_FLOOR ? S.floor_color : f.albedo);
        sum = c3(sum.r + a.r, sum.g + a.g, sum.b + a.b);
        cover += f.kind == FIRST_SURFACE ? 1.0f : 0.0f;
        wsum += 1.0f;
      }
      Q.albedo[3 * o_ix] = sum.r;
      Q.albedo[3 * o_ix + 1] = sum.g;
      Q.albedo[3 * o_ix + 2] = sum.b;
      Q.alpha[o_ix] = cover;
      Q.weight[o_ix] = wsum;
    }
  }
}

// The sampled passes with one wave per pixel: lane k traces the camera ray of sample first + k (in
// rounds of 64), and the samples' albedo and coverage are added to the pixel's sums in sample order
// (render.rs:558-640, passes.rs:209-235, 333-361), the same f32 additions as one lane adding them
// one by one. A pixel's samples are nearly one ray, so a wave walks the same nodes.
template <int BLK, int LDSN>
__global__ void __launch_bounds__(BLK) k_passes_sampled(const DevScene* __restrict__ scene,
                                                        const u32* __restrict__ cols, PassPlanes Q, u32 x0, u32 y0,
                                                        u32 w, u32 h, u32 ox, u32 oy, u32 ow, u32 first, u32 count) {
  __shared__ u32 lds[LDSN * BLK];
  u32* col = lds + threadIdx.x;
  const DevScene& S = *scene;
  const u32 lane = __lane_id();
  const u32 n = w * h;
  const u32 waves = gridDim.x * (BLK / 64);
  for (u32 i = blockIdx.x * (BLK / 64) + threadIdx.x / 64; i < n; i += waves) {
    u32 px = x0 + i % w, py = y0 + i / w;
    if (px >= S.width || py >= S.height) continue;
    size_t o_ix = (size_t)(py - oy) * ow + (px - ox);
    float sr = Q.albedo[3 * o_ix], sg = Q.albedo[3 * o_ix + 1], sb = Q.albedo[3 * o_ix + 2];
    float cover = Q.alpha[o_ix], wsum = Q.weight[o_ix];
    u64 key = pixel_key(S.seed, px, py);
    for (u32 base = 0; base < count; base += 64) {
      u32 k = base + lane;
      float ar = 0.0f, ag = 0.0f, ab = 0.0f, cv = 0.0f;
      if (k < count) {
        V3 o, d;
        camera_ray(S, cols, key, px, py, first + k, o, d);
        First f = first_hit<BLK, LDSN>(S, col, o, d);
        C3 a = f.kind == FIRST_BACKGROUND ? splat(0.0f) : (f.kind == FIRST_FLOOR ? S.floor_color : f.albedo);
        ar = a.r;
        ag = a.g;
        ab = a.b;
        cv = f.kind == FIRST_SURFACE ? 1.0f : 0.0f;
      }
      u32 m = count - base < 64u ? count - base : 64u;
      for (u32 j = 0; j < m; ++j) {
        sr = sr + u2f(__builtin_amdgcn_readlane(f2u(ar), j));
        sg = sg + u2f(__builtin_amdgcn_readlane(f2u(ag), j));
        sb = sb + u2f(__builtin_amdgcn_readlane(f2u(ab), j));
        cover = cover + u2f(__builtin_amdgcn_readlane(f2u(cv), j));
        wsum += 1.0f;
      }
    }
    if (lane == 0) {
      Q.albedo[3 * o_ix] = sr;
      Q.albedo[3 * o_ix + 1] = sg;
      Q.albedo[3 * o_ix + 2] = sb;
      Q.alpha[o_ix] = cover;
      Q.weight[o_ix] = wsum;
    }
  }
}

// Resolved passes (passes.rs:388-420): the albedo and coverage sums times the f32 reciprocal of
// their sample count.
__global__ void __launch_bounds__(256) k_passes_resolve(PassPlanes Q, u32 n, float* albedo_out, float* alpha_out) {
  u32 stride = gridDim.x * blockDim.x;
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    float w = Q.weight[i];
    float r = 0.0f, g = 0.0f, b = 0.0f, a = 0.0f;
    if (w > 0.0f) {
      float inv = 1.0f / w;
      r = Q.albedo[3 * (size_t)i] * inv;
      g = Q.albedo[3 * (size_t)i + 1] * inv;
      b = Q.albedo[3 * (size_t)i + 2] * inv;
      a = Q.alpha[i] * (1.0f / w);
    }
    if (albedo_out) {
      albedo_out[3 * (size_t)i] = r;
      albedo_out[3 * (size_t)i + 1] = g;
      albedo_out[3 * (size_t)i + 2] = b;
    }
    if (alpha_out) alpha_out[i] = a;
  }
}

// ================================================================== energy tables (energy.rs recipe)
namespace {
__device__ double g2_over_g1(float a2, V3 wo, V3 wi) {
  float lo = ggx_lambda(wo, a2), li = ggx_lambda(wi, a2);
  return (double)((1.0f + lo) / (1.0f + lo + li));
}
__device__ void jitter(Pcg32& rng, u32 s, float& a, float& b) {
  const u32 strata = 64;
  u32 i = s % strata, j = s / strata;
  float inv = 1.0f / (float)strata;
  a = fmin_r(((float)i + rng.next_f32()) * inv, 0.99999994f);
  b = fmin_r(((float)j + rng.next_f32()) * inv, 0.99999994f);
}
__device__ uint16_t quantise(double v) {
  double c = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
  return (uint16_t)__builtin_round(c * 65535.0);
}
}  // namespace

// One thread per table node: 1024 albedo nodes, 1024 Schlick-moment nodes, 2 x 17408 clear nodes.
__global__ void k_energy_nodes(uint16_t* albedo, uint16_t* schlick, uint16_t* clear_in, uint16_t* clear_out) {
  const u32 SLICE = 1024, CLEAR = 1024 * 17;
  const u64 SEED = 0xe9ae9a;
  const u32 N = 64 * 64;
  u32 n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= 2 * SLICE + 2 * CLEAR) return;
  u32 kind, m;
  if (n < SLICE) { kind = 0; m = n; }
  else if (n < 2 * SLICE) { kind = 1; m = n - SLICE; }
  else if (n < 2 * SLICE + CLEAR) { kind = 2; m = n - 2 * SLICE; }
  else { kind = 3; m = n - 2 * SLICE - CLEAR; }
  u32 rest = m % SLICE, k = m / SLICE;
  u32 jj = rest / 32, ii = rest % 32;
  float t = (float)ii / 31.0f;
  float mu = fmax_r(t * t, 1.0f / 2048.0f);
  float r = (float)jj / 31.0f;
  float alpha = r * r;
  V3 wo = v3(sqrt_(fmax_r(1.0f - mu * mu, 0.0f)), 0.0f, mu);
  float a2 = alpha * alpha;
  double sum = 0.0;
  double value;
  if (kind == 0) {
    if (alpha <= 0.0f) value = 1.0;
    else {
      Pcg32 rng = Pcg32::seeded(SEED, (u64)m);
      for (u32 s = 0; s < N; ++s) {
        float ua, ub;
        jitter(rng, s, ua, ub);
        V3 h = sample_vndf(wo, alpha, alpha, ua, ub);
        V3 wi = h * (2.0f * dot(wo, h)) - wo;
        if (wi.z > 0.0f) sum += g2_over_g1(a2, wo, wi);
      }
      value = sum / (double)N;
    }
    albedo[m] = quantise(value);
  } else if (kind == 1) {
    if (alpha <= 0.0f) {
      double mm = (double)(1.0f - mu);
      value = mm * mm * mm * mm * mm;
    } else {
      Pcg32 rng = Pcg32::seeded(SEED ^ 0x5c41, (u64)m + (1ull << 24));
      for (u32 s = 0; s < N; ++s) {
        float ua, ub;
        jitter(rng, s, ua, ub);
        V3 h = sample_vndf(wo, alpha, alpha, ua, ub);
        float c = dot(wo, h);
        V3 wi = h * (2.0f * c) - wo;
        if (wi.z > 0.0f) {
          double mm = (double)clampf(1.0f - c, 0.0f, 1.0f);
          sum += mm * mm * mm * mm * mm * g2_over_g1(a2, wo, wi);
        }
      }
      value = sum / (double)N;
    }
    schlick[m] = quantise(value);
  } else {
    bool entering = kind == 2;
    const float MAX_X = (E_MAX_IOR - 1.0f) / (E_MAX_IOR + 1.0f);
    float x = MAX_X * (float)k / 16.0f;
    float ior = (1.0f + x) / (1.0f - x);
    float eta = entering ? ior : 1.0f / ior;
    if (alpha <= 0.0f || fabs_(eta - 1.0f) < 1.0e-6f) value = 1.0;
    else {
      Pcg32 rng = Pcg32::seeded(SEED ^ 0xc1ea, (u64)m + (entering ? 0ull : (1ull << 20)));
      for (u32 s = 0; s < N; ++s) {
        float ua, ub;
        jitter(rng, s, ua, ub);
        V3 h = sample_vndf(wo, alpha, alpha, ua, ub);
        float cos = dot(wo, h);
        float f = fresnel_dielectric(cos, eta);
        V3 wr = h * (2.0f * cos) - wo;
        if (wr.z > 0.0f) sum += (double)f * g2_over_g1(a2, wo, wr);
        if (f < 1.0f) {
          V3 wt;
          if (refract_dir(wo, h, eta, wt) && wt.z < 0.0f) sum += (double)(1.0f - f) * g2_over_g1(a2, wo, wt);
        }
      }
      value = sum / (double)N;
    }
    (entering ? clear_in : clear_out)[m] = quantise(value);
  }
}

// E_avg of each roughness node: the midpoint rule over 4096 values of mu of the table's own lookup.
__global__ void k_energy_avg(const uint16_t* albedo, const uint16_t* schlick, float* albedo_avg, float* schlick_avg) {
  u32 j = threadIdx.x;
  if (j >= 64) return;
  const uint16_t* table = j < 32 ? albedo : schlick;
  u32 jj = j % 32;
  float r = (float)jj / 31.0f;
  float alpha = r * r;
  double sum = 0.0;
  for (u32 m = 0; m < 4096; ++m) {
    float mu = ((float)m + 0.5f) / 4096.0f;
    sum += (double)e_lookup(table, mu, alpha) * (double)mu;
  }
  float v = (float)(2.0 * sum / 4096.0);
  (j < 32 ? albedo_avg : schlick_avg)[jj] = v;
}

// ================================================================== launch wrappers
// Most blocks of a traversal launch: its grid strides over the live queue, so 2^17 blocks (16.8M
// lanes at 128) cover any pool while an iteration with few rays left does not pay for dispatching
// a block per slot of a 2^26 pool (about 110 us per launch; large render 49.5 -> 48.0 ms per step,
// 48.2 ms at 2^15 blocks and 48.8 at 2^14).
constexpr u32 TRAVERSAL_BLOCKS = 1u << 17;

static inline u32 grid_for(u32 n, u32 block, u32 cap) {
  u32 g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return g < cap ? g : cap;
}

hipError_t launch_pool_init(const PathState& P, const Launch& L, u32 work_end) {
  hipLaunchKernelGGL(k_pool_init, dim3(grid_for(L.pool, 256, 4096)), dim3(256), 0, L.stream, P, L.pool, work_end);
  return hipGetLastError();
}
hipError_t launch_regen(const PathState& P, const Launch& L, u32 parity, bool adj) {
  if (adj)
    hipLaunchKernelGGL(k_regen<true>, dim3(grid_for(L.pool, 256, 8192)), dim3(256), 0, L.stream, L.scenes, L.segs,
                       L.n_segs, P, parity);
  else
    hipLaunchKernelGGL(k_regen<false>, dim3(grid_for(L.pool, 256, 8192)), dim3(256), 0, L.stream, L.scenes, L.segs,
                       L.n_segs, P, parity);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bookkeep, dim3(1), dim3(64), 0, L.stream, P, parity, L.tail);
  return hipGetLastError();
}
hipError_t launch_finish(const PathState& P, const Launch& L, u32 parity) {
  u32 lanes = L.pool < L.tail ? L.pool : L.tail;
  if (L.ext) {
    u32 waves = lanes ? (lanes + L.drain_lanes - 1) / L.drain_lanes : 1u;
    hipLaunchKernelGGL((k_finish<64, 32, true>), dim3(waves < (1u << 16) ? waves : (1u << 16)), dim3(64), 0, L.stream,
                       L.scenes, P, parity, L.drain_lanes, L.park_at);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || L.park_at == 0) return e;
    hipLaunchKernelGGL(k_finish_wave, dim3(L.park_at), dim3(64), 0, L.stream, L.scenes, P, L.park_at);
  } else
    hipLaunchKernelGGL((k_finish<64, 32, false>), dim3(grid_for(lanes, 64, 1u << 16)), dim3(64), 0, L.stream, L.scenes,
                       P, parity, 64u, 0u);
  return hipGetLastError();
}

#define RPX_TRAVERSAL_SHAPES(X) X(0, 128, 24, 1) X(1, 128, 32, 1) X(2, 128, 16, 7) X(3, 256, 16, 1) X(4, 128, 16, 8)
#define RPX_FLAT_SHAPES(X) X(5, 128, 24, 1) X(6, 128, 16, 1) X(7, 256, 16, 1)

// Blocks of `kernel` the GPU keeps resident at once (all CUs), for the resident-lane kernels.
template <class K>
static u32 resident_blocks(K kernel, int block) {
  int dev = 0, cus = 0, per_cu = 0;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0);
  long long n = (long long)cus * (per_cu > 0 ? per_cu : 1);
  return n > 0 ? (u32)n : 1u;
}

hipError_t launch_extend(const PathState& P, const Launch& L, u32 parity) {
  if (L.budget && P.cont) {
    switch (L.shape) {
#define X(I, B, N, W)                                                                                             \
  case I:                                                                                                        \
    hipLaunchKernelGGL((k_extend_budget<B, N, W>), dim3(grid_for(L.pool, B, TRAVERSAL_BLOCKS)), dim3(B), 0, L.stream,     \
                       L.scenes, P, parity, L.budget);                                                          \
    hipLaunchKernelGGL((k_extend_resume<B, N, W>), dim3(grid_for(P.cont_cap, B, 1u << 16)), dim3(B), 0, L.stream, \
                       L.scenes, P);                                                                             \
    return hipGetLastError();
      RPX_TRAVERSAL_SHAPES(X)
#undef X
      default:
        break;  // the resident-lane shapes run unsplit
    }
  }
  switch (L.shape) {
#define X(I, B, N, W)                                                                                            \
  case I:                                                                                                       \
    hipLaunchKernelGGL((k_extend<B, N, W>), dim3(grid_for(L.pool, B, TRAVERSAL_BLOCKS)), dim3(B), 0, L.stream, L.scenes, P, \
                       parity);                                                                                 \
    break;
    RPX_TRAVERSAL_SHAPES(X)
#undef X
#define X(I, B, N, W)                                                                                       \
  case I: {                                                                                                 \
    static const u32 resident = resident_blocks(k_extend_flat<B, N, W>, B);                                \
    u32 grid = grid_for(L.pool, B, resident);                                                               \
    hipLaunchKernelGGL((k_extend_flat<B, N, W>), dim3(grid), dim3(B), 0, L.stream, L.scenes, P, parity);    \
    break;                                                                                                  \
  }
    RPX_FLAT_SHAPES(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
hipError_t launch_shade(const PathState& P, const Launch& L, u32 parity, bool adj) {
  if (adj && L.ext) return hipErrorInvalidValue;  // the replay does not follow media or sphere lights
  if (adj)
    hipLaunchKernelGGL((k_shade<true, false>), dim3(grid_for(L.pool, 256, 8192)), dim3(256), 0, L.stream, L.scenes, P,
                       parity);
  else if (L.ext && L.cau)
    hipLaunchKernelGGL((k_shade<false, true>), dim3(grid_for(L.pool, 256, 8192)), dim3(256), 0, L.stream, L.scenes, P,
                       parity);
  else if (L.ext)
    hipLaunchKernelGGL((k_shade<false, true, 2>), dim3(grid_for(L.pool, 256, 8192)), dim3(256), 0, L.stream, L.scenes,
                       P, parity);
  else
    hipLaunchKernelGGL((k_shade<false, false>), dim3(grid_for(L.pool, 256, 8192)), dim3(256), 0, L.stream, L.scenes, P,
                       parity);
  return hipGetLastError();
}
hipError_t launch_walk(const PathState& P, const Launch& L, u32 parity) {
  hipLaunchKernelGGL(k_walk, dim3(grid_for(L.pool, 256, 8192)), dim3(256), 0, L.stream, L.scenes, P, parity);
  return hipGetLastError();
}
hipError_t launch_shadow(const PathState& P, const Launch& L, u32 parity, bool adj) {
  if (adj) {
    hipLaunchKernelGGL((k_shadow<128, 24, 1, true>), dim3(grid_for(L.pool, 128, TRAVERSAL_BLOCKS)), dim3(128), 0, L.side,
                       L.scenes, P, parity);
    return hipGetLastError();
  }
  switch (L.shadow_shape) {
#define X(I, B, N, W)                                                                                             \
  case I:                                                                                                        \
    hipLaunchKernelGGL((k_shadow<B, N, W, false>), dim3(grid_for(L.pool, B, TRAVERSAL_BLOCKS)), dim3(B), 0, L.side, L.scenes, \
                       P, parity);                                                                               \
    break;
    RPX_TRAVERSAL_SHAPES(X)
#undef X
#define X(I, B, N, W)                                                                                       \
  case I: {                                                                                                 \
    static const u32 resident = resident_blocks(k_shadow_flat<B, N, W>, B);                                \
    u32 grid = grid_for(L.pool, B, resident);                                                               \
    hipLaunchKernelGGL((k_shadow_flat<B, N, W>), dim3(grid), dim3(B), 0, L.side, L.scenes, P, parity);      \
    break;                                                                                                  \
  }
    RPX_FLAT_SHAPES(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
hipError_t launch_splat(const PathState& P, const DevSegment& s, hipStream_t st) {
  hipLaunchKernelGGL(k_splat, dim3(grid_for(s.w * s.h, 256, 8192)), dim3(256), 0, st, P, s);
  return hipGetLastError();
}
hipError_t launch_splat_segments(const PathState& P, const DevSegment* segs, u32 n_segs, hipStream_t st) {
  hipLaunchKernelGGL(k_splat_segments, dim3(std::min<u32>(n_segs, 1u << 16)), dim3(64), 0, st, P, segs, n_segs);
  return hipGetLastError();
}
hipError_t launch_splat_kept(const KeptSplat* kept, u32 n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_splat_kept, dim3(std::min<u32>(n, 1u << 16)), dim3(64), 0, st, kept, n);
  return hipGetLastError();
}
hipError_t launch_tile_out(const PathState& P, float* out, u32 w, u32 h, u32 channels, hipStream_t st) {
  hipLaunchKernelGGL(k_tile_out, dim3(grid_for(w * h, 256, 8192)), dim3(256), 0, st, P, out, w, h, channels);
  return hipGetLastError();
}
hipError_t launch_trace_rays(const DevScene* scene, const float* rays, u32 n, u32* inst, u32* tri, float* t,
                             uint8_t* any, hipStream_t st, int wave_cap) {
  if (wave_cap > 0) {
    int cap = wave_cap < WAVE_STACK ? wave_cap : WAVE_STACK;
    hipLaunchKernelGGL(k_trace_rays_wave, dim3(n < 65536u ? n : 65536u), dim3(64), 0, st, scene, rays, n, cap, inst, tri,
                       t, any);
  } else {
    hipLaunchKernelGGL(k_trace_rays, dim3(grid_for(n, 128, 8192)), dim3(128), 0, st, scene, rays, n, inst, tri, t, any);
  }
  return hipGetLastError();
}
hipError_t launch_photons(const DevScene* scene, u32 photons, u32 max_steps, const PhotonBuffers& B,
                          const PhotonPaths& Q, u32 iteration, u32* host_count, hipStream_t st) {
  if (photons == 0) return hipSuccess;
  // The stored counter and the first queue start empty (fill kernels, not memsets).
  hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(64), 0, st, B.stored, 1u);
  hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(64), 0, st, Q.counts, 2u);
  hipLaunchKernelGGL(k_photon_emit, dim3(grid_for(photons, 256, 1u << 14)), dim3(256), 0, st, scene, photons, B, Q,
                     iteration);
  hipError_t e = hipGetLastError();
  const u32 grid_t = grid_for(photons, 128, 8192), grid_s = grid_for(photons, 256, 4096);
  u32 parity = 0;
  for (u32 step = 0; step < max_steps && e == hipSuccess; ++step) {
    hipLaunchKernelGGL((k_photon_trace<128, 24>), dim3(grid_t), dim3(128), 0, st, scene, Q, parity);
    hipLaunchKernelGGL(k_photon_scatter, dim3(grid_s), dim3(256), 0, st, scene, B, Q, parity, iteration,
                       step + 1 >= max_steps ? 1u : 0u);
    e = hipGetLastError();
    parity ^= 1u;
    // Every few steps: stop once no photon is in flight.
    if (e == hipSuccess && (step % 4 == 3 || step < 2)) {
      e = hipMemcpyAsync(host_count, Q.counts + parity, 4, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e == hipSuccess && *host_count == 0) break;
    }
  }
  return e;
}
hipError_t launch_passes(const DevScene* scene, const u32* sobol, const PassPlanes& Q, u32 x0, u32 y0, u32 w, u32 h,
                         u32 ox, u32 oy, u32 ow, u32 first, u32 count, u32 what, hipStream_t st) {
  // Many samples per pixel: a wave per pixel for them (two pixels per block), the exact planes
  // by a lane per pixel; a few (the tile entry points' one): a lane per pixel for everything.
  const bool by_wave = (what & PASSES_SAMPLED) && count >= 16;
  u32 per_lane = by_wave ? (what & PASSES_EXACT) : what;
  if (per_lane) {
    hipLaunchKernelGGL((k_passes<128, 16>), dim3(grid_for(w * h, 128, 1u << 16)), dim3(128), 0, st, scene, sobol, Q,
                       x0, y0, w, h, ox, oy, ow, first, count, per_lane);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (by_wave)
    hipLaunchKernelGGL((k_passes_sampled<128, 16>), dim3(grid_for(w * h, 2, 1u << 16)), dim3(128), 0, st, scene, sobol,
                       Q, x0, y0, w, h, ox, oy, ow, first, count);
  return hipGetLastError();
}
hipError_t launch_passes_resolve(const PassPlanes& Q, u32 n, float* albedo, float* alpha, hipStream_t st) {
  hipLaunchKernelGGL(k_passes_resolve, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, Q, n, albedo, alpha);
  return hipGetLastError();
}
hipError_t launch_energy(uint16_t* albedo, uint16_t* schlick, uint16_t* clear_in, uint16_t* clear_out, float* albedo_avg,
                         float* schlick_avg, hipStream_t st) {
  u32 n = 2 * 1024 + 2 * 17408;
  hipLaunchKernelGGL(k_energy_nodes, dim3((n + 63) / 64), dim3(64), 0, st, albedo, schlick, clear_in, clear_out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_energy_avg, dim3(1), dim3(64), 0, st, albedo, schlick, albedo_avg, schlick_avg);
  return hipGetLastError();
}

}  // namespace rpx
