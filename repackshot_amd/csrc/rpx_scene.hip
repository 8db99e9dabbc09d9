// Host-side scene preparation of the GPU tracer: the RSB subset reader, per-mesh 4-wide SAH BVHs
// (built in parallel on the host CPU, then flattened for HBM), the instance BVH, the lights'
// selection tables and the per-material BSDFs.
//
// Reference (paths inside engine/crates/rt-core/src): scene/rsb.rs:1197-1610 (reader),
// geometry/mesh.rs:120-226 (normals, inside_out), geometry/transform.rs:122-200 (inverse, aabb),
// geometry/mod.rs:326-384 (instance bounds, live instances), light.rs:656-994 (Lights::new),
// environment.rs:275-296 (ramp), integrator/mod.rs:237-442 (Prepared::new: background, floor,
// media). The BVH is this library's own (binned SAH, 4-wide, 128-B nodes): the nearest hit is
// the smallest (t, instance, triangle) whatever the tree, so any conservative tree gives the
// reference's answer.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include "rpx_scene.h"

namespace rpx {

// ------------------------------------------------------------------ RSB reader
namespace {
struct Reader {
  const uint8_t* p;
  size_t len, pos = 0;
  bool ok = true;
  size_t remaining() const { return len - pos; }
  u32 u() {
    if (remaining() < 4) {
      ok = false;
      return 0;
    }
    u32 v;
    std::memcpy(&v, p + pos, 4);
    pos += 4;
    return v;
  }
  float f() { return u2f(u()); }
  u32 opt_u(u32 d) { return remaining() >= 4 ? u() : d; }
  float opt_f(float d) { return remaining() >= 4 ? f() : d; }
  V3 vec() {
    float a = f(), b = f(), c = f();
    return v3(a, b, c);
  }
  C3 rgb() {
    float a = f(), b = f(), c = f();
    return c3(a, b, c);
  }
};
}  // namespace

bool parse_rsb(const uint8_t* bytes, size_t len, HostScene& s, std::string& err, bool meshes_may_follow) {
  if (!bytes) {
    err = "null pointer";
    return false;
  }
  if (len < 4 || std::memcmp(bytes, "RSB1", 4) != 0) {
    err = "not an RSB file (bad magic)";
    return false;
  }
  if (len < 8) {
    err = "truncated header";
    return false;
  }
  u32 header_len;
  std::memcpy(&header_len, bytes + 4, 4);
  size_t pos = 8 + (size_t)header_len;
  if (pos > len) {
    err = "truncated header";
    return false;
  }
  s = HostScene();
  bool cam = false, env = false, flr = false, rnd = false;
  while (pos < len) {
    if (pos + 8 > len) {
      err = "truncated section header";
      return false;
    }
    char tag[5] = {0};
    std::memcpy(tag, bytes + pos, 4);
    u32 n;
    std::memcpy(&n, bytes + pos + 4, 4);
    if (pos + 8 + (size_t)n > len) {
      err = "truncated section payload";
      return false;
    }
    Reader r{bytes + pos + 8, n};
    pos += 8 + (size_t)n;
    std::string t(tag);
    if (t == "CAMR") {
      if (cam) { err = "duplicate CAMR section"; return false; }
      cam = true;
      HostCamera& c = s.camera;
      c.position = r.vec();
      c.target = r.vec();
      c.up = r.vec();
      c.focal_mm = r.f();
      c.sensor_mm = r.f();
      c.fstop = r.f();
      c.focus_dist = r.f();
      c.ortho = r.u() != 0 ? 1u : 0u;
      c.ortho_width = r.f();
      c.projection = r.opt_u(0);
      (void)r.opt_f(360);
      (void)r.opt_f(180);
      c.exterior = r.opt_f(1);
      if (!r.ok) { err = "CAMR section too short"; return false; }
      if (c.focal_mm <= 0.0f || c.sensor_mm <= 0.0f) { err = "focal_mm and sensor_mm must be positive"; return false; }
      if (c.projection == 1 || c.projection == 2) { err = "panoramic cameras are not supported by the GPU tracer"; return false; }
    } else if (t == "MATL") {
      HostMaterial m;
      m.kind = r.u();
      m.base_color = r.rgb();
      m.metallic = r.f();
      m.roughness = r.f();
      m.ior = r.f();
      m.transmission = r.f();
      m.coat = r.f();
      m.coat_roughness = r.f();
      m.emission = r.rgb();
      m.emission_power = r.f();
      m.flags = r.opt_u(0);
      m.opacity = r.opt_f(1);
      m.anisotropy = r.opt_f(0);
      m.absorption_mm = r.opt_f(0);
      m.bevel_mm = r.opt_f(0);
      m.scatter_mm = r.opt_f(0);
      float a = r.opt_f(1), b = r.opt_f(1), c = r.opt_f(1);
      m.scatter_color = c3(a, b, c);
      m.scatter_g = r.opt_f(0);
      m.abbe = r.opt_f(0);
      if (!r.ok) { err = "MATL section too short"; return false; }
      if (r.remaining() >= 4) { err = "material texture blocks are not supported by the GPU tracer"; return false; }
      s.materials.push_back(m);
    } else if (t == "MESH") {
      HostMesh m;
      u32 vc = r.u(), ic = r.u();
      if (!r.ok || ((u64)vc * 6 + ic + 1) * 4 > r.remaining()) { err = "MESH section too short"; return false; }
      if (ic % 3 != 0) { err = "index_count is not a multiple of 3"; return false; }
      m.positions.resize((size_t)vc * 3);
      std::vector<float> nrm((size_t)vc * 3);
      for (size_t i = 0; i < (size_t)vc * 3; ++i) m.positions[i] = r.f();
      for (size_t i = 0; i < (size_t)vc * 3; ++i) nrm[i] = r.f();
      m.indices.resize(ic);
      for (u32 i = 0; i < ic; ++i) {
        m.indices[i] = r.u();
        if (m.indices[i] >= vc) { err = "vertex index out of range"; return false; }
      }
      m.material = r.u();
      if (r.remaining() >= 4 && r.u() != 0) { err = "UV sets are not supported by the GPU tracer"; return false; }
      for (float v : m.positions)
        if (!is_finite(v)) { err = "non-finite vertex position"; return false; }
      bool any = false;
      for (float v : nrm) {
        if (!is_finite(v)) { err = "non-finite vertex normal"; return false; }
        if (v != 0.0f) any = true;
      }
      if (any) m.normals = std::move(nrm);
      s.meshes.push_back(std::move(m));
    } else if (t == "INST") {
      HostInstance in;
      in.mesh = r.u();
      for (int i = 0; i < 12; ++i) in.rows[i] = r.f();
      in.material = r.u();
      in.flags = r.u();
      if (!r.ok) { err = "INST section too short"; return false; }
      s.instances.push_back(in);
    } else if (t == "LGHT") {
      HostLight l;
      l.kind = r.u();
      l.position = r.vec();
      l.direction = r.vec();
      l.up = r.vec();
      l.size_x = r.f();
      l.size_y = r.f();
      l.color = r.rgb();
      l.power = r.f();
      l.spread_deg = r.f();
      l.angle_deg = r.f();
      l.flags = r.u();
      if (!r.ok) { err = "LGHT section too short"; return false; }
      if (l.kind > 5) { err = "unknown light kind"; return false; }
      if (l.kind == 5) { err = "IES lights are not supported by the GPU tracer"; return false; }
      s.lights.push_back(l);
    } else if (t == "ENVR") {
      if (env) { err = "duplicate ENVR section"; return false; }
      env = true;
      s.env_kind = r.u();
      s.env_floor = r.f();
      s.env_horizon = r.f();
      s.env_zenith = r.f();
      s.env_strength = r.f();
      (void)r.f();  // rotation (no effect on the ramps)
      s.background_color = r.rgb();
      s.background_mode = r.u();
      if (!r.ok) { err = "ENVR section too short"; return false; }
      u32 hdri = r.opt_u(0xffffffffu);
      if (s.env_kind > 3) { err = "unknown environment kind"; return false; }
      if (s.env_kind == 2) { err = "sky environments are not supported by the GPU tracer"; return false; }
      if (s.env_kind == 3 && hdri != 0xffffffffu) { err = "HDRI environments are not supported by the GPU tracer"; return false; }
      if (s.env_kind == 3) s.env_kind = 0;  // an HDRI without its bitmap renders the dome
    } else if (t == "FLOR") {
      if (flr) { err = "duplicate FLOR section"; return false; }
      flr = true;
      s.floor_mode = r.u();
      s.floor_z = r.f();
      s.floor_color = r.rgb();
      if (!r.ok) { err = "FLOR section too short"; return false; }
      if (s.floor_mode > 2) { err = "unknown floor mode"; return false; }
      if (s.floor_mode == 2) { err = "the shadow catcher is not supported by the GPU tracer"; return false; }
    } else if (t == "RNDR") {
      if (rnd) { err = "duplicate RNDR section"; return false; }
      rnd = true;
      s.width = r.u();
      s.height = r.u();
      s.max_samples = r.u();
      s.seed = r.u();
      (void)r.f();  // max seconds
      s.exposure_ev = r.f();
      (void)r.u();  // tonemap
      if (!r.ok) { err = "RNDR section too short"; return false; }
      s.max_depth = r.opt_u(8);
      s.clamp_indirect = r.opt_f(0);
      (void)r.opt_u(0);
      (void)r.opt_u(0);
      s.caustic_photons = r.opt_u(0);
      s.caustic_radius = r.opt_f(0);
      (void)r.opt_f(1);
      s.pixel_filter = r.opt_u(0);
      u32 edge = r.opt_u(1);
      // E10 (rsb.rs:1105-1110): the adaptive bit asks the host for rounds (rpx_render_adaptive).
      s.adaptive = r.opt_u(0);
      s.noise_threshold = r.opt_f(0);
      if (s.caustic_photons > (1u << 22)) { err = "caustic_photons must be at most 4194304"; return false; }
      if (s.caustic_radius < 0.0f) { err = "caustic_radius must not be negative"; return false; }
      if (!(s.noise_threshold >= 0.0f && s.noise_threshold <= 1.0f)) {
        err = "noise_threshold must be between 0 and 1";
        return false;
      }
      if (s.pixel_filter == 3) { err = "the Blackman-Harris pixel filter is not supported by the GPU tracer"; return false; }
      if (edge > 1) { err = "edge supersampling is not supported by the GPU tracer"; return false; }
      if (s.width < 1 || s.width > 16384 || s.height < 1 || s.height > 16384) {
        err = "width and height must be between 1 and 16384";
        return false;
      }
    } else if (t == "LENS" || t == "BMAP" || t == "PATN" || t == "IMGS" || t == "PROG" || t == "IESP" ||
               t == "FXAM" || t == "SECT" || t == "GLAS" || t == "WETP" || t == "WTP2" || t == "WETF" ||
               t == "WETG" || t == "NEST") {
      err = "section " + t + " is not supported by the GPU tracer";
      return false;
    }
  }
  for (const HostMaterial& m : s.materials) {
    if (m.anisotropy > 0.0f) { err = "anisotropic materials are not supported by the GPU tracer"; return false; }
    if (m.opacity < 1.0f) { err = "cut-outs are not supported by the GPU tracer"; return false; }
    if (m.flags & MATERIAL_FLAG_THIN) { err = "thin walls are not supported by the GPU tracer"; return false; }
    if (m.abbe > 0.0f) { err = "dispersion is not supported by the GPU tracer"; return false; }
    if (m.bevel_mm > 0.0f) { err = "edge bevels are not supported by the GPU tracer"; return false; }
  }
  // A smooth, wholly clear solid without a medium of its own may be a thin wall, which the
  // reference probes for after a scattering medium's exit (integrator/mod.rs:289-299, 1822-1857);
  // scenes with both are outside what the GPU tracer restates.
  bool any_scatter = false, any_wall = false;
  for (const HostMaterial& m : s.materials) {
    DevMedium unused;
    bool scattering = scatter_medium_of(m, unused);
    any_scatter = any_scatter || scattering;
    bool metal = m.kind != MATERIAL_GLASS && m.metallic >= 1.0f;
    bool clear_solid = m.transmission > 0.0f && !metal && (m.flags & MATERIAL_FLAG_THIN) == 0 && m.kind != 2 &&
                       m.ior >= 1.0001f;
    any_wall = any_wall || (clear_solid && !scattering && (m.kind == MATERIAL_GLASS || m.metallic <= 0.0f) &&
                            m.transmission >= 1.0f && m.roughness <= 0.05f && m.coat <= 0.0f && m.opacity >= 1.0f);
  }
  if (any_scatter && any_wall) {
    err = "thin walls beside a scattering medium are not supported by the GPU tracer";
    return false;
  }
  for (const HostMesh& m : s.meshes)
    if (m.material >= s.materials.size()) { err = "material_id does not refer to a MATL section"; return false; }
  if (meshes_may_follow && s.meshes.empty()) return true;
  if (s.instances.empty()) {
    for (size_t i = 0; i < s.meshes.size(); ++i) {
      HostInstance in;
      in.mesh = (u32)i;
      in.material = s.meshes[i].material;
      s.instances.push_back(in);
    }
  }
  for (const HostInstance& in : s.instances) {
    if (in.mesh >= s.meshes.size()) { err = "mesh_id does not refer to a MESH section"; return false; }
    if (in.material >= s.materials.size()) { err = "material_id does not refer to a MATL section"; return false; }
  }
  return true;
}

// ------------------------------------------------------------------ transforms (transform.rs)
namespace {
struct Xf {
  float r[3][4];
};
Xf xf_from(const float* rows) {
  Xf x;
  std::memcpy(x.r, rows, 48);
  return x;
}
bool xf_finite(const Xf& x) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j)
      if (!is_finite(x.r[i][j])) return false;
  return true;
}
bool xf_identity(const Xf& x) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j)
      if (!(x.r[i][j] == (i == j ? 1.0f : 0.0f))) return false;
  return true;
}
V3 xf_point(const Xf& x, V3 p) {
  const auto& r = x.r;
  return v3(r[0][0] * p.x + r[0][1] * p.y + r[0][2] * p.z + r[0][3],
            r[1][0] * p.x + r[1][1] * p.y + r[1][2] * p.z + r[1][3],
            r[2][0] * p.x + r[2][1] * p.y + r[2][2] * p.z + r[2][3]);
}
bool xf_inverse(const Xf& x, Xf& out) {
  if (!xf_finite(x)) return false;
  const auto& r = x.r;
  float det = r[0][0] * (r[1][1] * r[2][2] - r[1][2] * r[2][1]) - r[0][1] * (r[1][0] * r[2][2] - r[1][2] * r[2][0]) +
              r[0][2] * (r[1][0] * r[2][1] - r[1][1] * r[2][0]);
  if (det == 0.0f || !is_finite(det)) return false;
  float inv_det = 1.0f / det;
  float c[3][3] = {
      {r[1][1] * r[2][2] - r[1][2] * r[2][1], r[0][2] * r[2][1] - r[0][1] * r[2][2], r[0][1] * r[1][2] - r[0][2] * r[1][1]},
      {r[1][2] * r[2][0] - r[1][0] * r[2][2], r[0][0] * r[2][2] - r[0][2] * r[2][0], r[0][2] * r[1][0] - r[0][0] * r[1][2]},
      {r[1][0] * r[2][1] - r[1][1] * r[2][0], r[0][1] * r[2][0] - r[0][0] * r[2][1], r[0][0] * r[1][1] - r[0][1] * r[1][0]},
  };
  Xf inv;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) inv.r[i][j] = c[i][j] * inv_det;
  float t0 = r[0][3], t1 = r[1][3], t2 = r[2][3];
  for (int i = 0; i < 3; ++i) inv.r[i][3] = -(inv.r[i][0] * t0 + inv.r[i][1] * t1 + inv.r[i][2] * t2);
  if (!xf_finite(inv)) return false;
  out = inv;
  return true;
}

struct Box {
  V3 lo{RPX_INF, RPX_INF, RPX_INF}, hi{-RPX_INF, -RPX_INF, -RPX_INF};
  void grow(V3 p) {
    lo = vmin(lo, p);
    hi = vmax(hi, p);
  }
  void grow(const Box& b) {
    lo = vmin(lo, b.lo);
    hi = vmax(hi, b.hi);
  }
  bool empty() const { return lo.x > hi.x || lo.y > hi.y || lo.z > hi.z; }
  V3 centroid() const { return (lo + hi) * 0.5f; }
  float area() const {
    if (empty()) return 0.0f;
    V3 d = hi - lo;
    return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x);
  }
};

// geometry/mod.rs:309-318
Box instance_bounds(const HostMesh& m, const Xf& x) {
  Box mb;
  mb.lo = v3(m.bmin[0], m.bmin[1], m.bmin[2]);
  mb.hi = v3(m.bmax[0], m.bmax[1], m.bmax[2]);
  if (xf_identity(x)) return mb;
  if (mb.empty()) return Box();
  Box out;
  for (int corner = 0; corner < 8; ++corner) {
    V3 p = v3((corner & 1) == 0 ? mb.lo.x : mb.hi.x, (corner & 2) == 0 ? mb.lo.y : mb.hi.y,
              (corner & 4) == 0 ? mb.lo.z : mb.hi.z);
    out.grow(xf_point(x, p));
  }
  V3 a = vmax(v3(fabs_(out.lo.x), fabs_(out.lo.y), fabs_(out.lo.z)), v3(fabs_(out.hi.x), fabs_(out.hi.y), fabs_(out.hi.z)));
  float scale = fmax_r(fmax_r(a.x, a.y), a.z);
  float pad = scale * 1.0e-5f;
  out.lo = out.lo - v3(pad, pad, pad);
  out.hi = out.hi + v3(pad, pad, pad);
  return out;
}

// ------------------------------------------------------------------ 4-wide SAH BVH builder
constexpr int BINS = 32;
constexpr int MAX_LEAF = 4;        // triangles (or instances) per leaf
constexpr int MEDIAN_DEPTH = 40;   // binary depth beyond which splits are at the median

struct BNode {
  Box box;
  u32 left = 0, right = 0, first = 0, count = 0;  // count > 0: leaf over order[first..first+count)
};

struct Bvh2 {
  std::vector<BNode> nodes;
  std::vector<u32> order;
};

void build_binary(const std::vector<Box>& boxes, int leaf_max, Bvh2& out) {
  size_t n = boxes.size();
  out.nodes.clear();
  out.order.resize(n);
  for (size_t i = 0; i < n; ++i) out.order[i] = (u32)i;
  if (n == 0) return;
  std::vector<V3> cent(n);
  for (size_t i = 0; i < n; ++i) cent[i] = boxes[i].centroid();
  out.nodes.reserve(2 * n / std::max(1, leaf_max) + 2);
  struct Task {
    u32 node, begin, end, depth;
  };
  std::vector<Task> stack;
  out.nodes.push_back(BNode());
  stack.push_back(Task{0, 0, (u32)n, 0});
  while (!stack.empty()) {
    Task t = stack.back();
    stack.pop_back();
    Box bb, cb;
    for (u32 i = t.begin; i < t.end; ++i) {
      bb.grow(boxes[out.order[i]]);
      cb.grow(cent[out.order[i]]);
    }
    BNode& node = out.nodes[t.node];
    node.box = bb;
    u32 count = t.end - t.begin;
    node.first = t.begin;
    node.count = count;
    if ((int)count <= 1) continue;
    u32 mid = t.begin + count / 2;
    bool split = true;
    if ((int)t.depth < MEDIAN_DEPTH) {
      // binned SAH over the three axes
      float best_cost = RPX_INF;
      int best_axis = -1, best_plane = -1;
      for (int axis = 0; axis < 3; ++axis) {
        float lo = comp(cb.lo, axis), hi = comp(cb.hi, axis);
        float extent = hi - lo;
        if (!(extent > 0.0f) || !is_finite(extent)) continue;
        float scale = (float)BINS / extent;
        Box bins[BINS];
        u32 cnt[BINS] = {};
        for (u32 i = t.begin; i < t.end; ++i) {
          u32 p = out.order[i];
          int b = (int)std::min<u32>(f_to_u32((comp(cent[p], axis) - lo) * scale), BINS - 1);
          cnt[b]++;
          bins[b].grow(boxes[p]);
        }
        float right_area[BINS];
        u32 right_cnt[BINS];
        Box acc;
        u32 c = 0;
        for (int b = BINS - 1; b > 0; --b) {
          acc.grow(bins[b]);
          c += cnt[b];
          right_area[b - 1] = acc.area();
          right_cnt[b - 1] = c;
        }
        acc = Box();
        c = 0;
        for (int b = 0; b < BINS - 1; ++b) {
          acc.grow(bins[b]);
          c += cnt[b];
          if (c == 0 || right_cnt[b] == 0) continue;
          float cost = acc.area() * (float)c + right_area[b] * (float)right_cnt[b];
          if (cost < best_cost) {
            best_cost = cost;
            best_axis = axis;
            best_plane = b;
          }
        }
      }
      if (best_axis >= 0) {
        float area = bb.area();
        float split_cost = (area > 0.0f && is_finite(area)) ? 1.0f + best_cost / area : 1.0f;
        if ((int)count <= leaf_max && (float)count <= split_cost) {
          split = false;
        } else {
          float lo = comp(cb.lo, best_axis);
          float scale = (float)BINS / (comp(cb.hi, best_axis) - lo);
          u32 i = t.begin, j = t.end;
          while (i < j) {
            u32 p = out.order[i];
            int b = (int)std::min<u32>(f_to_u32((comp(cent[p], best_axis) - lo) * scale), BINS - 1);
            if (b <= best_plane) {
              ++i;
            } else {
              --j;
              std::swap(out.order[i], out.order[j]);
            }
          }
          mid = (i == t.begin || i == t.end) ? t.begin + count / 2 : i;
        }
      } else if ((int)count <= leaf_max) {
        split = false;  // identical centroids: a leaf when small enough, else the median
      }
    }
    if (!split) continue;
    u32 l = (u32)out.nodes.size();
    out.nodes.push_back(BNode());
    out.nodes.push_back(BNode());
    out.nodes[t.node].left = l;
    out.nodes[t.node].right = l + 1;
    out.nodes[t.node].count = 0;
    stack.push_back(Task{l + 1, mid, t.end, t.depth + 1});
    stack.push_back(Task{l, t.begin, mid, t.depth + 1});
  }
}

// Collapses the binary tree into Node4s appended to `nodes`; leaves call `emit(first, count)`
// (indices into bvh.order) and get their reference back. Returns the root reference and the
// depth of the 4-wide tree.
template <class Emit>
u32 collapse4(const Bvh2& b, std::vector<Node4>& nodes, Emit emit, u32& depth_out) {
  depth_out = 0;
  if (b.nodes.empty()) return REF_EMPTY;
  struct Item {
    u32 bnode;
    u32 slot_node;  // Node4 index to patch (or ~0 for the root)
    int slot;
    u32 depth;
  };
  u32 root_ref = REF_EMPTY;
  std::vector<Item> work;
  work.push_back(Item{0, 0xffffffffu, 0, 0});
  while (!work.empty()) {
    Item it = work.back();
    work.pop_back();
    const BNode& bn = b.nodes[it.bnode];
    u32 ref;
    if (bn.count > 0) {
      ref = emit(bn.first, bn.count);
    } else {
      u32 kids[4] = {bn.left, bn.right, 0, 0};
      int len = 2;
      while (len < 4) {
        int widest = -1;
        float warea = -1.0f;
        for (int k = 0; k < len; ++k) {
          const BNode& kn = b.nodes[kids[k]];
          if (kn.count == 0 && kn.box.area() > warea) {
            warea = kn.box.area();
            widest = k;
          }
        }
        if (widest < 0) break;
        const BNode& op = b.nodes[kids[widest]];
        for (int k = len; k > widest + 1; --k) kids[k] = kids[k - 1];
        kids[widest] = op.left;
        kids[widest + 1] = op.right;
        ++len;
      }
      u32 idx = (u32)nodes.size();
      Node4 n4;
      for (int k = 0; k < 4; ++k) {
        n4.lo_x[k] = n4.lo_y[k] = n4.lo_z[k] = RPX_INF;
        n4.hi_x[k] = n4.hi_y[k] = n4.hi_z[k] = -RPX_INF;
        n4.child[k] = REF_EMPTY;
        n4.pad[k] = 0;
      }
      for (int k = 0; k < len; ++k) {
        const Box& kb = b.nodes[kids[k]].box;
        n4.lo_x[k] = kb.lo.x;
        n4.hi_x[k] = kb.hi.x;
        n4.lo_y[k] = kb.lo.y;
        n4.hi_y[k] = kb.hi.y;
        n4.lo_z[k] = kb.lo.z;
        n4.hi_z[k] = kb.hi.z;
      }
      nodes.push_back(n4);
      depth_out = std::max(depth_out, it.depth + 1);
      for (int k = len - 1; k >= 0; --k) work.push_back(Item{kids[k], idx, k, it.depth + 1});
      ref = idx;
    }
    if (it.slot_node == 0xffffffffu) root_ref = ref;
    else nodes[it.slot_node].child[it.slot] = ref;
  }
  return root_ref;
}
}  // namespace

// ------------------------------------------------------------------ lights (light.rs)
namespace {
constexpr float MM2_TO_M2 = 1.0e-6f;
constexpr float MM_TO_M = 1.0e-3f;

float spread_exponent(float spread_deg) {
  if (spread_deg >= 180.0f) return 0.0f;
  float half = 0.5f * fmax_r(spread_deg, 1.0f) * deg_to_rad();
  return -1.0f / flog2(fcos(half));
}
Frame light_frame(V3 normal, V3 up) {
  V3 z = normalize(normal);
  V3 y = normalize(up - z * dot(z, up));
  if (length_sq(y) < 0.5f) return frame_from_normal(z);
  Frame f;
  f.x = cross(y, z);
  f.y = y;
  f.z = z;
  return f;
}
}  // namespace

// integrator/medium.rs:99-141: a clear solid (transmissive, not a metal, not a thin wall, not an
// emitter, of index at least 1.0001) with a positive, finite scatter_mm; the absorption of its
// base colour over absorption_mm joins the walk's.
bool scatter_medium_of(const HostMaterial& m, DevMedium& out) {
  bool metal = m.kind != MATERIAL_GLASS && m.metallic >= 1.0f;
  bool clear_solid = m.transmission > 0.0f && !metal && (m.flags & MATERIAL_FLAG_THIN) == 0 && m.kind != 2 &&
                     m.ior >= 1.0001f;
  if (!(m.scatter_mm > 0.0f && is_finite(m.scatter_mm) && clear_solid)) return false;
  float sigma = 1.0f / m.scatter_mm;
  C3 albedo = c3(clampf(m.scatter_color.r, 0.0f, 1.0f), clampf(m.scatter_color.g, 0.0f, 1.0f),
                 clampf(m.scatter_color.b, 0.0f, 1.0f));
  C3 sigma_a = c3((1.0f - albedo.r) * sigma, (1.0f - albedo.g) * sigma, (1.0f - albedo.b) * sigma);
  if (m.absorption_mm > 0.0f) {
    C3 c = m.base_color;
    sigma_a = sigma_a + c3(-fln(clampf(c.r, 1.0e-4f, 1.0f)) / m.absorption_mm,
                           -fln(clampf(c.g, 1.0e-4f, 1.0f)) / m.absorption_mm,
                           -fln(clampf(c.b, 1.0e-4f, 1.0f)) / m.absorption_mm);
  }
  out.sigma_s = albedo * sigma;
  out.sigma_a = sigma_a;
  out.sigma_t = out.sigma_s + sigma_a;
  out.g = is_finite(m.scatter_g) ? clampf(m.scatter_g, -0.9f, 0.9f) : 0.0f;
  out.ior = m.ior;
  return true;
}

DevMaterial dev_material(const HostMaterial& m) {
  DevMaterial dm;
  dm.emissive = m.kind == 2 ? 1u : 0u;
  C3 e = m.emission * m.emission_power;
  dm.radiance = dm.emissive ? c3(fmax_r(e.r, 0.0f), fmax_r(e.g, 0.0f), fmax_r(e.b, 0.0f)) : splat(0.0f);
  // integrator/mod.rs:300-311 (medium_of); a scattering medium folds the absorption into its walk
  // (:419).
  bool clear = m.transmission > 0.0f && m.metallic < 1.0f && (m.flags & MATERIAL_FLAG_THIN) == 0;
  dm.has_medium = 0;
  dm.sigma = splat(0.0f);
  DevMedium scatter;
  if (scatter_medium_of(m, scatter)) {
    dm.has_medium = MEDIUM_SCATTERING;
  } else if (clear && m.absorption_mm > 0.0f && m.kind != 2) {
    C3 c = m.base_color;
    C3 sg = c3(-fln(clampf(c.r, 1.0e-4f, 1.0f)) / m.absorption_mm, -fln(clampf(c.g, 1.0e-4f, 1.0f)) / m.absorption_mm,
               -fln(clampf(c.b, 1.0e-4f, 1.0f)) / m.absorption_mm);
    if (sg.r > 0.0f || sg.g > 0.0f || sg.b > 0.0f) {
      dm.has_medium = MEDIUM_ABSORBING;
      dm.sigma = sg;
    }
  }
  dm.kind = m.kind;
  dm.flags = m.flags;
  dm.base = m.base_color;
  dm.metallic = m.metallic;
  dm.roughness = m.roughness;
  dm.ior = m.ior;
  dm.transmission = m.transmission;
  dm.coat = m.coat;
  dm.coat_roughness = m.coat_roughness;
  dm.absorption_mm = m.absorption_mm;
  return dm;
}

void material_sides(const EnergyTables& tables, const HostMaterial& m, Bsdf out[2]) {
  for (int side = 0; side < 2; ++side)
    out[side] = bsdf_surface(tables, m.kind, m.base_color, m.metallic, m.roughness, m.ior, m.transmission, m.coat,
                             m.coat_roughness, m.flags, m.absorption_mm, side == 0);
}

namespace {
// caustics.rs:104-119: glass or metal of roughness at most 0.25, not an emitter, not a scattering
// medium (anisotropic materials are refused at load).
bool is_caster(const HostMaterial& m) {
  float r = m.roughness;
  if (m.kind == 2 || r != r || r > 0.25f) return false;
  DevMedium unused;
  if (scatter_medium_of(m, unused)) return false;
  bool glass = m.kind == MATERIAL_GLASS || (m.transmission >= 0.5f && m.metallic < 0.5f);
  bool metal = m.kind != MATERIAL_GLASS && m.metallic >= 0.5f;
  return glass || metal;
}

// The world box of an instance: its mesh box's eight corners transformed (caustics.rs:122-143).
Box caustic_instance_box(const HostMesh& m, const float* rows) {
  Box out;
  if (m.bmin[0] > m.bmax[0] || m.bmin[1] > m.bmax[1] || m.bmin[2] > m.bmax[2]) return out;
  Xf x = xf_from(rows);
  for (u32 k = 0; k < 8; ++k) {
    V3 c = v3((k & 1) == 0 ? m.bmin[0] : m.bmax[0], (k & 2) == 0 ? m.bmin[1] : m.bmax[1],
              (k & 4) == 0 ? m.bmin[2] : m.bmax[2]);
    out.grow(xf_point(x, c));
  }
  return out;
}

void sphere_of_box(const Box& b, V3& c, float& r) {
  c = b.centroid();
  r = fmax_r(0.5f * length(b.hi - b.lo), 1.0e-3f);
}

// caustics.rs:171-178
float cone_one_minus_cos_host(float r, float d) {
  float sin2 = (r * r) / (d * d);
  float cos = sqrt_(fmax_r(1.0f - sin2, 0.0f));
  return sin2 / (1.0f + cos);
}

std::vector<float> cdf_of_weights(const std::vector<float>& weights) {
  std::vector<float> cdf;
  float total = -0.0f;
  for (float w : weights) total = total + w;
  if (!(total > 0.0f && is_finite(total))) return cdf;
  float acc = 0.0f;
  for (float w : weights) {
    acc += w;
    cdf.push_back(acc / total);
  }
  cdf.back() = 1.0f;
  return cdf;
}

void caustic_setup(const HostScene& s, const std::vector<HostMesh>& meshes, const Box& scene_bounds,
                   FlatScene& out) {
  DevScene& k = out.consts;
  k.caustics_on = 0;
  k.pmaps = nullptr;
  if (s.caustic_photons == 0) return;
  DevCaustics& c = k.caustics;
  std::vector<size_t> parts;
  std::vector<Box> boxes(s.instances.size());
  for (size_t i = 0; i < s.instances.size(); ++i) {
    const HostInstance& in = s.instances[i];
    bool caster = in.material < s.materials.size() && is_caster(s.materials[in.material]);
    if (in.mesh < meshes.size()) boxes[i] = caustic_instance_box(meshes[in.mesh], in.rows);
    if (caster) out.instances[i].flags |= INST_CASTER;
    if (caster && !boxes[i].empty()) parts.push_back(i);
  }
  if (parts.empty()) return;
  c.n_spheres = 0;
  if (parts.size() > CAUSTIC_MAX_SPHERES) {
    Box all;
    for (size_t i : parts) all.grow(boxes[i]);
    sphere_of_box(all, c.sphere_c[0], c.sphere_r[0]);
    c.n_spheres = 1;
  } else {
    for (size_t i : parts) {
      // caustics.rs:145-169: the box's centre and a hair beyond the farthest vertex.
      const HostInstance& in = s.instances[i];
      V3 centre = boxes[i].centroid();
      const HostMesh& m = meshes[in.mesh];
      Xf x = xf_from(in.rows);
      float r2 = 0.0f;
      for (size_t v = 0; v + 2 < m.positions.size(); v += 3)
        r2 = fmax_r(r2, length_sq(xf_point(x, v3(m.positions[v], m.positions[v + 1], m.positions[v + 2])) - centre));
      float r = sqrt_(r2) * 1.0001f + 1.0e-3f;
      V3 cc = centre;
      if (!(r > 0.0f && is_finite(r))) sphere_of_box(boxes[i], cc, r);
      c.sphere_c[c.n_spheres] = cc;
      c.sphere_r[c.n_spheres] = r;
      ++c.n_spheres;
    }
  }
  auto towards = [&](V3 x, const V3* normal) {
    float sum = -0.0f;
    for (u32 i = 0; i < c.n_spheres; ++i) {
      V3 d = c.sphere_c[i] - x;
      float dist = length(d);
      float facing = normal ? 0.05f + fmax_r(dot(*normal, d / fmax_r(dist, 1.0e-6f)), 0.0f) : 1.0f;
      float omega = dist <= c.sphere_r[i] ? PI : TAU * cone_one_minus_cos_host(c.sphere_r[i], dist);
      sum = sum + omega * facing;
    }
    return sum;
  };
  std::vector<float> weights;
  out.caustic_sources.clear();
  for (size_t i = 0; i < out.areas.size(); ++i) {
    const DevArea& a = out.areas[i];
    if (!a.glossy) continue;
    float w = luminance(a.radiance) * a.area * 1.0e-6f * towards(a.centre, &a.frame.z);
    if (w > 0.0f && is_finite(w)) {
      out.caustic_sources.push_back((0u << 24) | (u32)i);
      weights.push_back(w);
    }
  }
  for (size_t i = 0; i < out.suns.size(); ++i) {
    const DevSun& sn = out.suns[i];
    if (!sn.glossy) continue;
    C3 irradiance = sn.delta ? sn.radiance : sn.radiance * (sn.solid_angle * (1.0f - 0.5f * sn.one_minus_cos_max));
    float area = -0.0f;
    for (u32 j = 0; j < c.n_spheres; ++j) area = area + PI * c.sphere_r[j] * c.sphere_r[j] * 1.0e-6f;
    float w = luminance(irradiance) * area;
    if (w > 0.0f && is_finite(w)) {
      out.caustic_sources.push_back((1u << 24) | (u32)i);
      weights.push_back(w);
    }
  }
  for (size_t i = 0; i < out.emitters.size(); ++i) {
    const DevEmitter& m = out.emitters[i];
    V3 acc = v3(0.0f, 0.0f, 0.0f);
    for (u32 t = 0; t < m.count; ++t) {
      const float4* tp = &out.emit_tris[4 * (size_t)(m.first + t)];
      acc = acc + ((v3(tp[0].x, tp[0].y, tp[0].z) + v3(tp[1].x, tp[1].y, tp[1].z)) + v3(tp[2].x, tp[2].y, tp[2].z)) *
                      (1.0f / 3.0f);
    }
    V3 centroid = acc * (1.0f / (float)std::max<u32>(m.count, 1));
    float w = luminance(m.radiance) * m.area * 1.0e-6f * 0.5f * towards(centroid, nullptr);
    if (w > 0.0f && is_finite(w)) {
      out.caustic_sources.push_back((2u << 24) | (u32)i);
      weights.push_back(w);
    }
  }
  out.caustic_cdf = cdf_of_weights(weights);
  if (out.caustic_cdf.empty()) {
    out.caustic_sources.clear();
    return;
  }
  c.n_sources = (u32)out.caustic_sources.size();
  if (scene_bounds.empty()) {
    c.scene_centre = v3(0.0f, 0.0f, 0.0f);
    c.scene_radius = 1.0f;
  } else {
    sphere_of_box(scene_bounds, c.scene_centre, c.scene_radius);
  }
  // caustics.rs:180-197: the pixel footprint at the camera's target distance.
  const HostCamera& cam = s.camera;
  u32 w = std::max<u32>(s.width, 1), h = std::max<u32>(s.height, 1);
  float footprint;
  if (cam.ortho) footprint = fabs_(cam.ortho_width) / (float)w;
  else footprint = length(cam.target - cam.position) * cam.sensor_mm / (fmax_r(cam.focal_mm, 1.0e-3f) * (float)std::max(w, h));
  float radius_px = s.caustic_radius > 0.0f ? s.caustic_radius : 3.0f;
  c.radius = fmax_r(radius_px * footprint, 1.0e-4f);
  c.photons = s.caustic_photons;
  c.seed = s.seed;
  u32 depth = s.max_depth == 0 ? 8u : std::min<u32>(s.max_depth, 1024u);
  c.max_bounces = depth > 0 ? depth - 1 : 0;
  k.caustics_on = 1;
  k.ext = 1;
}
}  // namespace

bool flatten_scene(const HostScene& s, const EnergyTables& tables, FlatScene& out, std::string& err, int threads) {
  out = FlatScene();
  const size_t nm = s.meshes.size();
  // Mesh bounds and inside_out (mesh.rs:160-197).
  std::vector<HostMesh> meshes = s.meshes;
  for (HostMesh& m : meshes) {
    Box b;
    for (size_t i = 0; i + 2 < m.positions.size(); i += 3) b.grow(v3(m.positions[i], m.positions[i + 1], m.positions[i + 2]));
    m.bmin[0] = b.lo.x; m.bmin[1] = b.lo.y; m.bmin[2] = b.lo.z;
    m.bmax[0] = b.hi.x; m.bmax[1] = b.hi.y; m.bmax[2] = b.hi.z;
    m.inside_out = false;
    if (!m.normals.empty()) {
      float agreement = 0.0f;
      size_t nt = m.indices.size() / 3;
      for (size_t t = 0; t < nt; ++t) {
        u32 ia = m.indices[3 * t], ib = m.indices[3 * t + 1], ic = m.indices[3 * t + 2];
        V3 a = v3(m.positions[3 * ia], m.positions[3 * ia + 1], m.positions[3 * ia + 2]);
        V3 bb = v3(m.positions[3 * ib], m.positions[3 * ib + 1], m.positions[3 * ib + 2]);
        V3 c = v3(m.positions[3 * ic], m.positions[3 * ic + 1], m.positions[3 * ic + 2]);
        V3 winding = cross(bb - a, c - a);
        V3 n = v3(0, 0, 0);
        for (u32 id : {ia, ib, ic}) n = n + v3(m.normals[3 * id], m.normals[3 * id + 1], m.normals[3 * id + 2]);
        agreement += dot(normalize(winding), n);
      }
      m.inside_out = agreement < 0.0f;
    }
  }
  // Per-mesh BLAS, built in parallel.
  struct Blas {
    std::vector<Node4> nodes;
    std::vector<u32> tri_order;  // leaf order of the mesh's triangles
    std::vector<u32> leaf_first; // per emitted leaf (local order offset)
    u32 root = REF_EMPTY, depth = 0;
  };
  std::vector<Blas> blas(nm);
  {
    std::atomic<size_t> next{0};
    auto worker = [&]() {
      for (;;) {
        size_t mi = next.fetch_add(1);
        if (mi >= nm) return;
        const HostMesh& m = meshes[mi];
        size_t nt = m.indices.size() / 3;
        std::vector<Box> boxes(nt);
        for (size_t t = 0; t < nt; ++t) {
          Box b;
          for (int k = 0; k < 3; ++k) {
            u32 id = m.indices[3 * t + k];
            b.grow(v3(m.positions[3 * id], m.positions[3 * id + 1], m.positions[3 * id + 2]));
          }
          boxes[t] = b;
        }
        Bvh2 b2;
        build_binary(boxes, MAX_LEAF, b2);
        Blas& bl = blas[mi];
        // Leaves are emitted as contiguous ranges of the (local) leaf order; the references are
        // local and are rebased when the meshes are concatenated.
        auto emit = [&](u32 first, u32 count) -> u32 {
          u32 local_first = (u32)bl.tri_order.size();
          for (u32 k = 0; k < count; ++k) bl.tri_order.push_back(b2.order[first + k]);
          return make_leaf(local_first, count);
        };
        bl.root = collapse4(b2, bl.nodes, emit, bl.depth);
      }
    };
    int nth = std::max(1, std::min<int>(threads, (int)nm));
    std::vector<std::thread> pool;
    for (int t = 0; t < nth; ++t) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
  }
  // Concatenate: nodes, triangles, vertex data.
  std::vector<u32> node_base(nm), tri_base(nm);
  u64 total_tris = 0;
  for (size_t mi = 0; mi < nm; ++mi) {
    node_base[mi] = (u32)out.nodes.size();
    tri_base[mi] = (u32)(out.tris.size() / 3);
    const HostMesh& m = meshes[mi];
    const Blas& bl = blas[mi];
    if ((u64)tri_base[mi] + bl.tri_order.size() >= (1u << 28)) {
      err = "too many triangles for the leaf encoding";
      return false;
    }
    for (Node4 n : bl.nodes) {
      for (int k = 0; k < 4; ++k) {
        u32 c = n.child[k];
        if (c == REF_EMPTY) continue;
        if (c & REF_LEAF) {
          u32 first = (c & ~REF_LEAF) >> LEAF_COUNT_BITS, cnt = (c & 7u) + 1;
          n.child[k] = make_leaf(first + tri_base[mi], cnt);
        } else {
          n.child[k] = c + node_base[mi];
        }
      }
      out.nodes.push_back(n);
    }
    for (u32 t : bl.tri_order) {
      u32 ia = m.indices[3 * t], ib = m.indices[3 * t + 1], ic = m.indices[3 * t + 2];
      out.tris.push_back(make_float4(m.positions[3 * ia], m.positions[3 * ia + 1], m.positions[3 * ia + 2], u2f(t)));
      out.tris.push_back(make_float4(m.positions[3 * ib], m.positions[3 * ib + 1], m.positions[3 * ib + 2], 0.0f));
      out.tris.push_back(make_float4(m.positions[3 * ic], m.positions[3 * ic + 1], m.positions[3 * ic + 2], 0.0f));
    }
    total_tris += m.indices.size() / 3;
    DevMesh dm;
    dm.vertex_offset = (u32)(out.positions.size() / 3);
    dm.index_offset = (u32)out.indices.size();
    dm.has_normals = m.normals.empty() ? 0u : 1u;
    dm.inside_out = m.inside_out ? 1u : 0u;
    out.meshes.push_back(dm);
    out.positions.insert(out.positions.end(), m.positions.begin(), m.positions.end());
    if (m.normals.empty()) out.normals.insert(out.normals.end(), m.positions.size(), 0.0f);
    else out.normals.insert(out.normals.end(), m.normals.begin(), m.normals.end());
    out.indices.insert(out.indices.end(), m.indices.begin(), m.indices.end());
    for (size_t t = 0; t + 2 < m.indices.size(); t += 3) {
      float c[20] = {};
      for (int k = 0; k < 3; ++k) {
        size_t v = 3 * (size_t)m.indices[t + k];
        for (int a = 0; a < 3; ++a) {
          c[3 * k + a] = m.positions[v + a];
          c[9 + 3 * k + a] = m.normals.empty() ? 0.0f : m.normals[v + a];
        }
      }
      for (u32 q = 0; q < CORNER_FLOAT4S; ++q) out.corners.push_back(make_float4(c[4 * q], c[4 * q + 1], c[4 * q + 2], c[4 * q + 3]));
    }
    out.max_blas_depth = std::max(out.max_blas_depth, bl.depth);
  }
  out.triangles = total_tris;
  // Instances and the TLAS (geometry/mod.rs:326-384).
  std::vector<Box> inst_boxes;
  std::vector<u32> active;
  Box scene_bounds;
  for (size_t i = 0; i < s.instances.size(); ++i) {
    const HostInstance& in = s.instances[i];
    const HostMesh& m = meshes[in.mesh];
    Xf x = xf_from(in.rows), inv;
    bool has_inv = xf_inverse(x, inv);
    bool live = m.indices.size() >= 3 && has_inv;
    DevInstance d;
    std::memcpy(d.to_world, in.rows, 48);
    if (has_inv) std::memcpy(d.to_local, inv.r, 48);
    else {
      Xf id = xf_from(HostInstance().rows);
      std::memcpy(d.to_local, id.r, 48);
    }
    d.mesh = in.mesh;
    u32 mroot = blas[in.mesh].root;
    if (mroot != REF_EMPTY) {
      if (mroot & REF_LEAF) {
        u32 first = (mroot & ~REF_LEAF) >> LEAF_COUNT_BITS, cnt = (mroot & 7u) + 1;
        mroot = make_leaf(first + tri_base[in.mesh], cnt);
      } else {
        mroot += node_base[in.mesh];
      }
    }
    d.blas_root = mroot;
    d.material = in.material;
    d.flags = (xf_identity(x) ? INST_IDENTITY : 0u) | ((in.flags & 1u) ? INST_ONE_SIDED : 0u) |
              ((in.flags & 4u) ? INST_CAMERA_HIDDEN : 0u) | (live ? INST_ACTIVE : 0u);
    out.instances.push_back(d);
    if (live) {
      Box b = instance_bounds(m, x);
      scene_bounds.grow(b);
      inst_boxes.push_back(b);
      active.push_back((u32)i);
    }
  }
  {
    Bvh2 t2;
    build_binary(inst_boxes, 1, t2);
    auto emit = [&](u32 first, u32 count) -> u32 {
      (void)count;
      return make_leaf(active[t2.order[first]], 1);
    };
    u32 tlas_node_base = (u32)out.nodes.size();
    std::vector<Node4> tnodes;
    u32 root = collapse4(t2, tnodes, emit, out.max_tlas_depth);
    for (Node4 n : tnodes) {
      for (int k = 0; k < 4; ++k)
        if (n.child[k] != REF_EMPTY && !(n.child[k] & REF_LEAF)) n.child[k] += tlas_node_base;
      out.nodes.push_back(n);
    }
    if (root != REF_EMPTY && !(root & REF_LEAF)) root += tlas_node_base;
    out.consts.tlas_root = root;
  }
  if (out.nodes.empty()) out.nodes.push_back(Node4{});  // keep the array non-empty
  if (out.tris.empty()) out.tris.push_back(make_float4(0, 0, 0, 0));
  // Materials: the BSDF of each side and what the integrator reads of them.
  for (const HostMaterial& m : s.materials) {
    Bsdf sides[2];
    material_sides(tables, m, sides);
    out.bsdfs.push_back(sides[0]);
    out.bsdfs.push_back(sides[1]);
    out.materials.push_back(dev_material(m));
    DevMedium md{};
    if (scatter_medium_of(m, md)) out.consts.ext = 1;
    out.media.push_back(md);
  }
  // Lights (light.rs:656-994).
  V3 centre;
  float radius;
  if (scene_bounds.empty()) {
    centre = v3(0, 0, 0);
    radius = 1000.0f;
  } else {
    centre = scene_bounds.centroid();
    radius = fmax_r(0.5f * length(scene_bounds.hi - scene_bounds.lo), 1.0f);
  }
  float radius_m = radius * MM_TO_M;
  auto reach = [&](V3 p, const V3* n) -> float {
    V3 to_centre = centre - p;
    float d = length(to_centre) * MM_TO_M;
    float facing = n ? 0.05f + 0.95f * fmax_r(dot(*n, normalize(to_centre)), 0.0f) : 1.0f;
    d = fmax_r(d, radius_m);
    return facing / (d * d);
  };
  std::vector<float> weights;
  for (const HostLight& l : s.lights) {
    if (l.kind == 2) {
      // A sun (light.rs:755-791).
      V3 to_sun = -normalize(l.direction);
      if (length_sq(to_sun) < 0.5f) continue;
      float half = fmin_r(0.5f * fmax_r(l.angle_deg, 0.0f) * deg_to_rad(), 0.5f * PI);
      C3 color = c3(fmax_r(l.color.r, 0.0f), fmax_r(l.color.g, 0.0f), fmax_r(l.color.b, 0.0f));
      C3 irradiance = color * fmax_r(l.power, 0.0f);
      if (!(luminance(irradiance) > 0.0f)) continue;
      DevSun sun;
      sun.delta = half < 1.0e-5f ? 1u : 0u;
      float sh = fsin(0.5f * half);
      sun.one_minus_cos_max = 2.0f * sh * sh;
      float sin_max = fsin(half);
      sun.radiance = sun.delta ? irradiance : irradiance * (1.0f / (PI * sin_max * sin_max));
      weights.push_back(PI * luminance(irradiance));
      sun.entry = (u32)out.entries.size();
      out.entries.push_back(ENTRY_SUN - (i32)out.suns.size());
      sun.frame = frame_from_normal(to_sun);
      sun.cos_max = 1.0f - sun.one_minus_cos_max;
      sun.solid_angle = TAU * sun.one_minus_cos_max;
      sun.glossy = (l.flags & 1u) ? 1u : 0u;
      sun.camera = (l.flags & 2u) ? 1u : 0u;
      sun.shadow = (l.flags & 4u) ? 0u : 1u;
      sun.base = luminance(irradiance);
      out.suns.push_back(sun);
      out.consts.ext = 1;
      continue;
    }
    if (l.kind == 3 || l.kind == 4) {
      // light.rs:793-863 (no IES profile): a point (diameter 0) or a sphere, a spot masked by its
      // cone with a smoothstep over the soft edge.
      float power = fmax_r(l.power, 0.0f);
      C3 color = c3(fmax_r(l.color.r, 0.0f), fmax_r(l.color.g, 0.0f), fmax_r(l.color.b, 0.0f));
      C3 intensity = color * (power / (4.0f * PI));
      float rad = 0.5f * fabs_(l.size_x);
      if (!is_finite(rad)) rad = 0.0f;
      DevSphere sp;
      if (l.kind == 4) {
        V3 axis = normalize(l.direction);
        if (length_sq(axis) < 0.5f) continue;
        float half = 0.5f * clampf(l.spread_deg, 1.0f, 180.0f) * deg_to_rad();
        float soft = fmax_r(l.angle_deg, 0.0f) * deg_to_rad();
        float inner = fmax_r(half - soft, 0.0f);
        sp.axis = axis;
        sp.cos_outer = fcos(half);
        sp.cos_inner = fcos(inner);
      } else {
        sp.axis = v3(0.0f, 0.0f, -1.0f);
        sp.cos_outer = -2.0f;
        sp.cos_inner = -2.0f;
      }
      float flux = luminance(color) * power;
      if (!(flux > 0.0f)) continue;
      float r_m = rad * MM_TO_M;
      sp.radiance = rad > 0.0f ? intensity * (1.0f / (PI * r_m * r_m)) : splat(0.0f);
      sp.centre = l.position;
      sp.radius = rad;
      sp.intensity = intensity;
      sp.glossy = (l.flags & 1u) ? 1u : 0u;
      sp.camera = (l.flags & 2u) ? 1u : 0u;
      sp.shadow = (l.flags & 4u) ? 0u : 1u;
      sp.entry = (u32)out.entries.size();
      sp.base = luminance(intensity);
      sp.near2 = fmax_r(r_m * r_m, 1.0e-6f);
      float toward = 0.05f + 0.95f * sphere_mask(sp, centre);
      weights.push_back(flux * toward * reach(l.position, nullptr));
      out.entries.push_back(-2 - (i32)out.spheres.size());
      out.spheres.push_back(sp);
      out.consts.ext = 1;
      continue;
    }
    DevArea a;
    a.rect = l.kind == 0 ? 1u : 0u;
    a.glossy = (l.flags & 1u) ? 1u : 0u;
    a.camera = (l.flags & 2u) ? 1u : 0u;
    a.shadow = (l.flags & 4u) ? 0u : 1u;
    float power = fmax_r(l.power, 0.0f);
    a.half_x = 0.5f * fabs_(l.size_x);
    a.half_y = 0.5f * fabs_(l.size_y);
    a.radius = 0.5f * fabs_(l.size_x);
    if (a.rect) {
      a.area = 4.0f * a.half_x * a.half_y;
      a.bound = sqrt_(a.half_x * a.half_x + a.half_y * a.half_y);
    } else {
      a.area = PI * a.radius * a.radius;
      a.bound = a.radius;
    }
    a.frame = light_frame(l.direction, l.up);
    if (!(a.area > 0.0f) || length_sq(a.frame.z) < 0.5f) continue;
    C3 color = c3(fmax_r(l.color.r, 0.0f), fmax_r(l.color.g, 0.0f), fmax_r(l.color.b, 0.0f));
    a.radiance = color * (power / (PI * a.area * MM2_TO_M2));
    a.exponent = spread_exponent(l.spread_deg);
    float flux = luminance(color) * power * 2.0f / (a.exponent + 2.0f);
    if (!(flux > 0.0f)) continue;
    weights.push_back(flux * reach(l.position, &a.frame.z));
    a.entry = (u32)out.entries.size();
    out.entries.push_back((i32)out.areas.size());
    a.centre = l.position;
    a.base = luminance(a.radiance) * a.area * MM2_TO_M2;
    out.areas.push_back(a);
  }
  // Emissive instances whose material is flagged a light, sampled as area lights
  // (light.rs:866-953): world-space triangles of positive area, picked by area.
  out.instance_emitter.assign(s.instances.size(), -1);
  for (size_t i = 0; i < s.instances.size(); ++i) {
    const HostInstance& in = s.instances[i];
    if (in.material >= s.materials.size()) continue;
    const HostMaterial& mat = s.materials[in.material];
    C3 radiance = mat.kind == 2 ? c3(fmax_r(mat.emission.r * mat.emission_power, 0.0f),
                                     fmax_r(mat.emission.g * mat.emission_power, 0.0f),
                                     fmax_r(mat.emission.b * mat.emission_power, 0.0f))
                                : splat(0.0f);
    if (!(mat.flags & MATERIAL_FLAG_LIGHT) || !(luminance(radiance) > 0.0f)) continue;
    if (in.mesh >= meshes.size()) continue;
    Xf x = xf_from(in.rows), inv;
    if (!xf_inverse(x, inv)) continue;
    const HostMesh& m = meshes[in.mesh];
    DevEmitter em;
    em.first = (u32)(out.emit_tris.size() / 4);
    std::vector<float> tri_areas;
    std::vector<V3> corners;
    std::vector<V3> normals;
    V3 centroid = v3(0.0f, 0.0f, 0.0f);
    for (size_t t = 0; t + 2 < m.indices.size(); t += 3) {
      V3 w[3];
      for (int c = 0; c < 3; ++c) {
        const float* p = &m.positions[3 * (size_t)m.indices[t + c]];
        w[c] = xf_point(x, v3(p[0], p[1], p[2]));
      }
      V3 cr = cross(w[1] - w[0], w[2] - w[0]);
      float area = 0.5f * length(cr);
      if (!(area > 0.0f && is_finite(area))) continue;
      centroid = centroid + ((w[0] + w[1]) + w[2]) * (area / 3.0f);
      V3 n = normalize(cr);
      out.emit_tris.push_back(make_float4(w[0].x, w[0].y, w[0].z, 0.0f));
      out.emit_tris.push_back(make_float4(w[1].x, w[1].y, w[1].z, 0.0f));
      out.emit_tris.push_back(make_float4(w[2].x, w[2].y, w[2].z, 0.0f));
      out.emit_tris.push_back(make_float4(n.x, n.y, n.z, 0.0f));
      corners.push_back(w[0]);
      corners.push_back(w[1]);
      corners.push_back(w[2]);
      normals.push_back(n);
      tri_areas.push_back(area);
    }
    float area = -0.0f;
    for (float a : tri_areas) area = area + a;
    std::vector<float> cdf;
    {
      float total = -0.0f;
      for (float a : tri_areas) total = total + a;
      if (total > 0.0f && is_finite(total)) {
        float acc = 0.0f;
        for (float a : tri_areas) {
          acc += a;
          cdf.push_back(acc / total);
        }
        cdf.back() = 1.0f;
      }
    }
    if (cdf.empty()) {
      out.emit_tris.resize((size_t)em.first * 4);
      continue;
    }
    out.emit_cdf.insert(out.emit_cdf.end(), cdf.begin(), cdf.end());
    em.count = (u32)cdf.size();
    em.two_sided = (in.flags & 1u) ? 0u : 1u;
    float sides = em.two_sided ? 2.0f : 1.0f;
    float flux = luminance(radiance) * PI * area * MM2_TO_M2 * sides;
    em.centroid = centroid / area;
    float bound = 0.0f;
    for (const V3& c : corners) bound = fmax_r(bound, length(c - em.centroid));
    em.bound = bound;
    V3 mean = v3(0.0f, 0.0f, 0.0f);
    for (size_t t = 0; t < normals.size(); ++t) mean = mean + normals[t] * tri_areas[t];
    float mlen = length(mean);
    if (em.two_sided || mlen <= 1.0e-3f * area) {
      em.axis = v3(0.0f, 0.0f, 1.0f);
      em.cos_o = -1.0f;
      em.sin_o = 0.0f;
    } else {
      em.axis = mean / mlen;
      float cos_o = 1.0f;
      for (const V3& n : normals) cos_o = fmin_r(cos_o, dot(n, em.axis));
      cos_o = clampf(cos_o, -1.0f, 1.0f);
      em.cos_o = cos_o;
      em.sin_o = sqrt_(fmax_r(1.0f - cos_o * cos_o, 0.0f));
    }
    weights.push_back(flux * reach(em.centroid, nullptr));
    em.entry = (u32)out.entries.size();
    out.entries.push_back(ENTRY_MESH - (i32)out.emitters.size());
    out.instance_emitter[i] = (i32)out.emitters.size();
    em.area = area;
    em.radiance = radiance;
    em.base = luminance(radiance) * area * MM2_TO_M2;
    out.emitters.push_back(em);
    out.consts.ext = 1;
  }
  // Caustics (caustics.rs:104-345): casters, their bounding spheres, the photon sources by the
  // flux they send at the spheres, the first radius from the pixel footprint.
  caustic_setup(s, meshes, scene_bounds, out);
  float strength = fmax_r(s.env_strength, 0.0f);
  float ef = s.env_kind == 1 ? s.env_horizon : s.env_floor, eh = s.env_horizon, ez = s.env_kind == 1 ? s.env_horizon : s.env_zenith;
  DevScene& k = out.consts;
  k.env_floor = fmax_r(ef * strength, 0.0f);
  k.env_horizon = fmax_r(eh * strength, 0.0f);
  k.env_zenith = fmax_r(ez * strength, 0.0f);
  k.env_lower_mass = 0.5f * (k.env_floor + k.env_horizon);
  float upper = 0.5f * (k.env_horizon + k.env_zenith);
  k.env_total_mass = k.env_lower_mass + upper;
  float mean = 0.5f * k.env_total_mass;
  k.env_base = PI * mean;
  k.has_env_entry = 0;
  k.env_entry = 0;
  if (!(k.env_total_mass <= 0.0f)) {
    weights.push_back(4.0f * PI * PI * mean);
    k.env_entry = (u32)out.entries.size();
    k.has_env_entry = 1;
    out.entries.push_back(-1);
  }
  // build_cdf (light.rs:636-654)
  {
    float total = -0.0f;
    for (float w : weights) total = total + w;
    if (total > 0.0f && is_finite(total)) {
      float acc = 0.0f;
      for (float w : weights) {
        acc += w;
        out.cdf.push_back(acc / total);
      }
      out.cdf.back() = 1.0f;
    }
  }
  if (out.cdf.empty()) {
    out.entries.clear();
    k.has_env_entry = 0;
  }
  for (size_t e = 0; e < out.cdf.size(); ++e) out.flux_pmf.push_back(out.cdf[e] - (e == 0 ? 0.0f : out.cdf[e - 1]));
  k.n_entries = (u32)out.entries.size();
  k.per_point = (out.entries.size() >= 2 && out.entries.size() <= (size_t)MAX_ENTRIES) ? 1u : 0u;
  k.n_areas = (u32)out.areas.size();
  k.n_spheres = (u32)out.spheres.size();
  k.n_suns = (u32)out.suns.size();
  k.n_emitters = (u32)out.emitters.size();
  k.n_instances = (u32)out.instances.size();
  k.n_materials = (u32)out.materials.size();
  // Camera (scene/mod.rs:160-357): basis and the sensor of the longer side.
  const HostCamera& c = s.camera;
  V3 forward = normalize(c.target - c.position);
  if (length_sq(forward) == 0.0f) forward = v3(0, 1, 0);
  V3 right = normalize(cross(forward, c.up));
  if (length_sq(right) < 0.5f) {
    V3 axis = fabs_(forward.z) < 0.9f ? v3(0, 0, 1) : v3(1, 0, 0);
    right = normalize(cross(forward, axis));
  }
  V3 upv = cross(right, forward);
  u32 width = std::max<u32>(s.width, 1), height = std::max<u32>(s.height, 1);
  float aspect = (float)width / (float)height;
  k.cam_position = c.position;
  k.cam_forward = forward;
  k.cam_right = right;
  k.cam_up = upv;
  k.focal_mm = c.focal_mm;
  if (aspect >= 1.0f) {
    k.sensor_w = c.sensor_mm;
    k.sensor_h = c.sensor_mm / aspect;
  } else {
    k.sensor_w = c.sensor_mm * aspect;
    k.sensor_h = c.sensor_mm;
  }
  bool through_lens = !c.ortho && c.projection == 0;
  k.lens_radius = (c.fstop > 0.0f && c.focus_dist > 0.0f && through_lens) ? c.focal_mm / c.fstop * 0.5f : 0.0f;
  k.focus_dist = c.focus_dist;
  k.cam_fstop = c.fstop;
  k.cam_sensor_mm = c.sensor_mm;
  k.ortho = c.ortho;
  k.ortho_width = c.ortho_width;
  k.aspect = aspect;
  k.exterior = c.exterior;
  k.width = s.width;
  k.height = s.height;
  k.seed = s.seed;
  k.max_depth = s.max_depth == 0 ? 8u : std::min<u32>(s.max_depth, 1024u);
  k.pixel_filter = s.pixel_filter;
  u32 mode = s.background_mode;
  k.background = mode == 1 ? 1u : (mode == 3 ? 1u : (mode == 2 ? 2u : 0u));
  k.background_color = c3(fmax_r(s.background_color.r, 0.0f), fmax_r(s.background_color.g, 0.0f),
                          fmax_r(s.background_color.b, 0.0f));
  k.clamp_indirect = s.clamp_indirect;
  k.floor_mode = s.floor_mode;
  k.floor_z = s.floor_z;
  k.floor_color = s.floor_color;
  k.floor_bsdf = bsdf_lambert(s.floor_color);
  k.texture = nullptr;
  k.tex_material = NO_TEXTURE;
  k.tex_w = k.tex_h = 0;
  k.tex_origin_x = k.tex_origin_y = 0.0f;
  k.tex_scale_x = k.tex_scale_y = 1.0f;
  HostMaterial fallback;
  for (int side = 0; side < 2; ++side)
    k.fallback_bsdf[side] = bsdf_surface(tables, fallback.kind, fallback.base_color, fallback.metallic, fallback.roughness,
                                         fallback.ior, fallback.transmission, fallback.coat, fallback.coat_roughness,
                                         fallback.flags, fallback.absorption_mm, side == 0);
  return true;
}

}  // namespace rpx
