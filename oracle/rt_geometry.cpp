// ORACLE — test infrastructure only (see rt_core.hpp). Citations in rt_geometry.hpp.
#include "rt_geometry.hpp"
#include <algorithm>

namespace orc {

// ------------------------------------------------------------------ transform.rs:122-200
float Transform::determinant() const {
  const auto& r = rows;
  return r[0][0] * (r[1][1] * r[2][2] - r[1][2] * r[2][1]) -
         r[0][1] * (r[1][0] * r[2][2] - r[1][2] * r[2][0]) +
         r[0][2] * (r[1][0] * r[2][1] - r[1][1] * r[2][0]);
}

bool Transform::inverse(Transform& out) const {
  if (!is_finite()) return false;
  const auto& r = rows;
  float det = determinant();
  if (det == 0.0f || !std::isfinite(det)) return false;
  float inv_det = 1.0f / det;
  float c[3][3] = {
      {r[1][1] * r[2][2] - r[1][2] * r[2][1], r[0][2] * r[2][1] - r[0][1] * r[2][2],
       r[0][1] * r[1][2] - r[0][2] * r[1][1]},
      {r[1][2] * r[2][0] - r[1][0] * r[2][2], r[0][0] * r[2][2] - r[0][2] * r[2][0],
       r[0][2] * r[1][0] - r[0][0] * r[1][2]},
      {r[1][0] * r[2][1] - r[1][1] * r[2][0], r[0][1] * r[2][0] - r[0][0] * r[2][1],
       r[0][0] * r[1][1] - r[0][1] * r[1][0]},
  };
  Transform inv;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) inv.rows[i][j] = c[i][j] * inv_det;
  float t[3] = {r[0][3], r[1][3], r[2][3]};
  for (int i = 0; i < 3; ++i) {
    auto& row = inv.rows[i];
    row[3] = -(row[0] * t[0] + row[1] * t[1] + row[2] * t[2]);
  }
  if (!inv.is_finite()) return false;
  out = inv;
  return true;
}

Aabb Transform::aabb(const Aabb& b) const {
  if (b.is_empty()) return Aabb();
  Aabb out;
  for (int corner = 0; corner < 8; ++corner) {
    Vec3 p((corner & 1) == 0 ? b.min.x : b.max.x, (corner & 2) == 0 ? b.min.y : b.max.y,
           (corner & 4) == 0 ? b.min.z : b.max.z);
    out = out.extend(point(p));
  }
  return out;
}

// ------------------------------------------------------------------ triangle.rs:62-246
PreparedRay PreparedRay::make(const Ray& ray) {
  PreparedRay r;
  float d[3] = {ray.dir.x, ray.dir.y, ray.dir.z};
  float a[3];
  for (int i = 0; i < 3; ++i) {
    r.inv[i] = 1.0f / d[i];
    a[i] = std::fabs(d[i]);
  }
  int kz = (a[0] >= a[1] && a[0] >= a[2]) ? 0 : (a[1] >= a[2] ? 1 : 2);
  int kx = (kz + 1) % 3, ky = (kz + 2) % 3;
  if (d[kz] < 0.0f) std::swap(kx, ky);
  float sz = r.inv[kz];
  r.org[0] = ray.origin.x;
  r.org[1] = ray.origin.y;
  r.org[2] = ray.origin.z;
  for (int i = 0; i < 3; ++i) r.negative[i] = sign_negative(r.inv[i]);
  r.k[0] = kx;
  r.k[1] = ky;
  r.k[2] = kz;
  r.shear[0] = d[kx] * sz;
  r.shear[1] = d[ky] * sz;
  r.shear[2] = sz;
  r.t_min = ray.t_min;
  r.t_max = fmin_r(ray.t_max, F32_MAX);
  return r;
}

bool watertight(const PreparedRay& r, Vec3 v0, Vec3 v1, Vec3 v2, float t_max, Cull cull, EdgeHit& out) {
  int kx = r.k[0], ky = r.k[1], kz = r.k[2];
  float sx = r.shear[0], sy = r.shear[1], sz = r.shear[2];
  float ox = r.org[kx], oy = r.org[ky], oz = r.org[kz];
  float a_x = v0[kx] - ox, a_y = v0[ky] - oy, a_z = v0[kz] - oz;
  float b_x = v1[kx] - ox, b_y = v1[ky] - oy, b_z = v1[kz] - oz;
  float c_x = v2[kx] - ox, c_y = v2[ky] - oy, c_z = v2[kz] - oz;
  float ax = a_x - sx * a_z, ay = a_y - sy * a_z;
  float bx = b_x - sx * b_z, by = b_y - sy * b_z;
  float cx = c_x - sx * c_z, cy = c_y - sy * c_z;
  float u = cx * by - cy * bx;
  float v = ax * cy - ay * cx;
  float w = bx * ay - by * ax;
  bool negative = u < 0.0f || v < 0.0f || w < 0.0f;
  bool positive = u > 0.0f || v > 0.0f || w > 0.0f;
  bool rejected = cull == Cull::None ? (negative && positive) : negative;
  float det = u + v + w;
  if (rejected || det == 0.0f) return false;
  float az = sz * a_z, bz = sz * b_z, cz = sz * c_z;
  float t = (u * az + v * bz + w * cz) / det;
  if (t >= r.t_min && t <= t_max) {
    out.t = t;
    out.uvw[0] = u;
    out.uvw[1] = v;
    out.uvw[2] = w;
    out.det = det;
    return true;
  }
  return false;
}

u32 packet_hits(const PreparedRay& r, const TriPacket& p, float t_max, Cull cull, float t_out[4]) {
  int kx = r.k[0], ky = r.k[1], kz = r.k[2];
  float ox = r.org[kx], oy = r.org[ky], oz = r.org[kz];
  float sx = r.shear[0], sy = r.shear[1], sz = r.shear[2];
  u32 bits = 0;
  for (int l = 0; l < 4; ++l) {
    float a_x = p.v[0][kx][l] - ox, a_y = p.v[0][ky][l] - oy, a_z = p.v[0][kz][l] - oz;
    float b_x = p.v[1][kx][l] - ox, b_y = p.v[1][ky][l] - oy, b_z = p.v[1][kz][l] - oz;
    float c_x = p.v[2][kx][l] - ox, c_y = p.v[2][ky][l] - oy, c_z = p.v[2][kz][l] - oz;
    float ax = a_x - sx * a_z, ay = a_y - sy * a_z;
    float bx = b_x - sx * b_z, by = b_y - sy * b_z;
    float cx = c_x - sx * c_z, cy = c_y - sy * c_z;
    float u = cx * by - cy * bx;
    float v = ax * cy - ay * cx;
    float w = bx * ay - by * ax;
    bool negative = u < 0.0f || v < 0.0f || w < 0.0f;
    bool rejected = cull == Cull::None ? (negative && (u > 0.0f || v > 0.0f || w > 0.0f)) : negative;
    float det = u + v + w;
    float az = sz * a_z, bz = sz * b_z, cz = sz * c_z;
    float t = (u * az + v * bz + w * cz) / det;
    t_out[l] = t;
    bool inside = t >= r.t_min && t <= t_max;
    if (inside && !(rejected || det == 0.0f)) bits |= 1u << l;
  }
  return bits;
}

// ------------------------------------------------------------------ bvh.rs:24-505
namespace {
constexpr size_t BINS = 16;
constexpr u32 SAH_DEPTH_LIMIT = 32;
constexpr float TRAVERSAL_COST = 1.0f;
constexpr float PRIMITIVE_COST = 1.0f;

inline size_t bin_of(float c, float lo, float scale) {
  return std::min<size_t>((size_t)f_to_usize((c - lo) * scale), BINS - 1);
}

struct BinaryNode {
  Aabb bounds;
  u32 left = 0, right = 0, first = 0, count = 0;
};

struct Builder {
  const std::vector<Aabb>& bounds;
  const std::vector<Vec3>& centroids;
  BuildParams params;
  std::vector<u32> order;
  std::vector<BinaryNode> nodes;

  void range_bounds(size_t begin, size_t end, Aabb& b, Aabb& c) const {
    b = Aabb();
    c = Aabb();
    for (size_t i = begin; i < end; ++i) {
      u32 p = order[i];
      b = b.unite(bounds[p]);
      c = c.extend(centroids[p]);
    }
  }

  u32 build(size_t begin, size_t end, u32 depth) {
    Aabb b, cb;
    range_bounds(begin, end, b, cb);
    u32 index = (u32)nodes.size();
    BinaryNode n;
    n.bounds = b;
    n.first = (u32)begin;
    n.count = (u32)(end - begin);
    nodes.push_back(n);
    if (end - begin <= params.leaf_target) return index;
    size_t mid;
    if (!split(begin, end, b, cb, depth, mid)) return index;
    u32 left = build(begin, mid, depth + 1);
    u32 right = build(mid, end, depth + 1);
    nodes[index].left = left;
    nodes[index].right = right;
    nodes[index].count = 0;
    return index;
  }

  bool split(size_t begin, size_t end, const Aabb& b, const Aabb& cb, u32 depth, size_t& mid) {
    size_t count = end - begin;
    size_t median = begin + count / 2;
    if (depth >= SAH_DEPTH_LIMIT) { mid = median; return true; }
    int axis;
    size_t plane;
    float cost;
    if (!best_plane(begin, end, cb, axis, plane, cost)) {
      if (count > params.leaf_max) { mid = median; return true; }
      return false;
    }
    float area = b.surface_area();
    float split_cost = (area > 0.0f && std::isfinite(area)) ? TRAVERSAL_COST + PRIMITIVE_COST * cost / area
                                                             : TRAVERSAL_COST;
    if (count <= params.leaf_max && PRIMITIVE_COST * (float)count <= split_cost) return false;
    float lo = cb.min[axis];
    float scale = (float)BINS / (cb.max[axis] - lo);
    size_t i = begin, j = end;
    while (i < j) {
      float c = centroids[order[i]][axis];
      if (bin_of(c, lo, scale) <= plane) {
        ++i;
      } else {
        --j;
        std::swap(order[i], order[j]);
      }
    }
    mid = (i == begin || i == end) ? median : i;
    return true;
  }

  bool best_plane(size_t begin, size_t end, const Aabb& cb, int& best_axis, size_t& best_plane_out,
                  float& best_cost) const {
    u32 counts[3][BINS] = {};
    Aabb boxes[3][BINS];
    float scales[3] = {0, 0, 0};
    bool usable_axis[3] = {false, false, false};
    bool any = false;
    for (int axis = 0; axis < 3; ++axis) {
      float extent = cb.max[axis] - cb.min[axis];
      if (extent > 0.0f && std::isfinite(extent)) {
        usable_axis[axis] = true;
        scales[axis] = (float)BINS / extent;
        any = true;
      }
    }
    if (!any) return false;
    for (size_t k = begin; k < end; ++k) {
      u32 p = order[k];
      Vec3 c = centroids[p];
      const Aabb& bb = bounds[p];
      for (int axis = 0; axis < 3; ++axis) {
        if (usable_axis[axis]) {
          size_t bin = bin_of(c[axis], cb.min[axis], scales[axis]);
          counts[axis][bin] += 1;
          boxes[axis][bin] = boxes[axis][bin].unite(bb);
        }
      }
    }
    bool have = false;
    for (int axis = 0; axis < 3; ++axis) {
      if (!usable_axis[axis]) continue;
      float right_area[BINS] = {};
      u32 right_count[BINS] = {};
      Aabb acc;
      u32 n = 0;
      for (size_t plane = BINS - 1; plane-- > 0;) {
        acc = acc.unite(boxes[axis][plane + 1]);
        n += counts[axis][plane + 1];
        right_area[plane] = acc.surface_area();
        right_count[plane] = n;
      }
      acc = Aabb();
      n = 0;
      for (size_t plane = 0; plane < BINS - 1; ++plane) {
        acc = acc.unite(boxes[axis][plane]);
        n += counts[axis][plane];
        if (n == 0 || right_count[plane] == 0) continue;
        float cost = acc.surface_area() * (float)n + right_area[plane] * (float)right_count[plane];
        if (!have || cost < best_cost) {
          have = true;
          best_axis = axis;
          best_plane_out = plane;
          best_cost = cost;
        }
      }
    }
    return have;
  }
};

template <class F>
struct Collapser {
  const std::vector<BinaryNode>& bin;
  const std::vector<u32>& order;
  std::vector<WideNode> nodes;
  F& emit_leaf;

  static WideNode empty_node() {
    WideNode w;
    for (int a = 0; a < 3; ++a)
      for (int s = 0; s < WIDE; ++s) {
        w.bounds[a][0][s] = INF;
        w.bounds[a][1][s] = -INF;
      }
    for (int s = 0; s < WIDE; ++s) w.child[s] = BVH_EMPTY;
    return w;
  }

  u32 reference(u32 b) {
    const BinaryNode node = bin[b];
    if (node.count > 0) {
      std::vector<u32> prims(order.begin() + node.first, order.begin() + node.first + node.count);
      return emit_leaf(prims);
    }
    size_t index = nodes.size();
    nodes.push_back(empty_node());
    u32 kids[WIDE] = {};
    kids[0] = node.left;
    kids[1] = node.right;
    size_t len = 2;
    while (len < (size_t)WIDE) {
      bool have = false;
      size_t slot_w = 0;
      float area_w = 0.0f;
      for (size_t slot = 0; slot < len; ++slot) {
        const BinaryNode& kid = bin[kids[slot]];
        if (kid.count == 0) {
          float area = kid.bounds.surface_area();
          if (!have || area > area_w) {
            have = true;
            slot_w = slot;
            area_w = area;
          }
        }
      }
      if (!have) break;
      const BinaryNode opened = bin[kids[slot_w]];
      for (size_t k = len; k > slot_w + 1; --k) kids[k] = kids[k - 1];
      kids[slot_w] = opened.left;
      kids[slot_w + 1] = opened.right;
      ++len;
    }
    WideNode wide = empty_node();
    for (size_t slot = 0; slot < len; ++slot) {
      const Aabb& bb = bin[kids[slot]].bounds;
      for (int a = 0; a < 3; ++a) {
        wide.bounds[a][0][slot] = bb.min[a];
        wide.bounds[a][1][slot] = bb.max[a];
      }
      wide.child[slot] = reference(kids[slot]);
    }
    nodes[index] = wide;
    return (u32)index;
  }
};
}  // namespace

template <class F>
Bvh build_bvh(const std::vector<Aabb>& bounds, const std::vector<Vec3>& centroids, BuildParams params, F emit_leaf) {
  Bvh out;
  if (bounds.empty()) return out;
  Builder b{bounds, centroids, params, {}, {}};
  b.order.resize(bounds.size());
  for (size_t i = 0; i < bounds.size(); ++i) b.order[i] = (u32)i;
  b.build(0, bounds.size(), 0);
  Collapser<F> c{b.nodes, b.order, {}, emit_leaf};
  out.root = c.reference(0);
  out.nodes = std::move(c.nodes);
  return out;
}

// ------------------------------------------------------------------ mesh.rs:120-226
bool Mesh::create(const std::vector<Vec3>& positions, const std::vector<Vec3>& normals,
                  const std::vector<u32>& indices, u32 material_id, Mesh& m, std::string& err) {
  if (indices.size() % 3 != 0) { err = "index_count is not a multiple of 3"; return false; }
  for (u32 i : indices)
    if (i >= positions.size()) { err = "vertex index out of range"; return false; }
  if (!normals.empty() && normals.size() != positions.size()) { err = "normals are not one per vertex"; return false; }
  for (const Vec3& p : positions)
    if (!p.finite()) { err = "non-finite vertex position"; return false; }
  for (const Vec3& n : normals)
    if (!n.finite()) { err = "non-finite vertex normal"; return false; }
  m = Mesh();
  m.positions = positions;
  bool any_normal = false;
  for (const Vec3& n : normals)
    if (n != Vec3(0, 0, 0)) any_normal = true;
  if (any_normal) m.normals = normals;
  m.indices = indices;
  m.material_id = material_id;
  Aabb b;
  for (const Vec3& p : positions) b = b.extend(p);
  m.bounds = b;
  // normals_oppose_winding (mesh.rs:180-197)
  if (m.has_normals()) {
    float agreement = 0.0f;
    for (size_t t = 0; t < m.triangle_count(); ++t) {
      Vec3 a, bb, c;
      m.triangle(t, a, bb, c);
      Vec3 winding = (bb - a).cross(c - a);
      Vec3 n(0, 0, 0);
      for (int k = 0; k < 3; ++k) n = n + m.normals[m.indices[3 * t + k]];
      agreement += winding.normalize().dot(n);
    }
    m.inside_out = agreement < 0.0f;
  }
  // build_bvh (mesh.rs:205-226)
  size_t count = m.triangle_count();
  std::vector<Aabb> boxes(count);
  std::vector<Vec3> centroids(count);
  for (size_t i = 0; i < count; ++i) {
    Vec3 a, bb, c;
    m.triangle(i, a, bb, c);
    boxes[i] = Aabb().extend(a).extend(bb).extend(c);
    centroids[i] = boxes[i].centroid();
  }
  std::vector<TriPacket>& packets = m.packets;
  auto emit = [&](const std::vector<u32>& prims) -> u32 {
    size_t first = packets.size();
    for (size_t s = 0; s < prims.size(); s += 4) {
      size_t n = std::min<size_t>(4, prims.size() - s);
      TriPacket p{};
      for (int lane = 0; lane < 4; ++lane) {
        u32 id = prims[s + std::min<size_t>(lane, n - 1)];
        p.id[lane] = id;
        Vec3 cs[3];
        m.triangle(id, cs[0], cs[1], cs[2]);
        for (int corner = 0; corner < 3; ++corner) {
          p.v[corner][0][lane] = cs[corner].x;
          p.v[corner][1][lane] = cs[corner].y;
          p.v[corner][2][lane] = cs[corner].z;
        }
      }
      packets.push_back(p);
    }
    return leaf_ref(first, packets.size() - first);
  };
  m.bvh = build_bvh(boxes, centroids, BuildParams{4, 8}, emit);
  return true;
}

// ------------------------------------------------------------------ mod.rs:326-384
static Aabb instance_bounds(const Mesh& mesh, const Transform& t) {
  if (t.is_identity()) return mesh.bounds;
  Aabb b = t.aabb(mesh.bounds);
  float scale = b.min.abs().max(b.max.abs()).max_component();
  Vec3 pad = Vec3::splat(scale * 1.0e-5f);
  Aabb out;
  out.min = b.min - pad;
  out.max = b.max + pad;
  return out;
}

bool Geometry::build(std::vector<Mesh> ms, std::vector<Instance> insts, std::string& err) {
  meshes = std::move(ms);
  instances = std::move(insts);
  prepared.clear();
  std::vector<Aabb> boxes;
  std::vector<Vec3> cents;
  std::vector<u32> active;
  bounds = Aabb();
  for (size_t i = 0; i < instances.size(); ++i) {
    const Instance& inst = instances[i];
    if (inst.mesh >= meshes.size()) { err = "instance refers to a mesh that does not exist"; return false; }
    const Mesh& mesh = meshes[inst.mesh];
    Transform inv;
    bool has_inv = inst.transform.inverse(inv);
    bool live = mesh.triangle_count() > 0 && has_inv;
    PreparedInstance p;
    p.to_local = has_inv ? inv : Transform();
    p.identity = inst.transform.is_identity();
    p.cull = (inst.flags & INSTANCE_ONE_SIDED) ? Cull::Back : Cull::None;
    p.active = live;
    p.camera_hidden = (inst.flags & INSTANCE_NO_CAMERA) != 0;
    prepared.push_back(p);
    if (live) {
      Aabb b = instance_bounds(mesh, inst.transform);
      bounds = bounds.unite(b);
      boxes.push_back(b);
      cents.push_back(b.centroid());
      active.push_back((u32)i);
    }
  }
  tlas_instances.clear();
  auto emit = [&](const std::vector<u32>& prims) -> u32 {
    size_t first = tlas_instances.size();
    for (u32 p : prims) tlas_instances.push_back(active[p]);
    return leaf_ref(first, prims.size());
  };
  tlas = build_bvh(boxes, cents, BuildParams{1, 1}, emit);
  return true;
}

bool usable(const Ray& ray) { return ray.origin.finite() && ray.dir.finite() && ray.dir != Vec3(0, 0, 0); }

// ------------------------------------------------------------------ traverse.rs:41-802
namespace {
const float ROBUST_SCALE = 1.0f + 4.0f * F32_EPSILON;

inline float max_f32(float a, float b) { return a > b ? a : b; }
inline float min_f32(float a, float b) { return a < b ? a : b; }
inline u64 entry(float t, u32 ref) { return ((u64)f2u(t) << 32) | (u64)ref; }

struct Stack {
  std::vector<u64> s;
  size_t height() const { return s.size(); }
  void push(u64 e) { s.push_back(e); }
  bool pop(u64& e) {
    if (s.empty()) return false;
    e = s.back();
    s.pop_back();
    return true;
  }
  void truncate(size_t h) { if (s.size() > h) s.resize(h); }
};

struct Best {
  float t;
  u32 instance = 0xffffffffu, triangle = 0xffffffffu;
  void offer(float tt, u32 inst, u32 tri) {
    if (tt < t || (tt == t && (inst < instance || (inst == instance && tri < triangle)))) {
      t = tt;
      instance = inst;
      triangle = tri;
    }
  }
};

u32 node_test(const WideNode& node, const PreparedRay& r, float t_max, float near_out[WIDE]) {
  int nx = r.negative[0] ? 1 : 0, ny = r.negative[1] ? 1 : 0, nz = r.negative[2] ? 1 : 0;
  u32 bits = 0;
  for (int l = 0; l < WIDE; ++l) {
    float tx0 = (node.bounds[0][nx][l] - r.org[0]) * r.inv[0];
    float ty0 = (node.bounds[1][ny][l] - r.org[1]) * r.inv[1];
    float tz0 = (node.bounds[2][nz][l] - r.org[2]) * r.inv[2];
    float tx1 = (node.bounds[0][1 - nx][l] - r.org[0]) * r.inv[0];
    float ty1 = (node.bounds[1][1 - ny][l] - r.org[1]) * r.inv[1];
    float tz1 = (node.bounds[2][1 - nz][l] - r.org[2]) * r.inv[2];
    float near = max_f32(tz0, max_f32(ty0, max_f32(tx0, r.t_min)));
    float far = min_f32(tz1, min_f32(ty1, min_f32(tx1, t_max))) * ROBUST_SCALE;
    near_out[l] = near;
    if (near <= far) bits |= 1u << l;
  }
  return bits;
}

u32 push_ordered(const WideNode& node, u32 bits, const float near[WIDE], Stack& stack) {
  if ((bits & (bits - 1)) == 0) return node.child[__builtin_ctz(bits)];
  u64 keys[WIDE];
  int count = 0;
  while (bits) {
    int slot = __builtin_ctz(bits);
    bits &= bits - 1;
    u64 key = entry(near[slot], node.child[slot]);
    int j = count;
    while (j > 0 && keys[j - 1] < key) {
      keys[j] = keys[j - 1];
      --j;
    }
    keys[j] = key;
    ++count;
  }
  for (int i = 0; i < count - 1; ++i) stack.push(keys[i]);
  return (u32)keys[count - 1];
}

u32 push_any(const WideNode& node, u32 bits, const float near[WIDE], Stack& stack) {
  int first = __builtin_ctz(bits);
  bits &= bits - 1;
  while (bits) {
    int slot = __builtin_ctz(bits);
    bits &= bits - 1;
    stack.push(entry(near[slot], node.child[slot]));
  }
  return node.child[first];
}

// Walks `nodes` from `root`; `leaf(ref)` returns true to stop; `t_max()` is the search bound.
template <class TMax, class Leaf>
bool walk(const std::vector<WideNode>& nodes, u32 root, const PreparedRay& ray, bool ordered, TMax t_max,
          Leaf leaf, Stack& stack, TraversalStats* stats) {
  size_t base = stack.height();
  u32 current = root;
  for (;;) {
    bool descended = false;
    if ((current & BVH_LEAF) == 0) {
      if (current < nodes.size()) {
        const WideNode& node = nodes[current];
        float near[WIDE];
        if (stats) stats->nodes++;
        u32 bits = node_test(node, ray, t_max(), near);
        if (bits != 0) {
          current = ordered ? push_ordered(node, bits, near, stack) : push_any(node, bits, near, stack);
          descended = true;
        }
      }
    } else if (current != BVH_EMPTY && leaf(current, stack)) {
      stack.truncate(base);
      return true;
    }
    if (descended) continue;
    for (;;) {
      if (stack.height() <= base) return false;
      u64 e;
      if (!stack.pop(e)) return false;
      u32 ref = (u32)e;
      float t = u2f((u32)(e >> 32));
      if (t <= t_max() * ROBUST_SCALE) {
        current = ref;
        break;
      }
    }
  }
}
}  // namespace

bool Geometry::trace(const Ray& ray, bool camera, RayHit& out, TraversalStats* stats) const {
  if (!usable(ray)) return false;
  PreparedRay world = PreparedRay::make(ray);
  if (!world.has_interval()) return false;
  Best best{world.t_max};
  Stack stack;
  auto tmax_inst = [&]() { return best.t; };
  auto inst_leaf = [&](u32 ref, Stack& st) -> bool {
    size_t first, count;
    leaf_range(ref, first, count);
    for (size_t slot = first; slot < first + count; ++slot) {
      if (slot >= tlas_instances.size()) continue;
      u32 instance = tlas_instances[slot];
      const PreparedInstance& p = prepared[instance];
      if (!p.active) continue;
      if (camera && p.camera_hidden) continue;
      const Mesh& mesh = meshes[instances[instance].mesh];
      PreparedRay local = p.identity ? world : PreparedRay::make(p.to_local.ray(ray));
      Cull cull = p.cull;
      auto tri_leaf = [&](u32 r2, Stack&) -> bool {
        size_t f2, c2;
        leaf_range(r2, f2, c2);
        if (f2 + c2 > mesh.packets.size()) return false;
        for (size_t k = f2; k < f2 + c2; ++k) {
          const TriPacket& pk = mesh.packets[k];
          if (stats) stats->packets++;
          float ts[4];
          u32 bits = packet_hits(local, pk, best.t, cull, ts);
          while (bits) {
            int lane = __builtin_ctz(bits);
            bits &= bits - 1;
            best.offer(ts[lane], instance, pk.id[lane]);
          }
        }
        return false;
      };
      walk(mesh.bvh.nodes, mesh.bvh.root, local, true, tmax_inst, tri_leaf, st, stats);
    }
    return false;
  };
  walk(tlas.nodes, tlas.root, world, true, tmax_inst, inst_leaf, stack, stats);
  if (best.instance == 0xffffffffu) return false;
  out.instance = best.instance;
  out.triangle = best.triangle;
  out.t = best.t;
  return true;
}

bool Geometry::occluded(const Ray& ray, TraversalStats* stats) const {
  if (!usable(ray)) return false;
  PreparedRay world = PreparedRay::make(ray);
  if (!world.has_interval()) return false;
  Stack stack;
  float tmax_world = world.t_max;
  auto tmax_inst = [&]() { return tmax_world; };
  auto inst_leaf = [&](u32 ref, Stack& st) -> bool {
    size_t first, count;
    leaf_range(ref, first, count);
    for (size_t slot = first; slot < first + count; ++slot) {
      if (slot >= tlas_instances.size()) continue;
      u32 instance = tlas_instances[slot];
      const PreparedInstance& p = prepared[instance];
      if (!p.active) continue;
      const Mesh& mesh = meshes[instances[instance].mesh];
      PreparedRay local = p.identity ? world : PreparedRay::make(p.to_local.ray(ray));
      Cull cull = p.cull;
      float tmax_local = local.t_max;
      auto tmax_mesh = [&]() { return tmax_local; };
      auto tri_leaf = [&](u32 r2, Stack&) -> bool {
        size_t f2, c2;
        leaf_range(r2, f2, c2);
        if (f2 + c2 > mesh.packets.size()) return false;
        for (size_t k = f2; k < f2 + c2; ++k) {
          if (stats) stats->packets++;
          float ts[4];
          if (packet_hits(local, mesh.packets[k], tmax_local, cull, ts) != 0) return true;
        }
        return false;
      };
      if (walk(mesh.bvh.nodes, mesh.bvh.root, local, false, tmax_mesh, tri_leaf, st, stats)) return true;
    }
    return false;
  };
  return walk(tlas.nodes, tlas.root, world, false, tmax_inst, inst_leaf, stack, stats);
}

bool Geometry::brute_force(const Ray& ray, RayHit& nearest, bool& any) const {
  any = false;
  if (!usable(ray)) return false;
  PreparedRay world = PreparedRay::make(ray);
  if (!world.has_interval()) return false;
  Best best{world.t_max};
  for (size_t i = 0; i < instances.size(); ++i) {
    const PreparedInstance& p = prepared[i];
    if (!p.active) continue;
    const Mesh& mesh = meshes[instances[i].mesh];
    PreparedRay local = p.identity ? world : PreparedRay::make(p.to_local.ray(ray));
    for (size_t t = 0; t < mesh.triangle_count(); ++t) {
      Vec3 a, b, c;
      mesh.triangle(t, a, b, c);
      EdgeHit h;
      if (watertight(local, a, b, c, local.t_max, p.cull, h)) any = true;
      if (watertight(local, a, b, c, best.t, p.cull, h)) best.offer(h.t, (u32)i, (u32)t);
    }
  }
  if (best.instance == 0xffffffffu) return false;
  nearest.instance = best.instance;
  nearest.triangle = best.triangle;
  nearest.t = best.t;
  return true;
}

// ------------------------------------------------------------------ mod.rs:682-808
Vec3 Geometry::terminator_point(const Hit& hit) const {
  if (hit.instance >= prepared.size() || !prepared[hit.instance].active) return hit.point;
  const PreparedInstance& p = prepared[hit.instance];
  const Mesh& mesh = meshes[instances[hit.instance].mesh];
  if (hit.triangle >= mesh.triangle_count() || !mesh.has_normals()) return hit.point;
  Vec3 p0, p1, p2;
  mesh.triangle(hit.triangle, p0, p1, p2);
  Vec3 n0 = mesh.normals[mesh.indices[3 * hit.triangle]];
  Vec3 n1 = mesh.normals[mesh.indices[3 * hit.triangle + 1]];
  Vec3 n2 = mesh.normals[mesh.indices[3 * hit.triangle + 2]];
  float b0 = hit.bary[0], b1 = hit.bary[1], b2 = hit.bary[2];
  Vec3 pt = p0 * b0 + p1 * b1 + p2 * b2;
  Vec3 interpolated = n0 * b0 + n1 * b1 + n2 * b2;
  Vec3 world_normal = p.identity ? interpolated : p.to_local.transpose_vector(interpolated);
  float side = world_normal.dot(hit.geometric_normal) < 0.0f ? -1.0f : 1.0f;
  Vec3 lifted = pt;
  Vec3 corners[3] = {p0, p1, p2};
  Vec3 normals[3] = {n0, n1, n2};
  float weights[3] = {b0, b1, b2};
  for (int k = 0; k < 3; ++k) {
    Vec3 n = normals[k].normalize() * side;
    float below = (pt - corners[k]).dot(n);
    if (below < 0.0f) lifted = lifted - n * (below * weights[k]);
  }
  if (!lifted.finite()) return hit.point;
  if (p.identity) return lifted;
  return instances[hit.instance].transform.point(lifted);
}

bool Geometry::hit_details(const Ray& ray, const RayHit& hit, Hit& out) const {
  if (hit.instance >= prepared.size() || !prepared[hit.instance].active) return false;
  const PreparedInstance& p = prepared[hit.instance];
  const Instance& inst = instances[hit.instance];
  const Mesh& mesh = meshes[inst.mesh];
  if (hit.triangle >= mesh.triangle_count()) return false;
  Vec3 p0, p1, p2;
  mesh.triangle(hit.triangle, p0, p1, p2);
  Ray local = p.identity ? ray : p.to_local.ray(ray);
  PreparedRay r = PreparedRay::make(local);
  float bary[3];
  EdgeHit h;
  if (watertight(r, p0, p1, p2, F32_MAX, Cull::None, h)) {
    for (int k = 0; k < 3; ++k) bary[k] = h.uvw[k] / h.det;
  } else {
    for (int k = 0; k < 3; ++k) bary[k] = 1.0f / 3.0f;
  }
  float b0 = bary[0], b1 = bary[1], b2 = bary[2];
  Vec3 point_local = p0 * b0 + p1 * b1 + p2 * b2;
  auto to_normal = [&](Vec3 n) { return p.identity ? n : p.to_local.transpose_vector(n); };
  Vec3 winding = to_normal((p1 - p0).cross(p2 - p0)).normalize();
  Vec3 incoming = ray.dir.normalize();
  Vec3 geometric = winding.length_squared() > 0.0f ? winding : -incoming;
  bool front_face = winding.dot(ray.dir) < 0.0f;
  Vec3 smooth(0, 0, 0);
  if (mesh.has_normals()) {
    Vec3 n0 = mesh.normals[mesh.indices[3 * hit.triangle]];
    Vec3 n1 = mesh.normals[mesh.indices[3 * hit.triangle + 1]];
    Vec3 n2 = mesh.normals[mesh.indices[3 * hit.triangle + 2]];
    smooth = to_normal(n0 * b0 + n1 * b1 + n2 * b2);
  }
  Vec3 normal = smooth.length_squared() > 1.0e-12f ? smooth.normalize() : geometric;
  auto face = [&](Vec3 n) { return n.dot(ray.dir) > 0.0f ? -n : n; };
  out.instance = hit.instance;
  out.triangle = hit.triangle;
  out.t = hit.t;
  out.point = p.identity ? point_local : inst.transform.point(point_local);
  out.normal = face(normal);
  out.geometric_normal = face(geometric);
  out.bary[0] = b0;
  out.bary[1] = b1;
  out.bary[2] = b2;
  out.front_face = front_face;
  out.outside = front_face != mesh.inside_out;
  return true;
}

}  // namespace orc
