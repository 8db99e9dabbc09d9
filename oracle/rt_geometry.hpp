// ORACLE — test infrastructure only (see rt_core.hpp).
// Restates the reference's geometry queries (paths inside engine/crates/rt-core/src):
//   geometry/transform.rs:24-200   3x4 affine transforms
//   geometry/triangle.rs:41-246    PreparedRay, watertight test, TriPacket lanes
//   geometry/bvh.rs:24-505         binned SAH binary build (16 bins), collapse to wide nodes
//   geometry/mesh.rs:120-226       Mesh::new, inside_out, leaf packets
//   geometry/traverse.rs:41-802    stack walk, node_test, push_ordered/push_any, Best::offer
//   geometry/mod.rs:326-808        Geometry (instances, TLAS), hit_details, terminator_point
// The nearest hit is the smallest (t, instance, triangle); any tree gives the same answer, so
// this restatement of the reference's own tree serves as the checker and as the counter of the
// nodes and packets its traversal visits.
#pragma once
#include <vector>
#include <string>
#include "rt_core.hpp"

namespace orc {

struct Transform {
  float rows[3][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}};
  static Transform identity() { return Transform(); }
  Vec3 point(Vec3 p) const {
    const auto& r = rows;
    return Vec3(r[0][0] * p.x + r[0][1] * p.y + r[0][2] * p.z + r[0][3],
                r[1][0] * p.x + r[1][1] * p.y + r[1][2] * p.z + r[1][3],
                r[2][0] * p.x + r[2][1] * p.y + r[2][2] * p.z + r[2][3]);
  }
  Vec3 vector(Vec3 v) const {
    const auto& r = rows;
    return Vec3(r[0][0] * v.x + r[0][1] * v.y + r[0][2] * v.z,
                r[1][0] * v.x + r[1][1] * v.y + r[1][2] * v.z,
                r[2][0] * v.x + r[2][1] * v.y + r[2][2] * v.z);
  }
  Vec3 transpose_vector(Vec3 v) const {
    const auto& r = rows;
    return Vec3(r[0][0] * v.x + r[1][0] * v.y + r[2][0] * v.z,
                r[0][1] * v.x + r[1][1] * v.y + r[2][1] * v.z,
                r[0][2] * v.x + r[1][2] * v.y + r[2][2] * v.z);
  }
  Ray ray(const Ray& in) const {
    Ray out;
    out.origin = point(in.origin);
    out.dir = vector(in.dir);
    out.t_min = in.t_min;
    out.t_max = in.t_max;
    return out;
  }
  bool is_finite() const {
    for (auto& row : rows)
      for (float v : row)
        if (!std::isfinite(v)) return false;
    return true;
  }
  bool is_identity() const {
    Transform id;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 4; ++j)
        if (!(rows[i][j] == id.rows[i][j])) return false;
    return true;
  }
  float determinant() const;
  bool inverse(Transform& out) const;
  Aabb aabb(const Aabb& b) const;
};

// ---------------------------------------------------------------- triangle.rs
enum class Cull { None, Back };

struct PreparedRay {
  float org[3], inv[3];
  bool negative[3];
  int k[3];
  float shear[3];
  float t_min, t_max;
  static PreparedRay make(const Ray& ray);
  bool has_interval() const { return t_min <= t_max; }
};

struct EdgeHit { float t; float uvw[3]; float det; };
bool watertight(const PreparedRay& r, Vec3 v0, Vec3 v1, Vec3 v2, float t_max, Cull cull, EdgeHit& out);

struct TriPacket {
  float v[3][3][4];  // [corner][axis][lane]
  u32 id[4];
};
u32 packet_hits(const PreparedRay& r, const TriPacket& p, float t_max, Cull cull, float t_out[4]);

// ---------------------------------------------------------------- bvh.rs
constexpr u32 BVH_EMPTY = 0xffffffffu;
constexpr u32 BVH_LEAF = 1u << 31;
constexpr u32 COUNT_BITS = 3;
inline u32 leaf_ref(size_t first, size_t count) { return BVH_LEAF | ((u32)first << COUNT_BITS) | (u32)(count - 1); }
inline void leaf_range(u32 ref, size_t& first, size_t& count) {
  first = (ref & ~BVH_LEAF) >> COUNT_BITS;
  count = (ref & ((1u << COUNT_BITS) - 1)) + 1;
}

constexpr int WIDE = 8;  // the reference's native width on an AVX2 host
struct WideNode {
  float bounds[3][2][WIDE];
  u32 child[WIDE];
};

struct Bvh {
  std::vector<WideNode> nodes;
  u32 root = BVH_EMPTY;
};

struct BuildParams { size_t leaf_target, leaf_max; };
// Builds the wide tree over primitive boxes and centroids; `emit_leaf` receives the primitive
// indices of each leaf (tree order) and returns its reference.
template <class F>
Bvh build_bvh(const std::vector<Aabb>& bounds, const std::vector<Vec3>& centroids, BuildParams params, F emit_leaf);

// ---------------------------------------------------------------- mesh.rs
struct Mesh {
  std::vector<Vec3> positions;
  std::vector<Vec3> normals;  // empty: none (all-zero normals are dropped)
  std::vector<u32> indices;
  u32 material_id = 0;
  Aabb bounds;
  Bvh bvh;
  std::vector<TriPacket> packets;
  bool inside_out = false;

  static bool create(const std::vector<Vec3>& positions, const std::vector<Vec3>& normals,
                     const std::vector<u32>& indices, u32 material_id, Mesh& out, std::string& err);
  size_t triangle_count() const { return indices.size() / 3; }
  bool has_normals() const { return !normals.empty(); }
  void triangle(size_t t, Vec3& a, Vec3& b, Vec3& c) const {
    a = positions[indices[3 * t]];
    b = positions[indices[3 * t + 1]];
    c = positions[indices[3 * t + 2]];
  }
};

constexpr u32 INSTANCE_ONE_SIDED = 1;
constexpr u32 INSTANCE_NO_CAMERA = 4;

struct Instance {
  u32 mesh = 0;
  Transform transform;
  u32 material_id = 0;
  u32 flags = 0;
};

struct PreparedInstance {
  Transform to_local;
  bool identity = true;
  Cull cull = Cull::None;
  bool active = false;
  bool camera_hidden = false;
};

struct RayHit { u32 instance, triangle; float t; };

struct Hit {
  u32 instance, triangle;
  float t;
  Vec3 point, normal, geometric_normal;
  float bary[3];
  bool front_face, outside;
};

// Counters of the reference traversal: wide nodes tested and packets tested.
struct TraversalStats { u64 nodes = 0, packets = 0; };

struct Geometry {
  std::vector<Mesh> meshes;
  std::vector<Instance> instances;
  std::vector<PreparedInstance> prepared;
  Bvh tlas;
  std::vector<u32> tlas_instances;
  Aabb bounds;

  bool build(std::vector<Mesh> meshes, std::vector<Instance> instances, std::string& err);
  // Nearest hit. `camera` steps over instances hidden to the camera.
  bool trace(const Ray& ray, bool camera, RayHit& out, TraversalStats* stats = nullptr) const;
  bool occluded(const Ray& ray, TraversalStats* stats = nullptr) const;
  // Brute force over every triangle of every instance (the reference's test oracle).
  bool brute_force(const Ray& ray, RayHit& nearest, bool& any) const;
  bool hit_details(const Ray& ray, const RayHit& hit, Hit& out) const;
  Vec3 terminator_point(const Hit& hit) const;
};

bool usable(const Ray& ray);

}  // namespace orc
