// ORACLE — test infrastructure only (see rt_core.hpp).
// The reference's own geometry test generators, restated for the hit-parity tests (paths inside
// engine/crates/rt-core/src):
//   geometry/tests.rs:48-88    soup_geometry: 5,000 triangles, random and degenerate (points,
//                              repeated corners, collinear corners, slivers, exact duplicates)
//   geometry/tests.rs:92-140   random_rays: random origins and directions, rays aimed at random
//                              triangles of random instances, axis-aligned rays with signed zeros,
//                              random parameter intervals
// Both draw from PCG32 (rng.rs) in the reference's order, so the soup and the rays are the ones the
// reference's BVH-against-brute-force tests run on.
#include <cstring>
#include "rt_handle.hpp"

using namespace orc;

extern "C" {

// The soup's vertex positions (xyz, up to `cap_positions`) and its 15,000 indices. Returns the
// number of positions written (0 when `cap_positions` is too small).
uint32_t orc_soup(float* positions, uint32_t cap_positions, uint32_t* indices) {
  Pcg32 rng = Pcg32::seeded(2024, 7);
  auto f = [&]() { return rng.next_f32(); };
  std::vector<Vec3> pos;
  std::vector<u32> idx;
  while (idx.size() < 5000 * 3) {
    u32 kind = (u32)(f() * 10.0f);
    float cx = f() - 0.5f, cy = f() - 0.5f, cz = f() - 0.5f;
    Vec3 center = Vec3(cx, cy, cz) * 100.0f;
    auto corner = [&](float size) {
      float a = f() - 0.5f, b = f() - 0.5f, c = f() - 0.5f;
      return center + Vec3(a, b, c) * size;
    };
    u32 base = (u32)pos.size();
    if (kind == 0) {
      pos.push_back(center);
      pos.push_back(center);
      pos.push_back(center);
    } else if (kind == 1) {
      Vec3 a = corner(20.0f);
      Vec3 b = corner(20.0f);
      pos.push_back(a);
      pos.push_back(b);
      pos.push_back(a);
    } else if (kind == 2) {
      Vec3 a = corner(20.0f);
      Vec3 b = corner(20.0f);
      pos.push_back(a);
      pos.push_back(b);
      pos.push_back(a + (b - a) * 0.375f);
    } else if (kind == 3 && idx.size() >= 3) {
      // A copy of an earlier triangle's indices.
      u32 t = (u32)(idx.size() / 3);
      size_t at = (size_t)((u32)((float)t * 0.618f) % t) * 3;
      u32 c0 = idx[at], c1 = idx[at + 1], c2 = idx[at + 2];
      idx.push_back(c0);
      idx.push_back(c1);
      idx.push_back(c2);
      continue;
    } else if (kind == 4) {
      Vec3 a = corner(30.0f);
      Vec3 b = corner(30.0f);
      pos.push_back(a);
      pos.push_back(b);
      pos.push_back(a + (b - a) * 0.5f + Vec3(0.0f, 0.0f, 1.0e-4f));
    } else {
      Vec3 a = corner(15.0f);
      Vec3 b = corner(15.0f);
      Vec3 c = corner(15.0f);
      pos.push_back(a);
      pos.push_back(b);
      pos.push_back(c);
    }
    idx.push_back(base);
    idx.push_back(base + 1);
    idx.push_back(base + 2);
  }
  if (pos.size() > cap_positions) return 0;
  for (size_t i = 0; i < pos.size(); ++i) {
    positions[3 * i] = pos[i].x;
    positions[3 * i + 1] = pos[i].y;
    positions[3 * i + 2] = pos[i].z;
  }
  std::memcpy(indices, idx.data(), idx.size() * 4);
  return (u32)pos.size();
}

// `count` rays around the loaded scene's geometry, 8 floats each (origin, direction, t_min, t_max).
int orc_random_rays(void* p, uint32_t count, uint64_t seed, float* out) {
  const Geometry& g = static_cast<Handle*>(p)->scene.geometry;
  if (g.instances.empty()) return -1;
  Pcg32 rng = Pcg32::seeded(seed, 99);
  auto f = [&]() { return rng.next_f32(); };
  Aabb b = g.bounds;
  Vec3 ext = b.max - b.min;
  float size = ext.max_component();
  Vec3 lo = b.min - Vec3::splat(size * 0.5f);
  Vec3 span = ext + Vec3::splat(size);
  for (uint32_t n = 0; n < count; ++n) {
    float ox = f() * span.x, oy = f() * span.y, oz = f() * span.z;
    Vec3 origin = lo + Vec3(ox, oy, oz);
    Vec3 dir;
    u32 m4 = n % 4;
    if (m4 == 0 || m4 == 1) {
      size_t ni = g.instances.size();
      const Instance& inst = g.instances[(size_t)(f() * (float)ni) % ni];
      const Mesh& mesh = g.meshes[inst.mesh];
      size_t nt = mesh.triangle_count();
      size_t tri = (size_t)(f() * (float)nt) % nt;
      Vec3 p0, p1, p2;
      mesh.triangle(tri, p0, p1, p2);
      float u = f(), v = f();
      if (u + v > 1.0f) {
        u = 1.0f - u;
        v = 1.0f - v;
      }
      Vec3 target = p0 * (1.0f - u - v) + p1 * u + p2 * v;
      dir = inst.transform.point(target) - origin;
    } else if (m4 == 2 && n % 32 == 2) {
      size_t axis = (size_t)(f() * 3.0f) % 3;
      float sign = f() < 0.5f ? -1.0f : 1.0f;
      float zero = f() < 0.5f ? -0.0f : 0.0f;
      float d[3] = {zero, zero, zero};
      d[axis] = sign;
      dir = Vec3(d[0], d[1], d[2]);
    } else {
      float a = f() - 0.5f, bb = f() - 0.5f, c = f() - 0.5f;
      dir = Vec3(a, bb, c);
    }
    float t_min = 1.0e-3f, t_max = INF;
    u32 m7 = n % 7;
    if (m7 == 0) t_max = f() * 1.5f;
    else if (m7 == 1) t_min = f() * 0.5f;
    float* r = out + 8 * (size_t)n;
    r[0] = origin.x;
    r[1] = origin.y;
    r[2] = origin.z;
    r[3] = dir.x;
    r[4] = dir.y;
    r[5] = dir.z;
    r[6] = t_min;
    r[7] = t_max;
  }
  return 0;
}

}  // extern "C"
