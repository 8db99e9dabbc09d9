// ORACLE — test infrastructure only (see rt_core.hpp).
// Restates the reference's caustic photon map (paths inside engine/crates/rt-core/src):
//   caustics.rs:79-145    casters (glass or metal of roughness <= 0.25, not emissive, not a
//                         scattering medium), bounding spheres of the casters
//   caustics.rs:180-345   the setup: photon sources (glossy area lights and suns, emissive meshes)
//                         by the flux they send at the casters, the first radius from the pixel
//                         footprint, the radius schedule of Knaus and Zwicker, the photon streams
//   caustics.rs:345-560   emission (a point on the light, a direction into the casters' cones,
//                         the density of the whole mixture)
//   caustics.rs:581-735   tracing: through casters by BSDF sampling (flux, the adjoint shading
//                         normal correction), stored on the first receiver, roulette from 3 casters
//   caustics.rs:737-960   the map: a hashed grid of cells one diameter wide, buckets in trace
//                         order, the gather with the 2 (1 - d^2/r^2) / (pi r^2) kernel
// Pinned by the reference's closed-form caustic tests (caustics/tests.rs), restated in
// tests/test_oracle_caustics.py.
#pragma once
#include <algorithm>
#include <vector>
#include "rt_render.hpp"

namespace orc {

bool is_caster(const Material& m);

// The caustics of a scene with its prepared lights; false without photons, casters or sources.
bool caustic_setup(const Scene& scene, const Lights& lights, CausticSetup& out);

struct Photon {
  Vec3 position, wi;
  Rgb power;
  Vec3 normal;
  u32 bounces;
};

struct PhotonMap {
  u32 iteration = 0;
  float radius = 1, inv_r2 = 1, norm = 0, inv_cell = 1;
  std::vector<Photon> photons;  // in trace order
  std::vector<u32> order, starts;
  u32 mask = 0;
  u64 rays = 0;
  // The gathered radiance at `point` (normal `normal`) of photons that crossed at most
  // `max_bounces` casters: sum of f(wi) power kernel.
  template <class F>
  Rgb gather(Vec3 point, Vec3 normal, u32 max_bounces, F f) const;
};

// Photons [first, last) of pass `iteration` (the stored ones in index order, and the rays cast).
void trace_photons(const Scene& scene, const Prepared& prep, const CausticSetup& setup, u32 iteration, u32 first,
                   u32 last, std::vector<Photon>& out, u64& rays);
PhotonMap build_photon_map(const CausticSetup& setup, u32 iteration, std::vector<Photon> photons, u64 rays);
PhotonMap trace_photon_map(const Scene& scene, const Prepared& prep, u32 iteration, int threads);

inline void caustic_cell(Vec3 p, float inv_cell, int c[3]) {
  float v[3] = {p.x, p.y, p.z};
  for (int k = 0; k < 3; ++k) {
    float f = std::floor(v[k] * inv_cell);
    c[k] = std::isfinite(f) ? (int)clampf(f, -1.0e9f, 1.0e9f) : 0;
  }
}
inline u32 caustic_bucket(const int c[3], u32 mask) {
  u32 h = ((u32)c[0] * 73856093u) ^ ((u32)c[1] * 19349663u) ^ ((u32)c[2] * 83492791u);
  return h & mask;
}

// caustics.rs:900-950
template <class F>
Rgb PhotonMap::gather(Vec3 point, Vec3 normal, u32 max_bounces, F f) const {
  Rgb sum(0.0f, 0.0f, 0.0f);
  if (photons.empty()) return sum;
  float r = radius;
  int lo[3], hi[3];
  caustic_cell(point - Vec3::splat(r), inv_cell, lo);
  caustic_cell(point + Vec3::splat(r), inv_cell, hi);
  for (int cz = lo[2]; cz <= std::min(hi[2], lo[2] + 1); ++cz)
    for (int cy = lo[1]; cy <= std::min(hi[1], lo[1] + 1); ++cy)
      for (int cx = lo[0]; cx <= std::min(hi[0], lo[0] + 1); ++cx) {
        int cell[3] = {cx, cy, cz};
        u32 b = caustic_bucket(cell, mask);
        for (u32 k = starts[b]; k < starts[b + 1]; ++k) {
          const Photon& p = photons[order[k]];
          if (p.bounces > max_bounces || p.normal.dot(normal) < 0.5f) continue;
          float d2 = (p.position - point).length_squared() * inv_r2;
          int pc[3];
          caustic_cell(p.position, inv_cell, pc);
          if (d2 >= 1.0f || pc[0] != cx || pc[1] != cy || pc[2] != cz) continue;
          float kern = norm * (1.0f - d2);
          Rgb fv = f(p.wi);
          if (peak(fv) > 0.0f) sum += fv * p.power * kern;
        }
      }
  return sum;
}

}  // namespace orc
