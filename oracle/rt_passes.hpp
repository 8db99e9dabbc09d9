// ORACLE — test infrastructure only (see rt_core.hpp).
// Restates the reference's exact per-pixel passes, the guides of its denoiser (paths inside
// engine/crates/rt-core/src):
//   integrator/mod.rs:880-990  FirstHit, first_hit (the branches the supported scenes reach: no
//                              particles, patterns, programs, textures, weather or bevel)
//   passes.rs:76-250           PASS_* bits, coc_diameter, Exact::of, albedo_of, coverage_of
//   passes.rs:252-420          PassFilm (add_albedo, add_alpha, set_exact, resolve)
//   render.rs:495-640          centre_ray, render_tile_exact, render_tile_albedo (one camera ray
//                              per sample, no edge supersampling, no front glass)
// The albedo of a material carrying this build's differentiable texture is the texture's value at
// the hit point (the texture is an extension with no reference counterpart).
#pragma once
#include <vector>
#include "rt_render.hpp"

namespace orc {

constexpr u32 PASS_ALBEDO = 1u << 0, PASS_NORMAL = 1u << 1, PASS_DEPTH = 1u << 2, PASS_PART_ID = 1u << 3,
              PASS_MATERIAL_ID = 1u << 4, PASS_COC = 1u << 5, PASS_ALPHA = 1u << 6;
constexpr u32 PASS_ALL = (1u << 7) - 1;
constexpr u32 PASS_SAMPLED = PASS_ALBEDO | PASS_ALPHA;
constexpr u32 PASS_EXACT = PASS_ALL & ~PASS_SAMPLED;
constexpr u32 FLOOR_ID = 0xffffffffu;

enum FirstHitKind { FIRST_BACKGROUND = 0, FIRST_FLOOR = 1, FIRST_SURFACE = 2 };
struct FirstHit {
  int kind = FIRST_BACKGROUND;
  float t = 0;
  Vec3 normal;
  u32 instance = 0;
  Rgb albedo;
};

FirstHit first_hit(const Scene& scene, const Prepared& prep, const Ray& ray);

struct Exact {
  Vec3 normal;
  float depth = INF;
  u32 part_id = 0, material_id = 0;
  float coc = 0;
};
float coc_diameter(const Camera& camera, float depth, u32 long_side);
Exact exact_of(const Scene& scene, const Ray& ray, const FirstHit& hit);
Ray centre_ray(const Scene& scene, u32 x, u32 y);

// Resolved passes of a region: planes of the whole region, rows from the top.
struct Passes {
  u32 width = 0, height = 0, mask = 0;
  std::vector<float> albedo, normal, depth, coc, alpha;
  std::vector<u32> part_id, material_id;
};

struct PassFilm {
  u32 width = 0, height = 0, origin_x = 0, origin_y = 0, mask = 0;
  std::vector<float> albedo, albedo_weight, normal, depth, coc, alpha, alpha_weight;
  std::vector<u32> part_id, material_id;
  void init(u32 ox, u32 oy, u32 w, u32 h, u32 mask);
  bool index(u32 x, u32 y, size_t& i) const;
  void add_albedo(u32 x, u32 y, Rgb c);
  void add_alpha(u32 x, u32 y, float a);
  void set_exact(u32 x, u32 y, const Exact& e);
  Passes resolve() const;
};

// Returns the rays cast.
u64 render_tile_exact(const Scene& scene, const Prepared& prep, TileRect tile, PassFilm& film);
u64 render_tile_albedo(const Scene& scene, const Prepared& prep, TileRect tile, u32 sample_index, PassFilm& film);

}  // namespace orc
