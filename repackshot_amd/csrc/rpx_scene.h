// Device-resident scene layout of the GPU tracer and the host-side scene (parsed from the
// reference's RSB v1 bytes, BVHs built, lights prepared, BSDFs precomputed).
//
// HBM layout (all arrays 16-byte aligned, one allocation per array):
//   nodes      Node4[]       4-wide BVH nodes of every mesh (BLAS) followed by the instance BVH
//                            (TLAS): six planes x 4 children of f32 boxes + 4 child references,
//                            128 B per node (two 64-B cache lines, one load of 8 x 16 B)
//   tris       float4[3 n]   triangles in BLAS leaf order: v0 (w = mesh triangle id), v1, v2;
//                            full f32 so the watertight test sees the reference's bits
//   positions  float[3 v], normals float[3 v], indices u32[3 t]   per-mesh vertex data for
//                            hit details (barycentrics, shading normal, terminator lift)
//   instances  DevInstance[] world<->object transforms, BLAS root, material, flags
//   bsdfs      Bsdf[2 m]     per material: seen from outside / inside
//   materials  DevMaterial[] emission, medium absorption
//   areas      DevArea[]     rect and disk lights
// Child reference encoding: bit 31 set = leaf, then (first << 3) | (count - 1); all ones = empty.
#pragma once
#include <string>
#include <vector>
#include "rpx_bsdf.h"

namespace rpx {

constexpr u32 REF_EMPTY = 0xffffffffu;
constexpr u32 REF_LEAF = 1u << 31;
constexpr u32 LEAF_COUNT_BITS = 3;
RPX_HD u32 make_leaf(u32 first, u32 count) { return REF_LEAF | (first << LEAF_COUNT_BITS) | (count - 1); }

struct alignas(16) Node4 {
  float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
  u32 child[4];
  u32 pad[4];
};
static_assert(sizeof(Node4) == 128, "Node4 is two cache lines");

// INST_CASTER: a caster of caustic photons (glass or metal that focuses light).
constexpr u32 INST_IDENTITY = 1, INST_ONE_SIDED = 2, INST_CAMERA_HIDDEN = 4, INST_ACTIVE = 8, INST_CASTER = 16;

struct alignas(16) DevInstance {
  float to_local[12];  // world -> object rows
  float to_world[12];  // object -> world rows
  u32 mesh, blas_root, material, flags;
};

constexpr u32 CORNER_FLOAT4S = 5;

struct DevMesh {
  u32 vertex_offset, index_offset, has_normals, inside_out;
};

// What the integrator reads of a material, and the parameters its BSDF is built from (for
// textured materials, whose BSDF is built per hit, and for the adjoint's parameter derivatives).
// `has_medium`: 0 none, MEDIUM_ABSORBING (Beer-Lambert with `sigma`), MEDIUM_SCATTERING (a random
// walk in the scene's `media` entry of the material).
constexpr u32 MEDIUM_ABSORBING = 1, MEDIUM_SCATTERING = 2;
struct DevMaterial {
  u32 emissive, has_medium;
  C3 radiance;
  C3 sigma;
  u32 kind, flags;
  C3 base;
  float metallic, roughness, ior, transmission, coat, coat_roughness, absorption_mm;
};

// Differentiable parameters per material, in this order (the adjoint's gradient layout).
constexpr int N_MAT_PARAMS = 5;  // roughness, base colour r, g, b, ior

struct DevArea {
  u32 rect, glossy, camera, shadow;
  float half_x, half_y, radius, area;
  V3 centre;
  Frame frame;
  C3 radiance;
  float exponent, base, bound;
  u32 entry;
};

// Point, sphere and spot lights (light.rs:364-420, 793-863): a sphere of `radius` (0: a point)
// at `centre`, masked by a spot's cone about `axis` (cos_outer below -1: no cone).
struct DevSphere {
  V3 centre;
  float radius;
  V3 axis;
  float cos_outer, cos_inner;
  C3 radiance, intensity;
  float base, near2;
  u32 glossy, camera, shadow, entry;
};

// light.rs:364-373: the spot's mask at x (1 without a cone; the smoothstep across the soft edge).
RPX_HD float sphere_mask(const DevSphere& L, V3 x) {
  if (L.cos_outer < -1.0f) return 1.0f;
  V3 d = x - L.centre;
  float len2 = length_sq(d);
  if (!(len2 > 0.0f)) return 0.0f;
  float c = dot(d, L.axis) / sqrt_(len2);
  if (c <= L.cos_outer) return 0.0f;
  if (c >= L.cos_inner) return 1.0f;
  float t = (c - L.cos_outer) / (L.cos_inner - L.cos_outer);
  return t * t * (3.0f - 2.0f * t);
}
// light.rs:376-389: 1 - cos of the cone the sphere subtends from squared distance dist2 (0 when
// it is indistinguishable from a point); false from inside it.
RPX_HD bool sphere_cone(const DevSphere& L, float dist2, float& omc) {
  float r2 = L.radius * L.radius;
  if (!(dist2 - r2 > 0.0f)) return false;
  float sin2 = r2 / dist2;
  float v = sin2 / (1.0f + sqrt_(fmax_r(1.0f - sin2, 0.0f)));
  omc = v < 1.0e-7f ? 0.0f : v;
  return true;
}
// light.rs:392-408: the nearer crossing of the sphere in (RAY_EPSILON, t_max) from outside it.
RPX_HD bool sphere_intersect(const DevSphere& L, V3 o, V3 d, float t_max, float& t_out) {
  if (!(L.radius > 0.0f)) return false;
  V3 oc = o - L.centre;
  float r2 = L.radius * L.radius;
  if (length_sq(oc) <= r2) return false;
  float b = dot(oc, d);
  float disc = r2 - length_sq(oc - d * b);
  if (disc < 0.0f) return false;
  float t = -b - sqrt_(disc);
  if (!(t > RAY_EPSILON && t < t_max)) return false;
  t_out = t;
  return true;
}
// light.rs:411-420: the selection weight at x.
RPX_HD float sphere_weight(const DevSphere& L, V3 x) {
  float m = sphere_mask(L, x);
  if (m <= 0.0f) return 0.0f;
  float d2 = length_sq(x - L.centre) * 1.0e-6f;
  return L.base * m / fmax_r(d2, L.near2);
}

// A sun (light.rs:298-321, 755-791): a disc of half angle acos(cos_max) about frame.z (towards
// the sun), sampled uniformly in its cone; a delta direction below 1e-5 rad of half angle, whose
// radiance is then the irradiance.
struct DevSun {
  Frame frame;
  float cos_max, one_minus_cos_max, solid_angle, base;
  C3 radiance;
  u32 delta, glossy, camera, shadow, entry;
};
RPX_HD bool sun_contains(const DevSun& L, V3 dir) { return !L.delta && dot(dir, L.frame.z) >= L.cos_max; }

// An emissive instance sampled as an area light (light.rs:440-481, 866-953): triangles
// [first, first + count) of the scene's emitter triangle table (world corners and unit normal, 4 x
// float4 each) with their cumulative area fractions; the cone of its normals (cos_o <= -1: any).
struct DevEmitter {
  u32 first, count, two_sided, entry;
  float area, bound, cos_o, sin_o, base;
  C3 radiance;
  V3 centroid, axis;
};

// Entries of the light selection (i32): an area light k >= 0, the environment -1, a sphere light
// -2 - k, a sun ENTRY_SUN - k, an emissive mesh ENTRY_MESH - k.
constexpr i32 ENTRY_SUN = -(1 << 20), ENTRY_MESH = -(2 << 20);
enum EntryKind : u32 { E_AREA, E_ENV, E_SPHERE, E_SUN, E_MESH };
RPX_HD EntryKind entry_kind(i32 a, u32& k) {
  if (a >= 0) {
    k = (u32)a;
    return E_AREA;
  }
  if (a == -1) {
    k = 0;
    return E_ENV;
  }
  if (a > ENTRY_SUN) {
    k = (u32)(-2 - a);
    return E_SPHERE;
  }
  if (a > ENTRY_MESH) {
    k = (u32)(ENTRY_SUN - a);
    return E_SUN;
  }
  k = (u32)(ENTRY_MESH - a);
  return E_MESH;
}

// Caustics (caustics.rs:121-345): photons per sample index, the first gather radius (mm), the
// casters' bounding spheres, the photon sources (kind << 24 | index; kinds 0 area light, 1 sun,
// 2 emissive mesh) with their selection cdf, the scene's bounding sphere, the seed and the most
// casters a photon crosses.
constexpr u32 CAUSTIC_MAX_SPHERES = 16;
struct DevCaustics {
  u32 photons, n_spheres, n_sources, seed, max_bounces;
  float radius, scene_radius;
  V3 scene_centre;
  V3 sphere_c[CAUSTIC_MAX_SPHERES];
  float sphere_r[CAUSTIC_MAX_SPHERES];
  const u32* sources;
  const float* source_cdf;
};

// The photon map of one pass (caustics.rs:737-860): its stored photons in bucket order (trace
// order within a bucket) and where each bucket starts (mask + 2 entries), in device memory.
struct DevPhotonMap {
  u32 count, mask, iteration, pad;
  float radius, inv_r2, norm, inv_cell;
  const float4* position;  // xyz, casters crossed (bits)
  const float4* wi;
  const float4* power;
  const float4* normal;
  const u32* starts;
};

// The photon maps of the sample passes a render is running: pass `first + i` gathers from
// map[i], i < count.
constexpr u32 MAX_PASS_MAPS = 16;
struct DevPhotonMaps {
  u32 first, count, pad0, pad1;
  DevPhotonMap map[MAX_PASS_MAPS];
};

// caustics.rs:760-790: the grid cell of a point (cells one diameter wide) and its bucket.
RPX_HD void photon_cell(V3 p, float inv_cell, i32 c[3]) {
  float v[3] = {p.x, p.y, p.z};
  for (int k = 0; k < 3; ++k) {
    float f = floor_(v[k] * inv_cell);
    c[k] = is_finite(f) ? (i32)clampf(f, -1.0e9f, 1.0e9f) : 0;
  }
}
RPX_HD u32 photon_bucket(const i32 c[3], u32 mask) {
  return (((u32)c[0] * 73856093u) ^ ((u32)c[1] * 19349663u) ^ ((u32)c[2] * 83492791u)) & mask;
}

// A homogeneous scattering medium filling a clear solid (integrator/medium.rs:113-141): per mm
// coefficients, the Henyey-Greenstein anisotropy, and the solid's index.
struct DevMedium {
  C3 sigma_s, sigma_a, sigma_t;
  float g, ior;
};

constexpr int MAX_ENTRIES = 64;
constexpr int KEPT_WEIGHTS = 8;

// Everything the kernels read, passed by value as a kernel argument.
struct DevScene {
  const Node4* nodes;
  const float4* tris;
  const DevInstance* instances;
  const DevMesh* meshes;
  const float* positions;
  const float* normals;
  const u32* indices;
  // Per mesh triangle (mesh index_offset / 3 + triangle), CORNER_FLOAT4S float4s: the corners'
  // positions, then their normals (zero for a mesh without normals), packed, so that the hit
  // details read one record instead of the indices and then six scattered vertices.
  const float4* corners;
  const Bsdf* bsdfs;
  const DevMaterial* materials;
  const DevArea* areas;
  const float* cdf;       // [n_entries]
  const float* flux_pmf;  // [n_entries]
  const i32* entries;     // see entry_kind
  const DevSphere* spheres;
  const DevMedium* media;  // per material (read for MEDIUM_SCATTERING materials)
  const DevSun* suns;
  const DevEmitter* emitters;
  const float4* emit_tris;        // 4 per emitter triangle: corners a, b, c, unit normal
  const float* emit_cdf;          // per emitter triangle: cumulative area fraction within its emitter
  const i32* instance_emitter;    // per instance: its emitter, or -1
  u32 n_spheres, n_suns, n_emitters;
  // Caustics: on when the scene asks for photons and has casters and photon sources; the map of
  // the pass being rendered.
  u32 caustics_on;
  DevCaustics caustics;
  const DevPhotonMaps* pmaps;
  // The scene has sphere lights, suns, emissive meshes sampled as lights or scattering media (the
  // kernels' extended variant).
  u32 ext;
  EnergyTables tables;
  u32 tlas_root, n_instances, n_materials, n_areas, n_entries, per_point, has_env_entry, env_entry;
  // environment ramp
  float env_floor, env_horizon, env_zenith, env_lower_mass, env_total_mass, env_base;
  // camera
  V3 cam_position, cam_forward, cam_right, cam_up;
  float focal_mm, sensor_w, sensor_h, lens_radius, focus_dist, ortho_width, aspect, exterior;
  float cam_fstop, cam_sensor_mm;  // for the circle of confusion of the passes
  u32 ortho;
  // render settings
  u32 width, height, seed, max_depth, pixel_filter, background;
  float clamp_indirect;
  C3 background_color;
  u32 floor_mode;
  float floor_z;
  C3 floor_color;  // the floor's albedo pass
  Bsdf floor_bsdf;
  Bsdf fallback_bsdf[2];
  // An RGB albedo texture (f32, row-major, bilinear, repeating) replacing the base colour of
  // material `tex_material` (NO_TEXTURE: none), mapped planar in world x, y:
  // u = (x - tex_origin_x) * tex_scale_x, v = (y - tex_origin_y) * tex_scale_y.
  const float* texture;
  u32 tex_material, tex_w, tex_h;
  float tex_origin_x, tex_origin_y, tex_scale_x, tex_scale_y;
};
constexpr u32 NO_TEXTURE = 0xffffffffu;

// Bilinear texture coordinates of world point (x, y): the four texels (x0, y0), (x1, y0),
// (x0, y1), (x1, y1) and the weights of x1 and y1.
struct TexTap {
  u32 x0, y0, x1, y1;
  float ax, ay;
};
RPX_HD TexTap texture_tap(u32 w, u32 h, float ox, float oy, float sx, float sy, float x, float y) {
  float u = (x - ox) * sx, v = (y - oy) * sy;
  if (!is_finite(u)) u = 0.0f;
  if (!is_finite(v)) v = 0.0f;
  u = u - floor_(u);
  v = v - floor_(v);
  float fx = u * (float)w - 0.5f, fy = v * (float)h - 0.5f;
  float x0f = floor_(fx), y0f = floor_(fy);
  TexTap t;
  t.ax = fx - x0f;
  t.ay = fy - y0f;
  i32 ix = f_to_i32(x0f), iy = f_to_i32(y0f);
  t.x0 = (u32)(ix < 0 ? ix + (i32)w : (ix >= (i32)w ? ix - (i32)w : ix));
  t.y0 = (u32)(iy < 0 ? iy + (i32)h : (iy >= (i32)h ? iy - (i32)h : iy));
  t.x1 = t.x0 + 1 == w ? 0 : t.x0 + 1;
  t.y1 = t.y0 + 1 == h ? 0 : t.y0 + 1;
  return t;
}
// The texel weights of a tap, in the order (x0, y0), (x1, y0), (x0, y1), (x1, y1).
RPX_HD void texture_weights(const TexTap& t, float w[4]) {
  w[0] = (1.0f - t.ax) * (1.0f - t.ay);
  w[1] = t.ax * (1.0f - t.ay);
  w[2] = (1.0f - t.ax) * t.ay;
  w[3] = t.ax * t.ay;
}
RPX_HD C3 texture_value(const float* tex, u32 w, const TexTap& t) {
  float k[4];
  texture_weights(t, k);
  const float* a = tex + 3 * ((size_t)t.y0 * w + t.x0);
  const float* b = tex + 3 * ((size_t)t.y0 * w + t.x1);
  const float* c = tex + 3 * ((size_t)t.y1 * w + t.x0);
  const float* d = tex + 3 * ((size_t)t.y1 * w + t.x1);
  return c3(((a[0] * k[0] + b[0] * k[1]) + c[0] * k[2]) + d[0] * k[3],
            ((a[1] * k[0] + b[1] * k[1]) + c[1] * k[2]) + d[1] * k[3],
            ((a[2] * k[0] + b[2] * k[1]) + c[2] * k[2]) + d[2] * k[3]);
}

// The BSDF of one side of a material whose base colour is `base` (bsdf.rs:384-438).
RPX_HD Bsdf material_bsdf(const EnergyTables& t, const DevMaterial& m, C3 base, float roughness, float ior,
                          bool entering) {
  return bsdf_surface(t, m.kind, base, m.metallic, roughness, ior, m.transmission, m.coat, m.coat_roughness, m.flags,
                      m.absorption_mm, entering);
}

// ---------------------------------------------------------------- host side
struct HostMaterial {
  u32 kind = 0;
  C3 base_color{0.8f, 0.8f, 0.8f};
  float metallic = 0, roughness = 0.5f, ior = 1.5f, transmission = 0, coat = 0, coat_roughness = 0;
  C3 emission{0, 0, 0};
  float emission_power = 0;
  u32 flags = 0;
  float opacity = 1, anisotropy = 0, absorption_mm = 0, bevel_mm = 0, scatter_mm = 0;
  C3 scatter_color{1, 1, 1};
  float scatter_g = 0, abbe = 0;
};

struct HostLight {
  u32 kind = 0;
  V3 position{0, 0, 1000}, direction{0, 0, -1}, up{0, 1, 0};
  float size_x = 500, size_y = 500;
  C3 color{1, 1, 1};
  float power = 1, spread_deg = 180, angle_deg = 0.5f;
  u32 flags = 3;
};

struct HostMesh {
  std::vector<float> positions;  // xyz
  std::vector<float> normals;    // xyz, empty when none
  std::vector<u32> indices;
  u32 material = 0;
  bool inside_out = false;
  float bmin[3], bmax[3];
};

struct HostInstance {
  u32 mesh = 0;
  float rows[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  u32 material = 0, flags = 0;
};

struct HostCamera {
  V3 position{0, -1000, 0}, target{0, 0, 0}, up{0, 0, 1};
  float focal_mm = 50, sensor_mm = 36, fstop = 0, focus_dist = 1000;
  u32 ortho = 0;
  float ortho_width = 1000;
  u32 projection = 0;
  float exterior = 1;
};

struct HostScene {
  HostCamera camera;
  std::vector<HostMaterial> materials;
  std::vector<HostMesh> meshes;
  std::vector<HostInstance> instances;
  std::vector<HostLight> lights;
  u32 env_kind = 0;
  float env_floor = 0.1f, env_horizon = 0.3f, env_zenith = 0.6f, env_strength = 1;
  C3 background_color{0.05f, 0.05f, 0.05f};
  u32 background_mode = 0;
  u32 floor_mode = 0;
  float floor_z = 0;
  C3 floor_color{0.5f, 0.5f, 0.5f};
  u32 width = 640, height = 360, max_samples = 64, seed = 0, max_depth = 8, pixel_filter = 0;
  float exposure_ev = 0;
  u32 adaptive = 0;
  float noise_threshold = 0;
  float clamp_indirect = 0;
  u32 caustic_photons = 0;
  float caustic_radius = 0;
};

// Parses the RSB subset the tracer follows; false with a message for anything else. With
// `meshes_may_follow`, a file without MESH sections keeps its instances unchecked (an update that
// reuses the loaded meshes checks them).
bool parse_rsb(const uint8_t* bytes, size_t len, HostScene& out, std::string& err, bool meshes_may_follow = false);

// The flattened scene ready for upload (host arrays) plus its DevScene constants.
struct FlatScene {
  std::vector<Node4> nodes;
  std::vector<float4> tris;
  std::vector<DevInstance> instances;
  std::vector<DevMesh> meshes;
  std::vector<float> positions, normals;
  std::vector<u32> indices;
  std::vector<float4> corners;
  std::vector<Bsdf> bsdfs;
  std::vector<DevMaterial> materials;
  std::vector<DevArea> areas;
  std::vector<DevSphere> spheres;
  std::vector<DevSun> suns;
  std::vector<DevEmitter> emitters;
  std::vector<float4> emit_tris;
  std::vector<float> emit_cdf;
  std::vector<i32> instance_emitter;
  std::vector<u32> caustic_sources;
  std::vector<float> caustic_cdf;
  std::vector<DevMedium> media;
  std::vector<float> cdf, flux_pmf;
  std::vector<i32> entries;
  DevScene consts;  // pointers filled at upload
  u64 triangles = 0;
  u32 max_blas_depth = 0, max_tlas_depth = 0;
};

// Builds BVHs, lights and BSDFs (host tables of the energy compensation in `tables`).
bool flatten_scene(const HostScene& s, const EnergyTables& tables, FlatScene& out, std::string& err, int threads);

// The device record and the two BSDF sides (outside, inside) of a material.
DevMaterial dev_material(const HostMaterial& m);
// The scattering medium of a material; false when it has none (a positive, finite scatter_mm in
// a clear solid: integrator/medium.rs:99-141).
bool scatter_medium_of(const HostMaterial& m, DevMedium& out);
void material_sides(const EnergyTables& tables, const HostMaterial& m, Bsdf out[2]);

}  // namespace rpx
