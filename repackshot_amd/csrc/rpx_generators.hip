// Scene generators of this library's workloads, written as RSB v1 bytes (the reference's own scene
// interchange format, scene/rsb.rs:1-130 and the writer src/io/rsb.ts), so the same bytes feed the
// GPU tracer (rt_scene_load) and any other reader.
//
//   kind 0  the reference's example scene (example.rs:20-230): a level-2 icosphere on a floor quad,
//           a half-size second instance, one 5 W rect light, the studio dome, an 85 mm camera
//           at f/11, 512 x 384
//   kind 1  icosphere of 5,120 triangles (level 4, radius 1000 mm) as a rough GGX conductor
//           (F0 0.95/0.64/0.54, roughness 0.3) on a 2-triangle Lambert floor (albedo 0.8), a
//           1 x 1 m rect light of radiance 10 at 6500 K, 3 m up, a uniform dome of 0.5; pinhole
//           camera, fov 40 degrees; 256 x 256, 16 spp, depth 8, seed 0
//   kind 2  eight level-7 icospheres (327,680 triangles each, 2,621,440 in all) radially displaced
//           by seeded value noise (amplitude 0.05 r) at seeded places in a 6 x 6 m box: four rough
//           conductors (roughness 0.05..0.6), two smooth glass spheres (IOR 1.5), two Lambert
//           (albedo 0.2..0.9); three rect lights of seeded size 0.5..2 m and the studio dome;
//           1024 x 1024, 64 spp, depth 12
//   kind 4  the inverse-material fit's scene: kind 2's layout reduced to 200,000 triangles (eight
//           value-noise displaced latitude / longitude spheres of 25,000), every GGX material rough
//           (four conductors, two rough glass of IOR 1.5, two Lambert), a Lambert floor quad
//           (material 8, the one a base-colour texture goes on), three rect lights and the studio
//           dome; 512 x 512, 32 spp, depth 6
//   kind 5  caustics and media: a 200,000-triangle bottle of smooth glass (IOR 1.5) holding a
//           homogeneous scattering medium (sigma_s 0.5, sigma_a 0.05 per metre, g 0.7) on a Lambert
//           floor, lit by one sphere light (radius 50 mm, radiance 500); 1920 x 1080, 64 spp,
//           depth 16
//   kind 6  caustics: kind 5's bottle of clear glass under a glossy disc light (100 mm, radiance
//           500) aimed at it, 4,194,304 caustic photons per sample; 1920 x 1080, 64 spp, depth 16
//   queue   job j of a seeded render queue (rpx_write_queue_job): image 256, 512, 768 or 1024 px
//           square; 10k..500k triangles (log-uniform) in four value-noise displaced spheres
//           (latitude / longitude grids) on a floor; materials drawn as in kind 2; one of 12
//           cameras on a circle round the set; 1..4 rect lights and the studio dome; 16, 32 or
//           64 spp; depth 12
// All arithmetic is deterministic f32 (the sampler's fmath functions for angles), so a seed gives
// the same bytes on every host.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include "../../include/rpx.h"
#include "rpx_math.h"

namespace rpx {
namespace {

struct Writer {
  std::vector<uint8_t> b;
  void u(u32 v) {
    uint8_t t[4];
    std::memcpy(t, &v, 4);
    b.insert(b.end(), t, t + 4);
  }
  void f(float v) { u(f2u(v)); }
  void vec(V3 v) {
    f(v.x);
    f(v.y);
    f(v.z);
  }
  void rgb(C3 c) {
    f(c.r);
    f(c.g);
    f(c.b);
  }
  size_t begin(const char* tag) {
    b.insert(b.end(), tag, tag + 4);
    u(0);
    return b.size();
  }
  void end(size_t start) {
    u32 len = (u32)(b.size() - start);
    std::memcpy(&b[start - 4], &len, 4);
  }
};

struct Mat {
  u32 kind = 0;
  C3 base{0.8f, 0.8f, 0.8f};
  float metallic = 0, roughness = 0.5f, ior = 1.5f, transmission = 0;
  float scatter_mm = 0;
  C3 scatter_color{1, 1, 1};
  float scatter_g = 0;
};
struct MeshData {
  std::vector<V3> pos, nrm;
  std::vector<u32> idx;
  u32 material = 0;
};
struct Cam {
  V3 position, target, up{0, 0, 1};
  float focal = 50, sensor = 36, fstop = 0, focus = 1000, ortho_width = 1000;
};
struct Lgt {
  u32 kind = 0;  // 0 rect, 3 point / sphere
  V3 position, direction, up{0, 0, 1};
  float sx, sy;
  C3 color{1, 1, 1};
  float power, spread = 180, angle = 0;
  u32 flags = 3;
};
struct Inst {
  u32 mesh;
  float rows[12];
  u32 material, flags;
};

void write_camr(Writer& w, const Cam& c) {
  size_t s = w.begin("CAMR");
  w.vec(c.position);
  w.vec(c.target);
  w.vec(c.up);
  w.f(c.focal);
  w.f(c.sensor);
  w.f(c.fstop);
  w.f(c.focus);
  w.u(0);
  w.f(c.ortho_width);
  w.end(s);
}
void write_matl(Writer& w, const Mat& m) {
  size_t s = w.begin("MATL");
  w.u(m.kind);
  w.rgb(m.base);
  w.f(m.metallic);
  w.f(m.roughness);
  w.f(m.ior);
  w.f(m.transmission);
  w.f(0);  // coat
  w.f(0);  // coat roughness
  w.rgb(c3(0, 0, 0));
  w.f(0);  // emission power
  w.u(0);  // flags
  w.f(1);  // opacity
  w.f(0);  // anisotropy
  w.f(0);  // absorption
  w.f(0);  // bevel
  w.f(m.scatter_mm);
  w.rgb(m.scatter_color);
  w.f(m.scatter_g);
  w.f(0);  // abbe
  w.end(s);
}
void write_mesh(Writer& w, const MeshData& m) {
  size_t s = w.begin("MESH");
  w.u((u32)m.pos.size());
  w.u((u32)m.idx.size());
  for (V3 p : m.pos) w.vec(p);
  for (size_t i = 0; i < m.pos.size(); ++i) w.vec(m.nrm.empty() ? v3(0, 0, 0) : m.nrm[i]);
  for (u32 i : m.idx) w.u(i);
  w.u(m.material);
  w.end(s);
}
void write_inst(Writer& w, const Inst& in) {
  size_t s = w.begin("INST");
  w.u(in.mesh);
  for (float v : in.rows) w.f(v);
  w.u(in.material);
  w.u(in.flags);
  w.end(s);
}
void write_lght(Writer& w, const Lgt& l) {
  size_t s = w.begin("LGHT");
  w.u(l.kind);
  w.vec(l.position);
  w.vec(l.direction);
  w.vec(l.up);
  w.f(l.sx);
  w.f(l.sy);
  w.rgb(l.color);
  w.f(l.power);
  w.f(l.spread);
  w.f(l.angle);
  w.u(l.flags);
  w.end(s);
}
void write_envr(Writer& w, u32 kind, float fl, float hz, float zn, float strength = 1.0f) {
  size_t s = w.begin("ENVR");
  w.u(kind);
  w.f(fl);
  w.f(hz);
  w.f(zn);
  w.f(strength);
  w.f(0);  // rotation
  w.rgb(c3(0.05f, 0.05f, 0.05f));
  w.u(0);            // background: the environment
  w.u(0xffffffffu);  // hdri map
  w.u(0xffffffffu);  // backplate map
  w.u(0);            // fit
  w.f(3);            // turbidity
  w.f(0);            // sun azimuth
  w.f(45);           // sun elevation
  w.f(0);
  w.f(0);
  w.f(1);
  w.f(1);
  w.u(0);  // sky model
  w.end(s);
}
void write_flor(Writer& w) {
  size_t s = w.begin("FLOR");
  w.u(0);
  w.f(0);
  w.rgb(c3(0.5f, 0.5f, 0.5f));
  w.end(s);
}
void write_rndr(Writer& w, u32 width, u32 height, u32 spp, u32 seed, u32 depth, u32 photons = 0) {
  size_t s = w.begin("RNDR");
  w.u(width);
  w.u(height);
  w.u(spp);
  w.u(seed);
  w.f(0);  // max seconds
  w.f(0);  // exposure
  w.u(1);  // tonemap
  w.u(depth);
  w.f(0);  // clamp
  w.u(0);  // passes
  w.u(0);  // denoise
  w.u(photons);  // caustic photons
  w.f(0);        // caustic radius (0: three pixels)
  w.f(1);  // denoise strength
  w.u(0);  // pixel filter
  w.u(1);  // edge supersample
  w.end(s);
}

// example.rs:20-111 — the icosahedron subdivided `levels` times, vertices on the unit sphere
// (midpoints in order of first appearance).
void icosphere_dirs(u32 levels, std::vector<V3>& dirs, std::vector<u32>& faces) {
  float t = (1.0f + sqrt_(5.0f)) / 2.0f;
  const float raw[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                            {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  dirs.clear();
  for (auto& r : raw) dirs.push_back(normalize(v3(r[0], r[1], r[2])));
  const u32 f0[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                         {11, 10, 2}, {10, 7, 6}, {7, 1, 8},   {3, 9, 4},  {3, 4, 2},   {3, 2, 6}, {3, 6, 8},
                         {3, 8, 9},  {4, 9, 5},  {2, 4, 11},  {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
  faces.assign(&f0[0][0], &f0[0][0] + 60);
  for (u32 l = 0; l < levels; ++l) {
    std::unordered_map<u64, u32> mid;
    mid.reserve(faces.size() * 2);
    auto midpoint = [&](u32 a, u32 b) -> u32 {
      u64 key = ((u64)std::min(a, b) << 32) | std::max(a, b);
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      dirs.push_back(normalize(dirs[a] + dirs[b]));
      u32 id = (u32)dirs.size() - 1;
      mid.emplace(key, id);
      return id;
    };
    std::vector<u32> next;
    next.reserve(faces.size() * 4);
    for (size_t f = 0; f < faces.size(); f += 3) {
      u32 a = faces[f], b = faces[f + 1], c = faces[f + 2];
      u32 ab = midpoint(a, b);
      u32 bc = midpoint(b, c);
      u32 ca = midpoint(c, a);
      u32 q[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
      next.insert(next.end(), q, q + 12);
    }
    faces.swap(next);
  }
}

MeshData icosphere(V3 center, float radius, u32 levels, u32 material) {
  MeshData m;
  icosphere_dirs(levels, m.nrm, m.idx);
  for (V3 d : m.nrm) m.pos.push_back(center + d * radius);
  m.material = material;
  return m;
}

MeshData floor_quad(float z, float half, u32 material) {
  MeshData m;
  m.pos = {v3(-half, -half, z), v3(half, -half, z), v3(half, half, z), v3(-half, half, z)};
  m.nrm.assign(4, v3(0, 0, 1));
  m.idx = {0, 1, 2, 0, 2, 3};
  m.material = material;
  return m;
}

void identity_rows(float* r) {
  const float id[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  std::memcpy(r, id, 48);
}

// The colour of a black body (rt-cli/src/scenes.rs kelvin_rgb, src/core/light-math.ts:31-53),
// normalised to its largest channel.
C3 kelvin_rgb(float kelvin) {
  float t = clampf(kelvin, 1000.0f, 12000.0f) / 100.0f;
  float r = t <= 66.0f ? 255.0f : 329.69873f * fpow(t - 60.0f, -0.13320476f);
  float g = t <= 66.0f ? 99.4708f * fln(t) - 161.11957f : 288.12217f * fpow(t - 60.0f, -0.07551485f);
  float b = t >= 66.0f ? 255.0f : (t <= 19.0f ? 0.0f : 138.51773f * fln(t - 10.0f) - 305.0448f);
  C3 c = c3(clampf(r, 0.0f, 255.0f), clampf(g, 0.0f, 255.0f), clampf(b, 0.0f, 255.0f));
  float largest = fmax_r(fmax_r(fmax_r(c.r, c.g), c.b), 1.0f);
  return c / largest;
}

// The focal length (mm) of a 36 mm sensor for a horizontal field of view of `fov_deg`.
float focal_for_fov(float fov_deg) {
  float s, c;
  sin_cos(0.5f * fov_deg * deg_to_rad(), s, c);
  return 18.0f / (s / c);
}

// Seeded 3-D value noise in [-1, 1]: lattice values hashed from integer coordinates, smoothstep
// weights, trilinear interpolation.
float lattice(i32 x, i32 y, i32 z, u64 seed) {
  u64 h = mix64(seed ^ mix64(((u64)(u32)x << 42) ^ ((u64)(u32)y << 21) ^ (u64)(u32)z));
  return (float)(h >> 40) * (2.0f / 16777216.0f) - 1.0f;
}
float value_noise(V3 p, u64 seed) {
  float fx = floor_(p.x), fy = floor_(p.y), fz = floor_(p.z);
  i32 ix = f_to_i32(fx), iy = f_to_i32(fy), iz = f_to_i32(fz);
  float tx = p.x - fx, ty = p.y - fy, tz = p.z - fz;
  float sx = tx * tx * (3.0f - 2.0f * tx), sy = ty * ty * (3.0f - 2.0f * ty), sz = tz * tz * (3.0f - 2.0f * tz);
  float c[2][2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
  float x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * sx, x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * sx;
  float x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * sx, x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * sx;
  float y0 = x00 + (x10 - x00) * sy, y1 = x01 + (x11 - x01) * sy;
  return y0 + (y1 - y0) * sz;
}

// Smooth vertex normals: the normalised sum of the (area-weighted) normals of the faces around.
std::vector<V3> vertex_normals(const std::vector<V3>& pos, const std::vector<u32>& idx) {
  std::vector<V3> n(pos.size(), v3(0, 0, 0));
  for (size_t f = 0; f < idx.size(); f += 3) {
    V3 a = pos[idx[f]], b = pos[idx[f + 1]], c = pos[idx[f + 2]];
    V3 w = cross(b - a, c - a);
    for (int k = 0; k < 3; ++k) n[idx[f + k]] = n[idx[f + k]] + w;
  }
  for (V3& v : n) v = normalize(v);
  return n;
}

std::vector<uint8_t> example_scene() {
  Writer w;
  w.b = {'R', 'S', 'B', '1', 0, 0, 0, 0};
  V3 light_position = v3(250, -450, 650), camera_position = v3(560, -900, 400), target = v3(0, 0, 90);
  Cam cam;
  cam.position = camera_position;
  cam.target = target;
  cam.focal = 85;
  cam.sensor = 36;
  cam.fstop = 11;
  cam.focus = length(target - camera_position);
  cam.ortho_width = 600;
  write_camr(w, cam);
  Mat m0, m1, m2;
  m0.base = c3(0.92f, 0.38f, 0.22f);
  m0.roughness = 0.35f;
  m1.base = c3(0.3f, 0.3f, 0.32f);
  m1.roughness = 0.8f;
  m2.base = c3(0.2f, 0.42f, 0.85f);
  m2.roughness = 0.25f;
  write_matl(w, m0);
  write_matl(w, m1);
  write_matl(w, m2);
  write_mesh(w, icosphere(v3(0, 0, 100), 100, 2, 0));
  write_mesh(w, floor_quad(0, 900, 1));
  Inst i0{0, {}, 0, 0}, i1{1, {}, 1, 0}, i2{0, {}, 2, 0};
  identity_rows(i0.rows);
  identity_rows(i1.rows);
  const float small[12] = {0.5f, 0, 0, -330, 0, 0.5f, 0, 120, 0, 0, 0.5f, 0};
  std::memcpy(i2.rows, small, 48);
  write_inst(w, i0);
  write_inst(w, i1);
  write_inst(w, i2);
  Lgt l;
  l.position = light_position;
  l.direction = normalize(target - light_position);
  l.sx = 400;
  l.sy = 400;
  l.power = 5;
  write_lght(w, l);
  write_envr(w, 0, 0.05f, 0.12f, 0.3f);
  write_flor(w);
  write_rndr(w, 512, 384, 16, 1, 8);
  return w.b;
}

std::vector<uint8_t> config_plain(u32 seed) {
  Writer w;
  w.b = {'R', 'S', 'B', '1', 0, 0, 0, 0};
  Cam cam;
  cam.position = v3(0, -5200, 2400);
  cam.target = v3(0, 0, 800);
  cam.focal = focal_for_fov(40.0f);
  write_camr(w, cam);
  Mat metal;
  metal.base = c3(0.95f, 0.64f, 0.54f);
  metal.metallic = 1;
  metal.roughness = 0.3f;
  Mat lambert;
  lambert.base = c3(0.8f, 0.8f, 0.8f);
  lambert.ior = 1.0f;
  write_matl(w, metal);
  write_matl(w, lambert);
  write_mesh(w, icosphere(v3(0, 0, 1000), 1000, 4, 0));
  write_mesh(w, floor_quad(0, 5000, 1));
  Lgt l;
  l.position = v3(0, 0, 3000);
  l.direction = v3(0, 0, -1);
  l.up = v3(0, 1, 0);
  l.sx = 1000;
  l.sy = 1000;
  l.color = kelvin_rgb(6500.0f);
  // radiance 10 over 1 m^2: power = L * pi * A
  l.power = 10.0f * PI * 1.0f;
  write_lght(w, l);
  write_envr(w, 1, 0.5f, 0.5f, 0.5f);
  write_flor(w);
  write_rndr(w, 256, 256, 16, seed, 8);
  return w.b;
}

std::vector<uint8_t> config_large(u32 seed) {
  Writer w;
  w.b = {'R', 'S', 'B', '1', 0, 0, 0, 0};
  Pcg32 rng = Pcg32::seeded((u64)seed, 2);
  Cam cam;
  cam.position = v3(0, -7600, 3400);
  cam.target = v3(0, 0, 500);
  cam.focal = focal_for_fov(40.0f);
  write_camr(w, cam);
  // Materials 0..3 rough conductors, 4..5 smooth glass, 6..7 Lambert.
  for (int k = 0; k < 8; ++k) {
    Mat m;
    if (k < 4) {
      float a = rng.next_f32(), b = rng.next_f32(), c = rng.next_f32();
      m.base = c3(0.55f + 0.45f * a, 0.55f + 0.45f * b, 0.55f + 0.45f * c);
      m.metallic = 1;
      m.roughness = 0.05f + 0.55f * rng.next_f32();
    } else if (k < 6) {
      m.base = c3(1, 1, 1);
      m.transmission = 1;
      m.roughness = 0;
      m.ior = 1.5f;
    } else {
      float a = 0.2f + 0.7f * rng.next_f32();
      m.base = c3(a, a, a);
      m.ior = 1.0f;
      m.roughness = 1.0f;
    }
    write_matl(w, m);
  }
  std::vector<V3> dirs;
  std::vector<u32> faces;
  icosphere_dirs(7, dirs, faces);
  for (int k = 0; k < 8; ++k) {
    float radius = 450.0f + 350.0f * rng.next_f32();
    float cx = (-3000.0f + radius) + (6000.0f - 2.0f * radius) * rng.next_f32();
    float cy = (-3000.0f + radius) + (6000.0f - 2.0f * radius) * rng.next_f32();
    V3 centre = v3(cx, cy, radius * 1.05f);
    u64 nseed = mix64(((u64)seed << 8) | (u64)k);
    MeshData m;
    m.pos.resize(dirs.size());
    for (size_t i = 0; i < dirs.size(); ++i) {
      V3 d = dirs[i];
      float h = value_noise(d * 4.0f, nseed);
      m.pos[i] = centre + d * (radius * (1.0f + 0.05f * h));
    }
    m.idx = faces;
    m.nrm = vertex_normals(m.pos, m.idx);
    m.material = (u32)k;
    write_mesh(w, m);
  }
  for (int k = 0; k < 3; ++k) {
    Lgt l;
    float sx = 500.0f + 1500.0f * rng.next_f32(), sy = 500.0f + 1500.0f * rng.next_f32();
    float s, c;
    sin_cos(TAU * ((float)k / 3.0f + 0.15f * rng.next_f32()), s, c);
    float dist = 3500.0f + 1500.0f * rng.next_f32();
    l.position = v3(dist * c, dist * s, 2500.0f + 1500.0f * rng.next_f32());
    l.direction = normalize(v3(0, 0, 300) - l.position);
    l.sx = sx;
    l.sy = sy;
    l.color = kelvin_rgb(4500.0f + 3000.0f * rng.next_f32());
    // about radiance 8..16 over the panel's area in m^2
    l.power = (8.0f + 8.0f * rng.next_f32()) * PI * (sx * sy * 1.0e-6f);
    write_lght(w, l);
  }
  write_envr(w, 0, 0.1f, 0.3f, 0.6f);
  write_flor(w);
  write_rndr(w, 1024, 1024, 64, seed, 12);
  return w.b;
}

// The parameters of job `job` of the render queue with seed `seed`: every job draws from one
// PCG32 stream in job order.
struct QueueJob {
  u32 res, tris, spp, camera, lights;
  u64 geo_seed;
};
QueueJob queue_job(u32 seed, u32 job) {
  Pcg32 rng = Pcg32::seeded((u64)seed, 3);
  QueueJob q{};
  for (u32 j = 0; j <= job; ++j) {
    q.res = 256u * (1u + std::min<u32>(3u, (u32)(rng.next_f32() * 4.0f)));
    q.tris = (u32)(10000.0f * fpow(50.0f, rng.next_f32()));
    q.spp = 16u << std::min<u32>(2u, (u32)(rng.next_f32() * 3.0f));
    q.camera = std::min<u32>(11u, (u32)(rng.next_f32() * 12.0f));
    q.lights = 1u + std::min<u32>(3u, (u32)(rng.next_f32() * 4.0f));
    q.geo_seed = mix64(((u64)seed << 32) | (u64)j);
  }
  return q;
}

// A sphere of about `tris` triangles as a latitude / longitude grid (lat rows, 2 lat columns),
// its radius displaced by value noise of amplitude 0.05.
MeshData displaced_grid_sphere(V3 centre, float radius, u32 tris, u64 seed, u32 material) {
  u32 lat = std::max<u32>(2u, (u32)sqrt_((float)tris / 4.0f));
  u32 lon = 2 * lat;
  MeshData m;
  for (u32 i = 0; i <= lat; ++i) {
    float st, ct;
    sin_cos(PI * (float)i / (float)lat, st, ct);
    for (u32 j = 0; j < lon; ++j) {
      float sp, cp;
      sin_cos(TAU * (float)j / (float)lon, sp, cp);
      V3 d = v3(st * cp, st * sp, ct);
      float h = value_noise(d * 4.0f, seed);
      m.pos.push_back(centre + d * (radius * (1.0f + 0.05f * h)));
    }
  }
  for (u32 i = 0; i < lat; ++i)
    for (u32 j = 0; j < lon; ++j) {
      u32 a = i * lon + j, b = i * lon + (j + 1) % lon, c = a + lon, d = b + lon;
      u32 q[6] = {a, c, b, b, c, d};
      m.idx.insert(m.idx.end(), q, q + 6);
    }
  m.nrm = vertex_normals(m.pos, m.idx);
  m.material = material;
  return m;
}

std::vector<uint8_t> config_fit(u32 seed) {
  Writer w;
  w.b = {'R', 'S', 'B', '1', 0, 0, 0, 0};
  Pcg32 rng = Pcg32::seeded((u64)seed, 5);
  Cam cam;
  cam.position = v3(0, -7600, 3400);
  cam.target = v3(0, 0, 500);
  cam.focal = focal_for_fov(40.0f);
  write_camr(w, cam);
  // 0..3 rough conductors, 4..5 rough glass, 6..7 Lambert, 8 the floor.
  for (int k = 0; k < 8; ++k) {
    Mat m;
    if (k < 4) {
      float a = rng.next_f32(), b = rng.next_f32(), c = rng.next_f32();
      m.base = c3(0.55f + 0.35f * a, 0.55f + 0.35f * b, 0.55f + 0.35f * c);
      m.metallic = 1;
      m.roughness = 0.05f + 0.55f * rng.next_f32();
    } else if (k < 6) {
      m.base = c3(0.9f, 0.92f, 0.95f);
      m.transmission = 1;
      m.roughness = 0.05f + 0.25f * rng.next_f32();
      m.ior = 1.5f;
    } else {
      float a = 0.2f + 0.7f * rng.next_f32();
      m.base = c3(a, a * 0.9f, a * 0.8f);
      m.ior = 1.0f;
      m.roughness = 1.0f;
    }
    write_matl(w, m);
  }
  Mat floor_m;
  floor_m.base = c3(0.5f, 0.5f, 0.5f);
  floor_m.ior = 1.0f;
  floor_m.roughness = 1.0f;
  write_matl(w, floor_m);
  for (u32 k = 0; k < 8; ++k) {
    float radius = 450.0f + 350.0f * rng.next_f32();
    float cx = (-3000.0f + radius) + (6000.0f - 2.0f * radius) * rng.next_f32();
    float cy = (-3000.0f + radius) + (6000.0f - 2.0f * radius) * rng.next_f32();
    write_mesh(w, displaced_grid_sphere(v3(cx, cy, radius * 1.05f), radius, 25000, mix64(((u64)seed << 8) | k), k));
  }
  write_mesh(w, floor_quad(0, 4000, 8));
  for (int k = 0; k < 3; ++k) {
    Lgt l;
    float sx = 500.0f + 1500.0f * rng.next_f32(), sy = 500.0f + 1500.0f * rng.next_f32();
    float s, c;
    sin_cos(TAU * ((float)k / 3.0f + 0.15f * rng.next_f32()), s, c);
    float dist = 3500.0f + 1500.0f * rng.next_f32();
    l.position = v3(dist * c, dist * s, 2500.0f + 1500.0f * rng.next_f32());
    l.direction = normalize(v3(0, 0, 300) - l.position);
    l.sx = sx;
    l.sy = sy;
    l.color = kelvin_rgb(4500.0f + 3000.0f * rng.next_f32());
    l.power = (8.0f + 8.0f * rng.next_f32()) * PI * (sx * sy * 1.0e-6f);
    write_lght(w, l);
  }
  write_envr(w, 0, 0.1f, 0.3f, 0.6f);
  write_flor(w);
  write_rndr(w, 512, 512, 32, seed, 6);
  return w.b;
}

// A closed surface of revolution about the z axis through rings of (height, radius): `segments`
// around, smooth normals on the side, flat caps at both ends, wound outwards.
MeshData revolution(const std::vector<float>& z, const std::vector<float>& r, u32 segments, u32 material) {
  MeshData m;
  m.material = material;
  const u32 rings = (u32)z.size();
  std::vector<float> cs(segments), sn(segments);
  for (u32 j = 0; j < segments; ++j) sin_cos(TAU * (float)j / (float)segments, sn[j], cs[j]);
  for (u32 i = 0; i < rings; ++i)
    for (u32 j = 0; j < segments; ++j) m.pos.push_back(v3(r[i] * cs[j], r[i] * sn[j], z[i]));
  for (u32 i = 0; i + 1 < rings; ++i)
    for (u32 j = 0; j < segments; ++j) {
      u32 jn = (j + 1) % segments;
      u32 a = i * segments + j, b = i * segments + jn, c = (i + 1) * segments + jn, d = (i + 1) * segments + j;
      m.idx.insert(m.idx.end(), {a, b, c, a, c, d});
    }
  m.nrm = vertex_normals(m.pos, m.idx);
  // The caps: their own vertices (flat normals), a fan about the axis.
  for (int end = 0; end < 2; ++end) {
    u32 ring = end == 0 ? 0 : rings - 1;
    float nz = end == 0 ? -1.0f : 1.0f;
    u32 centre = (u32)m.pos.size();
    m.pos.push_back(v3(0, 0, z[ring]));
    m.nrm.push_back(v3(0, 0, nz));
    for (u32 j = 0; j < segments; ++j) {
      m.pos.push_back(v3(r[ring] * cs[j], r[ring] * sn[j], z[ring]));
      m.nrm.push_back(v3(0, 0, nz));
    }
    for (u32 j = 0; j < segments; ++j) {
      u32 a = centre + 1 + j, b = centre + 1 + (j + 1) % segments;
      if (end == 0) m.idx.insert(m.idx.end(), {centre, b, a});
      else m.idx.insert(m.idx.end(), {centre, a, b});
    }
  }
  return m;
}

// Caustics-and-media workload: a bottle (surface of revolution, 300 mm tall, its radius profile
// through 12 seeded control points, Catmull-Rom between them; 400 bands x 250 segments = 200,000
// side triangles and two 250-triangle caps) of smooth glass (IOR 1.5) filled with a homogeneous
// medium of sigma_s 0.5 and sigma_a 0.05 per metre (mean free path 1818.18 mm, single-scattering
// albedo 0.5 / 0.55) and Henyey-Greenstein g 0.7, standing 1 mm above a Lambert floor (albedo
// 0.8), lit by one sphere light of radius 50 mm and radiance 500 (power 500 pi r^2 4 pi W), no
// dome; 1920 x 1080, 64 spp, depth 16. With `caustics`: the bottle of clear glass (no medium)
// under a glossy disc light of the same diameter and radiance at the sphere light's place, aimed
// at the bottle, 4,194,304 caustic photons per sample (the most a scene may ask for).
std::vector<uint8_t> config_bottle(u32 seed, bool caustics = false) {
  Writer w;
  w.b = {'R', 'S', 'B', '1', 0, 0, 0, 0};
  Pcg32 rng = Pcg32::seeded((u64)seed, 5);
  Cam cam;
  cam.position = v3(0, -1150, 380);
  cam.target = v3(0, 0, 150);
  cam.focal = focal_for_fov(40.0f);
  write_camr(w, cam);
  Mat glass;
  glass.kind = 1;
  glass.base = c3(1, 1, 1);
  glass.roughness = 0;
  glass.ior = 1.5f;
  glass.transmission = 1;
  if (!caustics) {
    glass.scatter_mm = 1.0f / 0.55e-3f;
    float albedo = 0.5f / 0.55f;
    glass.scatter_color = c3(albedo, albedo, albedo);
    glass.scatter_g = 0.7f;
  }
  Mat floor_m;
  floor_m.base = c3(0.8f, 0.8f, 0.8f);
  floor_m.ior = 1.0f;
  floor_m.roughness = 1.0f;
  write_matl(w, glass);
  write_matl(w, floor_m);
  // A wine-bottle outline: body, shoulder, neck, lip (mm), each point moved by up to 12 %.
  const float shape[12] = {70, 74, 75, 75, 74, 72, 64, 44, 26, 20, 19, 21};
  float ctrl[12];
  for (int k = 0; k < 12; ++k) ctrl[k] = shape[k] * (1.0f + 0.12f * (2.0f * rng.next_f32() - 1.0f));
  const u32 bands = 400, segments = 250;
  const float height = 300.0f, base_z = 1.0f;
  std::vector<float> zs, rs;
  for (u32 i = 0; i <= bands; ++i) {
    float t = (float)i / (float)bands * 11.0f;
    int k = std::min(10, (int)t);
    float u = t - (float)k;
    float p0 = ctrl[k > 0 ? k - 1 : 0], p1 = ctrl[k], p2 = ctrl[k + 1], p3 = ctrl[k + 2 < 12 ? k + 2 : 11];
    float u2 = u * u, u3 = u2 * u;
    float radius = 0.5f * ((2.0f * p1) + (-p0 + p2) * u + (2.0f * p0 - 5.0f * p1 + 4.0f * p2 - p3) * u2 +
                           (-p0 + 3.0f * p1 - 3.0f * p2 + p3) * u3);
    zs.push_back(base_z + height * (float)i / (float)bands);
    rs.push_back(fmax_r(radius, 5.0f));
  }
  write_mesh(w, revolution(zs, rs, segments, 0));
  write_mesh(w, floor_quad(0, 3000, 1));
  Lgt l;
  l.kind = 3;
  l.position = v3(350, -250, 650);
  l.direction = v3(0, 0, -1);
  l.up = v3(0, 1, 0);
  l.sx = 100;  // diameter
  l.sy = 100;
  const float r_m = 0.05f;
  l.power = 500.0f * PI * r_m * r_m * 4.0f * PI;
  if (caustics) {
    // A disc light's power is its radiance times pi times its area (m^2).
    l.kind = 1;
    l.direction = normalize(v3(0, 0, 150) - l.position);
    l.up = v3(0, 0, 1);
    l.power = 500.0f * PI * (PI * r_m * r_m);
  }
  write_lght(w, l);
  write_envr(w, 0, 0.1f, 0.3f, 0.6f, 0.0f);
  write_flor(w);
  write_rndr(w, 1920, 1080, 64, seed, 16, caustics ? (1u << 22) : 0u);
  return w.b;
}

std::vector<uint8_t> config_queue_job(u32 seed, u32 job) {
  QueueJob q = queue_job(seed, job);
  Writer w;
  w.b = {'R', 'S', 'B', '1', 0, 0, 0, 0};
  Pcg32 rng = Pcg32::seeded(q.geo_seed, 4);
  Cam cam;
  float s, c;
  sin_cos(TAU * (float)q.camera / 12.0f, s, c);
  cam.position = v3(6000.0f * c, 6000.0f * s, 1500.0f + 300.0f * (float)(q.camera % 3));
  cam.target = v3(0, 0, 400);
  cam.focal = focal_for_fov(40.0f);
  write_camr(w, cam);
  // Materials 0..3 drawn as the large render's (conductor, glass, Lambert in 4 : 2 : 2), 4 the floor.
  for (int k = 0; k < 4; ++k) {
    Mat m;
    float pick = rng.next_f32();
    if (pick < 0.5f) {
      float a = rng.next_f32(), b = rng.next_f32(), cc = rng.next_f32();
      m.base = c3(0.55f + 0.45f * a, 0.55f + 0.45f * b, 0.55f + 0.45f * cc);
      m.metallic = 1;
      m.roughness = 0.05f + 0.55f * rng.next_f32();
    } else if (pick < 0.75f) {
      m.base = c3(1, 1, 1);
      m.transmission = 1;
      m.roughness = 0;
      m.ior = 1.5f;
    } else {
      float a = 0.2f + 0.7f * rng.next_f32();
      m.base = c3(a, a, a);
      m.ior = 1.0f;
      m.roughness = 1.0f;
    }
    write_matl(w, m);
  }
  Mat floor_m;
  floor_m.base = c3(0.8f, 0.8f, 0.8f);
  floor_m.ior = 1.0f;
  floor_m.roughness = 1.0f;
  write_matl(w, floor_m);
  for (u32 k = 0; k < 4; ++k) {
    float radius = 300.0f + 400.0f * rng.next_f32();
    float cx = (-2000.0f + radius) + (4000.0f - 2.0f * radius) * rng.next_f32();
    float cy = (-2000.0f + radius) + (4000.0f - 2.0f * radius) * rng.next_f32();
    write_mesh(w, displaced_grid_sphere(v3(cx, cy, radius * 1.05f), radius, q.tris / 4, mix64(q.geo_seed ^ k), k));
  }
  write_mesh(w, floor_quad(0, 6000, 4));
  for (u32 k = 0; k < q.lights; ++k) {
    Lgt l;
    float sx = 500.0f + 1500.0f * rng.next_f32(), sy = 500.0f + 1500.0f * rng.next_f32();
    float ls, lc;
    sin_cos(TAU * ((float)k / (float)q.lights + 0.2f * rng.next_f32()), ls, lc);
    float dist = 3000.0f + 1500.0f * rng.next_f32();
    l.position = v3(dist * lc, dist * ls, 2500.0f + 1500.0f * rng.next_f32());
    l.direction = normalize(v3(0, 0, 300) - l.position);
    l.sx = sx;
    l.sy = sy;
    l.color = kelvin_rgb(4500.0f + 3000.0f * rng.next_f32());
    l.power = (8.0f + 8.0f * rng.next_f32()) * PI * (sx * sy * 1.0e-6f);
    write_lght(w, l);
  }
  write_envr(w, 0, 0.1f, 0.3f, 0.6f);
  write_flor(w);
  write_rndr(w, q.res, q.res, q.spp, job, 12);
  return w.b;
}
}  // namespace
}  // namespace rpx

static uint64_t hand_out(std::vector<uint8_t>& bytes, uint8_t* buf, uint64_t cap) {
  if (!buf || cap < bytes.size()) return cap == 0 ? (uint64_t)bytes.size() : 0;
  std::memcpy(buf, bytes.data(), bytes.size());
  return bytes.size();
}

extern "C" uint64_t rpx_write_scene(uint32_t kind, uint32_t seed, uint8_t* buf, uint64_t cap) {
  // The last scene is kept: callers size the buffer with cap 0 and then fetch it.
  static std::mutex mutex;
  static uint32_t last_kind = 0xffffffffu, last_seed = 0;
  static std::vector<uint8_t> bytes;
  std::lock_guard<std::mutex> lock(mutex);
  if (kind != last_kind || seed != last_seed) {
    switch (kind) {
      case 0: bytes = rpx::example_scene(); break;
      case 1: bytes = rpx::config_plain(seed); break;
      case 2: bytes = rpx::config_large(seed); break;
      case 4: bytes = rpx::config_fit(seed); break;
      case 5: bytes = rpx::config_bottle(seed); break;
      case 6: bytes = rpx::config_bottle(seed, true); break;
      default: return 0;
    }
    last_kind = kind;
    last_seed = seed;
  }
  return hand_out(bytes, buf, cap);
}

extern "C" uint64_t rpx_write_queue_job(uint32_t seed, uint32_t job, uint8_t* buf, uint64_t cap) {
  static std::mutex mutex;
  static uint32_t last_seed = 0, last_job = 0xffffffffu;
  static std::vector<uint8_t> bytes;
  std::lock_guard<std::mutex> lock(mutex);
  if (seed != last_seed || job != last_job) {
    bytes = rpx::config_queue_job(seed, job);
    last_seed = seed;
    last_job = job;
  }
  return hand_out(bytes, buf, cap);
}
