// scene_synth.cpp — seeded synthetic scenes in the engine's GPU formats (host code, no GPU).
//
// Cluster records, the packed triangles, the cluster-ordered vertices, the quantized grid and the
// octahedral normals follow domain/geometry/include/domain/geometry/cluster.h:45-60,350-443 and
// domain/geometry/src/cluster.cpp:437-466; the camera follows core/math/include/core/math/math.h
// (look_at :564-584, perspective_reversed_z :586-600, inverse :336-370) and
// domain/gfx/include/domain/gfx/view_ray.h:26-32; the ray bias follows
// systems/renderer/include/systems/renderer/lighting.h:122-141.
#include "scene_synth.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  double next() { return double(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }
  float uniform(float a, float b) { return float(a + (b - a) * next()); }
};

uint32_t pcg_hash(uint32_t v) {
  const uint32_t state = v * 747796405u + 2891336453u;
  const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}

struct F3 {
  float x, y, z;
};
F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
F3 cross3(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
F3 norm3(F3 a) {
  const float l = std::sqrt(dot3(a, a));
  return l > 0 ? a * (1.0f / l) : F3{0, 0, 1};
}

// The inverse of scene.slang's decode_oct_normal (scene.slang:612-626, encode_oct_normal).
uint32_t encode_oct(F3 n) {
  const float l1 = std::fabs(n.x) + std::fabs(n.y) + std::fabs(n.z);
  float px = l1 > 0 ? n.x / l1 : 0.0f, py = l1 > 0 ? n.y / l1 : 0.0f;
  if (n.z < 0.0f) {
    const float qx = (1.0f - std::fabs(py)) * (px < 0.0f ? -1.0f : 1.0f);
    const float qy = (1.0f - std::fabs(px)) * (py < 0.0f ? -1.0f : 1.0f);
    px = qx;
    py = qy;
  }
  const int x = int(std::round(std::min(std::max(px, -1.0f), 1.0f) * 32767.0f));
  const int y = int(std::round(std::min(std::max(py, -1.0f), 1.0f) * 32767.0f));
  return (uint32_t(x) & 0xffffu) | ((uint32_t(y) & 0xffffu) << 16);
}

// A source mesh before clustering: positions, normals, UVs (two per vertex, or none), and triangle
// corner indices.
struct Source {
  std::vector<F3> pos, nrm;
  std::vector<float> uv;
  std::vector<uint32_t> idx;
};

// A float as an IEEE half, rounded to nearest even (what a half2 UV stream holds).
uint32_t float_to_half(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const float a = std::fabs(f);
  if (!(a < 65520.0f)) return sign | 0x7c00u;
  if (a < 6.103515625e-05f) return sign | uint32_t(std::nearbyint(a * 16777216.0f));
  const uint32_t ax = x & 0x7fffffffu;
  const uint32_t m = ax & 0x7fffffu;
  uint32_t h = (((ax >> 23) - 112u) << 10) | (m >> 13);
  const uint32_t rem = m & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
  return sign | h;
}

void icosphere(Source& m, F3 c, float radius, int subdivisions) {
  const float t = (1.0f + std::sqrt(5.0f)) * 0.5f;
  std::vector<F3> v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                       {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  std::vector<uint32_t> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6,
                             7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10,
                             8, 6, 7, 9, 8, 1};
  for (F3& p : v) p = norm3(p);
  for (int s = 0; s < subdivisions; ++s) {
    std::vector<uint32_t> g;
    std::vector<std::pair<uint64_t, uint32_t>> mid;
    auto midpoint = [&](uint32_t a, uint32_t b) {
      const uint64_t key = (uint64_t(std::min(a, b)) << 32) | std::max(a, b);
      for (auto& e : mid)
        if (e.first == key) return e.second;
      v.push_back(norm3((v[a] + v[b]) * 0.5f));
      mid.push_back({key, uint32_t(v.size() - 1)});
      return uint32_t(v.size() - 1);
    };
    for (size_t i = 0; i < f.size(); i += 3) {
      const uint32_t a = f[i], b = f[i + 1], c2 = f[i + 2];
      const uint32_t ab = midpoint(a, b), bc = midpoint(b, c2), ca = midpoint(c2, a);
      const uint32_t tris[12] = {a, ab, ca, b, bc, ab, c2, ca, bc, ab, bc, ca};
      g.insert(g.end(), tris, tris + 12);
    }
    f.swap(g);
  }
  const uint32_t base = uint32_t(m.pos.size());
  for (const F3& p : v) {
    m.pos.push_back(c + p * radius);
    m.nrm.push_back(p);
  }
  for (uint32_t i : f) m.idx.push_back(base + i);
}

// A box with each face split n x n (2 n^2 triangles a face), flat normals.
void split_box(Source& m, F3 c, F3 half, int n) {
  const F3 axes[6][3] = {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, {{-1, 0, 0}, {0, 0, 1}, {0, 1, 0}},
                         {{0, 1, 0}, {0, 0, 1}, {1, 0, 0}}, {{0, -1, 0}, {1, 0, 0}, {0, 0, 1}},
                         {{0, 0, 1}, {1, 0, 0}, {0, 1, 0}}, {{0, 0, -1}, {0, 1, 0}, {1, 0, 0}}};
  for (int face = 0; face < 6; ++face) {
    const F3 nrm = axes[face][0], u = axes[face][1], w = axes[face][2];
    const uint32_t base = uint32_t(m.pos.size());
    for (int j = 0; j <= n; ++j) {
      for (int i = 0; i <= n; ++i) {
        const float a = -1.0f + 2.0f * float(i) / float(n), b = -1.0f + 2.0f * float(j) / float(n);
        const F3 local = nrm + u * a + w * b;
        m.pos.push_back(c + F3{local.x * half.x, local.y * half.y, local.z * half.z});
        m.nrm.push_back(nrm);
      }
    }
    for (int j = 0; j < n; ++j) {
      for (int i = 0; i < n; ++i) {
        const uint32_t a = base + uint32_t(j * (n + 1) + i), b = a + 1, d = a + uint32_t(n + 1), e = d + 1;
        // Outward winding: u x w = normal for every face above.
        const uint32_t tris[6] = {a, b, e, a, e, d};
        m.idx.insert(m.idx.end(), tris, tris + 6);
      }
    }
  }
}

// A torus of `major` x `minor` quads, radius R, tube r, its tube wobbled by a per-mesh harmonic.
void torus(Source& m, float R, float r, int major, int minor, float wobble, int k1, int k2) {
  const uint32_t base = uint32_t(m.pos.size());
  const float two_pi = 6.283185307179586f;
  for (int i = 0; i <= major; ++i) {
    const float th = two_pi * float(i % major) / float(major);
    for (int j = 0; j <= minor; ++j) {
      const float ph = two_pi * float(j % minor) / float(minor);
      const float rr = r * (1.0f + wobble * std::sin(float(k1) * th + float(k2) * ph));
      const F3 ring{std::cos(th), 0.0f, std::sin(th)};
      const F3 n{std::cos(ph) * ring.x, std::sin(ph), std::cos(ph) * ring.z};
      m.pos.push_back(ring * R + n * rr);
      m.nrm.push_back(n);
    }
  }
  for (int i = 0; i < major; ++i) {
    for (int j = 0; j < minor; ++j) {
      const uint32_t a = base + uint32_t(i * (minor + 1) + j), b = a + 1;
      const uint32_t d = a + uint32_t(minor + 1), e = d + 1;
      const uint32_t tris[6] = {a, d, e, a, e, b};
      m.idx.insert(m.idx.end(), tris, tris + 6);
    }
  }
}

}  // namespace

struct ge_synth {
  std::vector<ge_cluster_desc> clusters;
  std::vector<float> vertices;
  std::vector<uint16_t> quantized;
  std::vector<ge_vertex_attributes> attributes;
  bool has_attributes = true;
  std::vector<uint32_t> triangles;
  std::vector<ge_mesh_desc> meshes;
  std::vector<ge_instance_desc> instances;
  std::vector<ge_resolve_material> materials;
  std::vector<uint32_t> cluster_materials;
  std::vector<ge_resolve_light> lights;
  std::vector<ge_texture_desc> textures;
  std::vector<ge_sampler_desc> samplers;
  std::vector<uint8_t> texels;
  // Frame.
  F3 eye{0, 0, 0}, target{0, 0, 1};
  float fov_deg = 60.0f;
  float sun[4] = {0, 1, 0, 3};
  float sky[3] = {0.55f, 0.70f, 0.90f};
  float ground[3] = {0.3f, 0.3f, 0.3f};
  uint32_t spp = 1, max_bounces = 4;
  float reach = 1.0f;
  uint32_t config = 0;
};

namespace {

// Appends a mesh: the source's triangles in clusters of `cluster_size` in source order, each
// cluster with its own block of vertices (cluster-ordered copies), quantized onto the mesh's own
// 16-bit grid (quantize_positions, cluster.cpp:437-466), with `material` per cluster.
uint32_t add_mesh(ge_synth& s, const Source& src, uint32_t cluster_size, const std::vector<uint32_t>& tri_material,
                  bool split_on_material = false) {
  ge_mesh_desc mesh{};
  mesh.first_cluster = uint32_t(s.clusters.size());
  const uint32_t vbase = uint32_t(s.vertices.size() / 3);
  const uint32_t ntri = uint32_t(src.idx.size() / 3);
  std::vector<F3> local_pos;
  for (uint32_t t0 = 0; t0 < ntri;) {
    // A cluster takes triangles in order until it holds `cluster_size` of them or one more would
    // take it past 255 vertices (local indices are bytes; cluster.h:350-358).
    uint32_t t1 = t0;
    {
      std::vector<uint32_t> probe;
      while (t1 < ntri && t1 - t0 < cluster_size &&
             !(split_on_material && tri_material[t1] != tri_material[t0])) {
        uint32_t fresh = 0;
        for (int k = 0; k < 3; ++k) {
          const uint32_t sv = src.idx[size_t(t1) * 3 + k];
          if (std::find(probe.begin(), probe.end(), sv) == probe.end()) ++fresh;
        }
        if (probe.size() + fresh > 255) break;
        for (int k = 0; k < 3; ++k) {
          const uint32_t sv = src.idx[size_t(t1) * 3 + k];
          if (std::find(probe.begin(), probe.end(), sv) == probe.end()) probe.push_back(sv);
        }
        ++t1;
      }
    }
    ge_cluster_desc cl{};
    cl.vertex_offset = uint32_t(s.vertices.size() / 3);
    cl.triangle_offset = uint32_t(s.triangles.size());
    std::vector<uint32_t> locals;
    for (uint32_t t = t0; t < t1; ++t) {
      uint32_t packed = 0;
      for (int k = 0; k < 3; ++k) {
        const uint32_t sv = src.idx[size_t(t) * 3 + k];
        uint32_t li = 0xffffffffu;
        for (uint32_t q = 0; q < locals.size(); ++q)
          if (locals[q] == sv) {
            li = q;
            break;
          }
        if (li == 0xffffffffu) {
          li = uint32_t(locals.size());
          locals.push_back(sv);
        }
        packed |= li << (8 * k);
      }
      s.triangles.push_back(packed);
    }
    F3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t sv : locals) {
      const F3 p = src.pos[sv];
      s.vertices.insert(s.vertices.end(), {p.x, p.y, p.z});
      const uint32_t uv = src.uv.empty() ? 0u
                                         : float_to_half(src.uv[size_t(sv) * 2]) |
                                               (float_to_half(src.uv[size_t(sv) * 2 + 1]) << 16);
      s.attributes.push_back({encode_oct(norm3(src.nrm[sv])), uv});
      lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
      hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
    }
    cl.vertex_count = uint32_t(locals.size());
    cl.triangle_count = t1 - t0;
    const F3 c = (lo + hi) * 0.5f;
    float rad = 0;
    for (uint32_t sv : locals) rad = std::max(rad, std::sqrt(dot3(src.pos[sv] - c, src.pos[sv] - c)));
    cl.center[0] = c.x;
    cl.center[1] = c.y;
    cl.center[2] = c.z;
    cl.radius = rad;
    cl.cone = 0x7f000000u;  // k_cone_none: never culled
    s.clusters.push_back(cl);
    s.cluster_materials.push_back(tri_material.empty() ? 0u : tri_material[t0]);
    t0 = t1;
  }
  mesh.cluster_count = uint32_t(s.clusters.size()) - mesh.first_cluster;
  // The mesh's grid (cluster.cpp:437-466).
  const uint32_t vend = uint32_t(s.vertices.size() / 3);
  F3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t v = vbase; v < vend; ++v) {
    const F3 p{s.vertices[v * 3], s.vertices[v * 3 + 1], s.vertices[v * 3 + 2]};
    lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
    hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
  }
  const float extent = std::max(std::max(hi.x - lo.x, hi.y - lo.y), hi.z - lo.z);
  const float scale = extent > 0.0f ? extent / 65535.0f : 1.0f;
  const float inverse_scale = 1.0f / scale;
  mesh.quant[0] = lo.x;
  mesh.quant[1] = lo.y;
  mesh.quant[2] = lo.z;
  mesh.quant[3] = scale;
  s.quantized.resize(size_t(vend) * 3);
  for (uint32_t v = vbase; v < vend; ++v) {
    const float axis[3] = {s.vertices[v * 3] - lo.x, s.vertices[v * 3 + 1] - lo.y, s.vertices[v * 3 + 2] - lo.z};
    for (int k = 0; k < 3; ++k) {
      const float grid = std::round(axis[k] * inverse_scale);
      const float clamped = grid < 0.0f ? 0.0f : (grid > 65535.0f ? 65535.0f : grid);
      s.quantized[size_t(v) * 3 + k] = uint16_t(uint32_t(clamped));
    }
  }
  s.meshes.push_back(mesh);
  return uint32_t(s.meshes.size() - 1);
}

ge_instance_desc make_instance(uint32_t mesh, const float rot[3][3], F3 t, uint32_t material_base, uint32_t first_pair) {
  ge_instance_desc in{};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) in.rows[r][c] = rot[r][c];
  in.rows[0][3] = t.x;
  in.rows[1][3] = t.y;
  in.rows[2][3] = t.z;
  in.mesh = mesh;
  in.material_base = material_base;
  in.first_pair = first_pair;
  in.scale_max = 1.0f;
  in.flags = GE_INSTANCE_STATIC_RT | 1u;  // static structure, uniform scale
  in.deform = GE_INVALID_DEFORM;
  return in;
}

// Shoemake 1992: a uniform rotation from three uniforms.
void random_rotation(Rng& rng, float m[3][3]) {
  const double u1 = rng.next(), u2 = rng.next(), u3 = rng.next();
  const double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
  const double w = a * std::sin(2 * M_PI * u2), x = a * std::cos(2 * M_PI * u2);
  const double y = b * std::sin(2 * M_PI * u3), z = b * std::cos(2 * M_PI * u3);
  m[0][0] = float(1 - 2 * (y * y + z * z));
  m[0][1] = float(2 * (x * y - z * w));
  m[0][2] = float(2 * (x * z + y * w));
  m[1][0] = float(2 * (x * y + z * w));
  m[1][1] = float(1 - 2 * (x * x + z * z));
  m[1][2] = float(2 * (y * z - x * w));
  m[2][0] = float(2 * (x * z - y * w));
  m[2][1] = float(2 * (y * z + x * w));
  m[2][2] = float(1 - 2 * (x * x + y * y));
}

ge_resolve_material make_material(float r, float g, float b, float roughness, float metallic) {
  ge_resolve_material m{};
  m.albedo[0] = r;
  m.albedo[1] = g;
  m.albedo[2] = b;
  m.albedo[3] = roughness;
  m.emissive[3] = metallic;
  m.albedo_texture = GE_NO_TEXTURE;
  m.metallic_roughness_texture = GE_NO_TEXTURE;
  m.normal_texture = GE_NO_TEXTURE;
  m.occlusion_texture = GE_NO_TEXTURE;
  m.emissive_texture = GE_NO_TEXTURE;
  m.uv_scale = 1.0f;
  m.normal_scale = 1.0f;
  m.occlusion_strength = 1.0f;
  m.samplers_mr_normal = 0xffffffffu;
  m.samplers_occlusion_emissive = 0xffffffffu;
  m.uv_transform[0] = 1.0f;
  m.uv_transform[3] = 1.0f;
  return m;
}

void set_sun(ge_synth& s, float x, float y, float z, float intensity) {
  const F3 d = norm3({x, y, z});
  s.sun[0] = d.x;
  s.sun[1] = d.y;
  s.sun[2] = d.z;
  s.sun[3] = intensity;
}

// `objects` icospheres (320 triangles) and split boxes (192) alternating, at uniform positions in a
// cube of side `extent`, in one mesh; object o takes material o mod `materials`. Plain materials are
// Lambert-like (roughness 1, dielectric); varied ones draw roughness U[0.05, 1] and are metal one
// time in five.
void build_objects(ge_synth& s, uint64_t seed, uint32_t objects, float extent, uint32_t materials, bool varied) {
  Rng rng(seed);
  Source src;
  for (uint32_t o = 0; o < objects; ++o) {
    const F3 c{rng.uniform(0, extent), rng.uniform(0, extent), rng.uniform(0, extent)};
    if (o % 2 == 0) {
      icosphere(src, c, rng.uniform(1.5f, 5.0f), 2);  // 320 triangles
    } else {
      split_box(src, c, {rng.uniform(1, 4), rng.uniform(1, 4), rng.uniform(1, 4)}, 4);  // 192 triangles
    }
    if (o < materials) {
      const float r = rng.uniform(0.1f, 0.9f), g = rng.uniform(0.1f, 0.9f), b = rng.uniform(0.1f, 0.9f);
      const float rough = varied ? rng.uniform(0.05f, 1.0f) : 1.0f;
      const float metal = varied ? (rng.next() < 0.2 ? 1.0f : 0.0f) : 0.0f;
      s.materials.push_back(make_material(r, g, b, rough, metal));
    }
  }
  // Clusters never straddle two objects: each object's run of triangles is clustered on its own
  // (objects are 320 or 192 triangles, so sequential clusters of 128 would straddle), into one mesh
  // whose grid is fit once over every object afterwards, as quantize_positions fits one per mesh.
  const uint32_t mesh = uint32_t(s.meshes.size());
  {
    ge_mesh_desc m{};
    m.first_cluster = uint32_t(s.clusters.size());
    s.meshes.push_back(m);
  }
  // Build clusters per object into the shared arrays, then quantize once over the whole mesh.
  size_t tri_cursor = 0;
  for (uint32_t o = 0; o < objects; ++o) {
    const uint32_t count = (o % 2 == 0) ? 320u : 192u;
    Source obj;
    std::vector<uint32_t> remap(src.pos.size(), 0xffffffffu);
    for (size_t t = tri_cursor; t < tri_cursor + count; ++t) {
      for (int k = 0; k < 3; ++k) {
        const uint32_t sv = src.idx[t * 3 + k];
        if (remap[sv] == 0xffffffffu) {
          remap[sv] = uint32_t(obj.pos.size());
          obj.pos.push_back(src.pos[sv]);
          obj.nrm.push_back(src.nrm[sv]);
        }
        obj.idx.push_back(remap[sv]);
      }
    }
    tri_cursor += count;
    const std::vector<uint32_t> mats(count, o % materials);
    const uint32_t before_meshes = uint32_t(s.meshes.size());
    add_mesh(s, obj, 128, mats);
    s.meshes.resize(before_meshes);  // fold the object's clusters into the one mesh
  }
  // One grid over every object (quantize_positions runs once per mesh).
  ge_mesh_desc& m = s.meshes[mesh];
  m.cluster_count = uint32_t(s.clusters.size()) - m.first_cluster;
  const uint32_t nv = uint32_t(s.vertices.size() / 3);
  F3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t v = 0; v < nv; ++v) {
    lo = {std::min(lo.x, s.vertices[v * 3]), std::min(lo.y, s.vertices[v * 3 + 1]), std::min(lo.z, s.vertices[v * 3 + 2])};
    hi = {std::max(hi.x, s.vertices[v * 3]), std::max(hi.y, s.vertices[v * 3 + 1]), std::max(hi.z, s.vertices[v * 3 + 2])};
  }
  const float span = std::max(std::max(hi.x - lo.x, hi.y - lo.y), hi.z - lo.z);
  const float scale = span > 0.0f ? span / 65535.0f : 1.0f;
  const float inv = 1.0f / scale;
  m.quant[0] = lo.x;
  m.quant[1] = lo.y;
  m.quant[2] = lo.z;
  m.quant[3] = scale;
  s.quantized.assign(size_t(nv) * 3, 0);
  for (uint32_t v = 0; v < nv; ++v) {
    const float axis[3] = {s.vertices[v * 3] - lo.x, s.vertices[v * 3 + 1] - lo.y, s.vertices[v * 3 + 2] - lo.z};
    for (int k = 0; k < 3; ++k) {
      const float grid = std::round(axis[k] * inv);
      const float clamped = grid < 0.0f ? 0.0f : (grid > 65535.0f ? 65535.0f : grid);
      s.quantized[size_t(v) * 3 + k] = uint16_t(uint32_t(clamped));
    }
  }
  const float identity[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  s.instances.push_back(make_instance(mesh, identity, {0, 0, 0}, 0, 0));
  set_sun(s, 0.3f, 0.8f, 0.5f, 3.0f);
  const float k = extent / 100.0f;  // the camera orbits the cube at its scale: a 22 degree orbit
  s.eye = {k * (50.0f + 125.0f * std::sin(0.3839724f)), k * 95.0f, k * (50.0f - 125.0f * std::cos(0.3839724f))};
  s.target = {k * 50, k * 45, k * 50};
  s.fov_deg = 60.0f;
  s.spp = 1;
  s.max_bounces = 4;
  s.reach = std::sqrt(3.0f) * 1.1f * extent;
}

void build_config1(ge_synth& s, uint64_t seed) { build_objects(s, seed, 256, 100.0f, 256, false); }

// Inverse rendering: 2,048 objects (524,288 triangles) in a 200 m cube, 1,024 varied materials, the
// sun and eight radius-limited point lights; 512 x 512 at 16 spp and 6 bounces.
void build_config4(ge_synth& s, uint64_t seed) {
  build_objects(s, seed, 2048, 200.0f, 1024, true);
  Rng rng(seed ^ 0x4c49u);
  for (uint32_t i = 0; i < 8; ++i) {
    ge_resolve_light light{};
    light.position_radius[0] = rng.uniform(20.0f, 180.0f);
    light.position_radius[1] = rng.uniform(20.0f, 180.0f);
    light.position_radius[2] = rng.uniform(20.0f, 180.0f);
    light.position_radius[3] = 80.0f;
    light.color_intensity[0] = rng.uniform(0.6f, 1.0f);
    light.color_intensity[1] = rng.uniform(0.6f, 1.0f);
    light.color_intensity[2] = rng.uniform(0.6f, 1.0f);
    light.color_intensity[3] = rng.uniform(200.0f, 800.0f);
    s.lights.push_back(light);
  }
  s.spp = 16;
  s.max_bounces = 6;
}

void build_config2(ge_synth& s, uint64_t seed, uint32_t count) {
  Rng rng(seed);
  const float box = 200.0f * std::cbrt(float(count) / 4096.0f);
  for (uint32_t k = 0; k < 16; ++k) {
    Source src;
    torus(src, rng.uniform(1.5f, 3.0f), rng.uniform(0.35f, 0.8f), 32, 16, rng.uniform(0.05f, 0.3f),
          int(1 + rng.next() * 5), int(1 + rng.next() * 4));  // 1,024 triangles
    add_mesh(s, src, 128, std::vector<uint32_t>(src.idx.size() / 3, 0u));
  }
  uint32_t first_pair = 0;
  for (uint32_t i = 0; i < count; ++i) {
    float rot[3][3];
    random_rotation(rng, rot);
    const F3 t{rng.uniform(0, box), rng.uniform(0, box), rng.uniform(0, box)};
    const uint32_t mesh = uint32_t(rng.next() * 16) & 15u;
    s.instances.push_back(make_instance(mesh, rot, t, i, first_pair));
    first_pair += s.meshes[mesh].cluster_count;
    const float roughness = rng.uniform(0.05f, 1.0f);
    const float metallic = rng.next() < 0.2 ? 1.0f : 0.0f;
    s.materials.push_back(make_material(rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f), roughness, metallic));
  }
  for (uint32_t l = 0; l < 64; ++l) {
    ge_resolve_light light{};
    light.position_radius[0] = rng.uniform(0, box);
    light.position_radius[1] = rng.uniform(0, box);
    light.position_radius[2] = rng.uniform(0, box);
    light.position_radius[3] = 0.175f * box;  // radius of influence
    light.color_intensity[0] = rng.uniform(0.6f, 1.0f);
    light.color_intensity[1] = rng.uniform(0.5f, 0.9f);
    light.color_intensity[2] = rng.uniform(0.4f, 0.8f);
    light.color_intensity[3] = rng.uniform(100.0f, 400.0f);
    light.spot[2] = 0.05f;  // the emitter's radius: shadow rays stop that far short
    s.lights.push_back(light);
  }
  set_sun(s, -0.4f, 0.75f, 0.35f, 3.0f);
  s.eye = {0.5f * box, 0.55f * box, -0.12f * box};
  s.target = {0.5f * box, 0.45f * box, 0.6f * box};
  s.fov_deg = 60.0f;
  s.spp = 2;
  s.max_bounces = 8;
  s.reach = std::sqrt(3.0f) * (box + 10.0f);
}

// PCG-hashed value noise, smoothstep-interpolated, summed over octaves (fBm).
float value_noise(float x, float z, uint32_t seed) {
  const float fx = std::floor(x), fz = std::floor(z);
  const int ix = int(fx), iz = int(fz);
  const float tx = x - fx, tz = z - fz;
  auto h = [&](int a, int b) {
    return float(pcg_hash(uint32_t(a) * 73856093u ^ uint32_t(b) * 19349663u ^ seed * 83492791u)) * (1.0f / 4294967296.0f);
  };
  const float sx = tx * tx * (3 - 2 * tx), sz = tz * tz * (3 - 2 * tz);
  const float a = h(ix, iz) + (h(ix + 1, iz) - h(ix, iz)) * sx;
  const float b = h(ix, iz + 1) + (h(ix + 1, iz + 1) - h(ix, iz + 1)) * sx;
  return a + (b - a) * sz;
}
float fbm(float x, float z, uint32_t seed) {
  float sum = 0, amp = 0.5f, freq = 1.0f;
  for (int o = 0; o < 6; ++o) {
    sum += amp * value_noise(x * freq, z * freq, seed + uint32_t(o));
    freq *= 2.0f;   // lacunarity
    amp *= 0.5f;    // gain
  }
  return sum;
}

void build_config3(ge_synth& s, uint64_t seed, uint32_t groups, uint32_t n) {
  Rng rng(seed);
  const float extent = 4096.0f;  // metres across the terrain
  const float cell = extent / float(n - 1);
  const float height = 220.0f;
  // The terrain: one mesh, one instance, rows of quads in clusters of 128 triangles (64 quads: a
  // 8 x 8 quad block, so each cluster is compact).
  {
    std::vector<float> hgt(size_t(n) * n);
    for (uint32_t j = 0; j < n; ++j)
      for (uint32_t i = 0; i < n; ++i)
        hgt[size_t(j) * n + i] = height * fbm(float(i) / 256.0f, float(j) / 256.0f, uint32_t(seed) ^ 0x51u);
    ge_mesh_desc mesh{};
    mesh.first_cluster = uint32_t(s.clusters.size());
    const uint32_t vstart = uint32_t(s.vertices.size() / 3);
    const uint32_t blocks = (n - 1 + 7) / 8;
    for (uint32_t bj = 0; bj < blocks; ++bj) {
      for (uint32_t bi = 0; bi < blocks; ++bi) {
        const uint32_t i0 = bi * 8, j0 = bj * 8;
        const uint32_t i1 = std::min(i0 + 8, n - 1), j1 = std::min(j0 + 8, n - 1);
        ge_cluster_desc cl{};
        cl.vertex_offset = uint32_t(s.vertices.size() / 3);
        cl.triangle_offset = uint32_t(s.triangles.size());
        const uint32_t w = i1 - i0 + 1;
        F3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t j = j0; j <= j1; ++j) {
          for (uint32_t i = i0; i <= i1; ++i) {
            const F3 p{float(i) * cell, hgt[size_t(j) * n + i], float(j) * cell};
            const float hx = hgt[size_t(j) * n + std::min(i + 1, n - 1)] - hgt[size_t(j) * n + (i > 0 ? i - 1 : 0)];
            const float hz = hgt[size_t(std::min(j + 1, n - 1)) * n + i] - hgt[size_t(j > 0 ? j - 1 : 0) * n + i];
            const F3 nrm = norm3({-hx, 2.0f * cell, -hz});
            s.vertices.insert(s.vertices.end(), {p.x, p.y, p.z});
            s.attributes.push_back({encode_oct(nrm), 0u});
            lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
            hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
          }
        }
        for (uint32_t j = 0; j < j1 - j0; ++j) {
          for (uint32_t i = 0; i < i1 - i0; ++i) {
            const uint32_t a = j * w + i, b = a + 1, d = a + w, e = d + 1;
            s.triangles.push_back(a | (d << 8) | (e << 16));
            s.triangles.push_back(a | (e << 8) | (b << 16));
          }
        }
        cl.vertex_count = w * (j1 - j0 + 1);
        cl.triangle_count = 2 * (i1 - i0) * (j1 - j0);
        const F3 c = (lo + hi) * 0.5f;
        cl.center[0] = c.x;
        cl.center[1] = c.y;
        cl.center[2] = c.z;
        cl.radius = std::sqrt(dot3(hi - c, hi - c));
        cl.cone = 0x7f000000u;
        s.clusters.push_back(cl);
        s.cluster_materials.push_back(0);
      }
    }
    mesh.cluster_count = uint32_t(s.clusters.size()) - mesh.first_cluster;
    const uint32_t vend = uint32_t(s.vertices.size() / 3);
    F3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t v = vstart; v < vend; ++v) {
      lo = {std::min(lo.x, s.vertices[v * 3]), std::min(lo.y, s.vertices[v * 3 + 1]), std::min(lo.z, s.vertices[v * 3 + 2])};
      hi = {std::max(hi.x, s.vertices[v * 3]), std::max(hi.y, s.vertices[v * 3 + 1]), std::max(hi.z, s.vertices[v * 3 + 2])};
    }
    const float ext = std::max(std::max(hi.x - lo.x, hi.y - lo.y), hi.z - lo.z);
    const float scale = ext > 0.0f ? ext / 65535.0f : 1.0f;
    const float inv = 1.0f / scale;
    mesh.quant[0] = lo.x;
    mesh.quant[1] = lo.y;
    mesh.quant[2] = lo.z;
    mesh.quant[3] = scale;
    s.quantized.resize(size_t(vend) * 3);
    for (uint32_t v = vstart; v < vend; ++v) {
      const float axis[3] = {s.vertices[v * 3] - lo.x, s.vertices[v * 3 + 1] - lo.y, s.vertices[v * 3 + 2] - lo.z};
      for (int k = 0; k < 3; ++k) {
        const float grid = std::round(axis[k] * inv);
        const float clamped = grid < 0.0f ? 0.0f : (grid > 65535.0f ? 65535.0f : grid);
        s.quantized[size_t(v) * 3 + k] = uint16_t(uint32_t(clamped));
      }
    }
    s.meshes.push_back(mesh);
    s.materials.push_back(make_material(0.35f, 0.3f, 0.22f, 0.9f, 0.0f));
    const float identity[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    s.instances.push_back(make_instance(0, identity, {0, 0, 0}, 0, 0));
  }
  // The cluster groups: unique 1,024-triangle meshes (8 clusters of 128) standing on the terrain.
  uint32_t first_pair = s.meshes[0].cluster_count;
  for (uint32_t g = 0; g < groups; ++g) {
    Source src;
    torus(src, rng.uniform(2.0f, 6.0f), rng.uniform(0.6f, 1.8f), 32, 16, rng.uniform(0.05f, 0.4f),
          int(1 + rng.next() * 5), int(1 + rng.next() * 4));
    const uint32_t mesh = add_mesh(s, src, 128, std::vector<uint32_t>(src.idx.size() / 3, 0u));
    float rot[3][3];
    random_rotation(rng, rot);
    const float x = rng.uniform(0.02f * extent, 0.98f * extent), z = rng.uniform(0.02f * extent, 0.98f * extent);
    const uint32_t ci = std::min(uint32_t(x / cell), n - 1), cj = std::min(uint32_t(z / cell), n - 1);
    const float y = height * fbm(float(ci) / 256.0f, float(cj) / 256.0f, uint32_t(seed) ^ 0x51u) + rng.uniform(2.0f, 9.0f);
    s.instances.push_back(make_instance(mesh, rot, {x, y, z}, 1 + (g & 63u), first_pair));
    first_pair += s.meshes[mesh].cluster_count;
  }
  for (uint32_t k = 0; k < 64; ++k)
    s.materials.push_back(make_material(rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f), rng.uniform(0.2f, 1.0f), 0.0f));
  set_sun(s, 0.45f, 0.55f, 0.3f, 3.0f);  // a low sun: long shadows
  s.eye = {0.5f * extent, 0.16f * extent, -0.05f * extent};
  s.target = {0.5f * extent, 0.0f, 0.55f * extent};
  s.fov_deg = 55.0f;
  s.spp = 1;
  s.max_bounces = 1;
  s.reach = std::sqrt(2.0f) * extent + 300.0f;
}

// --- camera (core/math/math.h) ---------------------------------------------------------------
struct M4 {
  float c[4][4];  // columns
  float& at(int r, int col) { return c[col][r]; }
  float at(int r, int col) const { return c[col][r]; }
};
M4 identity4() {
  M4 m{};
  for (int i = 0; i < 4; ++i) m.c[i][i] = 1.0f;
  return m;
}
M4 mul(const M4& a, const M4& b) {
  M4 r{};
  for (int col = 0; col < 4; ++col)
    for (int row = 0; row < 4; ++row) {
      float v = 0;
      for (int k = 0; k < 4; ++k) v += a.at(row, k) * b.at(k, col);
      r.at(row, col) = v;
    }
  return r;
}
M4 look_at(F3 eye, F3 target, F3 up) {
  const F3 f = norm3(target - eye);
  const F3 s = norm3(cross3(f, up));
  const F3 u = cross3(s, f);
  M4 m = identity4();
  m.at(0, 0) = s.x;
  m.at(0, 1) = s.y;
  m.at(0, 2) = s.z;
  m.at(1, 0) = u.x;
  m.at(1, 1) = u.y;
  m.at(1, 2) = u.z;
  m.at(2, 0) = -f.x;
  m.at(2, 1) = -f.y;
  m.at(2, 2) = -f.z;
  m.at(0, 3) = -dot3(s, eye);
  m.at(1, 3) = -dot3(u, eye);
  m.at(2, 3) = dot3(f, eye);
  return m;
}
M4 perspective_reversed_z(float fov_y, float aspect, float near) {
  const float t = 1.0f / std::tan(fov_y * 0.5f);
  M4 m = identity4();
  m.at(0, 0) = t / aspect;
  m.at(1, 1) = t;
  m.at(2, 2) = 0.0f;
  m.at(2, 3) = near;
  m.at(3, 2) = -1.0f;
  m.at(3, 3) = 0.0f;
  return m;
}
M4 inverse4(const M4& m) {  // cofactor expansion (math.h:336-370)
  const float a00 = m.at(0, 0), a01 = m.at(0, 1), a02 = m.at(0, 2), a03 = m.at(0, 3);
  const float a10 = m.at(1, 0), a11 = m.at(1, 1), a12 = m.at(1, 2), a13 = m.at(1, 3);
  const float a20 = m.at(2, 0), a21 = m.at(2, 1), a22 = m.at(2, 2), a23 = m.at(2, 3);
  const float a30 = m.at(3, 0), a31 = m.at(3, 1), a32 = m.at(3, 2), a33 = m.at(3, 3);
  const float b00 = a00 * a11 - a01 * a10, b01 = a00 * a12 - a02 * a10, b02 = a00 * a13 - a03 * a10;
  const float b03 = a01 * a12 - a02 * a11, b04 = a01 * a13 - a03 * a11, b05 = a02 * a13 - a03 * a12;
  const float b06 = a20 * a31 - a21 * a30, b07 = a20 * a32 - a22 * a30, b08 = a20 * a33 - a23 * a30;
  const float b09 = a21 * a32 - a22 * a31, b10 = a21 * a33 - a23 * a31, b11 = a22 * a33 - a23 * a32;
  const float det = b00 * b11 - b01 * b10 + b02 * b09 + b03 * b08 - b04 * b07 + b05 * b06;
  if (det > -1.0e-12f && det < 1.0e-12f) return identity4();
  const float inv = 1.0f / det;
  M4 r;
  r.at(0, 0) = (a11 * b11 - a12 * b10 + a13 * b09) * inv;
  r.at(0, 1) = (-a01 * b11 + a02 * b10 - a03 * b09) * inv;
  r.at(0, 2) = (a31 * b05 - a32 * b04 + a33 * b03) * inv;
  r.at(0, 3) = (-a21 * b05 + a22 * b04 - a23 * b03) * inv;
  r.at(1, 0) = (-a10 * b11 + a12 * b08 - a13 * b07) * inv;
  r.at(1, 1) = (a00 * b11 - a02 * b08 + a03 * b07) * inv;
  r.at(1, 2) = (-a30 * b05 + a32 * b02 - a33 * b01) * inv;
  r.at(1, 3) = (a20 * b05 - a22 * b02 + a23 * b01) * inv;
  r.at(2, 0) = (a10 * b10 - a11 * b08 + a13 * b06) * inv;
  r.at(2, 1) = (-a00 * b10 + a01 * b08 - a03 * b06) * inv;
  r.at(2, 2) = (a30 * b04 - a31 * b02 + a33 * b00) * inv;
  r.at(2, 3) = (-a20 * b04 + a21 * b02 - a23 * b00) * inv;
  r.at(3, 0) = (-a10 * b09 + a11 * b07 - a12 * b06) * inv;
  r.at(3, 1) = (a00 * b09 - a01 * b07 + a02 * b06) * inv;
  r.at(3, 2) = (-a30 * b03 + a31 * b01 - a32 * b00) * inv;
  r.at(3, 3) = (a20 * b03 - a21 * b01 + a22 * b00) * inv;
  return r;
}
void store(const M4& m, float out[16]) {
  for (int col = 0; col < 4; ++col)
    for (int row = 0; row < 4; ++row) out[col * 4 + row] = m.at(row, col);
}
uint32_t pack_unorm_rgba8(const float c[3]) {  // path_trace.h:135-141
  auto q = [](float v) {
    const double clamped = v < 0.0f ? 0.0 : (v > 1.0f ? 1.0 : double(v));
    return uint32_t(clamped * 255.0 + 0.5);
  };
  return q(c[0]) | (q(c[1]) << 8) | (q(c[2]) << 16) | (255u << 24);
}

}  // namespace

extern "C" {

uint64_t ge_synth_splitmix64(uint64_t* state) { return splitmix64(*state); }

namespace {

// ---- config 5: an alpha-tested forest ------------------------------------------------------------
float linear_to_srgb(float c) {
  c = c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c);
  return c <= 0.0031308f ? 12.92f * c : 1.055f * std::pow(c, 1.0f / 2.4f) - 0.055f;
}
uint8_t unorm8(float c) { return uint8_t(std::lround((c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c)) * 255.0f)); }

uint32_t add_texture(ge_synth& s, uint32_t w, uint32_t h, uint32_t format, const std::vector<uint8_t>& rgba) {
  ge_texture_desc t{};
  t.width = w;
  t.height = h;
  t.format = format;
  t.offset = s.texels.size();
  s.texels.insert(s.texels.end(), rgba.begin(), rgba.end());
  s.textures.push_back(t);
  return uint32_t(s.textures.size() - 1);
}

// A 256 x 256 sRGB atlas of 4 x 4 leaves, each a lobed ellipse at its own angle in a 64-texel cell,
// alpha 1 inside and 0 outside with a two-texel ramp between.
std::vector<uint8_t> leaf_atlas(uint32_t seed) {
  std::vector<uint8_t> out(256 * 256 * 4);
  for (uint32_t cell = 0; cell < 16; ++cell) {
    Rng rng(seed * 131u + cell);
    const float angle = rng.uniform(-0.6f, 0.6f) + 0.785f;
    const float a = rng.uniform(24.0f, 29.0f), b = rng.uniform(9.0f, 14.0f);
    const float lobes = float(int(rng.uniform(3.0f, 7.0f))), depth = rng.uniform(0.05f, 0.2f);
    const float tint = rng.uniform(0.75f, 1.2f);
    const float cs = std::cos(angle), sn = std::sin(angle);
    for (uint32_t y = 0; y < 64; ++y)
      for (uint32_t x = 0; x < 64; ++x) {
        const float px = float(x) + 0.5f - 32.0f, py = float(y) + 0.5f - 32.0f;
        const float u = cs * px + sn * py, v = -sn * px + cs * py;
        const float theta = std::atan2(v, u);
        const float r = std::sqrt((u / a) * (u / a) + (v / b) * (v / b)) / (1.0f + depth * std::sin(lobes * theta));
        const float alpha = std::min(std::max((1.0f - r) * a * 0.5f, 0.0f), 1.0f);
        const float vein = std::fabs(v) < 1.0f ? 0.7f : 1.0f;
        const float lin[3] = {0.10f * tint * vein, 0.28f * tint * vein, 0.05f * tint};
        uint8_t* t = &out[((size_t(cell / 4) * 64 + y) * 256 + size_t(cell % 4) * 64 + x) * 4];
        for (int k = 0; k < 3; ++k) t[k] = unorm8(linear_to_srgb(lin[k]));
        t[3] = unorm8(alpha);
      }
  }
  return out;
}

// Bark: sRGB colour in vertical streaks and a UNORM normal map of the same grooves.
void bark_textures(uint32_t seed, std::vector<uint8_t>& colour, std::vector<uint8_t>& normal) {
  const uint32_t n = 128;
  colour.assign(n * n * 4, 255);
  normal.assign(n * n * 4, 255);
  auto height = [&](float x, float y) {
    return 0.5f + 0.35f * std::sin(x * 0.39269908f + 2.0f * value_noise(x * 0.1f, y * 0.05f, seed)) +
           0.15f * value_noise(x * 0.3f, y * 0.3f, seed + 7u);
  };
  for (uint32_t y = 0; y < n; ++y)
    for (uint32_t x = 0; x < n; ++x) {
      const float h = height(float(x), float(y));
      const float lin[3] = {0.22f * (0.55f + 0.45f * h), 0.14f * (0.55f + 0.45f * h), 0.08f * (0.55f + 0.45f * h)};
      uint8_t* c = &colour[(size_t(y) * n + x) * 4];
      for (int k = 0; k < 3; ++k) c[k] = unorm8(linear_to_srgb(lin[k]));
      const float dx = (height(float(x) + 1.0f, float(y)) - height(float(x) - 1.0f, float(y))) * 2.0f;
      const float dy = (height(float(x), float(y) + 1.0f) - height(float(x), float(y) - 1.0f)) * 2.0f;
      const F3 nv = norm3({-dx, -dy, 1.0f});
      uint8_t* m = &normal[(size_t(y) * n + x) * 4];
      m[0] = unorm8(nv.x * 0.5f + 0.5f);
      m[1] = unorm8(nv.y * 0.5f + 0.5f);
      m[2] = unorm8(nv.z * 0.5f + 0.5f);
    }
}

// `trees` trees of 8 meshes (an 8-sided trunk of 96 triangles and 320 to 480 alpha-tested leaf
// cards of two triangles in a crown) on a jittered 6 m grid over gently rolling ground, under the
// sun and sky. Textures: a leaf atlas (alpha-tested through a UV transform onto one of six cells),
// bark colour with a normal map, ground colour with a metallic-roughness map.
void build_config5(ge_synth& s, uint64_t seed, uint32_t trees) {
  Rng rng(seed ^ 0xf0e57u);
  const uint32_t leaf_tex = add_texture(s, 256, 256, GE_FORMAT_R8G8B8A8_SRGB, leaf_atlas(uint32_t(seed)));
  std::vector<uint8_t> bark_c, bark_n;
  bark_textures(uint32_t(seed) + 3u, bark_c, bark_n);
  const uint32_t bark_tex = add_texture(s, 128, 128, GE_FORMAT_R8G8B8A8_SRGB, bark_c);
  const uint32_t bark_nrm = add_texture(s, 128, 128, GE_FORMAT_R8G8B8A8_UNORM, bark_n);
  std::vector<uint8_t> ground_c(256 * 256 * 4, 255), ground_mr(64 * 64 * 4, 255);
  for (uint32_t y = 0; y < 256; ++y)
    for (uint32_t x = 0; x < 256; ++x) {
      const float t = value_noise(float(x) * 0.06f, float(y) * 0.06f, uint32_t(seed) + 11u);
      const float lin[3] = {0.08f + 0.12f * t, 0.16f + 0.02f * t, 0.04f + 0.05f * t};
      for (int k = 0; k < 3; ++k) ground_c[(size_t(y) * 256 + x) * 4 + k] = unorm8(linear_to_srgb(lin[k]));
    }
  for (uint32_t y = 0; y < 64; ++y)
    for (uint32_t x = 0; x < 64; ++x) {
      uint8_t* m = &ground_mr[(size_t(y) * 64 + x) * 4];
      m[1] = unorm8(0.7f + 0.3f * value_noise(float(x) * 0.2f, float(y) * 0.2f, uint32_t(seed) + 13u));
      m[2] = 0;
    }
  const uint32_t ground_tex = add_texture(s, 256, 256, GE_FORMAT_R8G8B8A8_SRGB, ground_c);
  const uint32_t ground_mrt = add_texture(s, 64, 64, GE_FORMAT_R8G8B8A8_UNORM, ground_mr);
  s.samplers.push_back({GE_ADDRESS_REPEAT, GE_ADDRESS_REPEAT, GE_FILTER_LINEAR, GE_FILTER_LINEAR});
  s.samplers.push_back({GE_ADDRESS_CLAMP_TO_EDGE, GE_ADDRESS_CLAMP_TO_EDGE, GE_FILTER_LINEAR, GE_FILTER_LINEAR});
  s.samplers.push_back({GE_ADDRESS_MIRRORED_REPEAT, GE_ADDRESS_MIRRORED_REPEAT, GE_FILTER_NEAREST, GE_FILTER_NEAREST});

  ge_resolve_material ground = make_material(1.0f, 1.0f, 1.0f, 1.0f, 0.0f);
  ground.albedo_texture = ground_tex;
  ground.sampler = 0;
  ground.uv_scale = 0.25f;
  ground.metallic_roughness_texture = ground_mrt;
  ground.samplers_mr_normal = 0xffff0002u;  // the map through the nearest, mirrored sampler
  s.materials.push_back(ground);
  ge_resolve_material bark = make_material(1.0f, 1.0f, 1.0f, 0.85f, 0.0f);
  bark.albedo_texture = bark_tex;
  bark.sampler = 0;
  bark.normal_texture = bark_nrm;
  s.materials.push_back(bark);
  for (uint32_t k = 0; k < 6; ++k) {
    ge_resolve_material leaf = make_material(1.0f, 1.0f, 1.0f, 0.6f, 0.0f);
    leaf.flags = GE_MATERIAL_MASKED | GE_MATERIAL_UV_TRANSFORM;
    leaf.albedo_texture = leaf_tex;
    leaf.sampler = 1;
    leaf.uv_transform[0] = 0.25f;
    leaf.uv_transform[3] = 0.25f;
    leaf.uv_offset[0] = 0.25f * float(k % 4);
    leaf.uv_offset[1] = 0.25f * float(k / 4);
    leaf.uv_offset[2] = 0.5f;  // the cutoff
    leaf.uv_offset[3] = 1.0f;  // the base colour factor's alpha
    s.materials.push_back(leaf);
  }

  for (uint32_t m = 0; m < 8; ++m) {
    Source src;
    std::vector<uint32_t> mats;
    const float H = rng.uniform(8.0f, 14.0f), r0 = rng.uniform(0.25f, 0.4f);
    const int sides = 8, rings = 6;
    for (int j = 0; j <= rings; ++j) {
      const float y = 0.55f * H * float(j) / float(rings);
      const float r = r0 + (0.08f - r0) * float(j) / float(rings);
      for (int i = 0; i <= sides; ++i) {
        const float a = 6.2831853f * float(i) / float(sides);
        src.pos.push_back({r * std::cos(a), y, r * std::sin(a)});
        src.nrm.push_back({std::cos(a), 0.0f, std::sin(a)});
        src.uv.push_back(2.0f * float(i) / float(sides));
        src.uv.push_back(y * 0.5f);
      }
    }
    for (int j = 0; j < rings; ++j)
      for (int i = 0; i < sides; ++i) {
        const uint32_t a = uint32_t(j * (sides + 1) + i), b = a + 1, c = a + uint32_t(sides + 1), d = c + 1;
        src.idx.insert(src.idx.end(), {a, c, b, b, c, d});
        mats.insert(mats.end(), {1u, 1u});
      }
    const uint32_t cards = 320 + uint32_t(rng.next() * 160.0);
    for (uint32_t k = 0; k < cards; ++k) {
      F3 c;
      do {
        c = {rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1)};
      } while (dot3(c, c) > 1.0f);
      c = {c.x * 0.28f * H, 0.68f * H + c.y * 0.3f * H, c.z * 0.28f * H};
      float rot[3][3];
      random_rotation(rng, rot);
      const float half = 0.5f * rng.uniform(0.35f, 0.7f);
      const F3 ax{rot[0][0], rot[1][0], rot[2][0]}, ay{rot[0][1], rot[1][1], rot[2][1]}, an{rot[0][2], rot[1][2], rot[2][2]};
      const uint32_t base = uint32_t(src.pos.size());
      const float corner[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
      for (int q = 0; q < 4; ++q) {
        src.pos.push_back(c + ax * (half * corner[q][0]) + ay * (half * corner[q][1]));
        src.nrm.push_back(an);
        src.uv.push_back(0.5f + 0.5f * corner[q][0]);
        src.uv.push_back(0.5f - 0.5f * corner[q][1]);
      }
      src.idx.insert(src.idx.end(), {base, base + 1, base + 2, base, base + 2, base + 3});
      const uint32_t leaf = 2u + (k / 64u) % 6u;  // 64 cards (one cluster) a leaf material
      mats.insert(mats.end(), {leaf, leaf});
    }
    add_mesh(s, src, 128, mats, true);
  }

  const uint32_t per_side = uint32_t(std::ceil(std::sqrt(double(std::max(trees, 1u)))));
  const float spacing = 6.0f, side = spacing * float(per_side);
  auto terrain = [&](float x, float z) { return 1.5f * value_noise(x / 40.0f, z / 40.0f, uint32_t(seed) + 21u); };
  {
    Source g;
    const int n = 128;
    const float lo = -8.0f, step = (side + 16.0f) / float(n);
    for (int j = 0; j <= n; ++j)
      for (int i = 0; i <= n; ++i) {
        const float x = lo + step * float(i), z = lo + step * float(j);
        g.pos.push_back({x, terrain(x, z), z});
        const float dx = terrain(x + 0.5f, z) - terrain(x - 0.5f, z), dz = terrain(x, z + 0.5f) - terrain(x, z - 0.5f);
        g.nrm.push_back(norm3({-dx, 1.0f, -dz}));
        g.uv.push_back(x);
        g.uv.push_back(z);
      }
    std::vector<uint32_t> gm;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const uint32_t a = uint32_t(j * (n + 1) + i), b = a + 1, c = a + uint32_t(n + 1), d = c + 1;
        g.idx.insert(g.idx.end(), {a, c, b, b, c, d});
        gm.insert(gm.end(), {0u, 0u});
      }
    add_mesh(s, g, 128, gm, true);
  }
  uint32_t first_pair = 0;
  const float identity[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  s.instances.push_back(make_instance(8, identity, {0, 0, 0}, 0, first_pair));
  first_pair += s.meshes[8].cluster_count;
  for (uint32_t t = 0; t < trees; ++t) {
    const float x = spacing * (float(t % per_side) + 0.5f) + rng.uniform(-2.0f, 2.0f);
    const float z = spacing * (float(t / per_side) + 0.5f) + rng.uniform(-2.0f, 2.0f);
    const float yaw = rng.uniform(0.0f, 6.2831853f), scale = rng.uniform(0.8f, 1.25f);
    const float rot[3][3] = {{scale * std::cos(yaw), 0, scale * std::sin(yaw)},
                             {0, scale, 0},
                             {-scale * std::sin(yaw), 0, scale * std::cos(yaw)}};
    const uint32_t mesh = uint32_t(rng.next() * 8.0) & 7u;
    ge_instance_desc in = make_instance(mesh, rot, {x, terrain(x, z) - 0.1f, z}, 0, first_pair);
    in.scale_max = scale;
    s.instances.push_back(in);
    first_pair += s.meshes[mesh].cluster_count;
  }
  set_sun(s, -0.45f, 0.55f, 0.5f, 3.0f);
  s.sky[0] = 0.45f;
  s.sky[1] = 0.62f;
  s.sky[2] = 0.9f;
  s.ground[0] = 0.12f;
  s.ground[1] = 0.16f;
  s.ground[2] = 0.07f;
  s.eye = {0.5f * side, 32.0f, -12.0f};
  s.target = {0.5f * side, 4.0f, 0.3f * side};
  s.fov_deg = 55.0f;
  s.spp = 1;
  s.max_bounces = 4;
  s.reach = 1.5f * (side + 40.0f);
}

}  // namespace

ge_synth* ge_synth_create(uint32_t config, uint64_t seed, uint32_t detail, uint32_t terrain) {
  auto* s = new ge_synth();
  s->config = config;
  switch (config) {
    case 1: build_config1(*s, seed); break;
    case 2: build_config2(*s, seed, detail ? detail : 4096u); break;
    case 4: build_config4(*s, seed); break;
    case 3: build_config3(*s, seed, detail ? detail : 16384u, terrain ? std::max(terrain, 2u) : 4096u); break;
    case 5: build_config5(*s, seed, detail ? detail : 4096u); break;
    default: delete s; return nullptr;
  }
  if (s->quantized.size() & 1u) s->quantized.push_back(0);  // padded to an even count (cluster.h:436)
  return s;
}

ge_synth* ge_synth_from_triangles(const float* positions, uint32_t triangle_count, uint32_t cluster_size,
                                  int32_t attributes, uint64_t seed, const float* uvs) {
  if (!positions || triangle_count == 0 || cluster_size == 0 || cluster_size > 128) return nullptr;
  auto* s = new ge_synth();
  s->config = 0;
  Source src;
  for (uint32_t t = 0; t < triangle_count; ++t) {
    const F3 a{positions[t * 9], positions[t * 9 + 1], positions[t * 9 + 2]};
    const F3 b{positions[t * 9 + 3], positions[t * 9 + 4], positions[t * 9 + 5]};
    const F3 c{positions[t * 9 + 6], positions[t * 9 + 7], positions[t * 9 + 8]};
    const F3 n = norm3(cross3(b - a, c - a));
    for (const F3& p : {a, b, c}) {
      src.pos.push_back(p);
      src.nrm.push_back(n);
      src.idx.push_back(uint32_t(src.idx.size()));
    }
    if (uvs) src.uv.insert(src.uv.end(), uvs + size_t(t) * 6, uvs + size_t(t) * 6 + 6);
  }
  std::vector<uint32_t> mats(triangle_count, 0u);
  add_mesh(*s, src, cluster_size, mats);
  for (uint32_t c = 0; c < s->clusters.size(); ++c) s->cluster_materials[c] = c;  // a material per cluster
  s->has_attributes = attributes != 0;
  if (!s->has_attributes) s->attributes.clear();
  Rng rng(seed);
  for (uint32_t c = 0; c < s->clusters.size(); ++c)
    s->materials.push_back(make_material(rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f), 1.0f, 0.0f));
  const float identity[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  s->instances.push_back(make_instance(0, identity, {0, 0, 0}, 0, 0));
  set_sun(*s, 0.3f, 0.8f, 0.5f, 3.0f);
  F3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
  for (const F3& p : src.pos) {
    lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
    hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
  }
  const F3 c = (lo + hi) * 0.5f;
  const float r = std::max(1e-3f, std::sqrt(dot3(hi - c, hi - c)));
  s->target = c;
  s->eye = c + F3{0.35f, 0.45f, -1.0f} * (2.2f * r);
  s->reach = std::sqrt(dot3(hi, hi)) + r;
  if (s->quantized.size() & 1u) s->quantized.push_back(0);
  return s;
}

void ge_synth_destroy(ge_synth* s) { delete s; }

void ge_synth_scene(const ge_synth* s, ge_scene_desc* out) {
  std::memset(out, 0, sizeof(*out));
  out->clusters = s->clusters.data();
  out->cluster_count = uint32_t(s->clusters.size());
  out->vertex_count = uint32_t(s->vertices.size() / 3);
  out->vertices = s->vertices.data();
  out->quantized = s->quantized.data();
  out->attributes = s->has_attributes ? s->attributes.data() : nullptr;
  out->triangles = s->triangles.data();
  out->triangle_count = uint32_t(s->triangles.size());
  out->mesh_count = uint32_t(s->meshes.size());
  out->meshes = s->meshes.data();
  out->instances = s->instances.data();
  out->instance_count = uint32_t(s->instances.size());
  out->material_count = uint32_t(s->materials.size());
  out->materials = s->materials.data();
  out->cluster_materials = s->cluster_materials.data();
  out->lights = s->lights.empty() ? nullptr : s->lights.data();
  out->light_count = uint32_t(s->lights.size());
  out->texture_count = uint32_t(s->textures.size());
  out->textures = s->textures.empty() ? nullptr : s->textures.data();
  out->texels = s->texels.empty() ? nullptr : s->texels.data();
  out->texel_bytes = s->texels.size();
  out->samplers = s->samplers.empty() ? nullptr : s->samplers.data();
  out->sampler_count = uint32_t(s->samplers.size());
}

void ge_synth_set_materials(ge_synth* s, const ge_resolve_material* materials, uint32_t count,
                            const uint32_t* cluster_materials) {
  s->materials.assign(materials, materials + count);
  if (cluster_materials) s->cluster_materials.assign(cluster_materials, cluster_materials + s->clusters.size());
}

void ge_synth_set_textures(ge_synth* s, const ge_texture_desc* textures, uint32_t texture_count, const uint8_t* texels,
                           uint64_t texel_bytes, const ge_sampler_desc* samplers, uint32_t sampler_count) {
  s->textures.assign(textures, textures + texture_count);
  s->texels.assign(texels, texels + texel_bytes);
  s->samplers.assign(samplers, samplers + sampler_count);
}

void ge_synth_frame(const ge_synth* s, uint32_t width, uint32_t height, ge_path_trace_params* pt,
                    ge_ray_visibility_params* rv) {
  const M4 view = look_at(s->eye, s->target, {0, 1, 0});
  const M4 proj = perspective_reversed_z(s->fov_deg * 0.017453292519943295f, float(width) / float(height), 0.1f);
  M4 rotation = view;
  rotation.at(0, 3) = 0.0f;
  rotation.at(1, 3) = 0.0f;
  rotation.at(2, 3) = 0.0f;
  const M4 clip_to_ray = inverse4(mul(proj, rotation));
  const M4 view_proj = mul(proj, view);
  if (pt) {
    std::memset(pt, 0, sizeof(*pt));
    store(clip_to_ray, pt->clip_to_ray);
    pt->camera[0] = s->eye.x;
    pt->camera[1] = s->eye.y;
    pt->camera[2] = s->eye.z;
    std::memcpy(pt->sky, s->sky, 12);
    pt->sky[3] = 1.0f;
    std::memcpy(pt->sun, s->sun, 16);
    std::memcpy(pt->ground, s->ground, 12);
    pt->width = width;
    pt->height = height;
    pt->light_count = uint32_t(s->lights.size());
    pt->spp = s->spp;
    pt->sample_base = 0;
    pt->max_bounces = s->max_bounces;
    pt->seed = 0;
    pt->flags = 0;
    pt->ray_bias = s->reach * (1.0f / 262144.0f);  // k_shadow_bias_relative (lighting.h:141)
    pt->ray_bias_steps = 1.0f;                     // k_shadow_bias_steps (lighting.h:140)
    pt->background = pack_unorm_rgba8(s->sky);
  }
  if (rv) {
    std::memset(rv, 0, sizeof(*rv));
    store(view_proj, rv->view_proj);
    store(clip_to_ray, rv->clip_to_ray);
    rv->camera[0] = s->eye.x;
    rv->camera[1] = s->eye.y;
    rv->camera[2] = s->eye.z;
    rv->width = width;
    rv->height = height;
    rv->caster_base = 0xffffffffu;
  }
}

void ge_synth_set_camera(ge_synth* s, const float eye[3], const float target[3], float fov_y_deg) {
  s->eye = {eye[0], eye[1], eye[2]};
  s->target = {target[0], target[1], target[2]};
  s->fov_deg = fov_y_deg;
}

void ge_synth_move_instances(ge_synth* s, uint32_t first, uint32_t count, uint64_t seed, float step, float angle,
                             ge_instance_desc* out) {
  Rng rng(seed);
  for (uint32_t i = first; i < first + count && i < s->instances.size(); ++i) {
    ge_instance_desc& in = s->instances[i];
    const float a = angle * float(2.0 * rng.next() - 1.0);
    const float c = std::cos(a), sn = std::sin(a);
    const float spin[3][3] = {{c, 0, sn}, {0, 1, 0}, {-sn, 0, c}};  // about y
    float rows[3][3];
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) rows[r][k] = spin[r][0] * in.rows[0][k] + spin[r][1] * in.rows[1][k] + spin[r][2] * in.rows[2][k];
    for (int r = 0; r < 3; ++r) {
      for (int k = 0; k < 3; ++k) in.rows[r][k] = rows[r][k];
      in.rows[r][3] += step * float(2.0 * rng.next() - 1.0);
    }
    out[i - first] = in;
  }
}

void ge_synth_primary_rays(const ge_synth* s, uint32_t width, uint32_t height, uint32_t seed, ge_ray* rays) {
  ge_path_trace_params pt;
  ge_synth_frame(s, width, height, &pt, nullptr);
  const float* m = pt.clip_to_ray;
  for (uint32_t y = 0; y < height; ++y) {
    for (uint32_t x = 0; x < width; ++x) {
      const uint32_t h = pcg_hash(x * 73856093u ^ y * 19349663u ^ seed * 83492791u);
      const float jx = float(h & 0xffffu) * (1.0f / 65536.0f), jy = float(h >> 16) * (1.0f / 65536.0f);
      const float ndx = (float(x) + jx) / float(width) * 2.0f - 1.0f;
      const float ndy = 1.0f - (float(y) + jy) / float(height) * 2.0f;
      float d[3];
      for (int r = 0; r < 3; ++r) d[r] = ((m[r] * ndx + m[4 + r] * ndy) + m[8 + r] * 0.0f) + m[12 + r];
      const float inv = 1.0f / std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      ge_ray& ray = rays[size_t(y) * width + x];
      for (int r = 0; r < 3; ++r) {
        ray.origin[r] = pt.camera[r];
        ray.dir[r] = d[r] * inv;
      }
      ray.tmin = 0.0f;
      ray.tmax = 1e30f;
    }
  }
}

}  // extern "C"
