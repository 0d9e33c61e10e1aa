// rt_clusters.hip — the `.clusters` container reader: a built mesh from the engine's derived-data
// cache as the arrays ge_rt_scene_upload takes. Host code only.
//
// The container (domain/geometry/include/domain/geometry/cluster_file.h:14-30, 295-379): a 32-byte
// header (magic "CLST", version 1, flags, section count, total bytes, the content hash of every
// byte after the header), 24-byte section records {kind, element size, element count, offset},
// then the payloads at 16-byte aligned offsets. The checks follow the engine's reader
// (check_container, read_memory and check_cluster_ranges, domain/geometry/src/cluster_file.cpp:
// 419-515, 691-800, 1184-1215): structure before contents, every section inside the file, the
// hash, the required sections at their element sizes, every cluster's runs inside the streams and
// every triangle's corners inside its cluster, the quantized stream's length and its grid finite.
// Unknown kinds are skipped, as the format asks.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/ge_rt.h"

namespace {

// hash_bytes (core/hash/include/core/hash/hash.h:26-76): a word-at-a-time multiplicative hash
// seeded with the golden ratio, MurmurHash3's 64-bit finalizer at both ends, the tail folded with
// its length in the top byte.
constexpr uint64_t k_hash_seed = 0x9E3779B97F4A7C15ull;
uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}
uint64_t mix_word(uint64_t h, uint64_t word) {
  h ^= word;
  h *= 0x9FB21C651E98DF25ull;
  h ^= h >> 32;
  return h;
}
uint64_t hash_bytes(const uint8_t* p, uint64_t len) {
  uint64_t h = k_hash_seed ^ mix64(len);
  while (len >= 8) {
    uint64_t word;
    std::memcpy(&word, p, 8);
    h = mix_word(h, word);
    p += 8;
    len -= 8;
  }
  if (len > 0) {
    uint64_t word = 0;
    std::memcpy(&word, p, size_t(len));
    h = mix_word(h, word ^ (len << 56));
  }
  return mix64(h);
}

struct Header {
  char magic[4];
  uint32_t version, flags, section_count;
  uint64_t total_bytes, content_hash;
};
struct Section {
  uint32_t kind, element_size;
  uint64_t element_count, offset;
};
static_assert(sizeof(Header) == 32 && sizeof(Section) == 24, "the container's records");

// Section kinds (cluster_file.h:49-78): the ones this reader takes.
enum : uint32_t {
  kClusters = 1,
  kLod = 2,
  kVertices = 3,
  kAttributes = 4,
  kTriangles = 5,
  kVertexSource = 6,
  kLevelClusterCounts = 7,
  kClusterMaterial = 8,
  kMaterials = 9,
  kScalars = 12,
  kQuantized = 13,
  kKinds = 64,
};

// geometry::ClusterLodDesc (cluster_lod.h:26-33), ClusterFileScalars (cluster_file.h:313-319) and
// ClusterFileMaterial (cluster_file.h:367-379).
struct LodDesc {
  float own[4], parent[4];
  float own_error, parent_error;
  uint32_t level, group;
};
struct Scalars {
  uint32_t group_count, leaf_triangle_count, source_vertex_count, source_triangle_count;
  uint32_t pad[4];  // the quantization origin and step, as float bits
};
struct FileMaterial {
  float base_color[4];
  float metallic, roughness;
  int32_t base_color_image, normal_image;
  uint32_t metallic_roughness_image, occlusion_image, emissive_image;  // one-based, 0: none
  float emissive[3];
  float normal_scale;
  uint32_t alpha;  // mode in bits 0-1 (1: mask), double sided in bit 2, the cutoff unorm16 in 16-31
};
static_assert(sizeof(LodDesc) == 48 && sizeof(Scalars) == 32 && sizeof(FileMaterial) == 64, "records");

}  // namespace

struct ge_clusters {
  std::vector<ge_cluster_desc> clusters;  // the finest level's, in file order
  std::vector<float> vertices;
  std::vector<uint16_t> quantized;
  std::vector<ge_vertex_attributes> attributes;
  std::vector<uint32_t> triangles;
  std::vector<uint32_t> cluster_materials;
  std::vector<ge_resolve_material> materials;
  ge_mesh_desc mesh{};
  ge_instance_desc instance{};
  uint32_t all_clusters = 0, levels = 0;
};

namespace {

int32_t fail(char* error, size_t size, const std::string& msg) {
  if (error && size) {
    const size_t n = std::min(size - 1, msg.size());
    std::memcpy(error, msg.data(), n);
    error[n] = '\0';
  }
  return GE_RT_ERROR_INVALID_SCENE;
}

template <class T>
void copy_section(const uint8_t* bytes, const Section& s, std::vector<T>& out) {
  out.resize(size_t(s.element_count));
  if (s.element_count) std::memcpy(out.data(), bytes + s.offset, size_t(s.element_count) * sizeof(T));
}

// GpuScene::fill_material's factors (systems/renderer/src/gpu_scene_textures.cpp:477-540): albedo
// and roughness, the emissive factor and metallic, the normal scale, and the alpha mask packed as
// set_alpha_mask packs it (visibility_resolve.h:634-641). The images stay the caller's: with
// GE_CLUSTERS_IMAGES each slot names its image's index in the container's image list, otherwise
// none.
ge_resolve_material resolve_material(const FileMaterial& f, uint32_t flags) {
  ge_resolve_material m{};
  m.albedo[0] = f.base_color[0];
  m.albedo[1] = f.base_color[1];
  m.albedo[2] = f.base_color[2];
  m.albedo[3] = f.roughness;
  m.emissive[0] = f.emissive[0];
  m.emissive[1] = f.emissive[1];
  m.emissive[2] = f.emissive[2];
  m.emissive[3] = f.metallic;
  const bool images = (flags & GE_CLUSTERS_IMAGES) != 0;
  auto optional = [&](uint32_t slot) { return !images || slot == 0 ? GE_NO_TEXTURE : slot - 1u; };
  m.albedo_texture = images && f.base_color_image >= 0 ? uint32_t(f.base_color_image) : GE_NO_TEXTURE;
  m.normal_texture = images && f.normal_image >= 0 ? uint32_t(f.normal_image) : GE_NO_TEXTURE;
  m.metallic_roughness_texture = optional(f.metallic_roughness_image);
  m.occlusion_texture = optional(f.occlusion_image);
  m.emissive_texture = optional(f.emissive_image);
  m.sampler = 0;
  m.uv_scale = 1.0f;
  m.normal_scale = f.normal_scale;
  m.occlusion_strength = 1.0f;
  m.samplers_mr_normal = 0xffffffffu;
  m.samplers_occlusion_emissive = 0xffffffffu;
  m.uv_transform[0] = 1.0f;
  m.uv_transform[3] = 1.0f;
  if ((f.alpha & 3u) == 1u) {
    m.flags |= GE_MATERIAL_MASKED;
    m.uv_offset[2] = float(f.alpha >> 16) / 65535.0f;
    m.uv_offset[3] = f.base_color[3];
  }
  return m;
}

// The reader proper; ge_clusters_read turns a host allocation failure anywhere in it into
// GE_RT_ERROR_OUT_OF_MEMORY, and `c` frees a half-built result on every early return.
int32_t read_clusters(const uint8_t* bytes, uint64_t size, uint32_t flags, ge_clusters** out, char* error,
                      size_t error_size) {
  if (size < sizeof(Header))
    return fail(error, error_size, "cluster file is truncated: " + std::to_string(size) +
                                       " bytes, the 32-byte header does not fit");
  Header h;
  std::memcpy(&h, bytes, sizeof(h));
  if (std::memcmp(h.magic, "CLST", 4) != 0) return fail(error, error_size, "not a cluster file: the magic is not \"CLST\"");
  if (h.version != 1) return fail(error, error_size, "unsupported cluster file version " + std::to_string(h.version));
  if (h.total_bytes < sizeof(Header) || size < h.total_bytes)
    return fail(error, error_size, "cluster file is truncated: the header says " + std::to_string(h.total_bytes) + " bytes");
  if (sizeof(Header) + sizeof(Section) * uint64_t(h.section_count) > h.total_bytes)
    return fail(error, error_size, "cluster file is truncated: the section table does not fit");
  std::vector<Section> sections(h.section_count);
  if (h.section_count) std::memcpy(sections.data(), bytes + sizeof(Header), sizeof(Section) * h.section_count);
  for (const Section& s : sections) {
    if (s.element_count > 0xffffffffull)
      return fail(error, error_size, "cluster file section " + std::to_string(s.kind) + " holds too many elements");
    const uint64_t span = uint64_t(s.element_size) * s.element_count;
    if (s.offset < sizeof(Header) || s.offset > h.total_bytes || span > h.total_bytes - s.offset)
      return fail(error, error_size, "cluster file section " + std::to_string(s.kind) + " extends past the end of the file");
  }
  if (hash_bytes(bytes + sizeof(Header), h.total_bytes - sizeof(Header)) != h.content_hash)
    return fail(error, error_size, "cluster file content hash mismatch");
  const Section* found[kKinds] = {};
  for (const Section& s : sections)
    if (s.kind != 0 && s.kind < kKinds && !found[s.kind]) found[s.kind] = &s;
  std::string missing;
  auto require = [&](uint32_t kind, uint32_t element_size, const char* name) -> const Section* {
    const Section* s = found[kind];
    if (!s) {
      missing = std::string("cluster file has no ") + name + " section";
      return nullptr;
    }
    if (s->element_size != element_size) {
      missing = std::string("cluster file section ") + name + " has " + std::to_string(s->element_size) +
                "-byte elements, expected " + std::to_string(element_size);
      return nullptr;
    }
    return s;
  };
  const Section* s_clusters = require(kClusters, sizeof(ge_cluster_desc), "clusters");
  const Section* s_lod = s_clusters ? require(kLod, sizeof(LodDesc), "lod") : nullptr;
  const Section* s_vertices = s_lod ? require(kVertices, 12, "vertices") : nullptr;
  const Section* s_attributes = s_vertices ? require(kAttributes, sizeof(ge_vertex_attributes), "attributes") : nullptr;
  const Section* s_triangles = s_attributes ? require(kTriangles, 4, "triangles") : nullptr;
  const Section* s_source = s_triangles ? require(kVertexSource, 4, "vertex_source") : nullptr;
  const Section* s_levels = s_source ? require(kLevelClusterCounts, 4, "level_cluster_counts") : nullptr;
  const Section* s_scalars = s_levels ? require(kScalars, sizeof(Scalars), "scalars") : nullptr;
  if (!s_scalars) return fail(error, error_size, missing);
  if (s_scalars->element_count != 1) return fail(error, error_size, "cluster file scalars section holds other than one record");

  std::unique_ptr<ge_clusters> c(new (std::nothrow) ge_clusters());
  if (!c) return GE_RT_ERROR_OUT_OF_MEMORY;
  std::vector<ge_cluster_desc> all;
  std::vector<LodDesc> lod;
  copy_section(bytes, *s_clusters, all);
  copy_section(bytes, *s_lod, lod);
  c->vertices.resize(size_t(s_vertices->element_count) * 3);
  if (s_vertices->element_count) std::memcpy(c->vertices.data(), bytes + s_vertices->offset, c->vertices.size() * 4);
  copy_section(bytes, *s_attributes, c->attributes);
  copy_section(bytes, *s_triangles, c->triangles);
  Scalars sc;
  std::memcpy(&sc, bytes + s_scalars->offset, sizeof(sc));
  const uint64_t nv = s_vertices->element_count, nt = s_triangles->element_count;
  auto refuse = [&](const std::string& msg) { return fail(error, error_size, msg); };
  if (lod.size() != all.size()) return refuse("cluster file has " + std::to_string(lod.size()) + " lod records for " +
                                              std::to_string(all.size()) + " clusters");
  if (c->attributes.size() != nv) return refuse("cluster file attributes are not parallel to its vertices");
  // check_cluster_ranges (cluster_file.cpp:419-457).
  for (size_t k = 0; k < all.size(); ++k) {
    const ge_cluster_desc& cl = all[k];
    if (uint64_t(cl.vertex_offset) + cl.vertex_count > nv)
      return refuse("cluster file cluster " + std::to_string(k) + " names vertices past the stream");
    if (uint64_t(cl.triangle_offset) + cl.triangle_count > nt)
      return refuse("cluster file cluster " + std::to_string(k) + " names triangles past the stream");
    for (uint32_t t = 0; t < cl.triangle_count; ++t) {
      const uint32_t p = c->triangles[cl.triangle_offset + t];
      for (int j = 0; j < 3; ++j)
        if (((p >> (8 * j)) & 0xffu) >= cl.vertex_count)
          return refuse("cluster file cluster " + std::to_string(k) + " triangle " + std::to_string(t) +
                        " names a vertex outside the cluster");
    }
  }
  // The grid: the file's own, or quantize_positions' (domain/geometry/src/cluster.cpp:437-463).
  float origin[3] = {0.0f, 0.0f, 0.0f}, step = 1.0f;
  if (const Section* q = found[kQuantized]) {
    if (q->element_size != 2) return refuse("cluster file section quantized does not hold 2-byte coordinates");
    if (q->element_count != nv * 3 && q->element_count != nv * 3 + 1)
      return refuse("cluster file has " + std::to_string(q->element_count) + " quantized coordinates for " +
                    std::to_string(nv) + " vertices");
    copy_section(bytes, *q, c->quantized);
    std::memcpy(origin, sc.pad, 12);
    std::memcpy(&step, &sc.pad[3], 4);
  } else if (nv) {
    float lo[3] = {c->vertices[0], c->vertices[1], c->vertices[2]}, hi[3] = {lo[0], lo[1], lo[2]};
    for (uint64_t i = 1; i < nv; ++i)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], c->vertices[i * 3 + a]);
        hi[a] = std::max(hi[a], c->vertices[i * 3 + a]);
      }
    const float extent = std::max(std::max(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
    step = extent > 0.0f ? extent / 65535.0f : 1.0f;
    const float inverse = 1.0f / step;
    for (int a = 0; a < 3; ++a) origin[a] = lo[a];
    c->quantized.reserve(size_t(nv) * 3 + 1);
    for (uint64_t i = 0; i < nv; ++i)
      for (int a = 0; a < 3; ++a) {
        const float grid = std::round((c->vertices[i * 3 + a] - lo[a]) * inverse);
        const float clamped = grid < 0.0f ? 0.0f : (grid > 65535.0f ? 65535.0f : grid);
        c->quantized.push_back(uint16_t(uint32_t(clamped)));
      }
  }
  if (c->quantized.size() & 1u) c->quantized.push_back(0);
  const float extent = 65535.0f * step;
  if (!std::isfinite(origin[0]) || !std::isfinite(origin[1]) || !std::isfinite(origin[2]) || !std::isfinite(step) ||
      !(step >= 0.0f) || !std::isfinite(origin[0] + extent) || !std::isfinite(origin[1] + extent) ||
      !std::isfinite(origin[2] + extent))
    return refuse("cluster file's quantization grid is not finite");
  // Materials and the material map: optional (cluster_file.cpp:1184-1215). A mesh with none takes
  // one white, rough, dielectric material.
  std::vector<FileMaterial> fm;
  std::vector<uint32_t> map;
  if (const Section* s = found[kMaterials]) {
    if (s->element_size != sizeof(FileMaterial)) return refuse("cluster file section materials does not hold 64-byte records");
    copy_section(bytes, *s, fm);
  }
  if (const Section* s = found[kClusterMaterial]) {
    if (s->element_size != 4) return refuse("cluster file section cluster_material does not hold 4-byte indices");
    copy_section(bytes, *s, map);
    if (map.size() != all.size()) return refuse("cluster file's material map is not parallel to its clusters");
  }
  if (fm.empty()) {
    FileMaterial white{};
    for (float& v : white.base_color) v = 1.0f;
    white.roughness = 1.0f;
    white.base_color_image = white.normal_image = -1;
    white.normal_scale = 1.0f;
    fm.push_back(white);
  }
  for (uint32_t m : map)
    if (m >= fm.size()) return refuse("cluster file's material map names material " + std::to_string(m));
  for (const FileMaterial& f : fm) c->materials.push_back(resolve_material(f, flags));
  // The finest level: the clusters of lod level 0, in file order.
  uint32_t levels = 0;
  for (size_t k = 0; k < all.size(); ++k) {
    levels = std::max(levels, lod[k].level + 1);
    if (lod[k].level != 0) continue;
    c->clusters.push_back(all[k]);
    c->cluster_materials.push_back(map.empty() ? 0u : map[k]);
  }
  c->all_clusters = uint32_t(all.size());
  c->levels = levels;
  c->mesh.quant[0] = origin[0];
  c->mesh.quant[1] = origin[1];
  c->mesh.quant[2] = origin[2];
  c->mesh.quant[3] = step;
  c->mesh.first_cluster = 0;
  c->mesh.cluster_count = uint32_t(c->clusters.size());
  ge_instance_desc& in = c->instance;
  in.rows[0][0] = in.rows[1][1] = in.rows[2][2] = 1.0f;
  in.scale_max = 1.0f;
  in.flags = GE_INSTANCE_STATIC_RT | 1u;
  in.deform = GE_INVALID_DEFORM;
  *out = c.release();
  return GE_RT_OK;
}

}  // namespace

extern "C" {

int32_t ge_clusters_read(const uint8_t* bytes, uint64_t size, uint32_t flags, ge_clusters** out, char* error,
                         size_t error_size) {
  if (!out || (!bytes && size)) return GE_RT_ERROR_INVALID_ARGUMENT;
  *out = nullptr;
  try {
    return read_clusters(bytes, size, flags, out, error, error_size);
  } catch (const std::bad_alloc&) {
    static const char k_message[] = "out of host memory";
    if (error && error_size) {
      const size_t n = std::min(error_size - 1, sizeof(k_message) - 1);
      std::memcpy(error, k_message, n);
      error[n] = '\0';
    }
    return GE_RT_ERROR_OUT_OF_MEMORY;
  }
}

void ge_clusters_scene(const ge_clusters* c, ge_scene_desc* out) {
  if (!c || !out) return;
  std::memset(out, 0, sizeof(*out));
  out->clusters = c->clusters.data();
  out->cluster_count = uint32_t(c->clusters.size());
  out->vertex_count = uint32_t(c->vertices.size() / 3);
  out->vertices = c->vertices.data();
  out->quantized = c->quantized.data();
  out->attributes = c->attributes.empty() ? nullptr : c->attributes.data();
  out->triangles = c->triangles.data();
  out->triangle_count = uint32_t(c->triangles.size());
  out->mesh_count = 1;
  out->meshes = &c->mesh;
  out->instances = &c->instance;
  out->instance_count = 1;
  out->material_count = uint32_t(c->materials.size());
  out->materials = c->materials.data();
  out->cluster_materials = c->cluster_materials.data();
}

void ge_clusters_counts(const ge_clusters* c, uint32_t* all_clusters, uint32_t* levels) {
  if (!c) return;
  if (all_clusters) *all_clusters = c->all_clusters;
  if (levels) *levels = c->levels;
}

void ge_clusters_destroy(ge_clusters* c) { delete c; }

uint64_t ge_clusters_hash(const uint8_t* bytes, uint64_t size) { return hash_bytes(bytes, size); }

}  // extern "C"
