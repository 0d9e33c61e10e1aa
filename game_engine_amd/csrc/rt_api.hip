// rt_api.hip — the C ABI (include/ge_rt.h): context, scene upload with validation, the build
// orchestration (every mesh's LBVH in one segmented batch, then the instance level), and the
// traced passes. Errors are a status and a per-context message, never an exception across the ABI.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "rt_internal.h"
#include "rt_mlp.h"

using namespace ge;

struct ge_rt_cache {
  ge::RadianceCache net;
};

namespace {

// ---- the allocation-free shadow resolve's working memory ---------------------------------------------
// One array per GPU, part of the library's code object: the walk's ray counter in the first 256
// bytes, then GE_RT_SHADOW_STATIC_ENTRIES 16-byte ray records. Host code never uses the array's
// name as an address: each device's copy sits at its own address, looked up once per device with
// that device current (scratch_address) and passed to the kernels from there.
constexpr size_t k_scratch_head = 256;
constexpr size_t k_scratch_bytes = k_scratch_head + size_t(GE_RT_SHADOW_STATIC_ENTRIES) * 16;
static __device__ __align__(256) uint8_t scratch[k_scratch_bytes];

std::mutex g_scratch_mutex;
std::vector<uint8_t*> g_scratch_ptr;  // per device ordinal; null until looked up

// The address of `device`'s copy of the array; the caller has made `device` current.
hipError_t scratch_address(int device, uint8_t** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr.size() <= size_t(device)) g_scratch_ptr.resize(size_t(device) + 1, nullptr);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = static_cast<uint8_t*>(p);
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

// Scene versions come from one process-wide counter that only goes up, so a version names one
// state of one scene: a scene destroyed and another uploaded at the same address never share one.
std::atomic<uint64_t> g_version{0};
uint64_t next_version() { return g_version.fetch_add(1, std::memory_order_relaxed) + 1; }

int32_t set_error(ge_rt_ctx* ctx, int32_t code, const std::string& msg) {
  if (ctx) {
    ctx->error = msg;
    ctx->last_status = code;
  }
  return code;
}

// What an entry point returns when host code under it throws (std::bad_alloc from a container
// that could not grow, the only exception the backend's own code can raise): out of memory, with
// the message set where that itself allocates nothing new. Nothing crosses the C boundary.
int32_t host_failure(ge_rt_ctx* ctx) noexcept {
  if (ctx) {
    ctx->last_status = GE_RT_ERROR_OUT_OF_MEMORY;
    try {
      ctx->error = "out of host memory";
    } catch (...) {
      ctx->error.clear();
    }
  }
  return GE_RT_ERROR_OUT_OF_MEMORY;
}

#define GE_HIP(ctx, call)                                                                                   \
  do {                                                                                                      \
    const hipError_t e_ = (call);                                                                           \
    if (e_ != hipSuccess)                                                                                   \
      return set_error(ctx, e_ == hipErrorOutOfMemory ? GE_RT_ERROR_OUT_OF_MEMORY : GE_RT_ERROR_DEVICE,     \
                       std::string(#call) + ": " + hipGetErrorString(e_));                                 \
  } while (0)

#define GE_CHECK(ctx, cond, code, msg) \
  do {                                 \
    if (!(cond)) return set_error(ctx, code, msg); \
  } while (0)

template <class T>
hipError_t upload(DevBuf& buf, const T* host, size_t count, hipStream_t stream) {
  const size_t bytes = count * sizeof(T);
  hipError_t e = buf.reserve(bytes ? bytes : 256);
  if (e != hipSuccess || bytes == 0 || host == nullptr) return e;
  return hipMemcpyAsync(buf.ptr, host, bytes, hipMemcpyHostToDevice, stream);
}

// The build's pinned staging (ge_rt_ctx::pin): `bytes` in all for this build.
hipError_t pin_reserve(ge_rt_ctx* ctx, size_t bytes) {
  ctx->pin_used = 0;
  if (ctx->pin_cap >= bytes) return hipSuccess;
  if (ctx->pin) {
    const hipError_t e = hipHostFree(ctx->pin);
    if (e != hipSuccess) return e;
  }
  ctx->pin = nullptr;
  ctx->pin_cap = 0;
  const hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&ctx->pin), bytes);
  if (e == hipSuccess) ctx->pin_cap = bytes;
  return e;
}
constexpr size_t k_pin_align = 256;
size_t pin_piece(size_t bytes) { return (bytes + k_pin_align - 1) / k_pin_align * k_pin_align; }
// Copies `bytes` of host memory to `dev` through the next piece of the staging.
hipError_t staged_copy(ge_rt_ctx* ctx, void* dev, const void* host, size_t bytes, hipStream_t st) {
  if (bytes == 0) return hipSuccess;
  if (ctx->pin_used + pin_piece(bytes) > ctx->pin_cap) return hipErrorInvalidValue;  // sized at the build's start
  uint8_t* piece = ctx->pin + ctx->pin_used;
  ctx->pin_used += pin_piece(bytes);
  std::memcpy(piece, host, bytes);
  return hipMemcpyAsync(dev, piece, bytes, hipMemcpyHostToDevice, st);
}
template <class T>
hipError_t staged_upload(ge_rt_ctx* ctx, DevBuf& buf, const T* host, size_t count, hipStream_t st) {
  const hipError_t e = buf.reserve(count ? count * sizeof(T) : 256);
  if (e != hipSuccess) return e;
  return staged_copy(ctx, buf.ptr, host, count * sizeof(T), st);
}

// engine::relative_axis (core/math/world.h:181-195), operation for operation.
float relative_axis_host(int cell, float local, int eye_cell, float eye_local, float eye_residual) {
  const float cells = float(cell - eye_cell) * 64.0f;
  const float from_local = cells > 0.0f ? (cells - eye_local) + local : (cells + local) - eye_local;
  return from_local - eye_residual;
}

// The frame-to-mesh 3x4: the linear part's inverse by its adjugate in double, rounded once to
// float, and -inverse * translation in double.
void make_w2o(const ge_instance_desc& inst, const float T[3], float w2o[3][4]) {
  double m[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) m[r][c] = double(inst.rows[r][c]);
  const double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
  const double c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
  const double c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
  const double det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
  const double inv_det = det != 0.0 ? 1.0 / det : 0.0;
  double inv[3][3];
  inv[0][0] = c00 * inv_det;
  inv[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * inv_det;
  inv[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * inv_det;
  inv[1][0] = c01 * inv_det;
  inv[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * inv_det;
  inv[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * inv_det;
  inv[2][0] = c02 * inv_det;
  inv[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * inv_det;
  inv[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * inv_det;
  for (int r = 0; r < 3; ++r) {
    const double t = -(inv[r][0] * double(T[0]) + inv[r][1] * double(T[1]) + inv[r][2] * double(T[2]));
    for (int c = 0; c < 3; ++c) w2o[r][c] = float(inv[r][c]);
    w2o[r][3] = float(t);
  }
}

Bvh bvh_view(const SceneDev& s) {
  Bvh b{};
  b.tlas = s.tlas_nodes.as<ge_bvh_node>();
  b.tlas_slots = s.tlas_slots.as<uint32_t>();
  b.inst = s.inst_recs.as<InstRec>();
  b.nodes = s.blas_nodes.as<ge_bvh_node>();
  b.wtlas = s.tlas_wide.as<WideNode>();
  b.wnodes = s.blas_wide.as<WideNode>();
  b.tris = s.blas_tris.as<ge_bvh_triangle>();
  b.tlas_node_count = s.tlas_node_count;
  b.stack_cap = s.stack_cap;
  b.masks.inst = s.mask_inst.as<uint2>();
  b.masks.cluster_materials = s.cluster_materials.as<uint32_t>();
  b.masks.materials = s.materials.as<ge_resolve_material>();
  b.masks.clusters = s.clusters.as<ge_cluster_desc>();
  b.masks.triangles = s.triangles.as<uint32_t>();
  b.masks.attributes = s.has_attributes ? s.attributes.as<ge_vertex_attributes>() : nullptr;
  b.masks.tex = TexView{s.textures.as<ge_texture_desc>(), s.samplers.as<ge_sampler_desc>(), s.texels.as<uint8_t>(),
                        s.srgb.as<float>()};
  return b;
}

// The light grid over the box the lights' ranges reach (point and spot lights: a light at or
// beyond its radius adds nothing, light_sample's test). A cell's mask holds the lights whose
// sphere, grown by a thousandth of its radius and of a cell, meets the cell's box grown by a
// thousandth of a cell, so a point the shading places in the cell is within range only of lights
// in its mask; directional lights are in every cell. No light with a range: no grid.
hipError_t build_light_grid(SceneDev& s, hipStream_t stream) {
  s.light_grid = false;
  const uint32_t nl = std::min<uint32_t>(uint32_t(s.h_lights.size()), 64u);
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  bool ranged = false;
  auto radius_of = [](const ge_resolve_light& l) { return std::max(std::fabs(l.position_radius[3]), 1e-4f); };
  for (uint32_t i = 0; i < nl; ++i) {
    const ge_resolve_light& l = s.h_lights[i];
    if (l.direction_type[3] >= 1.5f) continue;
    if (!std::isfinite(l.direction_type[3])) return hipSuccess;
    for (int a = 0; a < 4; ++a)
      if (!std::isfinite(l.position_radius[a])) return hipSuccess;
    ranged = true;
    const float r = radius_of(l);
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], l.position_radius[a] - r);
      hi[a] = std::max(hi[a], l.position_radius[a] + r);
    }
  }
  if (!ranged || !(std::isfinite(lo[0]) && std::isfinite(lo[1]) && std::isfinite(lo[2]) && std::isfinite(hi[0]) &&
                   std::isfinite(hi[1]) && std::isfinite(hi[2])))
    return hipSuccess;
  // Grown by a thousandth of its largest extent: a point outside is that far from every range.
  double pad = 0.0;
  for (int a = 0; a < 3; ++a) pad = std::max(pad, double(hi[a]) - double(lo[a]));
  pad = 1e-3 * std::max(pad, 1e-3);
  for (int a = 0; a < 3; ++a) {
    lo[a] = float(double(lo[a]) - pad);
    hi[a] = float(double(hi[a]) + pad);
  }
  constexpr uint32_t G = k_light_grid;
  double cell[3];
  s.light_far = 0;
  for (uint32_t i = 0; i < nl; ++i)
    if (s.h_lights[i].direction_type[3] >= 1.5f) s.light_far |= 1ull << i;
  for (int a = 0; a < 3; ++a) {
    cell[a] = std::max(double(hi[a]) - double(lo[a]), 1e-3) / G;
    s.light_lo[a] = lo[a];
    s.light_inv[a] = float(1.0 / cell[a]);
  }
  std::vector<uint64_t> masks(size_t(G) * G * G, 0ull);
  for (uint32_t i = 0; i < nl; ++i) {
    const ge_resolve_light& l = s.h_lights[i];
    const bool everywhere = l.direction_type[3] >= 1.5f;
    const double r = radius_of(l);
    for (uint32_t z = 0; z < G; ++z)
      for (uint32_t y = 0; y < G; ++y)
        for (uint32_t x = 0; x < G; ++x) {
          bool in = everywhere;
          if (!in) {
            const uint32_t c[3] = {x, y, z};
            double d2 = 0.0, grow = 0.0;
            for (int a = 0; a < 3; ++a) grow = std::max(grow, 1e-3 * cell[a]);
            for (int a = 0; a < 3; ++a) {
              const double b0 = double(lo[a]) + c[a] * cell[a] - grow, b1 = double(lo[a]) + (c[a] + 1) * cell[a] + grow;
              const double p = l.position_radius[a];
              const double d = p < b0 ? b0 - p : (p > b1 ? p - b1 : 0.0);
              d2 += d * d;
            }
            const double reach = r * (1.0 + 1e-3) + grow;
            in = d2 < reach * reach;
          }
          if (in) masks[(size_t(z) * G + y) * G + x] |= 1ull << i;
        }
  }
  hipError_t e = s.light_cells.reserve(masks.size() * 8);
  if (e == hipSuccess) e = hipMemcpyAsync(s.light_cells.ptr, masks.data(), masks.size() * 8, hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e == hipSuccess) s.light_grid = true;
  return e;
}

SceneView scene_view(const SceneDev& s) {
  SceneView v{};
  v.clusters = s.clusters.as<ge_cluster_desc>();
  v.quantized = s.quantized.as<uint16_t>();
  v.attributes = s.has_attributes ? s.attributes.as<ge_vertex_attributes>() : nullptr;
  v.triangles = s.triangles.as<uint32_t>();
  v.meshes = s.meshes.as<ge_mesh_desc>();
  v.instances = s.instances.as<ge_instance_desc>();
  v.materials = s.materials.as<ge_resolve_material>();
  v.cluster_materials = s.cluster_materials.as<uint32_t>();
  v.lights = s.lights.as<ge_resolve_light>();
  v.light_count = s.light_count;
  v.light_cells = s.light_grid ? s.light_cells.as<uint64_t>() : nullptr;
  v.light_far = s.light_far;
  for (int a = 0; a < 3; ++a) {
    v.light_lo[a] = s.light_lo[a];
    v.light_inv[a] = s.light_inv[a];
  }
  v.features = (s.any_masked ? k_scene_masked : 0u) | (s.any_textured ? k_scene_textured : 0u);
  v.tex = TexView{s.textures.as<ge_texture_desc>(), s.samplers.as<ge_sampler_desc>(), s.texels.as<uint8_t>(),
                  s.srgb.as<float>()};
  return v;
}

// Whether the walks and the shading need the textured and alpha-tested forms (scene_view's flag).
bool full_scene(const SceneView& v) { return v.features != 0u; }

// Every texture slot a material reads names a texture and its sampler a sampler of the scene.
int32_t validate_materials(ge_rt_ctx* ctx, const SceneDev& s, const ge_resolve_material* m, uint32_t count,
                           uint32_t first) {
  auto slot_ok = [&](uint32_t texture, uint32_t sampler) {
    return texture == GE_NO_TEXTURE || (texture < s.texture_count && sampler < s.sampler_count);
  };
  for (uint32_t k = 0; k < count; ++k) {
    const ge_resolve_material& mat = m[k];
    const std::string name = "material " + std::to_string(first + k);
    auto slot = [&](uint32_t word, uint32_t shift) {
      const uint32_t v = (word >> shift) & 0xffffu;
      return v == GE_SAME_SAMPLER ? mat.sampler : v;
    };
    if (!slot_ok(mat.albedo_texture, mat.sampler) ||
        !slot_ok(mat.metallic_roughness_texture, slot(mat.samplers_mr_normal, 0)) ||
        !slot_ok(mat.normal_texture, slot(mat.samplers_mr_normal, 16)) ||
        !slot_ok(mat.emissive_texture, slot(mat.samplers_occlusion_emissive, 16)))
      return set_error(ctx, GE_RT_ERROR_INVALID_SCENE, name + " names a texture or sampler the scene does not hold");
    if (mat.flags & GE_MATERIAL_STOCHASTIC)
      return set_error(ctx, GE_RT_ERROR_UNSUPPORTED, name + " asks for stochastic (hex-tiled) sampling");
    if ((mat.flags >> 24) != 0u) return set_error(ctx, GE_RT_ERROR_UNSUPPORTED, name + " declares foliage optics");
  }
  return GE_RT_OK;
}

// Recomputes the scene's alpha-tested and textured flags and, when it has masked materials, each
// instance's mask record (its mesh's first cluster, its material base, bit 31 when one of its
// clusters' materials is masked), uploaded synchronously.
hipError_t refresh_masks(SceneDev& s, hipStream_t stream) {
  s.any_masked = false;
  s.any_textured = false;
  for (const ge_resolve_material& m : s.h_materials) {
    s.any_masked = s.any_masked || (m.flags & GE_MATERIAL_MASKED) != 0;
    s.any_textured = s.any_textured || m.albedo_texture != GE_NO_TEXTURE || m.metallic_roughness_texture != GE_NO_TEXTURE ||
                     m.normal_texture != GE_NO_TEXTURE || m.emissive_texture != GE_NO_TEXTURE;
  }
  if (!s.any_masked || s.instance_count == 0) return hipSuccess;
  std::vector<uint2> rec(s.instance_count);
  for (uint32_t i = 0; i < s.instance_count; ++i) {
    const ge_instance_desc& in = s.h_instances[i];
    bool masked = false;
    for (uint32_t cm : s.mesh_materials[in.mesh])
      masked = masked || (s.h_materials[cm + in.material_base].flags & GE_MATERIAL_MASKED) != 0;
    rec[i] = make_uint2(s.h_meshes[in.mesh].first_cluster, in.material_base | (masked ? 0x80000000u : 0u));
  }
  hipError_t e = s.mask_inst.reserve(rec.size() * sizeof(uint2));
  if (e == hipSuccess) e = hipMemcpyAsync(s.mask_inst.ptr, rec.data(), rec.size() * sizeof(uint2), hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  return e;
}

// The walks' rule for masked candidates from a query's flags (GE_RT_ALPHA_MASKS, GE_RT_CULL_MASKED).
uint32_t alpha_rule(const SceneDev& s, uint32_t flags) {
  if (!s.any_masked) return k_alpha_opaque;
  if (flags & GE_RT_ALPHA_MASKS) return k_alpha_test;
  return (flags & GE_RT_CULL_MASKED) ? k_alpha_cull : k_alpha_opaque;
}

// The engine's pair numbering, pair -> (instance, cluster within its mesh): instance i owns pairs
// [first_pair, first_pair + its clusters). Rebuilt by every structure build after an upload or an
// instance update, so that a shadow resolve finds it current; a numbering the table cannot hold
// (two instances claiming one pair, pairs past 2^24) is kept as `pairs_error` for the resolve to
// report. Synchronous.
hipError_t refresh_pairs(SceneDev& s) {
  s.pairs_error.clear();
  s.pair_count = 0;
  uint64_t count = 0;
  for (const ge_instance_desc& in : s.h_instances)
    count = std::max<uint64_t>(count, uint64_t(in.first_pair) + s.h_meshes[in.mesh].cluster_count);
  if (count > (1u << 24)) {
    s.pairs_error = "pairs past 2^24 do not fit a visibility word";
    s.pairs_dirty = false;
    return hipSuccess;
  }
  std::vector<uint32_t> table(size_t(count) * 2, GE_RT_MISS);
  for (uint32_t i = 0; i < s.instance_count && s.pairs_error.empty(); ++i) {
    const ge_instance_desc& in = s.h_instances[i];
    for (uint32_t c = 0; c < s.h_meshes[in.mesh].cluster_count; ++c) {
      uint32_t* e = &table[(size_t(in.first_pair) + c) * 2];
      if (e[0] != GE_RT_MISS) {
        s.pairs_error = "instances " + std::to_string(e[0]) + " and " + std::to_string(i) + " claim pair " +
                        std::to_string(in.first_pair + c);
        break;
      }
      e[0] = i;
      e[1] = c;
    }
  }
  s.pairs_dirty = false;
  if (!s.pairs_error.empty()) return hipSuccess;
  hipError_t e = s.pairs.reserve(std::max<size_t>(table.size() * 4, 8));
  if (e == hipSuccess && !table.empty()) e = hipMemcpy(s.pairs.ptr, table.data(), table.size() * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    s.pairs_dirty = true;
    return e;
  }
  s.pair_count = uint32_t(count);
  return hipSuccess;
}

int32_t validate_instances(ge_rt_ctx* ctx, const SceneDev& s, const ge_instance_desc* in, uint32_t count) {
  for (uint32_t i = 0; i < count; ++i) {
    const ge_instance_desc& inst = in[i];
    if (inst.mesh >= s.mesh_count)
      return set_error(ctx, GE_RT_ERROR_INVALID_SCENE, "instance " + std::to_string(i) + " names mesh " +
                                                           std::to_string(inst.mesh) + " of " + std::to_string(s.mesh_count));
    if (inst.deform != GE_INVALID_DEFORM)
      return set_error(ctx, GE_RT_ERROR_UNSUPPORTED, "instance " + std::to_string(i) + " is deformed");
    if (s.h_meshes[inst.mesh].cluster_count &&
        uint64_t(s.mesh_max_material[inst.mesh]) + inst.material_base >= s.material_count)
      return set_error(ctx, GE_RT_ERROR_INVALID_SCENE,
                       "instance " + std::to_string(i) + " reaches past the material array");
  }
  return GE_RT_OK;
}


// The visible list's entries against the scene's instances: each {instance, cluster} names an
// instance in range and a cluster of its mesh, and none repeats.
int32_t check_visible(ge_rt_ctx* ctx, const SceneDev& s, const uint32_t* e, uint32_t count, int32_t code) {
  std::vector<uint64_t> keys(count);
  for (uint32_t j = 0; j < count; ++j) {
    const uint32_t i = e[2 * j], c = e[2 * j + 1];
    if (i >= s.instance_count)
      return set_error(ctx, code, "visible entry " + std::to_string(j) + " names instance " + std::to_string(i) + " of " +
                                      std::to_string(s.instance_count));
    const ge_mesh_desc& mesh = s.h_meshes[s.h_instances[i].mesh];
    if (c < mesh.first_cluster || c - mesh.first_cluster >= mesh.cluster_count)
      return set_error(ctx, code, "visible entry " + std::to_string(j) + " names cluster " + std::to_string(c) +
                                      ", not one of instance " + std::to_string(i) + "'s mesh");
    keys[j] = (uint64_t(i) << 32) | c;
  }
  std::sort(keys.begin(), keys.end());
  const auto rep = std::adjacent_find(keys.begin(), keys.end());
  if (rep != keys.end())
    return set_error(ctx, code, "the visible list names instance " + std::to_string(*rep >> 32) + "'s cluster " +
                                    std::to_string(uint32_t(*rep)) + " twice");
  return GE_RT_OK;
}

// The frame structures' regions, laid out again when the instances changed: one for each instance
// that is not static, as large as its mesh's own structure, after the meshes' regions in the pools;
// its root slot mesh_count + k; its clusters' visible entries from geom_base in entry_of.
int32_t layout_frames(ge_rt_ctx* ctx, SceneDev& s, hipStream_t st) {
  const uint32_t ni = s.instance_count;
  s.frame_of.assign(ni, 0xffffffffu);
  s.h_geom_base.assign(std::max(ni, 1u), 0xffffffffu);
  s.frame_inst.clear();
  s.frame_tri_base.clear();
  s.frame_node_base.clear();
  std::vector<uint32_t> root_of(std::max(ni, 1u), 0u);
  uint64_t tri = s.tri_pool, node = s.node_pool, geom = 0;
  for (uint32_t i = 0; i < ni; ++i) {
    const ge_instance_desc& in = s.h_instances[i];
    root_of[i] = in.mesh;
    if (in.flags & GE_INSTANCE_STATIC_RT) continue;
    const uint32_t k = uint32_t(s.frame_inst.size());
    s.frame_of[i] = k;
    s.frame_inst.push_back(i);
    s.frame_tri_base.push_back(uint32_t(tri));
    s.frame_node_base.push_back(uint32_t(node));
    s.h_geom_base[i] = uint32_t(geom);
    root_of[i] = s.mesh_count + k;
    tri += s.mesh_tri_count[in.mesh];
    node += s.mesh_node_count[in.mesh];
    geom += s.h_meshes[in.mesh].cluster_count;
    GE_CHECK(ctx, tri < (1ull << 31) && node < (1ull << 31) && geom < (1ull << 31), GE_RT_ERROR_INVALID_SCENE,
             "the frame's structures would hold more than 2^31 triangles");
  }
  s.frame_tri_pool = uint32_t(tri - s.tri_pool);
  s.frame_node_pool = uint32_t(node - s.node_pool);
  s.h_entry_of.assign(size_t(std::max<uint64_t>(geom, 1)), 0xffffffffu);
  const size_t nf = s.frame_inst.size();
  s.frame_clusters.assign(nf, {});
  s.frame_tri_count.assign(nf, 0u);
  s.frame_node_count.assign(nf, 0u);
  GE_HIP(ctx, upload(s.root_of, root_of.data(), root_of.size(), st));
  GE_HIP(ctx, upload(s.geom_base, s.h_geom_base.data(), s.h_geom_base.size(), st));
  s.frame_layout_dirty = false;
  return GE_RT_OK;
}

// Each frame structure's clusters from the visible list, in list order, and each cluster's entry.
int32_t resolve_visible(ge_rt_ctx* ctx, SceneDev& s, hipStream_t st) {
  const uint32_t n = uint32_t(s.h_visible.size() / 2);
  const int32_t r = check_visible(ctx, s, s.h_visible.data(), n, GE_RT_ERROR_INVALID_SCENE);  // the instances may have changed since
  if (r != GE_RT_OK) return r;
  for (auto& v : s.frame_clusters) v.clear();
  std::fill(s.h_entry_of.begin(), s.h_entry_of.end(), 0xffffffffu);
  for (uint32_t j = 0; j < n; ++j) {
    const uint32_t i = s.h_visible[2 * j], c = s.h_visible[2 * j + 1];
    const uint32_t k = s.frame_of[i];
    if (k == 0xffffffffu) continue;  // a static instance traces its mesh's structure
    s.h_entry_of[s.h_geom_base[i] + (c - s.h_meshes[s.h_instances[i].mesh].first_cluster)] = j;
    s.frame_clusters[k].push_back(c);
  }
  for (size_t k = 0; k < s.frame_clusters.size(); ++k) {
    uint64_t t = 0;
    for (uint32_t c : s.frame_clusters[k]) t += s.h_clusters[c].triangle_count;
    s.frame_tri_count[k] = uint32_t(t);
    s.frame_node_count[k] = t == 0 ? 0u : (t == 1 ? 1u : uint32_t(t - 1));
  }
  GE_HIP(ctx, staged_upload(ctx, s.entry_of, s.h_entry_of.data(), s.h_entry_of.size(), st));
  return GE_RT_OK;
}

// Whether instance i has geometry to trace: its mesh's triangles, or its frame structure's.
bool instance_live(const SceneDev& s, uint32_t i) {
  const uint32_t k = s.frame_of[i];
  return k == 0xffffffffu ? s.mesh_tri_count[s.h_instances[i].mesh] > 0 : s.frame_tri_count[k] > 0;
}

}  // namespace

extern "C" {

// One-wave workgroups a persistent walk launches per CU: as many as its four SIMDs hold at the
// walks' occupancy (20 at 5 waves a SIMD). On the 4M-triangle frame 20 runs 1.7% faster than 32
// (whose extra waves start only as the first finish), 24 alike, 16 3% slower.
#ifndef GE_PERSISTENT_WAVES_PER_CU
#if GE_WALK_WAVES_PER_EU > 0
#define GE_PERSISTENT_WAVES_PER_CU (4u * GE_WALK_WAVES_PER_EU)
#else
#define GE_PERSISTENT_WAVES_PER_CU 20u
#endif
#endif

int32_t ge_rt_create(ge_rt_ctx** out, int32_t device_ordinal) try {
  if (!out) return GE_RT_ERROR_INVALID_ARGUMENT;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return GE_RT_ERROR_DEVICE;
  if (device_ordinal < 0 || device_ordinal >= count) return GE_RT_ERROR_INVALID_ARGUMENT;
  auto* ctx = new (std::nothrow) ge_rt_ctx();
  if (!ctx) return GE_RT_ERROR_OUT_OF_MEMORY;
  ctx->device = device_ordinal;
  if (hipSetDevice(device_ordinal) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
      hipEventCreate(&ctx->bev0) != hipSuccess || hipEventCreate(&ctx->bev1) != hipSuccess ||
      ctx->counters.reserve(128) != hipSuccess || hipMemset(ctx->counters.ptr, 0, 128) != hipSuccess) {
    ge_rt_destroy(ctx);
    return GE_RT_ERROR_DEVICE;
  }
  for (auto& kind : ctx->stage_ev)
    for (hipEvent_t& ev : kind)
      if (hipEventCreate(&ev) != hipSuccess) {
        ge_rt_destroy(ctx);
        return GE_RT_ERROR_DEVICE;
      }
  if (scratch_address(device_ordinal, &ctx->static_scratch) != hipSuccess) {
    ge_rt_destroy(ctx);
    return GE_RT_ERROR_DEVICE;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_ordinal) == hipSuccess && cus > 0)
    ctx->persistent_blocks = uint32_t(cus) * GE_PERSISTENT_WAVES_PER_CU;
  *out = ctx;
  return GE_RT_OK;
} catch (...) {
  return host_failure(nullptr);
}

void ge_rt_destroy(ge_rt_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (DevBuf* b : {&ctx->build.keys[0], &ctx->build.keys[1], &ctx->build.vals[0], &ctx->build.vals[1], &ctx->build.hist,
                    &ctx->build.scan_tmp, &ctx->build.node_box, &ctx->build.node_parent, &ctx->build.seg_cbox,
                    &ctx->build.seg_scale, &ctx->build.depth, &ctx->build.buckets, &ctx->build.bucket_count,
                    &ctx->build.flag, &ctx->build.cl_out, &ctx->build.cl_part, &ctx->build.cl_seg, &ctx->build.links,
                    &ctx->counters})
    b->release();
  if (ctx->host_readback) (void)hipHostFree(ctx->host_readback);
  if (ctx->pin) (void)hipHostFree(ctx->pin);
  for (DevBuf& b : ctx->io) b.release();
  ctx->resolve_rays.release();
  for (DevBuf& b : ctx->tmp) b.release();
  for (DevBuf* b : {&ctx->wave.state[0], &ctx->wave.state[1], &ctx->wave.hits, &ctx->wave.nee, &ctx->wave.result,
                    &ctx->wave.queue[0], &ctx->wave.queue[1], &ctx->wave.counts, &ctx->wave.shadow, &ctx->wave.blocked,
                    &ctx->wave.sgrad, &ctx->wave.ts, &ctx->wave.adj_a, &ctx->wave.adj_r, &ctx->wave.lpath})
    b->release();
  if (ctx->wave.host_counts) (void)hipHostFree(ctx->wave.host_counts);
  if (ctx->wave.counted) (void)hipEventDestroy(ctx->wave.counted);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->bev0) (void)hipEventDestroy(ctx->bev0);
  if (ctx->bev1) (void)hipEventDestroy(ctx->bev1);
  for (auto& kind : ctx->stage_ev)
    for (hipEvent_t ev : kind)
      if (ev) (void)hipEventDestroy(ev);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int32_t ge_rt_last_error(const ge_rt_ctx* ctx, char* buffer, size_t size) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  if (buffer && size) {
    const size_t n = std::min(size - 1, ctx->error.size());
    std::memcpy(buffer, ctx->error.data(), n);
    buffer[n] = 0;
  }
  return ctx->last_status;
} catch (...) {
  return host_failure(nullptr);
}

int32_t ge_rt_synchronize(ge_rt_ctx* ctx) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_HIP(ctx, hipSetDevice(ctx->device));
  GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_scene_upload(ge_rt_ctx* ctx, const ge_scene_desc* d, ge_rt_scene** out) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, d && out, GE_RT_ERROR_INVALID_ARGUMENT, "null scene or output");
  *out = nullptr;
  GE_CHECK(ctx, (d->clusters || !d->cluster_count) && (d->vertices || !d->vertex_count) &&
                    (d->quantized || !d->vertex_count) && (d->triangles || !d->triangle_count) &&
                    (d->meshes || !d->mesh_count) && (d->instances || !d->instance_count) &&
                    (d->materials || !d->material_count) && (d->cluster_materials || !d->cluster_count) &&
                    (d->lights || !d->light_count),
           GE_RT_ERROR_INVALID_ARGUMENT, "a scene array is null with a non-zero count");
  GE_HIP(ctx, hipSetDevice(ctx->device));
  auto* scene = new (std::nothrow) ge_rt_scene();
  if (scene) scene->dev.version = next_version();
  GE_CHECK(ctx, scene, GE_RT_ERROR_OUT_OF_MEMORY, "host allocation");
  SceneDev& s = scene->dev;
  s.cluster_count = d->cluster_count;
  s.vertex_count = d->vertex_count;
  s.triangle_count = d->triangle_count;
  s.mesh_count = d->mesh_count;
  s.instance_count = d->instance_count;
  s.material_count = d->material_count;
  s.light_count = d->light_count;
  s.has_attributes = d->attributes != nullptr;
  // Validation: every offset and count inside its array, every local index inside its cluster.
  auto fail = [&](int32_t code, const std::string& msg) {
    delete scene;
    return set_error(ctx, code, msg);
  };
  for (uint32_t c = 0; c < d->cluster_count; ++c) {
    const ge_cluster_desc& cl = d->clusters[c];
    if (uint64_t(cl.vertex_offset) + cl.vertex_count > d->vertex_count ||
        uint64_t(cl.triangle_offset) + cl.triangle_count > d->triangle_count)
      return fail(GE_RT_ERROR_INVALID_SCENE, "cluster " + std::to_string(c) + " reaches past the vertex or triangle array");
    if (cl.triangle_count > 256 || cl.vertex_count > 256)
      return fail(GE_RT_ERROR_INVALID_SCENE, "cluster " + std::to_string(c) + " holds more than 256 triangles or vertices");
    for (uint32_t t = 0; t < cl.triangle_count; ++t) {
      const uint32_t p = d->triangles[cl.triangle_offset + t];
      if ((p & 0xffu) >= cl.vertex_count || ((p >> 8) & 0xffu) >= cl.vertex_count || ((p >> 16) & 0xffu) >= cl.vertex_count)
        return fail(GE_RT_ERROR_INVALID_SCENE, "cluster " + std::to_string(c) + " triangle " + std::to_string(t) +
                                                   " names a vertex outside the cluster");
    }
  }
  s.h_clusters.assign(d->clusters, d->clusters + d->cluster_count);
  s.h_meshes.assign(d->meshes, d->meshes + d->mesh_count);
  s.mesh_tri_count.assign(d->mesh_count, 0);
  s.mesh_tri_base.assign(d->mesh_count, 0);
  s.mesh_node_base.assign(d->mesh_count, 0);
  s.mesh_node_count.assign(d->mesh_count, 0);
  uint64_t tri_pool = 0, node_pool = 0;
  for (uint32_t m = 0; m < d->mesh_count; ++m) {
    const ge_mesh_desc& mesh = d->meshes[m];
    if (uint64_t(mesh.first_cluster) + mesh.cluster_count > d->cluster_count || mesh.cluster_count >= (1u << 24))
      return fail(GE_RT_ERROR_INVALID_SCENE, "mesh " + std::to_string(m) + " reaches past the cluster array");
    if (mesh.deform_pool || mesh.deform_slots || mesh.templates || mesh.skin)
      return fail(GE_RT_ERROR_UNSUPPORTED, "mesh " + std::to_string(m) + " carries deformation streams");
    uint64_t n = 0;
    for (uint32_t c = mesh.first_cluster; c < mesh.first_cluster + mesh.cluster_count; ++c) n += d->clusters[c].triangle_count;
    if (n >= (1u << 30)) return fail(GE_RT_ERROR_INVALID_SCENE, "mesh " + std::to_string(m) + " is too large");
    s.mesh_tri_count[m] = uint32_t(n);
    s.mesh_tri_base[m] = uint32_t(tri_pool);
    s.mesh_node_base[m] = uint32_t(node_pool);
    s.mesh_node_count[m] = n == 0 ? 0u : (n == 1 ? 1u : uint32_t(n - 1));
    tri_pool += n;
    node_pool += s.mesh_node_count[m];
  }
  if (tri_pool >= (1ull << 31) || node_pool >= (1ull << 31))
    return fail(GE_RT_ERROR_INVALID_SCENE, "the scene's meshes hold more than 2^31 triangles");
  s.tri_pool = uint32_t(tri_pool);
  s.node_pool = uint32_t(node_pool);
  GE_CHECK(ctx, (d->textures || !d->texture_count) && (d->samplers || !d->sampler_count) && (d->texels || !d->texel_bytes),
           GE_RT_ERROR_INVALID_ARGUMENT, "a texture array is null with a non-zero count");
  for (uint32_t t = 0; t < d->texture_count; ++t) {
    const ge_texture_desc& td = d->textures[t];
    if (td.width == 0 || td.height == 0 || td.width > 32768 || td.height > 32768 ||
        (td.format != GE_FORMAT_R8G8B8A8_UNORM && td.format != GE_FORMAT_R8G8B8A8_SRGB) || (td.offset & 3u) ||
        td.offset + uint64_t(td.width) * td.height * 4 > d->texel_bytes)
      return fail(GE_RT_ERROR_INVALID_SCENE, "texture " + std::to_string(t) +
                                                 " is empty, larger than 32768 texels a side, of a format other than "
                                                 "RGBA8 (UNORM or sRGB), unaligned or past the texel pool");
  }
  for (uint32_t k = 0; k < d->sampler_count; ++k) {
    const ge_sampler_desc& sd = d->samplers[k];
    if (sd.address_u > GE_ADDRESS_CLAMP_TO_BORDER || sd.address_v > GE_ADDRESS_CLAMP_TO_BORDER ||
        sd.mag > GE_FILTER_LINEAR || sd.min > GE_FILTER_LINEAR)
      return fail(GE_RT_ERROR_INVALID_SCENE, "sampler " + std::to_string(k) + " has an unknown address mode or filter");
  }
  s.texture_count = d->texture_count;
  s.sampler_count = d->sampler_count;
  {
    const int32_t st = validate_materials(ctx, s, d->materials, d->material_count, 0);
    if (st != GE_RT_OK) {
      const std::string msg = ctx->error;
      return fail(st, msg);
    }
  }
  s.h_materials.assign(d->materials, d->materials + d->material_count);
  if (d->light_count) s.h_lights.assign(d->lights, d->lights + d->light_count);
  {
    // The largest cluster material of each mesh: an instance's materials are that plus its base.
    s.mesh_max_material.assign(d->mesh_count, 0u);
    for (uint32_t m = 0; m < d->mesh_count; ++m) {
      const ge_mesh_desc& mesh = d->meshes[m];
      for (uint32_t c = mesh.first_cluster; c < mesh.first_cluster + mesh.cluster_count; ++c)
        s.mesh_max_material[m] = std::max(s.mesh_max_material[m], d->cluster_materials[c]);
    }
    const int32_t st = validate_instances(ctx, s, d->instances, d->instance_count);
    if (st != GE_RT_OK) {
      const std::string msg = ctx->error;
      return fail(st, msg);
    }
    s.mesh_materials.assign(d->mesh_count, {});
    for (uint32_t m = 0; m < d->mesh_count; ++m) {
      std::vector<uint32_t>& list = s.mesh_materials[m];
      const ge_mesh_desc& mesh = d->meshes[m];
      for (uint32_t c = mesh.first_cluster; c < mesh.first_cluster + mesh.cluster_count; ++c) list.push_back(d->cluster_materials[c]);
      std::sort(list.begin(), list.end());
      list.erase(std::unique(list.begin(), list.end()), list.end());
    }
  }
  s.h_instances.assign(d->instances, d->instances + d->instance_count);
  const hipStream_t st = ctx->stream;
  const size_t qn = (size_t(d->vertex_count) * 3 + 1) & ~size_t(1);
  std::vector<ge_mesh_desc> meshes(d->meshes, d->meshes + d->mesh_count);
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = upload(s.clusters, d->clusters, d->cluster_count, st);
  if (e == hipSuccess) e = upload(s.vertices, d->vertices, size_t(d->vertex_count) * 3, st);
  if (e == hipSuccess) e = upload(s.quantized, d->quantized, qn, st);
  if (e == hipSuccess) e = upload(s.attributes, d->attributes, s.has_attributes ? d->vertex_count : 0, st);
  if (e == hipSuccess) e = upload(s.triangles, d->triangles, d->triangle_count, st);
  for (ge_mesh_desc& m : meshes) m.quantized = reinterpret_cast<uint64_t>(s.quantized.ptr);
  if (e == hipSuccess) e = upload(s.meshes, meshes.data(), meshes.size(), st);
  if (e == hipSuccess) e = upload(s.instances, d->instances, d->instance_count, st);
  if (e == hipSuccess) e = upload(s.materials, d->materials, d->material_count, st);
  if (e == hipSuccess) e = upload(s.cluster_materials, d->cluster_materials, d->cluster_count, st);
  if (e == hipSuccess) e = upload(s.lights, d->lights, d->light_count, st);
  if (e == hipSuccess) e = upload(s.textures, d->textures, d->texture_count, st);
  if (e == hipSuccess) e = upload(s.samplers, d->samplers, d->sampler_count, st);
  if (e == hipSuccess) e = upload(s.texels, d->texels, size_t(d->texel_bytes), st);
  float srgb[256];
  for (uint32_t c = 0; c < 256; ++c) {
    const double v = double(c) / 255.0;
    srgb[c] = float(v <= 0.04045 ? v / 12.92 : std::pow((v + 0.055) / 1.055, 2.4));
  }
  if (e == hipSuccess) e = upload(s.srgb, srgb, 256, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess) e = refresh_masks(s, st);
  if (e == hipSuccess) e = build_light_grid(s, st);
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? GE_RT_ERROR_OUT_OF_MEMORY : GE_RT_ERROR_DEVICE,
                std::string("scene upload: ") + hipGetErrorString(e));
  *out = scene;
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

void ge_rt_scene_destroy(ge_rt_ctx* ctx, ge_rt_scene* scene) {
  if (!scene) return;
  if (ctx) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->primal.scene == scene) ctx->primal.valid = false;
  }
  SceneDev& s = scene->dev;
  for (DevBuf* b : {&s.clusters, &s.vertices, &s.quantized, &s.attributes, &s.triangles, &s.meshes, &s.instances,
                    &s.materials, &s.cluster_materials, &s.lights, &s.tri_prefix, &s.tri_mesh_of, &s.blas_nodes,
                    &s.blas_tris, &s.blas_morton, &s.mesh_roots, &s.mesh_height, &s.tlas_nodes, &s.tlas_slots,
                    &s.tlas_morton, &s.inst_recs, &s.inst_boxes, &s.live_ids, &s.tlas_root, &s.tlas_height, &s.pairs,
                    &s.textures, &s.samplers, &s.texels, &s.srgb, &s.mask_inst, &s.blas_wide, &s.tlas_wide,
                    &s.light_cells})
    b->release();
  delete scene;
}

int32_t ge_rt_scene_update_instances(ge_rt_ctx* ctx, ge_rt_scene* scene, uint32_t first, uint32_t count,
                                     const ge_instance_desc* instances) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, scene && (instances || !count), GE_RT_ERROR_INVALID_ARGUMENT, "null scene or instances");
  SceneDev& s = scene->dev;
  GE_CHECK(ctx, uint64_t(first) + count <= s.instance_count, GE_RT_ERROR_INVALID_ARGUMENT, "instance range out of bounds");
  GE_HIP(ctx, hipSetDevice(ctx->device));
  const int32_t st = validate_instances(ctx, s, instances, count);
  if (st != GE_RT_OK) return st;
  // A moved instance changes only its record; one that changes mesh or pair range changes the pair
  // table and the list of instances with geometry, and one that changes materials the mask table.
  bool topology = false, materials = false, frames = false;
  for (uint32_t k = 0; k < count; ++k) {
    const ge_instance_desc& was = s.h_instances[first + k];
    topology = topology || was.mesh != instances[k].mesh || was.first_pair != instances[k].first_pair;
    materials = materials || was.mesh != instances[k].mesh || was.material_base != instances[k].material_base;
    frames = frames || was.mesh != instances[k].mesh || ((was.flags ^ instances[k].flags) & GE_INSTANCE_STATIC_RT);
  }
  std::memcpy(&s.h_instances[first], instances, size_t(count) * sizeof(ge_instance_desc));
  if (topology) {
    s.pairs_dirty = true;
    s.live_dirty = true;
  }
  if (frames) s.frame_layout_dirty = true;  // an instance changed mesh or static form: regions laid out again
  if (count) {
    s.rec_lo = std::min(s.rec_lo, first);
    s.rec_hi = std::max(s.rec_hi, first + count);
  }
  s.version = next_version();
  GE_HIP(ctx, hipMemcpyAsync(s.instances.as<ge_instance_desc>() + first, &s.h_instances[first],
                             size_t(count) * sizeof(ge_instance_desc), hipMemcpyHostToDevice, ctx->stream));
  if (materials) GE_HIP(ctx, refresh_masks(s, ctx->stream));
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

}  // extern "C"

namespace {

// The pool region a download names: mesh `sel`'s structure, or with GE_RT_FRAME_STRUCTURE the frame
// structure of instance sel & ~GE_RT_FRAME_STRUCTURE; false when there is none.
bool structure_of(const SceneDev& s, uint32_t sel, uint32_t& node_base, uint32_t& node_count, uint32_t& tri_base,
                  uint32_t& tri_count) {
  if (sel & GE_RT_FRAME_STRUCTURE) {
    const uint32_t i = sel & ~GE_RT_FRAME_STRUCTURE;
    if (i >= s.frame_of.size() || s.frame_of[i] == 0xffffffffu) return false;
    const uint32_t k = s.frame_of[i];
    node_base = s.frame_node_base[k];
    node_count = s.frame_node_count[k];
    tri_base = s.frame_tri_base[k];
    tri_count = s.frame_tri_count[k];
    return true;
  }
  if (sel >= s.mesh_count) return false;
  node_base = s.mesh_node_base[sel];
  node_count = s.mesh_node_count[sel];
  tri_base = s.mesh_tri_base[sel];
  tri_count = s.mesh_tri_count[sel];
  return true;
}

}  // namespace

extern "C" {

int32_t ge_rt_scene_set_visible(ge_rt_ctx* ctx, ge_rt_scene* scene, const uint32_t* entries, uint32_t count) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, scene && (entries || !count), GE_RT_ERROR_INVALID_ARGUMENT, "null scene or entries");
  SceneDev& s = scene->dev;
  const int32_t r = check_visible(ctx, s, entries, count, GE_RT_ERROR_INVALID_ARGUMENT);
  if (r != GE_RT_OK) return r;
  s.h_visible.assign(entries, entries + 2 * size_t(count));
  s.has_visible = true;
  s.visible_dirty = true;
  s.version = next_version();
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_bvh_build(ge_rt_ctx* ctx, ge_rt_scene* scene, uint32_t flags, const ge_frame_eye* eye, const uint32_t* meshes,
                        uint32_t count) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, scene, GE_RT_ERROR_INVALID_ARGUMENT, "null scene");
  GE_CHECK(ctx, flags <= GE_RT_BUILD_REFIT, GE_RT_ERROR_INVALID_ARGUMENT, "unknown build flags");
  SceneDev& s = scene->dev;
  GE_CHECK(ctx, flags == GE_RT_BUILD_FULL || s.built, GE_RT_ERROR_INVALID_SCENE, "a partial build needs a full one first");
  const bool subset = flags == GE_RT_BUILD_SUBSET || flags == GE_RT_BUILD_REFIT;
  GE_CHECK(ctx, !subset || meshes || !count, GE_RT_ERROR_INVALID_ARGUMENT, "null mesh list");
  GE_HIP(ctx, hipSetDevice(ctx->device));
  const hipStream_t st = ctx->stream;

  // ---- the frame's structures: their regions (when the instances changed), their clusters (when
  // the visible list did), and which instances have geometry ----
  bool frames_changed = false;
  if (s.frame_layout_dirty) {
    const int32_t r = layout_frames(ctx, s, st);
    if (r != GE_RT_OK) return r;
    s.rec_all = true;
    s.visible_dirty = true;
  }
  {
    // The staging for every upload below: the entry table, the batch's packed arrays (five per
    // structure, four per cluster, each padded to 64 words), the instance records and translations,
    // the live list and the instance structure's three words.
    const size_t ni = s.instance_count, nstruct = size_t(s.mesh_count) + s.frame_inst.size();
    const size_t nclusters = size_t(s.cluster_count) + s.h_entry_of.size();
    const size_t bytes = pin_piece(s.h_entry_of.size() * 4) + pin_piece((5 * nstruct + 4 * nclusters + 9 * 64) * 4) +
                         pin_piece(ni * sizeof(InstRec)) + pin_piece(ni * 12) + pin_piece(ni * 4) + pin_piece(12) +
                         pin_piece(12);
    GE_HIP(ctx, pin_reserve(ctx, bytes));
  }
  GE_CHECK(ctx, s.frame_inst.empty() || s.has_visible, GE_RT_ERROR_INVALID_SCENE,
           "instance " + std::to_string(s.frame_inst.empty() ? 0u : s.frame_inst[0]) +
               " is not static (InstanceDesc::flags bit 7): its structure needs the frame's visible list");
  if (s.visible_dirty) {
    const int32_t r = resolve_visible(ctx, s, st);
    if (r != GE_RT_OK) return r;
    frames_changed = true;
  }
  const uint32_t ni = s.instance_count;
  std::vector<uint32_t> live;
  for (uint32_t i = 0; i < ni; ++i)
    if (instance_live(s, i)) live.push_back(i);
  if (live != s.h_live) s.live_dirty = true;
  GE_CHECK(ctx, flags != GE_RT_BUILD_REFIT || !s.live_dirty, GE_RT_ERROR_INVALID_SCENE,
           "a refit keeps the instance structure's slots: an instance given another mesh, or whose visible clusters came "
           "or went, needs a rebuild");
  s.eye = eye ? *eye : ge_frame_eye{};
  const uint32_t nf = uint32_t(s.frame_inst.size());
  const uint32_t roots = s.mesh_count + nf;
  const size_t node_all = size_t(std::max(s.node_pool + s.frame_node_pool, 1u)),
               tri_all = size_t(std::max(s.tri_pool + s.frame_tri_pool, 1u));
  // A pool that has to grow loses what it held: every structure is built again into it.
  const bool regrow = s.blas_nodes.bytes < node_all * sizeof(ge_bvh_node) || s.blas_wide.bytes < node_all * sizeof(WideNode) ||
                      s.blas_tris.bytes < tri_all * sizeof(ge_bvh_triangle) || s.blas_morton.bytes < tri_all * 4 ||
                      s.mesh_roots.bytes < size_t(std::max(roots, 1u)) * 24 || s.mesh_height.bytes < size_t(std::max(roots, 1u)) * 4;
  GE_HIP(ctx, s.blas_nodes.reserve(node_all * sizeof(ge_bvh_node)));
  GE_HIP(ctx, s.blas_wide.reserve(node_all * sizeof(WideNode)));
  GE_HIP(ctx, s.blas_tris.reserve(tri_all * sizeof(ge_bvh_triangle)));
  GE_HIP(ctx, s.blas_morton.reserve(tri_all * 4));
  GE_HIP(ctx, s.mesh_roots.reserve(size_t(std::max(roots, 1u)) * 24));
  GE_HIP(ctx, s.mesh_height.reserve(size_t(std::max(roots, 1u)) * 4));
  GE_HIP(ctx, ctx->build.flag.reserve(4));
  const size_t readback = size_t(roots) + 2;  // the flag, the instance structure's height, the structures'
  if (ctx->host_readback_cap < readback) {
    if (ctx->host_readback) GE_HIP(ctx, hipHostFree(ctx->host_readback));
    ctx->host_readback = nullptr;
    ctx->host_readback_cap = 0;
    GE_HIP(ctx, hipHostMalloc(reinterpret_cast<void**>(&ctx->host_readback), readback * 4));
    ctx->host_readback_cap = readback;
  }
  uint32_t* flag = ctx->build.flag.as<uint32_t>();
  GE_HIP(ctx, hipEventRecord(ctx->bev0, st));
  GE_HIP(ctx, hipMemsetAsync(flag, 0, 4, st));

  // ---- the segments: meshes, then frame structures ----
  struct Seg {
    uint32_t count, node, out, root, mesh;
    const std::vector<uint32_t>* clusters;  // a frame structure's clusters; null: the mesh's own
  };
  std::vector<Seg> segs;
  std::vector<uint32_t> list;
  if (flags == GE_RT_BUILD_FULL || regrow) {
    for (uint32_t m = 0; m < s.mesh_count; ++m) list.push_back(m);
  } else if (subset) {
    for (uint32_t k = 0; k < count; ++k) {
      GE_CHECK(ctx, meshes[k] < s.mesh_count, GE_RT_ERROR_INVALID_ARGUMENT, "mesh index out of range");
      list.push_back(meshes[k]);
    }
    std::sort(list.begin(), list.end());
    list.erase(std::unique(list.begin(), list.end()), list.end());
  }
  for (uint32_t m : list)
    if (s.mesh_tri_count[m] > 0) segs.push_back({s.mesh_tri_count[m], s.mesh_node_base[m], s.mesh_tri_base[m], m, m, nullptr});
  if (frames_changed || flags == GE_RT_BUILD_FULL || regrow)
    for (uint32_t k = 0; k < nf; ++k)
      if (s.frame_tri_count[k] > 0)
        segs.push_back({s.frame_tri_count[k], s.frame_node_base[k], s.frame_tri_base[k], s.mesh_count + k,
                        s.h_instances[s.frame_inst[k]].mesh, &s.frame_clusters[k]});
  // Structures too large for a one-workgroup build take the device-wide passes over a batch of
  // triangle records; the others are built one workgroup each. Either kind reads its clusters on
  // the device from a table of segments (a mesh's own clusters, or a frame structure's list).
  uint32_t nseg_global = 0, n = 0, ncl_global = 0;
  std::vector<uint32_t> seg_first, seg_count, seg_node, seg_out, seg_root, seg_pre;
  std::vector<GlobalSeg> gsegs;
  std::vector<BlockSeg>& bsegs = ctx->host_bsegs;
  std::vector<uint32_t>& lists = ctx->host_blists;
  bsegs.clear();
  lists.clear();
  for (const Seg& g : segs) {
    const ge_mesh_desc& mesh = s.h_meshes[g.mesh];
    uint32_t first = mesh.first_cluster, ncl = mesh.cluster_count, is_list = 0;
    if (g.clusters) {
      first = uint32_t(lists.size());
      ncl = uint32_t(g.clusters->size());
      is_list = 1;
      lists.insert(lists.end(), g.clusters->begin(), g.clusters->end());
    }
    if (g.count <= k_block_build_max) {
      bsegs.push_back({g.count, g.node, g.out, g.root, first, ncl, mesh.first_cluster, is_list});
      continue;
    }
    ++nseg_global;
    seg_first.push_back(n);
    seg_count.push_back(g.count);
    seg_node.push_back(g.node);
    seg_out.push_back(g.out);
    seg_root.push_back(g.root);
    seg_pre.push_back(ncl_global);
    gsegs.push_back({first, is_list, mesh.first_cluster, 0u});
    ncl_global += ncl;
    n += g.count;
  }
  seg_pre.push_back(ncl_global);
  std::vector<uint32_t>& meta = ctx->host_meta;  // every small array of the batch, packed, one upload
  meta.clear();
  auto pack_words = [&](const uint32_t* p, size_t count) {
    const size_t at = meta.size();
    meta.insert(meta.end(), p, p + count);
    while (meta.size() % 64) meta.push_back(0);
    return at;
  };
  auto pack = [&](const std::vector<uint32_t>& v) { return pack_words(v.data(), v.size()); };
  static_assert(sizeof(BlockSeg) == 32 && sizeof(GlobalSeg) == 16, "whole words a segment");
  if (n > 0 || !bsegs.empty()) {
    const size_t o_sf = pack(seg_first), o_sc = pack(seg_count), o_sn = pack(seg_node), o_so = pack(seg_out),
                 o_sr = pack(seg_root), o_sp = pack(seg_pre),
                 o_gs = pack_words(reinterpret_cast<const uint32_t*>(gsegs.data()), gsegs.size() * 4),
                 o_bs = pack_words(reinterpret_cast<const uint32_t*>(bsegs.data()), bsegs.size() * 8), o_li = pack(lists);
    GE_HIP(ctx, staged_upload(ctx, ctx->tmp[0], meta.data(), meta.size(), st));
    const uint32_t* md = ctx->tmp[0].as<uint32_t>();
    GE_HIP(ctx, launch_block_meshes(st, s, uint32_t(bsegs.size()), reinterpret_cast<const BlockSeg*>(md + o_bs), md + o_li,
                                    s.blas_nodes.as<ge_bvh_node>(), s.blas_tris.as<ge_bvh_triangle>(),
                                    s.blas_morton.as<uint32_t>(), s.mesh_roots.as<float>(), s.mesh_height.as<uint32_t>(),
                                    s.blas_wide.as<WideNode>()));
    if (n > 0) {
      GE_HIP(ctx, ctx->tmp[1].reserve(size_t(n) * sizeof(ge_bvh_triangle)));  // records in batch order
      GE_HIP(ctx, ctx->tmp[2].reserve(size_t(n) * 24));                       // boxes
      GE_HIP(ctx, ctx->tmp[3].reserve(size_t(n) * 4));                        // segment of each primitive
      GE_HIP(ctx, ctx->tmp[4].reserve(size_t(n) * 4));                        // sorted order
      GE_HIP(ctx, ctx->tmp[5].reserve(size_t(n) * 4));                        // sorted codes
      // The scan scratch at the size the sort's scan needs later, so no reserve frees it under the
      // cluster scan queued before.
      GE_HIP(ctx, ctx->build.scan_tmp.reserve(size_t(n) / 1024 + size_t(ncl_global) / 256 + 65536 + 256));
      GE_HIP(ctx, launch_triangle_records(st, s, ctx->build, nseg_global, ncl_global, md + o_sp,
                                          reinterpret_cast<const GlobalSeg*>(md + o_gs), md + o_li,
                                          ctx->tmp[1].as<ge_bvh_triangle>(), ctx->tmp[2].as<float>(), ctx->tmp[3].as<uint32_t>()));
      GE_HIP(ctx, build_segments(ctx->build, st, n, ctx->tmp[2].as<float>(), nseg_global, ctx->tmp[3].as<uint32_t>(),
                                 md + o_sf, md + o_sc, md + o_sn, s.blas_nodes.as<ge_bvh_node>(), ctx->tmp[4].as<uint32_t>(),
                                 ctx->tmp[5].as<uint32_t>(), s.mesh_roots.as<float>(), s.mesh_height.as<uint32_t>(), md + o_sr,
                                 uint32_t(node_all), s.blas_wide.as<WideNode>(), nseg_global, n,
                                 seg_count.data(), flag, ncl_global));
      GE_HIP(ctx, launch_gather_triangles(st, n, ctx->tmp[4].as<uint32_t>(), ctx->tmp[3].as<uint32_t>(), md + o_sf, md + o_so,
                                          ctx->tmp[1].as<ge_bvh_triangle>(), s.blas_tris.as<ge_bvh_triangle>(),
                                          ctx->tmp[5].as<uint32_t>(), s.blas_morton.as<uint32_t>()));
    }
  }

  // ---- the instances: the records the walks read, recomputed where they changed ----
  const bool eye_moved = std::memcmp(&s.eye, &s.rec_eye, sizeof(ge_frame_eye)) != 0;
  if (flags == GE_RT_BUILD_FULL || s.rec_all || eye_moved || s.h_recs.size() != ni) {
    s.h_recs.assign(ni, InstRec{});
    s.h_T.assign(size_t(std::max(ni, 1u)) * 3, 0.0f);
    s.rec_lo = 0;
    s.rec_hi = ni;
    s.live_dirty = true;
  }
  const uint32_t rlo = std::min(s.rec_lo, ni), rhi = std::min(s.rec_hi, ni);
  for (uint32_t i = rlo; i < rhi; ++i) {
    const ge_instance_desc& in = s.h_instances[i];
    for (int r = 0; r < 3; ++r)
      s.h_T[size_t(i) * 3 + r] =
          relative_axis_host(in.cell[r], in.rows[r][3], s.eye.cell[r], s.eye.local[r], s.eye.residual[r]);
    InstRec& rec = s.h_recs[i];
    make_w2o(in, &s.h_T[size_t(i) * 3], rec.w2o);
    const uint32_t k = s.frame_of[i];
    rec.node_base = k == 0xffffffffu ? s.mesh_node_base[in.mesh] : s.frame_node_base[k];
    rec.tri_base = k == 0xffffffffu ? s.mesh_tri_base[in.mesh] : s.frame_tri_base[k];
    rec.first_pair = in.first_pair;
    rec.mask = instance_ray_mask(in.flags);
  }
  GE_HIP(ctx, s.inst_recs.reserve(size_t(std::max(ni, 1u)) * sizeof(InstRec)));
  GE_HIP(ctx, s.inst_T.reserve(size_t(std::max(ni, 1u)) * 12));
  if (rhi > rlo) {
    GE_HIP(ctx, staged_copy(ctx, s.inst_recs.as<InstRec>() + rlo, s.h_recs.data() + rlo, size_t(rhi - rlo) * sizeof(InstRec), st));
    GE_HIP(ctx, staged_copy(ctx, s.inst_T.as<float>() + size_t(rlo) * 3, s.h_T.data() + size_t(rlo) * 3,
                            size_t(rhi - rlo) * 12, st));
  }
  if (s.live_dirty) {
    s.h_live = live;
    GE_HIP(ctx, staged_upload(ctx, s.live_ids, s.h_live.data(), s.h_live.size(), st));
  }
  s.rec_lo = 0xffffffffu;
  s.rec_hi = 0;
  s.rec_all = false;
  s.live_dirty = false;
  s.rec_eye = s.eye;
  const uint32_t nl = uint32_t(s.h_live.size());
  s.tlas_prims = nl;
  s.tlas_node_count = nl == 0 ? 0u : (nl == 1 ? 1u : nl - 1);
  GE_HIP(ctx, s.tlas_nodes.reserve(size_t(std::max(s.tlas_node_count, 1u)) * sizeof(ge_bvh_node)));
  GE_HIP(ctx, s.tlas_wide.reserve(size_t(std::max(s.tlas_node_count, 1u)) * sizeof(WideNode)));
  GE_HIP(ctx, s.tlas_slots.reserve(size_t(std::max(nl, 1u)) * 4));
  GE_HIP(ctx, s.tlas_morton.reserve(size_t(std::max(nl, 1u)) * 4));
  GE_HIP(ctx, s.tlas_root.reserve(32));
  GE_HIP(ctx, s.tlas_height.reserve(4));
  const uint32_t* root_of = s.root_of.as<uint32_t>();
  const bool refit = flags == GE_RT_BUILD_REFIT && nl >= 2;
  if (refit) {
    // The slots' boxes from the instances' current placement, then the structure refitted.
    std::vector<uint32_t>& tmeta = ctx->host_tmeta;
    tmeta = {0u, nl, 0u};
    GE_HIP(ctx, staged_upload(ctx, ctx->tmp[7], tmeta.data(), tmeta.size(), st));
    GE_HIP(ctx, ctx->tmp[6].reserve(size_t(nl) * 24 + 256));
    GE_HIP(ctx, ctx->tmp[8].reserve(size_t(nl) * 4));
    float* boxes = ctx->tmp[6].as<float>();
    GE_HIP(ctx, launch_instance_boxes(st, nl, s.tlas_slots.as<uint32_t>(), s.instances.as<ge_instance_desc>(),
                                      s.inst_T.as<float>(), s.mesh_roots.as<float>(), root_of, boxes,
                                      ctx->tmp[8].as<uint32_t>()));
    GE_HIP(ctx, refit_segment(ctx->build, st, nl, boxes, ctx->tmp[7].as<uint32_t>(), s.tlas_nodes.as<ge_bvh_node>(),
                              s.tlas_wide.as<WideNode>(), s.tlas_root.as<float>(), s.tlas_height.as<uint32_t>(), flag));
  } else if (nl > 0) {
    std::vector<uint32_t>& tmeta = ctx->host_tmeta;
    tmeta = {0u, nl, 0u};  // seg_first, seg_count, node_base
    GE_HIP(ctx, ctx->tmp[6].reserve(size_t(nl) * 24 + 256));
    float* boxes = ctx->tmp[6].as<float>();
    const float* d_T = s.inst_T.as<float>();
    GE_HIP(ctx, staged_upload(ctx, ctx->tmp[7], tmeta.data(), tmeta.size(), st));
    GE_HIP(ctx, ctx->tmp[8].reserve(size_t(nl) * 12));
    uint32_t* seg_of = ctx->tmp[8].as<uint32_t>();
    uint32_t* perm = seg_of + nl;
    uint32_t* codes = perm + nl;
    const uint32_t* tm = ctx->tmp[7].as<uint32_t>();
    GE_HIP(ctx, launch_instance_boxes(st, nl, s.live_ids.as<uint32_t>(), s.instances.as<ge_instance_desc>(), d_T,
                                      s.mesh_roots.as<float>(), root_of, boxes, seg_of));
    GE_HIP(ctx, build_segments(ctx->build, st, nl, boxes, 1, seg_of, tm, tm + 1, tm + 2, s.tlas_nodes.as<ge_bvh_node>(), perm,
                               codes, s.tlas_root.as<float>(), s.tlas_height.as<uint32_t>(), nullptr, s.tlas_node_count,
                               s.tlas_wide.as<WideNode>(), nl > k_block_build_max ? 1u : 0u,
                               nl > k_block_build_max ? nl : 0u, &nl, flag));
    GE_HIP(ctx, launch_gather_slots(st, nl, perm, s.live_ids.as<uint32_t>(), s.tlas_slots.as<uint32_t>()));
    GE_HIP(ctx, hipMemcpyAsync(s.tlas_morton.ptr, codes, size_t(nl) * 4, hipMemcpyDeviceToDevice, st));
  }
  GE_HIP(ctx, hipEventRecord(ctx->bev1, st));
  uint32_t* rb = ctx->host_readback;
  rb[1] = 0;
  GE_HIP(ctx, hipMemcpyAsync(rb, flag, 4, hipMemcpyDeviceToHost, st));
  if (nl > 0) GE_HIP(ctx, hipMemcpyAsync(rb + 1, s.tlas_height.ptr, 4, hipMemcpyDeviceToHost, st));
  if (roots) GE_HIP(ctx, hipMemcpyAsync(rb + 2, s.mesh_height.ptr, size_t(roots) * 4, hipMemcpyDeviceToHost, st));
  GE_HIP(ctx, hipStreamSynchronize(st));
  float ms = 0.0f;
  GE_HIP(ctx, hipEventElapsedTime(&ms, ctx->bev0, ctx->bev1));
  ctx->last_build_ms = ms;
  if (rb[0]) s.built = false;
  GE_CHECK(ctx, !rb[0], GE_RT_ERROR_UNSUPPORTED,
           "a structure deeper than 254 levels, or too unbalanced for the build's tiles");
  s.visible_dirty = false;
  // The traversal stack holds at most three deferred children per wide level (two binary levels) of
  // each structure plus the sentinel between them.
  s.h_mesh_height.assign(rb + 2, rb + 2 + roots);
  s.tlas_height_value = rb[1];
  uint32_t blas_max = 0;
  for (uint32_t m = 0; m < s.mesh_count; ++m)
    if (s.mesh_tri_count[m] > 0) blas_max = std::max(blas_max, s.h_mesh_height[m]);
  for (uint32_t k = 0; k < nf; ++k)
    if (s.frame_tri_count[k] > 0) blas_max = std::max(blas_max, s.h_mesh_height[s.mesh_count + k]);
  s.built = false;
  const uint32_t need = 3 * ((s.tlas_height_value + 1) / 2) + 3 * ((blas_max + 1) / 2) + 2;
  GE_CHECK(ctx, need <= uint32_t(k_stack), GE_RT_ERROR_UNSUPPORTED,
           "the structures are " + std::to_string(s.tlas_height_value) + " + " + std::to_string(blas_max) +
               " levels deep, past the traversal stack of " + std::to_string(k_stack));
  s.stack_cap = need;
  if (s.pairs_dirty) GE_HIP(ctx, refresh_pairs(s));
  s.built = true;
  s.version = next_version();
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_bvh_info(ge_rt_ctx* ctx, const ge_rt_scene* scene, uint32_t mesh, uint32_t* node_count, uint32_t* prim_count) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, scene && scene->dev.built, GE_RT_ERROR_INVALID_SCENE, "no structure built");
  const SceneDev& s = scene->dev;
  if (mesh == 0xffffffffu) {
    if (node_count) *node_count = s.tlas_node_count;
    if (prim_count) *prim_count = s.tlas_prims;
    return GE_RT_OK;
  }
  uint32_t nb, nn, tb, nt;
  GE_CHECK(ctx, structure_of(s, mesh, nb, nn, tb, nt), GE_RT_ERROR_INVALID_ARGUMENT,
           "no such structure (mesh index out of range, or an instance without a frame structure)");
  if (node_count) *node_count = nn;
  if (prim_count) *prim_count = nt;
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_bvh_download(ge_rt_ctx* ctx, const ge_rt_scene* scene, uint32_t mesh, ge_bvh_node* nodes,
                           ge_bvh_triangle* triangles, uint32_t* instance_slots, uint32_t* morton) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, scene && scene->dev.built, GE_RT_ERROR_INVALID_SCENE, "no structure built");
  const SceneDev& s = scene->dev;
  GE_HIP(ctx, hipSetDevice(ctx->device));
  if (mesh == 0xffffffffu) {
    if (nodes && s.tlas_node_count)
      GE_HIP(ctx, hipMemcpy(nodes, s.tlas_nodes.ptr, size_t(s.tlas_node_count) * sizeof(ge_bvh_node), hipMemcpyDeviceToHost));
    if (instance_slots && s.tlas_prims)
      GE_HIP(ctx, hipMemcpy(instance_slots, s.tlas_slots.ptr, size_t(s.tlas_prims) * 4, hipMemcpyDeviceToHost));
    if (morton && s.tlas_prims) GE_HIP(ctx, hipMemcpy(morton, s.tlas_morton.ptr, size_t(s.tlas_prims) * 4, hipMemcpyDeviceToHost));
    return GE_RT_OK;
  }
  uint32_t nb, nn, tb, np;
  GE_CHECK(ctx, structure_of(s, mesh, nb, nn, tb, np), GE_RT_ERROR_INVALID_ARGUMENT,
           "no such structure (mesh index out of range, or an instance without a frame structure)");
  if (nodes && nn)
    GE_HIP(ctx, hipMemcpy(nodes, s.blas_nodes.as<ge_bvh_node>() + nb, size_t(nn) * sizeof(ge_bvh_node), hipMemcpyDeviceToHost));
  if (triangles && np)
    GE_HIP(ctx, hipMemcpy(triangles, s.blas_tris.as<ge_bvh_triangle>() + tb, size_t(np) * sizeof(ge_bvh_triangle),
                          hipMemcpyDeviceToHost));
  if (morton && np) GE_HIP(ctx, hipMemcpy(morton, s.blas_morton.as<uint32_t>() + tb, size_t(np) * 4, hipMemcpyDeviceToHost));
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_bvh_download_wide(ge_rt_ctx* ctx, const ge_rt_scene* scene, uint32_t mesh, ge_bvh_wide_node* wide) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, scene && scene->dev.built, GE_RT_ERROR_INVALID_SCENE, "no structure built");
  GE_CHECK(ctx, wide != nullptr, GE_RT_ERROR_INVALID_ARGUMENT, "no output");
  const SceneDev& s = scene->dev;
  GE_HIP(ctx, hipSetDevice(ctx->device));
  if (mesh == 0xffffffffu) {
    if (s.tlas_node_count)
      GE_HIP(ctx, hipMemcpy(wide, s.tlas_wide.ptr, size_t(s.tlas_node_count) * sizeof(WideNode), hipMemcpyDeviceToHost));
    return GE_RT_OK;
  }
  uint32_t nb, nn, tb, np;
  GE_CHECK(ctx, structure_of(s, mesh, nb, nn, tb, np), GE_RT_ERROR_INVALID_ARGUMENT,
           "no such structure (mesh index out of range, or an instance without a frame structure)");
  if (nn) GE_HIP(ctx, hipMemcpy(wide, s.blas_wide.as<WideNode>() + nb, size_t(nn) * sizeof(WideNode), hipMemcpyDeviceToHost));
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

}  // extern "C"

namespace {

// Host-pointer staging: device pointer for a caller buffer, copied in when `in` is set.
template <class T>
int32_t stage(ge_rt_ctx* ctx, int slot, const T* host, size_t count, bool copy_in, T** dev) {
  GE_HIP(ctx, ctx->io[slot].reserve(std::max<size_t>(count * sizeof(T), 256)));
  *dev = ctx->io[slot].as<T>();
  if (copy_in && count) GE_HIP(ctx, hipMemcpyAsync(*dev, host, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  return GE_RT_OK;
}

// Opens a dispatch's timing: its begin event, and no stage launches timed yet.
hipError_t begin_timing(ge_rt_ctx* ctx) {
  ctx->stage_n[0] = ctx->stage_n[1] = 0;
  ctx->stage_ms[0] = ctx->stage_ms[1] = 0.0;
  ctx->stage_launches[0] = ctx->stage_launches[1] = 0;
  return hipEventRecord(ctx->ev0, ctx->stream);
}

int32_t finish(ge_rt_ctx* ctx, bool sync) {
  GE_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  ctx->timing_pending = true;
  if (sync) GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GE_RT_OK;
}

int32_t ready(ge_rt_ctx* ctx, const ge_rt_scene* scene) {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, scene && scene->dev.built, GE_RT_ERROR_INVALID_SCENE, "no structure built: call ge_rt_bvh_build");
  GE_HIP(ctx, hipSetDevice(ctx->device));
  return GE_RT_OK;
}

// What both shadow resolves check before they touch the device: the parameters, the extent, and
// the pointers; `pixels` and `entries` receive the visibility buffer's size and the answers'.
int32_t resolve_checks(ge_rt_ctx* ctx, const ge_rt_scene* scene, const ge_shadow_resolve_params* params,
                       const uint64_t* visibility, const uint32_t* out, uint64_t* pixels, uint64_t* entries) {
  int32_t r = ready(ctx, scene);
  if (r != GE_RT_OK) return r;
  GE_CHECK(ctx, params, GE_RT_ERROR_INVALID_ARGUMENT, "null parameters");
  GE_CHECK(ctx, (params->flags & ~(GE_SHADOW_HALF_RATE | GE_SHADOW_JITTER | GE_SHADOW_ALPHA_MASKS)) == 0,
           GE_RT_ERROR_UNSUPPORTED, "unknown shadow resolve flags");
  *pixels = uint64_t(params->width) * params->height;
  const bool half = params->flags & GE_SHADOW_HALF_RATE;
  *entries = half ? uint64_t((params->width + 1) / 2) * ((params->height + 1) / 2) : *pixels;
  GE_CHECK(ctx, *pixels < (1ull << 31), GE_RT_ERROR_INVALID_ARGUMENT, "more than 2^31 pixels");
  GE_CHECK(ctx, (visibility && out) || *entries == 0, GE_RT_ERROR_INVALID_ARGUMENT, "null visibility or output");
  return GE_RT_OK;
}

}  // namespace

extern "C" {

int32_t ge_rt_trace_closest(ge_rt_ctx* ctx, const ge_rt_scene* scene, const ge_ray* rays, uint32_t count, uint32_t mask,
                            ge_hit* hits, uint32_t flags) try {
  int32_t r = ready(ctx, scene);
  if (r != GE_RT_OK) return r;
  GE_CHECK(ctx, (rays && hits) || count == 0, GE_RT_ERROR_INVALID_ARGUMENT, "null rays or hits");
  const bool dev = flags & GE_RT_DEVICE_PTR;
  const ge_ray* d_rays = rays;
  ge_hit* d_hits = hits;
  if (!dev) {
    ge_ray* dr;
    if ((r = stage(ctx, 0, rays, count, true, &dr)) != GE_RT_OK) return r;
    d_rays = dr;
    if ((r = stage(ctx, 1, hits, count, false, &d_hits)) != GE_RT_OK) return r;
  }
  GE_HIP(ctx, begin_timing(ctx));
  GE_CHECK(ctx, (flags & (GE_RT_ALPHA_MASKS | GE_RT_CULL_MASKED)) != (GE_RT_ALPHA_MASKS | GE_RT_CULL_MASKED),
           GE_RT_ERROR_INVALID_ARGUMENT, "GE_RT_ALPHA_MASKS and GE_RT_CULL_MASKED together");
  GE_HIP(ctx, launch_trace_closest(ctx->stream, bvh_view(scene->dev), d_rays, count, mask, d_hits,
                                   (flags & GE_RT_COUNT) ? ctx->counters.as<unsigned long long>() : nullptr,
                                   ctx->counters.as<uint32_t>() + 16, ctx->persistent_blocks,
                                   alpha_rule(scene->dev, flags),
                                   GeomMap{scene->dev.frame_inst.empty() ? nullptr : scene->dev.geom_base.as<uint32_t>(),
                                           scene->dev.entry_of.as<uint32_t>()}));
  if ((r = finish(ctx, false)) != GE_RT_OK) return r;
  if (!dev) {
    GE_HIP(ctx, hipMemcpyAsync(hits, d_hits, size_t(count) * sizeof(ge_hit), hipMemcpyDeviceToHost, ctx->stream));
    GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_trace_any(ge_rt_ctx* ctx, const ge_rt_scene* scene, const ge_ray* rays, uint32_t count, uint32_t mask,
                        uint32_t* occluded, uint32_t flags) try {
  int32_t r = ready(ctx, scene);
  if (r != GE_RT_OK) return r;
  GE_CHECK(ctx, (rays && occluded) || count == 0, GE_RT_ERROR_INVALID_ARGUMENT, "null rays or output");
  const bool dev = flags & GE_RT_DEVICE_PTR;
  const size_t words = (size_t(count) + 31) / 32;
  const ge_ray* d_rays = rays;
  uint32_t* d_bits = occluded;
  if (!dev) {
    ge_ray* dr;
    if ((r = stage(ctx, 0, rays, count, true, &dr)) != GE_RT_OK) return r;
    d_rays = dr;
    if ((r = stage(ctx, 1, occluded, words, false, &d_bits)) != GE_RT_OK) return r;
  }
  GE_HIP(ctx, begin_timing(ctx));
  GE_CHECK(ctx, !(flags & GE_RT_CULL_MASKED), GE_RT_ERROR_INVALID_ARGUMENT, "an any-hit query does not cull");
  GE_HIP(ctx, launch_trace_any(ctx->stream, bvh_view(scene->dev), d_rays, count, mask, d_bits,
                               (flags & GE_RT_COUNT) ? ctx->counters.as<unsigned long long>() : nullptr,
                               ctx->counters.as<uint32_t>() + 16, ctx->persistent_blocks,
                               alpha_rule(scene->dev, flags)));
  if ((r = finish(ctx, false)) != GE_RT_OK) return r;
  if (!dev) {
    GE_HIP(ctx, hipMemcpyAsync(occluded, d_bits, words * 4, hipMemcpyDeviceToHost, ctx->stream));
    GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_ray_visibility(ge_rt_ctx* ctx, const ge_rt_scene* scene, const ge_ray_visibility_params* params,
                             uint64_t* words, uint32_t flags) try {
  int32_t r = ready(ctx, scene);
  if (r != GE_RT_OK) return r;
  GE_CHECK(ctx, params, GE_RT_ERROR_INVALID_ARGUMENT, "null params");
  GE_CHECK(ctx, params->masks[3] == 0, GE_RT_ERROR_UNSUPPORTED, "the cross-fade (MaskTables::fade) is not supported");
  GE_CHECK(ctx, params->width > 0 && params->height > 0, GE_RT_ERROR_INVALID_ARGUMENT, "empty extent");
  const size_t n = size_t(params->width) * params->height;
  const bool dev = flags & GE_RT_DEVICE_PTR;
  uint64_t* out = dev ? (words ? words : reinterpret_cast<uint64_t*>(params->output)) : nullptr;
  GE_CHECK(ctx, !dev || out, GE_RT_ERROR_INVALID_ARGUMENT, "null output");
  GE_CHECK(ctx, dev || words, GE_RT_ERROR_INVALID_ARGUMENT, "null output");
  if (!dev && (r = stage(ctx, 0, words, n, false, &out)) != GE_RT_OK) return r;
  GE_HIP(ctx, begin_timing(ctx));
  // A frame with mask tables decides masked candidates by their masks, one without culls them
  // (trace_main, ray_visibility.slang:130-177).
  const uint32_t alpha = alpha_rule(scene->dev, params->masks[0] != 0 ? GE_RT_ALPHA_MASKS : GE_RT_CULL_MASKED);
  GE_HIP(ctx, launch_ray_visibility(ctx->stream, bvh_view(scene->dev), *params, out,
                                    (flags & GE_RT_COUNT) ? ctx->counters.as<unsigned long long>() : nullptr,
                                    ctx->counters.as<uint32_t>() + 16, ctx->persistent_blocks, alpha));
  if ((r = finish(ctx, false)) != GE_RT_OK) return r;
  if (!dev) {
    GE_HIP(ctx, hipMemcpyAsync(words, out, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

}  // extern "C"

namespace {

// A path-trace dispatch's checks, its parameter block with the scene's addresses (as the
// renderer's GpuScene fills them), and the scene view over the block's lights.
int32_t path_setup(ge_rt_ctx* ctx, const ge_rt_scene* scene, const ge_path_trace_params* params,
                   ge_path_trace_params* p, SceneView* svl) {
  int32_t r = ready(ctx, scene);
  if (r != GE_RT_OK) return r;
  GE_CHECK(ctx, params, GE_RT_ERROR_INVALID_ARGUMENT, "null params");
  GE_CHECK(ctx, params->width > 0 && params->height > 0, GE_RT_ERROR_INVALID_ARGUMENT, "empty extent");
  GE_CHECK(ctx, params->sky_params == 0, GE_RT_ERROR_UNSUPPORTED, "the physical sky (sky_params) is not supported");
  GE_CHECK(ctx, params->ground_detail == 0, GE_RT_ERROR_UNSUPPORTED, "ground detail is not supported");
  GE_CHECK(ctx, params->light_count <= scene->dev.light_count, GE_RT_ERROR_INVALID_ARGUMENT,
           "light_count exceeds the scene's uploaded lights");
  *p = *params;
  const SceneView sv = scene_view(scene->dev);
  p->clusters = reinterpret_cast<uint64_t>(sv.clusters);
  p->mesh = reinterpret_cast<uint64_t>(sv.meshes);
  p->triangles = reinterpret_cast<uint64_t>(sv.triangles);
  p->instances = reinterpret_cast<uint64_t>(sv.instances);
  p->attributes = reinterpret_cast<uint64_t>(sv.attributes);
  p->materials = reinterpret_cast<uint64_t>(sv.materials);
  p->cluster_materials = reinterpret_cast<uint64_t>(sv.cluster_materials);
  p->lights = reinterpret_cast<uint64_t>(sv.lights);
  *svl = sv;
  svl->light_count = params->light_count;
  return GE_RT_OK;
}

// A path-trace dispatch's parameters as an adjoint's primal pass would see them: the addresses the
// backend fills in or the caller's outputs cleared.
ge_path_trace_params primal_params(const ge_path_trace_params* params) {
  ge_path_trace_params k = *params;
  k.accum = k.output = k.coverage = 0;
  k.clusters = k.mesh = k.triangles = k.instances = k.attributes = k.materials = k.cluster_materials = 0;
  k.visible = k.lights = k.hit_entries = k.extra = 0;
  return k;
}

// The accumulator and coverage a path-trace dispatch adds into: the caller's device pointers (or
// the block's), or staging copies of host buffers (copied in when resuming at sample_base > 0).
int32_t path_outputs(ge_rt_ctx* ctx, const ge_path_trace_params* params, float* accum, uint32_t* coverage,
                     uint32_t flags, float4** d_accum, uint32_t** d_cov) {
  const size_t n = size_t(params->width) * params->height;
  if (flags & GE_RT_DEVICE_PTR) {
    *d_accum = reinterpret_cast<float4*>(accum ? accum : reinterpret_cast<float*>(params->accum));
    *d_cov = coverage ? coverage : reinterpret_cast<uint32_t*>(params->coverage);
    GE_CHECK(ctx, *d_accum && *d_cov, GE_RT_ERROR_INVALID_ARGUMENT, "null accumulator or coverage");
    return GE_RT_OK;
  }
  GE_CHECK(ctx, accum && coverage, GE_RT_ERROR_INVALID_ARGUMENT, "null accumulator or coverage");
  const bool resume = params->sample_base != 0;
  float* da;
  int32_t r;
  if ((r = stage(ctx, 0, accum, n * 4, resume, &da)) != GE_RT_OK) return r;
  *d_accum = reinterpret_cast<float4*>(da);
  return stage(ctx, 1, coverage, n, resume, d_cov);
}

int32_t path_copy_out(ge_rt_ctx* ctx, const ge_path_trace_params* params, float* accum, uint32_t* coverage,
                      uint32_t flags, const float4* d_accum, const uint32_t* d_cov) {
  if (flags & GE_RT_DEVICE_PTR) return GE_RT_OK;
  const size_t n = size_t(params->width) * params->height;
  GE_HIP(ctx, hipMemcpyAsync(accum, d_accum, n * 16, hipMemcpyDeviceToHost, ctx->stream));
  GE_HIP(ctx, hipMemcpyAsync(coverage, d_cov, n * 4, hipMemcpyDeviceToHost, ctx->stream));
  GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GE_RT_OK;
}

}  // namespace

extern "C" {

int32_t ge_rt_path_trace(ge_rt_ctx* ctx, const ge_rt_scene* scene, const ge_path_trace_params* params, float* accum,
                         uint32_t* coverage, uint32_t flags) try {
  ge_path_trace_params p;
  SceneView svl;
  int32_t r = path_setup(ctx, scene, params, &p, &svl);
  if (r != GE_RT_OK) return r;
  float4* d_accum;
  uint32_t* d_cov;
  if ((r = path_outputs(ctx, params, accum, coverage, flags, &d_accum, &d_cov)) != GE_RT_OK) return r;
  p.accum = reinterpret_cast<uint64_t>(d_accum);
  p.coverage = reinterpret_cast<uint64_t>(d_cov);
  GE_HIP(ctx, begin_timing(ctx));
  unsigned long long* counters = (flags & GE_RT_COUNT) ? ctx->counters.as<unsigned long long>() : nullptr;
  ctx->primal.valid = false;
  if (flags & GE_RT_PT_MEGAKERNEL) {
    GE_HIP(ctx, launch_path_trace(ctx->stream, bvh_view(scene->dev), svl, p, d_accum, d_cov, counters, full_scene(svl)));
  } else {
    GE_HIP(ctx, wavefront_path_trace(ctx, bvh_view(scene->dev), svl, p, d_accum, d_cov, counters));
    // Every path's radiance is in the wavefront's result buffer until the next wavefront dispatch.
    ctx->primal.valid = true;
    ctx->primal.scene = scene;
    ctx->primal.version = scene->dev.version;
    ctx->primal.params = primal_params(params);
  }
  if ((r = finish(ctx, false)) != GE_RT_OK) return r;
  return path_copy_out(ctx, params, accum, coverage, flags, d_accum, d_cov);
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_scene_update_materials(ge_rt_ctx* ctx, ge_rt_scene* scene, uint32_t first, uint32_t count,
                                     const ge_resolve_material* materials) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, scene && (materials || !count), GE_RT_ERROR_INVALID_ARGUMENT, "null scene or materials");
  SceneDev& s = scene->dev;
  GE_CHECK(ctx, uint64_t(first) + count <= s.material_count, GE_RT_ERROR_INVALID_ARGUMENT, "material range out of bounds");
  const int32_t st = validate_materials(ctx, s, materials, count, first);
  if (st != GE_RT_OK) return st;
  if (!count) return GE_RT_OK;
  GE_HIP(ctx, hipSetDevice(ctx->device));
  if (std::memcmp(&s.h_materials[first], materials, size_t(count) * sizeof(ge_resolve_material)) == 0) return GE_RT_OK;
  s.version = next_version();
  std::memcpy(&s.h_materials[first], materials, size_t(count) * sizeof(ge_resolve_material));
  GE_HIP(ctx, hipMemcpyAsync(s.materials.as<ge_resolve_material>() + first, materials,
                             size_t(count) * sizeof(ge_resolve_material), hipMemcpyHostToDevice, ctx->stream));
  GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  GE_HIP(ctx, refresh_masks(s, ctx->stream));
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_scene_update_lights(ge_rt_ctx* ctx, ge_rt_scene* scene, uint32_t first, uint32_t count,
                                  const ge_resolve_light* lights) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, scene && (lights || !count), GE_RT_ERROR_INVALID_ARGUMENT, "null scene or lights");
  SceneDev& s = scene->dev;
  GE_CHECK(ctx, uint64_t(first) + count <= s.light_count, GE_RT_ERROR_INVALID_ARGUMENT, "light range out of bounds");
  if (!count) return GE_RT_OK;
  GE_HIP(ctx, hipSetDevice(ctx->device));
  if (std::memcmp(&s.h_lights[first], lights, size_t(count) * sizeof(ge_resolve_light)) == 0) return GE_RT_OK;
  s.version = next_version();
  // The light grid reads lights 0..63's positions, radii and kinds: an update of colours and
  // intensities alone (the inverse-rendering loop's) keeps it.
  bool placed = false;
  for (uint32_t i = 0; i < count && first + i < 64u && !placed; ++i) {
    const ge_resolve_light& was = s.h_lights[first + i];
    placed = std::memcmp(was.position_radius, lights[i].position_radius, sizeof(was.position_radius)) != 0 ||
             std::memcmp(&was.direction_type[3], &lights[i].direction_type[3], sizeof(float)) != 0;
  }
  std::memcpy(&s.h_lights[first], lights, size_t(count) * sizeof(ge_resolve_light));
  GE_HIP(ctx, hipMemcpyAsync(s.lights.as<ge_resolve_light>() + first, lights, size_t(count) * sizeof(ge_resolve_light),
                             hipMemcpyHostToDevice, ctx->stream));
  GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (placed) GE_HIP(ctx, build_light_grid(s, ctx->stream));
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_path_adjoint(ge_rt_ctx* ctx, const ge_rt_scene* scene, const ge_path_trace_params* params,
                           const float* d_image, const ge_resolve_material* sampling, float* image,
                           double* grad_materials, double* grad_lights, uint32_t flags) try {
  int32_t r = ready(ctx, scene);
  if (r != GE_RT_OK) return r;
  GE_CHECK(ctx, params, GE_RT_ERROR_INVALID_ARGUMENT, "null params");
  GE_CHECK(ctx, params->width > 0 && params->height > 0 && params->spp > 0, GE_RT_ERROR_INVALID_ARGUMENT,
           "empty extent or no samples");
  GE_CHECK(ctx, !(params->flags & GE_PT_ALPHA_MASKS), GE_RT_ERROR_UNSUPPORTED, "alpha-tested materials are not supported");
  GE_CHECK(ctx, params->sky_params == 0 && params->ground_detail == 0, GE_RT_ERROR_UNSUPPORTED,
           "the physical sky and ground detail are not supported");
  GE_CHECK(ctx, params->light_count <= scene->dev.light_count, GE_RT_ERROR_INVALID_ARGUMENT,
           "light_count exceeds the scene's uploaded lights");
  GE_CHECK(ctx, params->max_bounces <= 15, GE_RT_ERROR_UNSUPPORTED, "the adjoint records at most 15 bounces");
  GE_CHECK(ctx, !d_image || (grad_materials && grad_lights), GE_RT_ERROR_INVALID_ARGUMENT,
           "a loss gradient needs both gradient outputs");
  const SceneDev& s = scene->dev;
  for (const ge_resolve_material& m : s.h_materials)
    GE_CHECK(ctx, m.specular[0] == 0.0f && m.specular[1] == 0.0f && m.specular[2] == 0.0f && m.specular[3] == 0.0f,
             GE_RT_ERROR_UNSUPPORTED, "the adjoint covers materials without their own specular record");
  GE_CHECK(ctx, !s.any_masked && !s.any_textured, GE_RT_ERROR_UNSUPPORTED,
           "the adjoint covers untextured, opaque materials");
  const size_t n = size_t(params->width) * params->height;
  const size_t nm = size_t(s.material_count) * 4, nl = size_t(params->light_count) + 1;
  const bool dev = flags & GE_RT_DEVICE_PTR;
  const float* dd = d_image;
  const ge_resolve_material* dsamp = sampling;
  float* dimg = image;
  double* dgm = grad_materials;
  double* dgl = grad_lights;
  if (!dev) {
    float* t;
    if (d_image) {
      if ((r = stage(ctx, 0, d_image, n * 4, true, &t)) != GE_RT_OK) return r;
      dd = t;
    }
    if (sampling) {
      ge_resolve_material* sm;
      if ((r = stage(ctx, 1, sampling, size_t(s.material_count), true, &sm)) != GE_RT_OK) return r;
      dsamp = sm;
    }
    if (image && (r = stage(ctx, 2, image, n * 4, false, &dimg)) != GE_RT_OK) return r;
    if (d_image) {
      if ((r = stage(ctx, 3, grad_materials, nm, false, &dgm)) != GE_RT_OK) return r;
      if ((r = stage(ctx, 4, grad_lights, nl, false, &dgl)) != GE_RT_OK) return r;
    }
  }
  if (d_image) {
    GE_HIP(ctx, hipMemsetAsync(dgm, 0, nm * 8, ctx->stream));
    GE_HIP(ctx, hipMemsetAsync(dgl, 0, nl * 8, ctx->stream));
  }
  ge_path_trace_params p = *params;
  SceneView sv = scene_view(s);
  sv.light_count = params->light_count;
  GE_HIP(ctx, begin_timing(ctx));
  if (flags & GE_RT_PT_MEGAKERNEL) {
    GE_HIP(ctx, launch_path_adjoint(ctx->stream, bvh_view(s), sv, p, dsamp, reinterpret_cast<const float4*>(dd),
                                    reinterpret_cast<float4*>(dimg), d_image ? dgm : nullptr, d_image ? dgl : nullptr));
  } else {
    // The wavefront: a primal pass (the detached estimate of every path, and the image), then with
    // d_image a gradient pass that replays the same paths.
    const uint64_t npaths = uint64_t(p.width) * p.height * p.spp;
    GE_CHECK(ctx, npaths < (1ull << 31), GE_RT_ERROR_INVALID_ARGUMENT, "too many paths in one dispatch");
    p.clusters = reinterpret_cast<uint64_t>(sv.clusters);
    p.mesh = reinterpret_cast<uint64_t>(sv.meshes);
    p.triangles = reinterpret_cast<uint64_t>(sv.triangles);
    p.instances = reinterpret_cast<uint64_t>(sv.instances);
    p.attributes = reinterpret_cast<uint64_t>(sv.attributes);
    p.materials = reinterpret_cast<uint64_t>(sv.materials);
    p.cluster_materials = reinterpret_cast<uint64_t>(sv.cluster_materials);
    p.lights = reinterpret_cast<uint64_t>(sv.lights);
    AdjointHook h{};
    h.sampling = dsamp;
    const ge_path_trace_params key = primal_params(params);
    const bool reuse = (flags & GE_RT_ADJOINT_REUSE_PRIMAL) && !sampling && d_image && !image && ctx->primal.valid &&
                       ctx->primal.scene == scene && ctx->primal.version == s.version &&
                       std::memcmp(&ctx->primal.params, &key, sizeof key) == 0;
    ctx->primal.valid = false;
    if (!reuse) {
      uint32_t* cov = nullptr;
      if (dimg) {
        GE_HIP(ctx, ctx->tmp[8].reserve(std::max<size_t>(n * 4, 256)));
        cov = ctx->tmp[8].as<uint32_t>();
        GE_HIP(ctx, hipMemsetAsync(dimg, 0, n * 16, ctx->stream));  // the image holds this dispatch's samples alone
      }
      GE_HIP(ctx, wavefront_path_trace(ctx, bvh_view(s), sv, p, reinterpret_cast<float4*>(dimg), cov, nullptr, nullptr, &h));
    }
    if (d_image) {
      Wavefront& w = ctx->wave;
      GE_HIP(ctx, w.lpath.reserve(npaths * 16));
      GE_HIP(ctx, hipMemcpyAsync(w.lpath.ptr, w.result.ptr, npaths * 16, hipMemcpyDeviceToDevice, ctx->stream));
      h.grad = true;
      h.d_image = reinterpret_cast<const float4*>(dd);
      h.lpath = w.lpath.as<float4>();
      h.grad_materials = dgm;
      h.grad_lights = dgl;
      GE_HIP(ctx, wavefront_path_trace(ctx, bvh_view(s), sv, p, nullptr, nullptr, nullptr, nullptr, &h));
    }
  }
  if ((r = finish(ctx, false)) != GE_RT_OK) return r;
  if (!dev) {
    if (image) GE_HIP(ctx, hipMemcpyAsync(image, dimg, n * 16, hipMemcpyDeviceToHost, ctx->stream));
    if (d_image) {
      GE_HIP(ctx, hipMemcpyAsync(grad_materials, dgm, nm * 8, hipMemcpyDeviceToHost, ctx->stream));
      GE_HIP(ctx, hipMemcpyAsync(grad_lights, dgl, nl * 8, hipMemcpyDeviceToHost, ctx->stream));
    }
    GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_tonemap(ge_rt_ctx* ctx, const ge_path_trace_params* params, const float* accum, const uint32_t* coverage,
                      uint32_t* output, uint32_t flags) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, params && accum && coverage && output, GE_RT_ERROR_INVALID_ARGUMENT, "null argument");
  GE_CHECK(ctx, params->sky_params == 0, GE_RT_ERROR_UNSUPPORTED, "the physical sky (sky_params) is not supported");
  GE_HIP(ctx, hipSetDevice(ctx->device));
  int32_t r;
  const size_t n = size_t(params->width) * params->height;
  const bool dev = flags & GE_RT_DEVICE_PTR;
  const float* da = accum;
  const uint32_t* dc = coverage;
  uint32_t* dout = output;
  if (!dev) {
    float* a;
    uint32_t* c;
    if ((r = stage(ctx, 0, accum, n * 4, true, &a)) != GE_RT_OK) return r;
    if ((r = stage(ctx, 1, coverage, n, true, &c)) != GE_RT_OK) return r;
    if ((r = stage(ctx, 2, output, n, false, &dout)) != GE_RT_OK) return r;
    da = a;
    dc = c;
  }
  GE_HIP(ctx, launch_tonemap(ctx->stream, *params, reinterpret_cast<const float4*>(da), dc, dout));
  if (!dev) GE_HIP(ctx, hipMemcpyAsync(output, dout, n * 4, hipMemcpyDeviceToHost, ctx->stream));
  GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_shadow_pass(ge_rt_ctx* ctx, const ge_rt_scene* scene, const ge_shadow_params* params,
                          const ge_shadow_receiver* receivers, uint32_t* blocked, uint32_t flags) try {
  int32_t r = ready(ctx, scene);
  if (r != GE_RT_OK) return r;
  GE_CHECK(ctx, params && ((receivers && blocked) || params->count == 0), GE_RT_ERROR_INVALID_ARGUMENT, "null argument");
  const uint32_t n = params->count;
  const size_t words = (size_t(n) + 31) / 32;
  const bool dev = flags & GE_RT_DEVICE_PTR;
  const ge_shadow_receiver* dr = receivers;
  uint32_t* db = blocked;
  if (!dev) {
    ge_shadow_receiver* rr;
    if ((r = stage(ctx, 0, receivers, n, true, &rr)) != GE_RT_OK) return r;
    dr = rr;
    if ((r = stage(ctx, 1, blocked, words, false, &db)) != GE_RT_OK) return r;
  }
  GE_HIP(ctx, begin_timing(ctx));
  GE_CHECK(ctx, !(flags & GE_RT_CULL_MASKED), GE_RT_ERROR_INVALID_ARGUMENT, "a shadow ray does not cull");
  GE_HIP(ctx, launch_shadow_pass(ctx->stream, bvh_view(scene->dev), scene_view(scene->dev), *params, dr, db,
                                 (flags & GE_RT_COUNT) ? ctx->counters.as<unsigned long long>() : nullptr,
                                 alpha_rule(scene->dev, flags)));
  if ((r = finish(ctx, false)) != GE_RT_OK) return r;
  if (!dev) {
    GE_HIP(ctx, hipMemcpyAsync(blocked, db, words * 4, hipMemcpyDeviceToHost, ctx->stream));
    GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_shadow_resolve(ge_rt_ctx* ctx, const ge_rt_scene* scene, const ge_shadow_resolve_params* params,
                             const uint64_t* visibility, uint32_t* out, uint32_t flags) try {
  uint64_t pixels = 0, entries = 0;
  int32_t r = resolve_checks(ctx, scene, params, visibility, out, &pixels, &entries);
  if (r != GE_RT_OK || entries == 0) return r;
  SceneDev& s = const_cast<SceneDev&>(scene->dev);
  if (s.pairs_dirty) GE_HIP(ctx, refresh_pairs(s));
  GE_CHECK(ctx, s.pairs_error.empty(), GE_RT_ERROR_INVALID_SCENE, s.pairs_error);
  const bool dev = flags & GE_RT_DEVICE_PTR;
  const uint64_t* dw = visibility;
  uint32_t* dout = out;
  if (!dev) {
    uint64_t* w;
    if ((r = stage(ctx, 0, visibility, size_t(pixels), true, &w)) != GE_RT_OK) return r;
    dw = w;
    if ((r = stage(ctx, 1, out, size_t(entries), false, &dout)) != GE_RT_OK) return r;
  }
  GE_HIP(ctx, ctx->resolve_rays.reserve(size_t(entries) * 16));
  GE_HIP(ctx, begin_timing(ctx));
  GE_HIP(ctx, launch_shadow_resolve(ctx, bvh_view(s), scene_view(s), s.eye, *params, dw,
                                    s.pairs.as<uint2>(), s.pair_count, ctx->resolve_rays.as<float4>(), dout,
                                    (flags & GE_RT_COUNT) ? ctx->counters.as<unsigned long long>() : nullptr,
                                    ctx->counters.as<uint32_t>() + 16, ctx->persistent_blocks,
                                    alpha_rule(s, (params->flags & GE_SHADOW_ALPHA_MASKS) ? GE_RT_ALPHA_MASKS : 0u)));
  if ((r = finish(ctx, false)) != GE_RT_OK) return r;
  if (!dev) {
    GE_HIP(ctx, hipMemcpyAsync(out, dout, size_t(entries) * 4, hipMemcpyDeviceToHost, ctx->stream));
    GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_shadow_resolve_noalloc(ge_rt_ctx* ctx, const ge_rt_scene* scene, const ge_shadow_resolve_params* params,
                                     const uint64_t* visibility, uint32_t* out, uint32_t flags) try {
  uint64_t pixels = 0, entries = 0;
  int32_t r = resolve_checks(ctx, scene, params, visibility, out, &pixels, &entries);
  if (r != GE_RT_OK || entries == 0) return r;
  GE_CHECK(ctx, flags & GE_RT_DEVICE_PTR, GE_RT_ERROR_INVALID_ARGUMENT,
           "the allocation-free shadow resolve takes device pointers only (GE_RT_DEVICE_PTR)");
  GE_CHECK(ctx, entries <= GE_RT_SHADOW_STATIC_ENTRIES, GE_RT_ERROR_INVALID_ARGUMENT,
           std::to_string(entries) + " entries exceed the allocation-free shadow resolve's " +
               std::to_string(GE_RT_SHADOW_STATIC_ENTRIES));
  const SceneDev& s = scene->dev;
  GE_CHECK(ctx, !s.pairs_dirty, GE_RT_ERROR_INVALID_SCENE,
           "instances changed since the last ge_rt_bvh_build, which rebuilds the pair table");
  GE_CHECK(ctx, s.pairs_error.empty(), GE_RT_ERROR_INVALID_SCENE, s.pairs_error);
  GE_CHECK(ctx, ctx->static_scratch, GE_RT_ERROR_DEVICE, "no allocation-free working memory on this device");
  GE_HIP(ctx, begin_timing(ctx));
  GE_HIP(ctx, launch_shadow_resolve(ctx, bvh_view(s), scene_view(s), s.eye, *params, visibility, s.pairs.as<uint2>(),
                                    s.pair_count, reinterpret_cast<float4*>(ctx->static_scratch + k_scratch_head), out,
                                    (flags & GE_RT_COUNT) ? ctx->counters.as<unsigned long long>() : nullptr,
                                    reinterpret_cast<uint32_t*>(ctx->static_scratch), ctx->persistent_blocks,
                                    alpha_rule(s, (params->flags & GE_SHADOW_ALPHA_MASKS) ? GE_RT_ALPHA_MASKS : 0u)));
  return finish(ctx, false);
} catch (...) {
  return host_failure(ctx);
}

// ---- the radiance cache ----------------------------------------------------------------------------
int32_t ge_rt_cache_create(ge_rt_ctx* ctx, const float box_min[3], const float box_max[3], uint32_t seed,
                           ge_rt_cache** out) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, out && box_min && box_max, GE_RT_ERROR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  for (int k = 0; k < 3; ++k)
    GE_CHECK(ctx, std::isfinite(box_min[k]) && std::isfinite(box_max[k]) && box_max[k] > box_min[k],
             GE_RT_ERROR_INVALID_ARGUMENT, "the cache's box is empty or not finite");
  GE_HIP(ctx, hipSetDevice(ctx->device));
  auto* c = new (std::nothrow) ge_rt_cache();
  GE_CHECK(ctx, c, GE_RT_ERROR_OUT_OF_MEMORY, "out of host memory");
  for (int k = 0; k < 3; ++k) {
    c->net.box.lo[k] = box_min[k];
    c->net.box.inv[k] = 1.0f / (box_max[k] - box_min[k]);
  }
  const hipError_t e = c->net.init(ctx->stream, seed);
  if (e != hipSuccess) {
    c->net.release();
    delete c;
    GE_HIP(ctx, e);
  }
  *out = c;
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

void ge_rt_cache_destroy(ge_rt_ctx* ctx, ge_rt_cache* cache) {
  if (!cache) return;
  if (ctx) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
  }
  cache->net.release();
  delete cache;
}

int32_t ge_rt_cache_weights(ge_rt_ctx* ctx, ge_rt_cache* cache, float* get, const float* set) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, cache, GE_RT_ERROR_INVALID_ARGUMENT, "null cache");
  GE_HIP(ctx, hipSetDevice(ctx->device));
  if (get) GE_HIP(ctx, cache->net.get_weights(ctx->stream, get));
  if (set) {
    for (uint32_t i = 0; i < GE_CACHE_WEIGHTS; ++i)
      GE_CHECK(ctx, std::isfinite(set[i]), GE_RT_ERROR_INVALID_ARGUMENT, "weights must be finite");
    GE_HIP(ctx, cache->net.set_weights(ctx->stream, set));
    GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_cache_gradient(ge_rt_ctx* ctx, ge_rt_cache* cache, float* out) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, cache && out, GE_RT_ERROR_INVALID_ARGUMENT, "null argument");
  GE_HIP(ctx, hipSetDevice(ctx->device));
  GE_HIP(ctx, cache->net.get_gradient(ctx->stream, out));
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_cache_infer(ge_rt_ctx* ctx, ge_rt_cache* cache, const ge_cache_query* queries, uint32_t n,
                          float* radiance, uint32_t flags) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, cache && (n == 0 || (queries && radiance)), GE_RT_ERROR_INVALID_ARGUMENT, "null argument");
  if (n == 0) return GE_RT_OK;
  GE_HIP(ctx, hipSetDevice(ctx->device));
  const bool dev = flags & GE_RT_DEVICE_PTR;
  const ge_cache_query* dq = queries;
  float* dout = radiance;
  int32_t r;
  if (!dev) {
    ge_cache_query* t;
    if ((r = stage(ctx, 0, queries, n, true, &t)) != GE_RT_OK) return r;
    dq = t;
    if ((r = stage(ctx, 1, radiance, size_t(n) * 3, false, &dout)) != GE_RT_OK) return r;
  }
  GE_HIP(ctx, begin_timing(ctx));
  GE_HIP(ctx, cache->net.infer(ctx->stream, reinterpret_cast<const float*>(dq), sizeof(ge_cache_query) / 4, n,
                               nullptr, dout));
  if ((r = finish(ctx, false)) != GE_RT_OK) return r;
  if (!dev) GE_HIP(ctx, hipMemcpyAsync(radiance, dout, size_t(n) * 12, hipMemcpyDeviceToHost, ctx->stream));
  GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

namespace {

int32_t read_loss(ge_rt_ctx* ctx, ge_rt_cache* cache, double* loss) {
  if (!loss) return GE_RT_OK;
  GE_HIP(ctx, hipMemcpyAsync(cache->net.host_loss, cache->net.loss.ptr, 8, hipMemcpyDeviceToHost, ctx->stream));
  GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *loss = *cache->net.host_loss;
  return GE_RT_OK;
}

bool finite_rate(float lr) { return std::isfinite(lr) && lr >= 0.0f; }

// Traces params' paths through the wavefront, recording a sample at every hit vertex into the
// cache's own buffer (capacity samples); *count receives how many vertices there were.
int32_t collect_samples(ge_rt_ctx* ctx, const ge_rt_scene* scene, ge_rt_cache* cache,
                        const ge_path_trace_params* params, uint32_t capacity, uint32_t* count) {
  ge_path_trace_params p;
  SceneView svl;
  int32_t r = path_setup(ctx, scene, params, &p, &svl);
  if (r != GE_RT_OK) return r;
  RadianceCache& net = cache->net;
  GE_HIP(ctx, net.samples.reserve(std::max<size_t>(size_t(capacity) * sizeof(ge_cache_sample), 256)));
  GE_HIP(ctx, net.sample_weights.reserve(std::max<size_t>(size_t(capacity) * 16, 256)));
  CacheHook hook{};
  hook.mode = CacheHook::kRecord;
  hook.box = net.box;
  hook.cache = &net;
  hook.samples = net.samples.as<ge_cache_sample>();
  hook.sample_weights = net.sample_weights.as<float4>();
  hook.sample_capacity = capacity;
  ctx->primal.valid = false;
  GE_HIP(ctx, wavefront_path_trace(ctx, bvh_view(scene->dev), svl, p, nullptr, nullptr, nullptr, &hook));
  GE_HIP(ctx, read_cache_counts(ctx, nullptr, count));
  net.kept = std::min(*count, capacity);
  return GE_RT_OK;
}

// `steps` optimizer steps over interleaved slices of the first n of the cache's samples: step s
// trains on samples s, s + k, s + 2k, ... so that every step sees every part of the frame.
hipError_t train_slices(ge_rt_ctx* ctx, RadianceCache& net, uint32_t n, uint32_t steps, float lr) {
  const uint32_t k = std::min(std::max(steps, 1u), std::max(n, 1u));
  const uint32_t per = n / k;
  if (per == 0) return hipSuccess;
  for (uint32_t st = 0; st < k; ++st) {
    const hipError_t e = net.train(ctx->stream, net.samples.as<ge_cache_sample>() + st, k, per, lr);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace

int32_t ge_rt_cache_train(ge_rt_ctx* ctx, ge_rt_cache* cache, const ge_cache_sample* samples, uint32_t n,
                          float learning_rate, double* loss, uint32_t flags) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, cache && (n == 0 || samples), GE_RT_ERROR_INVALID_ARGUMENT, "null argument");
  GE_CHECK(ctx, finite_rate(learning_rate), GE_RT_ERROR_INVALID_ARGUMENT, "the learning rate must be finite and >= 0");
  if (loss) *loss = 0.0;
  if (n == 0) return GE_RT_OK;
  GE_HIP(ctx, hipSetDevice(ctx->device));
  const ge_cache_sample* ds = samples;
  if (!(flags & GE_RT_DEVICE_PTR)) {
    ge_cache_sample* t;
    const int32_t r = stage(ctx, 0, samples, n, true, &t);
    if (r != GE_RT_OK) return r;
    ds = t;
  }
  GE_HIP(ctx, cache->net.train(ctx->stream, ds, 1, n, learning_rate));
  GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return read_loss(ctx, cache, loss);
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_cache_samples(ge_rt_ctx* ctx, ge_rt_cache* cache, ge_cache_sample* out, uint32_t capacity,
                            uint32_t* count, uint32_t flags) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, cache && count && (capacity == 0 || out), GE_RT_ERROR_INVALID_ARGUMENT, "null argument");
  GE_HIP(ctx, hipSetDevice(ctx->device));
  const uint32_t n = std::min(cache->net.kept, capacity);
  *count = n;
  if (n) {
    const hipMemcpyKind kind = (flags & GE_RT_DEVICE_PTR) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    GE_HIP(ctx, hipMemcpyAsync(out, cache->net.samples.ptr, size_t(n) * sizeof(ge_cache_sample), kind, ctx->stream));
  }
  GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_cache_collect(ge_rt_ctx* ctx, const ge_rt_scene* scene, ge_rt_cache* cache,
                            const ge_path_trace_params* params, ge_cache_sample* samples, uint32_t capacity,
                            uint32_t* count, uint32_t flags) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, cache && count && (capacity == 0 || samples), GE_RT_ERROR_INVALID_ARGUMENT, "null argument");
  GE_CHECK(ctx, !(flags & GE_RT_PT_MEGAKERNEL), GE_RT_ERROR_UNSUPPORTED, "the cache's paths run on the wavefront");
  *count = 0;
  int32_t r = collect_samples(ctx, scene, cache, params, capacity, count);
  if (r != GE_RT_OK) return r;
  const size_t kept = std::min(*count, capacity);
  if (kept) {
    const hipMemcpyKind kind = (flags & GE_RT_DEVICE_PTR) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    GE_HIP(ctx, hipMemcpyAsync(samples, cache->net.samples.ptr, kept * sizeof(ge_cache_sample), kind, ctx->stream));
  }
  GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_cache_update(ge_rt_ctx* ctx, const ge_rt_scene* scene, ge_rt_cache* cache,
                           const ge_path_trace_params* params, uint32_t samples, uint32_t steps, float learning_rate,
                           double* loss, uint32_t flags) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, cache && steps > 0, GE_RT_ERROR_INVALID_ARGUMENT, "null cache or no steps");
  GE_CHECK(ctx, finite_rate(learning_rate), GE_RT_ERROR_INVALID_ARGUMENT, "the learning rate must be finite and >= 0");
  GE_CHECK(ctx, !(flags & GE_RT_PT_MEGAKERNEL), GE_RT_ERROR_UNSUPPORTED, "the cache's paths run on the wavefront");
  if (loss) *loss = 0.0;
  uint32_t count = 0;
  int32_t r = collect_samples(ctx, scene, cache, params, samples, &count);
  if (r != GE_RT_OK) return r;
  GE_HIP(ctx, train_slices(ctx, cache->net, std::min(count, samples), steps, learning_rate));
  return read_loss(ctx, cache, loss);
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_path_trace_cached(ge_rt_ctx* ctx, const ge_rt_scene* scene, ge_rt_cache* cache,
                                const ge_path_trace_params* params, uint32_t depth,
                                const ge_cache_train_params* train, float* accum, uint32_t* coverage, double* loss,
                                uint32_t flags) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, cache, GE_RT_ERROR_INVALID_ARGUMENT, "null cache");
  GE_CHECK(ctx, !(flags & GE_RT_PT_MEGAKERNEL), GE_RT_ERROR_UNSUPPORTED, "the cached tracer runs on the wavefront");
  const bool training = train && train->samples > 0 && train->steps > 0;
  GE_CHECK(ctx, !train || finite_rate(train->learning_rate), GE_RT_ERROR_INVALID_ARGUMENT,
           "the learning rate must be finite and >= 0");
  if (loss) *loss = 0.0;
  ge_path_trace_params p;
  SceneView svl;
  int32_t r = path_setup(ctx, scene, params, &p, &svl);
  if (r != GE_RT_OK) return r;
  const uint64_t npaths = uint64_t(params->width) * params->height * params->spp;
  GE_CHECK(ctx, npaths < (1ull << 30), GE_RT_ERROR_INVALID_ARGUMENT, "too many paths in one dispatch");
  float4* d_accum;
  uint32_t* d_cov;
  if ((r = path_outputs(ctx, params, accum, coverage, flags, &d_accum, &d_cov)) != GE_RT_OK) return r;
  p.accum = reinterpret_cast<uint64_t>(d_accum);
  p.coverage = reinterpret_cast<uint64_t>(d_cov);
  RadianceCache& net = cache->net;
  CacheHook hook{};
  hook.mode = CacheHook::kCached;
  hook.depth = depth;
  hook.box = net.box;
  hook.cache = &net;
  if (training) {
    uint32_t stride = train->stride;
    if (stride == 0) {
      const double want = double(train->samples) / std::max(double(net.hit_rate), 1e-3);
      stride = uint32_t(std::max(1.0, std::floor(double(npaths) / want)));
    }
    hook.train_stride = stride;
    hook.train_phase = params->sample_base % stride;
    hook.sample_capacity = train->samples;
    GE_HIP(ctx, net.samples.reserve(size_t(train->samples) * sizeof(ge_cache_sample)));
    GE_HIP(ctx, net.sample_weights.reserve(size_t(train->samples) * 16));
    hook.samples = net.samples.as<ge_cache_sample>();
    hook.sample_weights = net.sample_weights.as<float4>();
  }
  // A query where each path ends, and one more where each training path does.
  const uint64_t qcap = npaths + (training ? npaths / hook.train_stride + 1 : 0);
  GE_HIP(ctx, net.queries.reserve(std::max<size_t>(qcap * sizeof(ge_cache_query), 256)));
  GE_HIP(ctx, net.query_weights.reserve(std::max<size_t>(qcap * 16, 256)));
  hook.queries = net.queries.as<ge_cache_query>();
  hook.query_weights = net.query_weights.as<float4>();
  hook.query_capacity = uint32_t(qcap);
  GE_HIP(ctx, begin_timing(ctx));
  unsigned long long* counters = (flags & GE_RT_COUNT) ? ctx->counters.as<unsigned long long>() : nullptr;
  ctx->primal.valid = false;
  GE_HIP(ctx, wavefront_path_trace(ctx, bvh_view(scene->dev), svl, p, d_accum, d_cov, counters, &hook));
  if (training) {
    uint32_t found = 0;
    GE_HIP(ctx, read_cache_counts(ctx, nullptr, &found));
    const uint64_t started = npaths / hook.train_stride + (npaths % hook.train_stride > hook.train_phase ? 1 : 0);
    if (started) net.hit_rate = float(double(found) / double(started));
    net.kept = std::min(found, train->samples);
    GE_HIP(ctx, train_slices(ctx, net, std::min(found, train->samples), train->steps, train->learning_rate));
  }
  if ((r = finish(ctx, false)) != GE_RT_OK) return r;
  if (training && (r = read_loss(ctx, cache, loss)) != GE_RT_OK) return r;
  return path_copy_out(ctx, params, accum, coverage, flags, d_accum, d_cov);
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_stats(ge_rt_ctx* ctx, ge_rt_counters* out) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_CHECK(ctx, out, GE_RT_ERROR_INVALID_ARGUMENT, "null output");
  GE_HIP(ctx, hipSetDevice(ctx->device));
  GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->timing_pending) {
    float ms = 0.0f;
    GE_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
    for (int k = 0; k < 2; ++k) {
      double sum = 0.0;
      for (uint32_t i = 0; i < ctx->stage_n[k]; ++i) {
        GE_HIP(ctx, hipEventElapsedTime(&ms, ctx->stage_ev[k][2 * i], ctx->stage_ev[k][2 * i + 1]));
        sum += ms;
      }
      ctx->stage_ms[k] = sum;
      ctx->stage_launches[k] = ctx->stage_n[k];
    }
    ctx->timing_pending = false;
  }
  unsigned long long c[11];  // [8] is the walks' head
  GE_HIP(ctx, hipMemcpy(c, ctx->counters.ptr, sizeof(c), hipMemcpyDeviceToHost));
  out->mask_tests = c[9];
  out->any_mask_tests = c[10];
  out->closest_rays = c[0];
  out->any_rays = c[1];
  out->nodes = c[2] + c[5];
  out->triangles = c[3] + c[6];
  out->instances = c[4] + c[7];
  out->any_nodes = c[5];
  out->any_triangles = c[6];
  out->any_instances = c[7];
  out->last_kernel_ms = ctx->last_kernel_ms;
  out->last_build_ms = ctx->last_build_ms;
  out->closest_ms = ctx->stage_ms[0];
  out->any_ms = ctx->stage_ms[1];
  out->closest_launches = ctx->stage_launches[0];
  out->any_launches = ctx->stage_launches[1];
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

int32_t ge_rt_reset_stats(ge_rt_ctx* ctx) try {
  if (!ctx) return GE_RT_ERROR_INVALID_ARGUMENT;
  GE_HIP(ctx, hipSetDevice(ctx->device));
  GE_HIP(ctx, hipMemsetAsync(ctx->counters.ptr, 0, 88, ctx->stream));  // the walks' head is cleared with them
  GE_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GE_RT_OK;
} catch (...) {
  return host_failure(ctx);
}

}  // extern "C"
