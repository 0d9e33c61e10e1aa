// rt_internal.h — host-side state of the backend: the context, the scene on the device, device
// buffers, and the launchers each kernel file exports to rt_api.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ge_rt.h"
#include "rt_device.h"

namespace ge {

// A device allocation that grows and never shrinks (a scratch buffer reused across calls).
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipError_t reserve(size_t need) {
    if (need <= bytes) return hipSuccess;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    if (need == 0) return hipSuccess;
    const hipError_t e = hipMalloc(&ptr, need);
    if (e == hipSuccess) bytes = need;
    return e;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// Everything a segmented LBVH build needs besides its inputs. One per context, reused.
struct BuildScratch {
  DevBuf keys[2], vals[2], hist, scan_tmp, node_box, node_parent, seg_cbox, seg_scale;
  DevBuf depth;                   // per node of the device-wide build, for the levels above the first
  DevBuf buckets, bucket_count;   // the nodes each tile hands up to the next level
  DevBuf flag;                    // the build's refusal flag (a tile over capacity, too deep)
  DevBuf cl_out, cl_part, cl_seg; // the device-wide batch's clusters: first triangle, centroid bounds, segment
  DevBuf links;                   // Karras' child links and counts, 16 B a batch entry, until the fit writes the records
};

// A one-workgroup build over triangles (at most k_block_build_max): its `count` triangles are those
// of clusters first .. first + ncl - 1 (list 0: a mesh's own) or lists[first .. first + ncl - 1]
// (list 1: a frame structure's), in that order; cbase is the mesh's first cluster (a record's
// cluster is numbered within its mesh). Its nodes go to nodes[node ..], its sorted triangle
// records and codes to tris[out ..] and morton[out ..], its root box and height to roots[6 root ..]
// and heights[root].
struct BlockSeg {
  uint32_t count, node, out, root, first, ncl, cbase, list;
};

struct SceneDev {
  // The engine's arrays, as uploaded.
  DevBuf clusters, vertices, quantized, attributes, triangles, meshes, instances, materials,
      cluster_materials, lights;
  std::vector<ge_cluster_desc> h_clusters;
  std::vector<ge_mesh_desc> h_meshes;
  std::vector<ge_instance_desc> h_instances;
  std::vector<uint32_t> mesh_max_material;  // per mesh: its clusters' largest material index
  std::vector<ge_resolve_material> h_materials;
  uint32_t cluster_count = 0, vertex_count = 0, triangle_count = 0, mesh_count = 0, instance_count = 0,
           material_count = 0, light_count = 0;
  bool has_attributes = false;
  // Per mesh: its triangles' place in the pools and its cluster-triangle prefix.
  std::vector<uint32_t> mesh_tri_count, mesh_tri_base, mesh_node_base, mesh_node_count;
  DevBuf tri_prefix;   // per cluster: the mesh-local index of its first triangle
  DevBuf tri_mesh_of;  // per cluster: its mesh
  uint32_t tri_pool = 0, node_pool = 0;
  // The built structures.
  DevBuf blas_nodes, blas_tris, blas_morton, mesh_roots, mesh_height;  // roots: 6 floats per mesh
  DevBuf tlas_nodes, tlas_slots, tlas_morton, inst_recs, inst_boxes, live_ids, tlas_root, tlas_height;
  DevBuf inst_T;  // per instance: its translation in the frame's space (3 floats)
  // The instance records as last uploaded, kept between builds: entries [rec_lo, rec_hi) changed
  // since (all of them when rec_all, or when the frame's eye moves), and the list of instances with
  // geometry when live_dirty.
  std::vector<InstRec> h_recs;
  std::vector<float> h_T;
  std::vector<uint32_t> h_live;
  uint32_t rec_lo = 0xffffffffu, rec_hi = 0;
  bool rec_all = true, live_dirty = true;
  ge_frame_eye rec_eye{};
  DevBuf blas_wide, tlas_wide;  // the 4-wide collapse of each, at the binary nodes' indices
  uint32_t tlas_prims = 0, tlas_node_count = 0;
  std::vector<uint32_t> h_mesh_height;
  uint32_t tlas_height_value = 0;
  uint32_t stack_cap = 0;  // traversal stack entries a lane needs: both heights + 2
  bool built = false;
  // The engine's pair numbering, pair -> (instance, cluster within its mesh), for visibility words;
  // rebuilt on first use after an upload or an instance update.
  DevBuf pairs;
  uint32_t pair_count = 0;
  bool pairs_dirty = true;
  std::string pairs_error;  // why the pair table could not be built (empty: it is current)
  ge_frame_eye eye{};
  // Bumped by every change the traced image can see (materials, lights, instances, structures).
  uint64_t version = 0;
  std::vector<ge_resolve_light> h_lights;
  DevBuf light_cells;               // k_light_grid^3 masks (empty: no grid)
  bool light_grid = false;
  uint64_t light_far = 0;
  float light_lo[3] = {0, 0, 0}, light_inv[3] = {0, 0, 0};
  // The bindless textures (level 0), samplers, the sRGB decode table, and per instance its mesh's
  // first cluster and material base, bit 31 set when a cluster's material is alpha-tested.
  DevBuf textures, samplers, texels, srgb, mask_inst;
  uint32_t texture_count = 0, sampler_count = 0;
  bool any_masked = false, any_textured = false;
  std::vector<std::vector<uint32_t>> mesh_materials;  // per mesh: its clusters' distinct material indices
  // The frame's structures (the engine's per-frame cluster chain): each instance that is not static
  // (InstanceDesc::flags bit 7 clear) traces a structure built over the clusters the frame's visible
  // list gives it. Frame structure k (the k-th such instance in index order) has a fixed region of
  // the pools as large as its mesh's own structure, after the meshes' regions, and root slot
  // mesh_count + k; geom_base[instance] is where its clusters' visible entries start in entry_of
  // (~0 for a static instance).
  bool has_visible = false;        // a visible list was given
  bool visible_dirty = false;      // given since the last build
  std::vector<uint32_t> h_visible;  // the list as given: {instance, cluster} pairs
  std::vector<uint32_t> frame_of;   // per instance: its frame structure, ~0 if static
  std::vector<uint32_t> frame_inst, frame_tri_base, frame_node_base, frame_tri_count, frame_node_count;
  std::vector<std::vector<uint32_t>> frame_clusters;  // per frame structure: its clusters (scene-wide), in list order
  std::vector<uint32_t> h_geom_base, h_entry_of;
  uint32_t frame_tri_pool = 0, frame_node_pool = 0;  // the regions' total, after the meshes'
  bool frame_layout_dirty = true;  // instances changed: the regions and root slots are laid out again
  DevBuf root_of;                  // per instance: its structure's root slot (its mesh, or mesh_count + frame)
  DevBuf geom_base, entry_of;
};

// What the context's last wavefront path-trace dispatch traced, so an adjoint can reuse its
// per-path radiance (`result`) as its primal pass.
struct PrimalKey {
  bool valid = false;
  const void* scene = nullptr;
  uint64_t version = 0;
  ge_path_trace_params params{};  // with the output addresses cleared
};

// A hit's geometry index as the API reports it (hit_visible_entry, scene.slang:718-724): the
// cluster within the mesh for a static instance; the cluster's visible entry for an instance that
// traces the frame's structure (base[instance] != ~0: its clusters' entries start there in `entry`).
struct GeomMap {
  const uint32_t* base;
  const uint32_t* entry;
};
__device__ __forceinline__ uint32_t geom_of(const GeomMap& g, uint32_t instance, uint32_t local_cluster) {
  const uint32_t b = g.base ? g.base[instance] : 0xffffffffu;
  return b == 0xffffffffu ? local_cluster : g.entry[b + local_cluster];
}

// ---- the wavefront path tracer's records (rt_wavefront.hip) ----
struct PathState {  // 64 bytes: one live path
  float o[3];
  uint32_t pad;
  float d[3];
  uint32_t rng;
  float T[3];      // throughput
  uint32_t bounce;
  float L[3];      // radiance so far
  uint32_t flags;  // bit 0: the camera ray hit; bit 1: the path ended
};
struct PathHit {  // 20 bytes: the closest hit of the path's ray
  float t, u, v;
  uint32_t inst, prim;
};
struct NeeRec {  // 32 bytes: a vertex's next-event estimation, resolved after its shadow rays
  float T[3];
  uint32_t first;
  uint32_t count;  // ~0: the ray missed, nothing to resolve
  uint32_t pad[3];
};
struct ShadowRec {  // 48 bytes: one shadow ray and what its light adds if nothing blocks it
  float o[3];
  float tmax;
  float d[3];
  uint32_t pad0;
  float c[3];
  uint32_t pad1;
};
struct ShadowGrad {  // 32 bytes: what one unblocked shadow ray adds to the gradient (the adjoint)
  float ga[3];    // delta T df/dalbedo cos radiance, for the vertex's material
  float gr;       // delta . T df/droughness cos radiance
  float gl;       // delta . T f cos dradiance/dintensity, for light `light`
  uint32_t light; // 0 the sun, 1 + i light i
  uint32_t pad[2];
};
struct Wavefront {
  DevBuf state[2], hits, nee, result, queue[2], counts, shadow, blocked;
  DevBuf sgrad;               // one ShadowGrad a shadow record (the adjoint)
  DevBuf ts, adj_a, adj_r;    // per path (the adjoint): sampling throughput; dlnf/dalbedo + material, dlnf/droughness + going on
  DevBuf lpath;               // per path (the adjoint): the primal pass's radiance
  uint64_t shadow_capacity = 0;
  uint32_t* host_counts = nullptr;  // pinned
  hipEvent_t counted = nullptr;     // the shading's counts have reached host_counts
};

// The path-replay adjoint on the wavefront (wavefront_path_trace's `adjoint`): directions, lobe
// choices, pdfs and roulette from `sampling` (null: the scene's materials), BSDF values from the
// scene's. With grad, the pass replays the primal pass's paths and adds each vertex's gradient:
// its next-event term, and (df/f) times what the path gathered after it, L - prefix (path replay
// backpropagation: no stored vertices once the primal radiance `lpath` is known).
struct AdjointHook {
  const ge_resolve_material* sampling;
  bool grad;
  const float4* d_image;    // dLoss/dI per pixel
  const float4* lpath;      // per path: the primal pass's radiance
  double* grad_materials;   // per material: albedo rgb, roughness
  double* grad_lights;      // the sun, then each light
};

}  // namespace ge

struct ge_rt_scene {
  ge::SceneDev dev;
};

struct ge_rt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, bev0 = nullptr, bev1 = nullptr;
  std::string error;
  int32_t last_status = 0;
  ge::BuildScratch build;
  ge::DevBuf io[6];    // host-pointer staging
  ge::DevBuf tmp[9];   // build batches: metadata, records, boxes, segments, order, codes, instances
  // A build's host arrays, copied to the device asynchronously: kept until the next build, which
  // starts after the last one's closing synchronization.
  std::vector<uint32_t> host_meta, host_tmeta, host_blists;
  std::vector<ge::BlockSeg> host_bsegs;
  // Pinned host memory for the build's end: the refusal flag, the instance structure's height and
  // every mesh's height, read back with one synchronization.
  uint32_t* host_readback = nullptr;
  size_t host_readback_cap = 0;
  // Pinned staging for a build's uploads (copies from pageable memory block the host while the
  // runtime stages them): sized once at the build's start, then handed out piece by piece; the
  // build's closing synchronization ends every copy before the next build writes it again.
  uint8_t* pin = nullptr;
  size_t pin_cap = 0, pin_used = 0;
  // 8 x u64 (closest rays, any rays, closest-hit nodes, triangles, instances, any-hit nodes,
  // triangles, instances), then the persistent walks' head counter at byte 64
  ge::DevBuf counters;
  double last_kernel_ms = 0.0, last_build_ms = 0.0;
  bool timing_pending = false;
  // Begin/end event pairs around the wavefront's traversal launches: [0] closest-hit, [1] any-hit.
  hipEvent_t stage_ev[2][2 * GE_RT_STAGE_TIMED] = {};
  uint32_t stage_n[2] = {0, 0};
  double stage_ms[2] = {0.0, 0.0};
  uint32_t stage_launches[2] = {0, 0};
  ge::Wavefront wave;
  ge::PrimalKey primal;
  ge::DevBuf resolve_rays;  // the visibility-buffer shadow pass's ray records
  // The device address of this device's allocation-free working array (the shadow resolve's
  // counter, then GE_RT_SHADOW_STATIC_ENTRIES ray records), looked up once per device at creation.
  uint8_t* static_scratch = nullptr;
  uint32_t persistent_blocks = 5120;  // one-wave workgroups a persistent walk launches: CUs x 20
};

namespace ge {

// ---- launchers (rt_bvh.hip) ----
// Builds one LBVH per segment over `n` primitive boxes (6 floats each, lo xyz then hi xyz, grouped
// by segment in order; seg_of[i] names primitive i's segment). Segment s covers primitives
// [seg_first[s], seg_first[s] + seg_count[s]); its nodes go to nodes[node_base[s] ..], its sorted
// order (batch indices) to perm[seg_first[s] ..], its sorted codes to morton[seg_first[s] ..], and
// its root box and height to roots[6 r ..] and heights[r], r = root_index[s] (or s when null).
// `node_pool` bounds every node_base[s] + node count. All pointers are device pointers.
// Segments of at most k_block_build_max primitives are built by one workgroup each, in LDS, in one
// launch; larger ones by device-wide passes. The caller puts the larger ones first: segments
// [0, nseg_global), primitives [0, n_global), take the device-wide passes; h_seg_count holds their
// counts on the host. A structure the build refuses (deeper than 254 levels, or too unbalanced for
// a tile's capacity) sets the device word `flag`, which the caller clears before and reads after.
// With nparts, the device-wide segments' centroid bounds come from the per-cluster bounds that
// launch_triangle_records (over nparts clusters) left in the scratch, not from the boxes.
constexpr uint32_t k_block_build_max = 1024;
hipError_t build_segments(BuildScratch& scratch, hipStream_t stream, uint32_t n, const float* boxes,
                          uint32_t nseg, const uint32_t* d_seg_of, const uint32_t* d_seg_first,
                          const uint32_t* d_seg_count, const uint32_t* d_node_base, ge_bvh_node* nodes,
                          uint32_t* perm, uint32_t* morton, float* roots, uint32_t* heights,
                          const uint32_t* d_root_index, uint32_t node_pool, WideNode* wide, uint32_t nseg_global,
                          uint32_t n_global, const uint32_t* h_seg_count, uint32_t* flag, uint32_t nparts = 0);
// The refit of one structure of n >= 2 leaves (GE_RT_BUILD_REFIT): its records' topology kept,
// the leaves' boxes given in slot order (6 floats each), every node's box, the records, the wide
// nodes and the root's box and height recomputed. d_seg_tables holds {0, n, 0} on the device.
hipError_t refit_segment(BuildScratch& scratch, hipStream_t stream, uint32_t n, const float* slot_boxes,
                         const uint32_t* d_seg_tables, ge_bvh_node* nodes, WideNode* wide, float* root,
                         uint32_t* height, uint32_t* flag);
// Builds the BlockSeg segments, one workgroup each, in one launch.
hipError_t launch_block_meshes(hipStream_t stream, const SceneDev& s, uint32_t nseg, const BlockSeg* segs,
                               const uint32_t* lists, ge_bvh_node* nodes, ge_bvh_triangle* tris, uint32_t* morton,
                               float* roots, uint32_t* heights, WideNode* wide);
// The device-wide batch's segments over triangles: segment s's clusters are clusters first + j (list
// 0: a mesh's own) or lists[first + j] (list 1: a frame structure's), j < pre[s + 1] - pre[s], its
// triangles at batch index pre-scan on; cbase is its mesh's first cluster.
struct GlobalSeg {
  uint32_t first, list, cbase, pad;
};
// Triangle records, boxes and segment ids of the batch's ncl clusters (nseg segments, `pre` their
// first clusters, nseg + 1 entries), and each cluster's centroid bounds and segment in scratch for
// build_segments (nparts = ncl).
hipError_t launch_triangle_records(hipStream_t stream, const SceneDev& s, BuildScratch& scratch, uint32_t nseg,
                                   uint32_t ncl, const uint32_t* pre, const GlobalSeg* segs, const uint32_t* lists,
                                   ge_bvh_triangle* tris, float* boxes, uint32_t* seg_of);
hipError_t launch_gather_triangles(hipStream_t stream, uint32_t n, const uint32_t* perm, const uint32_t* seg_of,
                                   const uint32_t* seg_first, const uint32_t* seg_out, const ge_bvh_triangle* in,
                                   ge_bvh_triangle* out, const uint32_t* morton_in, uint32_t* morton_out);
hipError_t launch_instance_boxes(hipStream_t stream, uint32_t n, const uint32_t* live_ids, const ge_instance_desc* inst,
                                 const float* translations, const float* roots, const uint32_t* root_of,
                                 float* boxes, uint32_t* seg_of);
hipError_t launch_gather_slots(hipStream_t stream, uint32_t n, const uint32_t* perm, const uint32_t* live_ids,
                               uint32_t* slots);

// ---- launchers (rt_kernels.hip) ----
struct SceneView {
  const ge_cluster_desc* clusters;
  const uint16_t* quantized;
  const ge_vertex_attributes* attributes;  // null: flat shading
  const uint32_t* triangles;
  const ge_mesh_desc* meshes;
  const ge_instance_desc* instances;
  const ge_resolve_material* materials;
  const uint32_t* cluster_materials;
  const ge_resolve_light* lights;
  uint32_t light_count;
  uint32_t features;  // k_scene_masked | k_scene_textured
  TexView tex;        // the scene's textures and samplers (none: null)
  // Lights 0..63 by cell of a k_light_grid^3 grid over the box their ranges reach: bit i of a cell's
  // mask is set when light i may reach a point of the cell (a light without a range everywhere).
  // A point outside the box takes light_far (the lights without a range); light_cells null: every
  // light.
  const uint64_t* light_cells;
  uint64_t light_far;
  float light_lo[3], light_inv[3];  // the box's low corner, cells per unit length
};
constexpr uint32_t k_light_grid = 16;
// SceneView::features: a material is alpha-tested (the walks' mask records are built), a material
// reads a texture.
constexpr uint32_t k_scene_masked = 1u, k_scene_textured = 2u;

// Records the begin (end = 0) or end (end = 1) event of a traversal launch of `kind` (0 closest-hit,
// 1 any-hit) on the context's stream, for ge_rt_stats; launches past the event pool go untimed.
hipError_t stage_mark(ge_rt_ctx* ctx, int kind, int end);

// The path-replay adjoint (rt_kernels.hip): per pixel the detached estimator's sample sum into
// `image` (when not null) and, with `d_image`, the loss gradient added into the double sums.
hipError_t launch_path_adjoint(hipStream_t stream, const Bvh& bvh, const SceneView& sv, const ge_path_trace_params& p,
                               const ge_resolve_material* sampling, const float4* d_image, float4* image,
                               double* grad_materials, double* grad_lights);

// The sun shadow pass over a visibility buffer (rt_shadow.hip): `rays` holds one 16-byte record an
// entry, `out` the entries' answers. The any-hit walk is timed as a stage of kind 1.
hipError_t launch_shadow_resolve(ge_rt_ctx* ctx, const Bvh& bvh, const SceneView& sv, const ge_frame_eye& eye,
                                 const ge_shadow_resolve_params& p, const uint64_t* words, const uint2* pairs,
                                 uint32_t pair_count, float4* rays, uint32_t* out, unsigned long long* counters,
                                 uint32_t* head, uint32_t max_blocks, uint32_t alpha);

// Persistent walks: `head` is a device counter the launch clears, `max_blocks` the grid's cap.
// `alpha` (k_alpha_*) is the walks' rule for masked candidates; `full` compiles textured materials
// and the masked rule into the path tracer.
hipError_t launch_trace_closest(hipStream_t stream, const Bvh& bvh, const ge_ray* rays, uint32_t n, uint32_t mask,
                                ge_hit* hits, unsigned long long* counters, uint32_t* head, uint32_t max_blocks,
                                uint32_t alpha, GeomMap geom);
hipError_t launch_trace_any(hipStream_t stream, const Bvh& bvh, const ge_ray* rays, uint32_t n, uint32_t mask,
                            uint32_t* bits, unsigned long long* counters, uint32_t* head, uint32_t max_blocks,
                            uint32_t alpha);
hipError_t launch_ray_visibility(hipStream_t stream, const Bvh& bvh, const ge_ray_visibility_params& p, uint64_t* words,
                                 unsigned long long* counters, uint32_t* head, uint32_t max_blocks, uint32_t alpha);
hipError_t launch_path_trace(hipStream_t stream, const Bvh& bvh, const SceneView& sv, const ge_path_trace_params& p,
                             float4* accum, uint32_t* coverage, unsigned long long* counters, bool full);
hipError_t launch_tonemap(hipStream_t stream, const ge_path_trace_params& p, const float4* accum,
                          const uint32_t* coverage, uint32_t* out);
// The wavefront form of launch_path_trace (rt_wavefront.hip); synchronizes once a bounce to size
// the next launches.
// With a hook (rt_mlp.h) the paths also feed the radiance cache.
struct CacheHook;
hipError_t wavefront_path_trace(ge_rt_ctx* ctx, const Bvh& bvh, const SceneView& sv, const ge_path_trace_params& p,
                                float4* accum, uint32_t* coverage, unsigned long long* counters,
                                const CacheHook* hook = nullptr, const AdjointHook* adjoint = nullptr);
hipError_t launch_shadow_pass(hipStream_t stream, const Bvh& bvh, const SceneView& sv, const ge_shadow_params& p,
                              const ge_shadow_receiver* rec, uint32_t* bits, unsigned long long* counters,
                              uint32_t alpha);

}  // namespace ge
