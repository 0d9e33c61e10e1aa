/*
 * ge_rt.h — the C ABI of the MI355X ray-tracing backend.
 *
 * This is the drop-in boundary for the renderer's ray-traced hot path: BVH construction over the
 * cluster-geometry triangles, closest-hit and any-hit traversal, the reference path tracer and
 * the ray-traced shadow pass. On the engine side these are today Vulkan calls (the driver's
 * acceleration-structure builds and `RayQuery` traversal) driven by
 * `systems/renderer/src/reference.cpp` and the resolve's shadow query; here they are plain C entry
 * points over plain pointers and sizes, no exceptions, `int32_t` status (0 = ok) and a per-context
 * error string, which is the engine's own `bool + std::string* error` convention
 * (`systems/renderer/include/systems/renderer/reference.h:111-146`) flattened for a C ABI.
 *
 * Every record the GPU reads keeps the engine's byte layout, so the host can hand over the arrays
 * it already builds for its GPU scene unchanged. The layouts are pinned by the static assertions
 * at the end of this file, against the engine's size tables
 * (`domain/gfx/tests/size_table.cpp:41,122,125,388,397,452`, `domain/geometry/tests/size_table.cpp:13`).
 *
 * Threading: a context is owned by one host thread; calls on a context are serialized and run on
 * the context's own HIP stream. Nothing here allocates on a caller's stream.
 */
#ifndef GE_RT_H
#define GE_RT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------------- */
enum {
  GE_RT_OK = 0,
  GE_RT_ERROR_INVALID_ARGUMENT = 1, /* a null pointer, a zero size where one is needed, a bad index */
  GE_RT_ERROR_INVALID_SCENE = 2,    /* offsets or counts out of range, no BVH built yet */
  GE_RT_ERROR_OUT_OF_MEMORY = 3,
  GE_RT_ERROR_DEVICE = 4,           /* a HIP call failed; the text names it */
  GE_RT_ERROR_UNSUPPORTED = 5       /* a feature flag this backend does not implement */
};

/* Memory flag: every caller pointer of the call is a device pointer of the context's device.
 * Without it, caller pointers are host memory and the call copies in and out. With it, the call
 * returns once its kernels are queued on the context's stream: device outputs are complete after
 * ge_rt_synchronize (or ge_rt_stats, which waits for the stream). */
#define GE_RT_DEVICE_PTR 0x1u
/* Trace flag: count internal nodes popped, triangle tests and instance entries into the context's
 * counters (ge_rt_stats). Off in timed runs. */
#define GE_RT_COUNT 0x2u
/* Path-trace flag: run the integrator as one kernel, one path per lane, instead of the default
 * wavefront (a kernel per stage over compacted queues). Both give the same radiance to the bit. */
#define GE_RT_PT_MEGAKERNEL 0x4u
/* Adjoint flag: with the scene's own sampling (`sampling` NULL), reuse the per-path radiance of
 * the context's last ge_rt_path_trace when it traced the same parameters over the same scene,
 * materials, lights and structures (the adjoint's primal pass is that dispatch, to the bit);
 * otherwise trace the primal pass as usual. `image` must then be NULL. */
#define GE_RT_ADJOINT_REUSE_PRIMAL 0x8u
/* Query flags for alpha-tested materials (a material with GE_MATERIAL_MASKED; every other triangle
 * is opaque). Without either flag a query treats masked triangles as opaque, as a ray that forces
 * opacity does (RAY_FLAG_FORCE_OPAQUE, path_trace.slang:196-203).
 * GE_RT_ALPHA_MASKS: a masked candidate is kept where its mask keeps the hit at level 0
 * (pt_candidate_keeps, path_trace.slang:117-135; mask_keeps_level0, alpha_mask.slang:66-75).
 * GE_RT_CULL_MASKED: masked candidates are left out (RAY_FLAG_CULL_NON_OPAQUE, the camera ray of a
 * frame without masks, path_trace.slang:280-281). ge_rt_trace_closest takes either,
 * ge_rt_trace_any and ge_rt_shadow_pass GE_RT_ALPHA_MASKS. */
#define GE_RT_ALPHA_MASKS 0x10u
#define GE_RT_CULL_MASKED 0x20u

/* ---- the engine's GPU records (byte layouts mirrored, sizes asserted below) ----------------- */

/* geometry::ClusterDesc, 48 bytes (domain/geometry/include/domain/geometry/cluster.h:45-60). */
typedef struct ge_cluster_desc {
  uint32_t vertex_offset;   /* first vertex, scene-wide */
  uint32_t triangle_offset; /* first packed triangle, scene-wide */
  uint32_t vertex_count;
  uint32_t triangle_count;
  float center[3];
  float radius;
  float cone_apex[3];
  uint32_t cone;
} ge_cluster_desc;

/* geometry::VertexAttributes, 8 bytes (cluster.h:75-78): octahedral normal in two snorm16 and a UV
 * in two halves. */
typedef struct ge_vertex_attributes {
  uint32_t normal_oct;
  uint32_t uv_half2;
} ge_vertex_attributes;

/* gfx::MeshDesc, 64 bytes (domain/gfx/shaders/scene.slang:46-55). `quantized` is an address the
 * backend fills in from the uploaded stream; the deform addresses must be 0 (rigid scenes). */
typedef struct ge_mesh_desc {
  float quant[4]; /* xyz grid origin, w grid step */
  uint64_t quantized;
  uint32_t first_cluster;
  uint32_t cluster_count;
  uint64_t deform_pool;
  uint64_t deform_slots;
  uint64_t templates;
  uint64_t skin;
} ge_mesh_desc;

/* gfx::InstanceDesc, 96 bytes (scene.slang:61-80). rows = mesh space to the cell's frame,
 * (linear row r | translation r). */
typedef struct ge_instance_desc {
  float rows[3][4];
  int32_t cell[3];
  uint32_t mesh;
  uint32_t material_base;
  uint32_t first_pair;
  float scale_max;
  uint32_t flags;
  uint32_t deform;
  float bounds_padding;
  uint32_t terrain;
  uint32_t band;
} ge_instance_desc;

/* InstanceDesc::flags bit 7 (scene.slang:694): the instance traces its mesh's own structure and a
 * hit's geometry index is the cluster within the mesh. An instance without it traces the frame's
 * structure of the clusters the visible list gives it (ge_rt_scene_set_visible). */
#define GE_INSTANCE_STATIC_RT 128u
/* InstanceDesc::flags bit 6, k_instance_untraced (scene.slang:473-492): never met by a light's ray.
 * The instance's top-level mask is then 0xfe instead of 0xff (instance_ray_mask), so camera rays
 * (mask 0xfe) meet it and bounce and shadow rays (the caster bit 0x01 alone) pass it by. */
#define GE_INSTANCE_UNTRACED 64u
#define GE_INVALID_DEFORM 0xffffffffu

/* gfx::FrameEye, 48 bytes (scene.slang:96-104): the frame's origin. All zero is the world's. */
typedef struct ge_frame_eye {
  int32_t cell[3];
  uint32_t masks;
  float local[3];
  float pad1;
  float residual[3];
  float pad2;
} ge_frame_eye;

/* gfx::ResolveMaterial, 128 bytes (domain/gfx/shaders/material.slang:71-90). */
typedef struct ge_resolve_material {
  float albedo[4];   /* rgb linear, w roughness */
  float emissive[4]; /* rgb linear factor, w metallic */
  uint32_t albedo_texture;
  uint32_t sampler;
  float uv_scale;
  uint32_t flags;
  uint32_t metallic_roughness_texture;
  uint32_t normal_texture;
  float normal_scale;
  uint32_t occlusion_texture;
  uint32_t emissive_texture;
  float occlusion_strength;
  uint32_t samplers_mr_normal;
  uint32_t samplers_occlusion_emissive;
  float uv_transform[4];
  float uv_offset[4];
  float specular[4]; /* rgb F0 - 0.04, w 1 - F90; zero: the 4 % and 1 */
} ge_resolve_material;
#define GE_NO_TEXTURE 0xffffffffu
/* A per-slot sampler half that borrows the material's `sampler` (material.slang:67-68, 92-97):
 * samplers_mr_normal holds the metallic-roughness slot's in bits 0-15 and the normal map's in
 * 16-31, samplers_occlusion_emissive the emissive map's in bits 16-31. */
#define GE_SAME_SAMPLER 0xffffu
/* ResolveMaterial::flags (material.slang:35-59; scene.slang:360). Level 0 is all a ray samples, so
 * GE_MATERIAL_MIPPED changes nothing here. GE_MATERIAL_MASKED: alpha-tested, alpha =
 * uv_offset.w times the albedo texture's alpha, kept where alpha >= uv_offset.z. Ground detail and
 * water are the resolve's and take no part in the path tracer without their tables; the stochastic
 * flag (131072) and foliage optics (bits 24-31) are refused with GE_RT_ERROR_UNSUPPORTED. */
#define GE_MATERIAL_MIPPED 1u
#define GE_MATERIAL_NORMAL_RG 2u
#define GE_MATERIAL_UV_TRANSFORM 4u
#define GE_MATERIAL_MASKED 32u
#define GE_MATERIAL_STOCHASTIC 131072u

/* One texture of the scene's bindless array (g_textures, material.slang:30; the GpuScene's bindless
 * set, gpu_scene.h:198-204), at level 0 only: width x height RGBA8 texels, rows top to bottom,
 * starting `offset` bytes into the scene's texel pool. 24 bytes. */
typedef struct ge_texture_desc {
  uint32_t width;
  uint32_t height;
  uint32_t format; /* GE_FORMAT_*: gfx::Format's numbering (domain/gfx/include/domain/gfx/rhi.h:240-246) */
  uint32_t reserved;
  uint64_t offset;
} ge_texture_desc;
#define GE_FORMAT_R8G8B8A8_UNORM 37u
#define GE_FORMAT_R8G8B8A8_SRGB 43u /* rgb through the sRGB transfer function, alpha linear */

/* One sampler of the bindless array (g_samplers, material.slang:31): gfx::SamplerDesc
 * (domain/gfx/include/domain/gfx/bindless.h:149-157) as level 0 reads it. A level-0 read
 * (SampleLevel(..., 0)) has a level of detail of 0 and so filters with `mag`; `min` is kept for
 * the record. Address modes and filters take gfx's numbering (rhi.h:352-359); the border of
 * GE_ADDRESS_CLAMP_TO_BORDER is transparent black. 16 bytes. */
typedef struct ge_sampler_desc {
  uint32_t address_u;
  uint32_t address_v;
  uint32_t mag;
  uint32_t min;
} ge_sampler_desc;
#define GE_FILTER_NEAREST 0u
#define GE_FILTER_LINEAR 1u
#define GE_ADDRESS_REPEAT 0u
#define GE_ADDRESS_MIRRORED_REPEAT 1u
#define GE_ADDRESS_CLAMP_TO_EDGE 2u
#define GE_ADDRESS_CLAMP_TO_BORDER 3u

/* gfx::ResolveLight, 64 bytes (material.slang:150-157). */
typedef struct ge_resolve_light {
  float position_radius[4]; /* xyz position (directional: towards the light), w radius */
  float color_intensity[4];
  float direction_type[4]; /* xyz spot direction, w 0 point, 1 spot, 2 directional */
  float spot[4];           /* x cos inner, y cos outer, z emitter radius */
} ge_resolve_light;

/* gfx::PathTraceParams, 368 bytes (domain/gfx/include/domain/gfx/path_trace.h:51-131 and
 * domain/gfx/shaders/path_trace.slang:55-98). The host fills it exactly as it fills the engine's
 * block; the scene addresses (clusters .. lights, hit_entries, extra) are written by the backend
 * from the scene handle at dispatch, and `accum`, `output`, `coverage` are the caller's device
 * pointers when GE_RT_DEVICE_PTR is given and ignored otherwise. */
typedef struct ge_path_trace_params {
  float clip_to_ray[16]; /* column-major: element (row r, col c) at [c * 4 + r] */
  float camera[4];
  float sky[4];
  float sun[4]; /* xyz normalized direction towards the light, w intensity */
  float ground[4];
  uint64_t accum;
  uint64_t output;
  uint64_t clusters;
  uint64_t mesh;
  uint64_t triangles;
  uint64_t instances;
  uint64_t attributes;
  uint64_t materials;
  uint64_t cluster_materials;
  uint64_t visible;
  uint64_t lights;
  uint32_t width;
  uint32_t height;
  uint32_t light_count;
  uint32_t scene;
  uint32_t spp;
  uint32_t sample_base;
  uint32_t max_bounces;
  uint32_t seed;
  uint32_t flags;
  float ray_bias;
  uint64_t coverage;
  uint32_t background;
  float ray_bias_steps;
  uint64_t ground_detail;
  uint64_t sky_params;
  ge_frame_eye eye;
  int32_t terrain_uv_x0_mm;
  int32_t terrain_uv_z0_mm;
  float terrain_uv_per_mm;
  uint32_t extra_vertices;
  uint64_t hit_entries;
  uint64_t extra;
} ge_path_trace_params;

/* PathTraceParams::flags (path_trace.h:35-48). */
#define GE_PT_UNIFORM_SKY 1u
#define GE_PT_PIXEL_CENTER 2u
#define GE_PT_DITHER 4u
/* The frame has alpha-tested materials: every ray decides its masked candidates by their masks
 * (path_trace.slang:110-135, 278-292). Without it the camera ray leaves masked triangles out and
 * every later ray stops on them. */
#define GE_PT_ALPHA_MASKS 8u
#define GE_PT_ROULETTE_DEPTH 3u

/* gfx::RayVisibilityParams, 256 bytes (domain/gfx/include/domain/gfx/ray_visibility.h:39-61,
 * ray_visibility.slang:39-57). `output` is the caller's device pointer under GE_RT_DEVICE_PTR. */
typedef struct ge_ray_visibility_params {
  float view_proj[16];   /* column-major */
  float clip_to_ray[16]; /* column-major */
  float camera[4];
  uint64_t output;
  uint64_t instance_base;
  uint32_t width;
  uint32_t height;
  uint32_t scene;
  uint32_t caster_base;
  uint64_t visible;
  uint64_t instances;
  uint64_t meshes;
  uint64_t triangles;
  uint64_t clusters;
  uint64_t hit_entries;
  /* MaskTables (cluster_cull.h:154-167): materials, cluster_materials, attributes, fade. Non-zero
   * `materials` marks a frame with alpha-tested materials (ray_visibility.h:32-36): masked
   * candidates are decided by their masks from the scene's own tables; zero leaves them out as the
   * cull flag does. The cross-fade (`fade`) must be 0. */
  uint64_t masks[4];
} ge_ray_visibility_params;

/* ---- the backend's own records --------------------------------------------------------------- */

/* One ray of the test and bench surface: the fields of a RayDesc (origin, TMin, direction, TMax).
 * 32 bytes. Hits are accepted for tmin <= t <= tmax, as a Vulkan ray query accepts them. */
typedef struct ge_ray {
  float origin[3];
  float tmin;
  float dir[3];
  float tmax;
} ge_ray;

/* A committed closest hit: what `CommittedRayT`, `CommittedInstanceID`, `CommittedGeometryIndex`,
 * `CommittedPrimitiveIndex` and `CommittedTriangleBarycentrics` return, plus the visibility id
 * `pair << 8 | triangle` (ray_visibility.slang:88-93). `instance_id` is ~0 on a miss. 32 bytes. */
typedef struct ge_hit {
  float t;
  float u; /* barycentric weight of the triangle's second corner */
  float v; /* barycentric weight of the third corner */
  uint32_t instance_id;
  uint32_t geometry;  /* cluster index within the instance's mesh (a static instance), or the visible
                         entry of the cluster (an instance tracing the frame's structure) */
  uint32_t primitive; /* triangle index within the cluster */
  uint32_t key;       /* pair << 8 | triangle */
  uint32_t pad;
} ge_hit;
#define GE_RT_MISS 0xffffffffu

/* The shadow pass's frame: one sun ray per receiver, as the resolve's shadow_visibility and
 * shadow_ray_blocked ask it (domain/gfx/shaders/visibility_resolve.slang:516-605,742-755). */
typedef struct ge_shadow_params {
  float sun[4];       /* xyz normalized direction towards the light */
  float offset;       /* world units along the geometric normal (ray_offset's floor) */
  float offset_steps; /* steps of the receiver's own grid (ray_offset) */
  float t_max;        /* 1e30 for the sun */
  uint32_t count;     /* receivers */
  uint32_t pad[4];
} ge_shadow_params;

/* The shadow pass over a visibility buffer: each covered pixel's surface reconstructed from its
 * visibility word as the resolve does it (reconstruct_screen, material.slang:386-458), one sun ray
 * from it (shadow_visibility, visibility_resolve.slang:742-755), or one ray per 2x2 block of one
 * pair (half_sun_block, resolve_half_sun.slang:35-96). The frame is the one the structure was
 * built for (ge_rt_bvh_build's eye). */
typedef struct ge_shadow_resolve_params {
  float view_proj[16];     /* column-major, the projection the visibility words were drawn with */
  float camera[4];         /* xyz: the eye in the frame's space (zero for the renderer's views) */
  float sun[4];            /* xyz: unit direction towards the sun */
  uint32_t width, height;  /* the visibility buffer's pixels */
  float shadow_bias;       /* ray_offset's floor (FrameLighting::shadow_bias) */
  float shadow_bias_steps; /* ray_offset's grid steps (FrameLighting::shadow_bias_steps) */
  uint32_t flags;          /* GE_SHADOW_HALF_RATE, GE_SHADOW_JITTER, GE_SHADOW_ALPHA_MASKS */
  uint32_t seed;           /* the jitter's seed */
  uint32_t pad[2];
} ge_shadow_resolve_params;
/* One ray per 2x2 block whose four words hold one pair (the top-left pixel's), as half_sun_block. */
#define GE_SHADOW_HALF_RATE 0x1u
/* The pixel's point at a per-pixel hash offset (rng_seed(x, y, 0, seed), two draws) instead of its
 * centre: the stochastic sampling of one ray per pixel. The resolve itself uses the centre. */
#define GE_SHADOW_JITTER 0x2u
/* Alpha-tested casters (k_shadow_alpha_masks, visibility_resolve.slang:357, 528-597): each masked
 * candidate of the sun's ray is decided by its mask at level 0 and the first one kept ends the
 * search. Without it masked casters are solid. */
#define GE_SHADOW_ALPHA_MASKS 0x4u
/* An entry with no ray: no surface, a pair with no instance, a surface facing away from the sun by
 * both normals, or (half rate) a block of more than one pair or at the picture's edge. */
#define GE_SHADOW_NOT_TAKEN 0xffffffffu

/* One receiver of the shadow pass: a surface point in the frame's space, the geometric normal
 * facing the eye, and the instance whose grid sets the offset. 32 bytes. */
typedef struct ge_shadow_receiver {
  float position[3];
  uint32_t instance_id;
  float geometric[3];
  uint32_t pad;
} ge_shadow_receiver;

/* Traversal counters (GE_RT_COUNT) and the last dispatch's kernel times. */
typedef struct ge_rt_counters {
  uint64_t closest_rays;   /* closest-hit queries issued */
  uint64_t any_rays;       /* any-hit queries issued */
  uint64_t nodes;          /* internal BVH nodes popped (TLAS + BLAS), both query kinds */
  uint64_t triangles;      /* ray/triangle tests, both query kinds */
  uint64_t instances;      /* TLAS leaves entered (ray transformed into a BLAS), both query kinds */
  double last_kernel_ms;   /* the last traced dispatch's device time, hipEvent on the ctx stream */
  double last_build_ms;    /* the last BVH build's device time */
  uint64_t any_nodes;      /* the part of nodes, triangles and instances that any-hit queries made */
  uint64_t any_triangles;
  uint64_t any_instances;
  /* The last dispatch's traversal launches, timed with hipEvents on the ctx stream, summed, and how
   * many launches the sums cover (the first GE_RT_STAGE_TIMED of each kind are timed): in closest_*
   * the wavefront path tracer's walk launches (one a bounce: its closest-hit rays together with the
   * previous bounce's shadow rays, the last one shadow rays alone); in any_* the shadow resolve's
   * sun walk. */
  double closest_ms;
  double any_ms;
  uint32_t closest_launches;
  uint32_t any_launches;
  /* Masked candidates whose mask was evaluated (GE_RT_ALPHA_MASKS walks), closest-hit then any-hit. */
  uint64_t mask_tests;
  uint64_t any_mask_tests;
} ge_rt_counters;
#define GE_RT_STAGE_TIMED 32

/* The built structure, read back for topology parity. Node and triangle records are the backend's
 * 64-byte and 48-byte formats. */
typedef struct ge_bvh_node {
  float lo0[3];
  float hi0[3];
  float lo1[3];
  float hi1[3];
  uint32_t child[2]; /* bit 31 set: a leaf, the rest its primitive slot; else an internal node */
  uint32_t parent;   /* ~0 at the root */
  uint32_t count;    /* primitives below this node */
} ge_bvh_node;
#define GE_BVH_LEAF 0x80000000u
#define GE_BVH_NONE 0xffffffffu

/* A node of the structures the walks take: two binary levels collapsed into one 4-wide node (an
 * even-depth binary node's children's children, a leaf child taken as it is), 64 bytes. Its child
 * boxes sit on a per-axis grid of 8-bit plane indices: plane q of axis a lies at
 * origin[a] + q 2^(e_a - 127), e_a = (meta >> 8a) & 0xff; byte c of lo[a] (hi[a]) is slot c's low
 * (high) plane, its box's plane rounded outward to the grid (one step further where that left less
 * than 2^-10 of a step). Bit 24 + c of meta is set when slot c holds a child, bit 28 + c when that
 * child is a leaf (a triangle, or an instance in the top level). child[c]: bit 31 set a
 * leaf (its primitive slot), GE_BVH_NONE an empty slot, else the binary index of an even-depth node,
 * whose wide node sits at that same index. No counterpart in the engine: its acceleration
 * structures are the Vulkan driver's and opaque (acceleration.h:46-110). */
typedef struct ge_bvh_wide_node {
  float origin[3];
  uint32_t meta;
  uint32_t child[4];
  uint32_t lo[3];
  uint32_t hi[3];
  uint32_t reserved[2];
} ge_bvh_wide_node;

typedef struct ge_bvh_triangle {
  float v0[3];
  float v1[3];
  float v2[3];
  uint32_t prim; /* cluster-within-mesh << 8 | triangle-within-cluster */
  uint32_t index; /* the triangle's index within its mesh (sort order's tie break) */
  uint32_t pad;
} ge_bvh_triangle;

/* The scene, as the engine's GpuScene holds it (systems/renderer/include/systems/renderer/gpu_scene.h
 * :171-191): global arrays with the offsets baked into each cluster. Copied in from host memory. */
typedef struct ge_scene_desc {
  const ge_cluster_desc* clusters;
  uint32_t cluster_count;
  uint32_t vertex_count;
  const float* vertices;     /* fp32 xyz per vertex, mesh space, what the AS builds read */
  const uint16_t* quantized; /* 3 per vertex, padded to an even count: the shading grid */
  const ge_vertex_attributes* attributes; /* per vertex, or NULL for flat shading */
  const uint32_t* triangles; /* packed i0 | i1 << 8 | i2 << 16 */
  uint32_t triangle_count;
  uint32_t mesh_count;
  const ge_mesh_desc* meshes;
  const ge_instance_desc* instances;
  uint32_t instance_count;
  uint32_t material_count;
  const ge_resolve_material* materials;
  const uint32_t* cluster_materials; /* per cluster */
  const ge_resolve_light* lights;
  uint32_t light_count;
  /* The bindless textures and samplers the materials name (level 0; none: 0 and NULL). */
  uint32_t texture_count;
  const ge_texture_desc* textures;
  const uint8_t* texels; /* the pool the textures' offsets point into */
  uint64_t texel_bytes;
  const ge_sampler_desc* samplers;
  uint32_t sampler_count;
  uint32_t pad;
} ge_scene_desc;

typedef struct ge_rt_ctx ge_rt_ctx;
typedef struct ge_rt_scene ge_rt_scene;

/* ---- entry points -------------------------------------------------------------------------- */

/* A context on HIP device `device_ordinal`, with its own stream. Replaces the engine's device and
 * pipeline creation in ReferenceRenderer::create (systems/renderer/src/reference.cpp:81-157). */
int32_t ge_rt_create(ge_rt_ctx** out, int32_t device_ordinal);
/* Idempotent; null is fine (ReferenceRenderer::destroy, reference.h:132). */
void ge_rt_destroy(ge_rt_ctx* ctx);
/* The last error's text, NUL-terminated and cut to `size`. Returns the status it described. */
int32_t ge_rt_last_error(const ge_rt_ctx* ctx, char* buffer, size_t size);
/* Waits for the context's stream. */
int32_t ge_rt_synchronize(ge_rt_ctx* ctx);

/* Uploads a scene (the GpuScene buffers: clusters, meshes, triangles, instances, attributes,
 * materials, cluster_materials, lights; gpu_scene.h:1693-1804). Validates offsets and counts. */
int32_t ge_rt_scene_upload(ge_rt_ctx* ctx, const ge_scene_desc* desc, ge_rt_scene** out);
void ge_rt_scene_destroy(ge_rt_ctx* ctx, ge_rt_scene* scene);

/* Builds the structures: one LBVH per mesh over its clusters' triangles (the engine's BLAS with one
 * geometry per cluster: create_blas/build_blas, domain/gfx/include/domain/gfx/acceleration.h:46-110,
 * domain/gfx/src/acceleration.cpp:134-169) and one LBVH over the instances in the frame of `eye`
 * (create_tlas/write_instances/build_tlas, acceleration.cpp:196-225, with the eye-relative 3x4 of
 * domain/gfx/shaders/tlas_references.slang:75-104). `eye` may be NULL (the world's origin).
 * flags: GE_RT_BUILD_FULL builds everything; GE_RT_BUILD_SUBSET rebuilds the meshes named in
 * `meshes[0..count)` and then the top level; GE_RT_BUILD_TOP rebuilds the top level only (for
 * instances that moved: re-upload them with ge_rt_scene_update_instances first).
 * GE_RT_BUILD_REFIT (no counterpart in the reference, whose structures are the driver's; the
 * per-frame update's cheaper option) rebuilds the meshes named, as GE_RT_BUILD_SUBSET does, then
 * refits the top level: the topology, sorted slots and codes of its last build are kept and its
 * boxes and wide nodes recomputed from the instances' current placement in the frame of `eye`. Hits
 * are a rebuild's (the walk's answer does not depend on the tree); the node visits grow as
 * instances move away from where the structure was built. It needs the same instances holding
 * geometry as at that build: an instance given another mesh since returns
 * GE_RT_ERROR_INVALID_SCENE. */
#define GE_RT_BUILD_FULL 0u
#define GE_RT_BUILD_SUBSET 1u
#define GE_RT_BUILD_TOP 2u
#define GE_RT_BUILD_REFIT 3u
int32_t ge_rt_bvh_build(ge_rt_ctx* ctx, ge_rt_scene* scene, uint32_t flags, const ge_frame_eye* eye,
                        const uint32_t* meshes, uint32_t count);
/* Replaces instance records [first, first + count) (the per-frame placement the engine's
 * tlas_references pass writes); follow with ge_rt_bvh_build(GE_RT_BUILD_TOP). */
int32_t ge_rt_scene_update_instances(ge_rt_ctx* ctx, ge_rt_scene* scene, uint32_t first,
                                     uint32_t count, const ge_instance_desc* instances);
/* Replaces material records [first, first + count) (the GpuScene's material buffer update); the
 * structures are unaffected. Host pointer. */
int32_t ge_rt_scene_update_materials(ge_rt_ctx* ctx, ge_rt_scene* scene, uint32_t first, uint32_t count,
                                     const ge_resolve_material* materials);
/* Replaces light records [first, first + count). Host pointer. Also rebuilds, on the host, the
 * scene's grid of per-cell light masks over lights 0..63 (16^3 cells, 32 KiB uploaded) that the
 * shading uses to skip lights out of range; an update equal to the records held does nothing. */
int32_t ge_rt_scene_update_lights(ge_rt_ctx* ctx, ge_rt_scene* scene, uint32_t first, uint32_t count,
                                  const ge_resolve_light* lights);
/* The frame's visible list (PathTraceParams::visible: cluster_cull's output, scene.slang:663-741):
 * `count` entries {instance, cluster}, `cluster` scene-wide and within the instance's mesh. From the
 * next ge_rt_bvh_build on, each instance without GE_INSTANCE_STATIC_RT traces a structure of this
 * frame built over its entries' clusters in list order, as the engine's per-frame cluster chain
 * builds one (cluster_acceleration.h:82-200); such an instance with no entry is met by no ray. Its
 * hits name the entry as their geometry (hit_visible_entry then hit_pair), and their keys stay
 * pair << 8 | triangle. Entries of static instances are checked and change nothing: those trace
 * their mesh's own structure. Host pointer; the list is copied. An instance out of range, a cluster
 * outside its instance's mesh or a repeated {instance, cluster} returns GE_RT_ERROR_INVALID_ARGUMENT
 * and keeps the previous list. A build with such instances and no list returns
 * GE_RT_ERROR_INVALID_SCENE. */
int32_t ge_rt_scene_set_visible(ge_rt_ctx* ctx, ge_rt_scene* scene, const uint32_t* entries, uint32_t count);
/* With ge_rt_bvh_info, ge_rt_bvh_download and ge_rt_bvh_download_wide, `mesh` =
 * GE_RT_FRAME_STRUCTURE | instance names the frame's structure of that (non-static) instance. */
#define GE_RT_FRAME_STRUCTURE 0x80000000u
/* Sizes of the built structure of mesh `mesh` (or of the top level for mesh == ~0). */
int32_t ge_rt_bvh_info(ge_rt_ctx* ctx, const ge_rt_scene* scene, uint32_t mesh, uint32_t* node_count,
                       uint32_t* prim_count);
/* Copies the built structure out to host memory: nodes (node_count), sorted primitive slots
 * (prim_count: triangles for a mesh, instance ids for the top level), 30-bit Morton codes in sorted
 * order (prim_count). Any output may be NULL. */
int32_t ge_rt_bvh_download(ge_rt_ctx* ctx, const ge_rt_scene* scene, uint32_t mesh,
                           ge_bvh_node* nodes, ge_bvh_triangle* triangles, uint32_t* instance_slots,
                           uint32_t* morton);
/* Copies out the 4-wide nodes of mesh `mesh` (~0: the top level): node_count entries (ge_rt_bvh_info)
 * at the binary nodes' indices, of which those at even depth (the root's is 0) are the walks' nodes
 * and the others unspecified. */
int32_t ge_rt_bvh_download_wide(ge_rt_ctx* ctx, const ge_rt_scene* scene, uint32_t mesh, ge_bvh_wide_node* wide);

/* Closest-hit queries (RayQuery<RAY_FLAG_NONE> committed hit, path_trace.slang:268-298; the tie
 * between hits at one t goes to the larger pair << 8 | triangle, ray_visibility.slang:13-30).
 * `mask` is the ray's 8-bit instance mask (0xfe camera, 0x01 caster). */
int32_t ge_rt_trace_closest(ge_rt_ctx* ctx, const ge_rt_scene* scene, const ge_ray* rays,
                            uint32_t count, uint32_t mask, ge_hit* hits, uint32_t flags);
/* Any-hit queries (RayQuery<FORCE_OPAQUE | ACCEPT_FIRST_HIT_AND_END_SEARCH>, path_trace.slang:180-207,
 * visibility_resolve.slang:599-603): bit i of occluded[i / 32] is set when anything lies in
 * [tmin, tmax]. A ray with tmax <= 0 is never occluded (occluded(), path_trace.slang:181). */
int32_t ge_rt_trace_any(ge_rt_ctx* ctx, const ge_rt_scene* scene, const ge_ray* rays, uint32_t count,
                        uint32_t mask, uint32_t* occluded, uint32_t flags);

/* The traced primary visibility pass (trace_main, domain/gfx/shaders/ray_visibility.slang:106-217):
 * one ray per pixel centre, the 64-bit word asuint(depth) << 32 | pair << 8 | triangle, 0 empty.
 * `words` receives width * height words (host memory, or device under GE_RT_DEVICE_PTR). */
int32_t ge_rt_ray_visibility(ge_rt_ctx* ctx, const ge_rt_scene* scene,
                             const ge_ray_visibility_params* params, uint64_t* words, uint32_t flags);

/* One batch of the reference path tracer (path_trace_main, path_trace.slang:463-495), replacing one
 * iteration of ReferenceRenderer::render's batch loop (reference.cpp:323-347). Adds params->spp
 * samples numbered from params->sample_base into accum (float4 per pixel: rgb sum, w count) and
 * ORs the primary coverage into coverage (u32 per pixel). With GE_RT_DEVICE_PTR both are device
 * pointers (null takes the block's own `accum` / `coverage` address). */
int32_t ge_rt_path_trace(ge_rt_ctx* ctx, const ge_rt_scene* scene, const ge_path_trace_params* params,
                         float* accum, uint32_t* coverage, uint32_t flags);
/* The path-replay adjoint of the path tracer, for inverse rendering. No counterpart in the reference
 * (its path tracer has no derivatives); it replays path_trace_main's paths (path_trace.slang:257-495)
 * from their counter RNG (path_trace.h:17-21) with sampling detached: directions, lobe choices,
 * pdfs and roulette come from `sampling` (material_count records; null: the scene's own), the
 * BSDF values from the scene's materials, so with equal tables the estimate is path_trace's.
 *   image (optional, 4 floats a pixel): the samples' radiance sum and their count, as accum;
 *   d_image (optional, 4 floats a pixel, the fourth unused): dLoss/dI for I = sum / spp; with it,
 *   grad_materials (material_count x 4 doubles: albedo r, g, b, roughness) and grad_lights
 *   (1 + light_count doubles: the sun's intensity sun[3], then each light's color_intensity[3])
 *   receive dLoss/dparameter, summed over every path in double.
 * Materials with their own specular record and max_bounces > 15 return GE_RT_ERROR_UNSUPPORTED.
 * By default it runs on the wavefront in two passes: the detached estimate of every path, then a
 * replay that adds each vertex's gradient from its next-event terms and (df/f) times what the path
 * gathered after it (path replay backpropagation). GE_RT_PT_MEGAKERNEL replays one path a lane with
 * its vertices stored and swept backwards; both give the same gradient up to float order. */
int32_t ge_rt_path_adjoint(ge_rt_ctx* ctx, const ge_rt_scene* scene, const ge_path_trace_params* params,
                           const float* d_image, const ge_resolve_material* sampling, float* image,
                           double* grad_materials, double* grad_lights, uint32_t flags);
/* The tonemap (tonemap_main, path_trace.slang:511-543): accum + coverage -> packed RGBA8. */
int32_t ge_rt_tonemap(ge_rt_ctx* ctx, const ge_path_trace_params* params, const float* accum,
                      const uint32_t* coverage, uint32_t* output, uint32_t flags);

/* The ray-traced sun shadow pass: one any-hit ray per receiver, origin = position + geometric *
 * ray_offset (shadow_visibility, visibility_resolve.slang:742-755), bit i of blocked[i / 32] set
 * when the sun is blocked (shadow_ray_blocked, :516-605). */
int32_t ge_rt_shadow_pass(ge_rt_ctx* ctx, const ge_rt_scene* scene, const ge_shadow_params* params,
                          const ge_shadow_receiver* receivers, uint32_t* blocked, uint32_t flags);

/* The sun shadow pass over a visibility buffer of width x height 64-bit words (ge_rt_ray_visibility's
 * or the rasterizer's: depth bits << 32 | pair << 8 | triangle, zero where empty). out[i] is
 * pair << 1 | blocked, or GE_SHADOW_NOT_TAKEN; i is the pixel (width x height entries), or with
 * GE_SHADOW_HALF_RATE the 2x2 block ((width + 1) / 2 x (height + 1) / 2 entries). Pairs map to
 * (instance, cluster) through the instances' first_pair (the engine's pair numbering). Replaces the
 * resolve's sun query (visibility_resolve.slang:742-755, 898-989) and the half-rate sun pass
 * (resolve_half_sun.slang:35-110). */
int32_t ge_rt_shadow_resolve(ge_rt_ctx* ctx, const ge_rt_scene* scene, const ge_shadow_resolve_params* params,
                             const uint64_t* visibility, uint32_t* out, uint32_t flags);

/* The allocation-free shadow resolve. No counterpart in the reference: it is ge_rt_shadow_resolve for
 * callers that cannot allocate device memory (code under HIP graph capture, tight per-frame
 * loops, harnesses that forbid hipMalloc). Its working memory, the walk's ray counter and a 16-byte
 * ray record for each entry, lives in a fixed array the library holds on each GPU it runs on (256
 * bytes + GE_RT_SHADOW_STATIC_ENTRIES records, 32 MiB, kept until the process ends) instead of the
 * context's growable buffer, and it calls no hipMalloc, hipFree or other allocating routine. The
 * answers are ge_rt_shadow_resolve's. Device pointers only (flags must hold GE_RT_DEVICE_PTR);
 * more than GE_RT_SHADOW_STATIC_ENTRIES entries return GE_RT_ERROR_INVALID_ARGUMENT; the scene's
 * pair table must be current, as every ge_rt_bvh_build leaves it (instances given another mesh or
 * pair range since the last build return GE_RT_ERROR_INVALID_SCENE; instances that only move keep
 * it current). Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (contexts on different streams) race; calls on one context are ordered
 * by its stream. */
#define GE_RT_SHADOW_STATIC_ENTRIES (1u << 21)
int32_t ge_rt_shadow_resolve_noalloc(ge_rt_ctx* ctx, const ge_rt_scene* scene, const ge_shadow_resolve_params* params,
                                     const uint64_t* visibility, uint32_t* out, uint32_t flags);

/* ---- The radiance cache ---------------------------------------------------------------------------
 * No counterpart in the reference: its GI cache is a world-space hash grid filled by a compute pass
 * (domain/gfx/include/domain/gfx/gi_cache.h:1-30, domain/gfx/shaders/gi_update.slang:85-115). Here
 * a small MLP predicts the radiance a surface point scatters toward a direction (everything but its
 * own emission), trained online on the path tracer's own path suffixes, and the cached tracer ends
 * its paths into it. The network: a 12-band sin/cos encoding of position (3), direction (2) and
 * normal (2), plus albedo (3), roughness (1) and a constant 1, padded to GE_CACHE_INPUTS; five
 * hidden ReLU layers of GE_CACHE_WIDTH; GE_CACHE_OUTPUTS output rows of which the first three are
 * RGB; no biases. BF16 weights and activations, FP32 accumulation (CDNA4 matrix cores), FP32 master
 * weights and Adam (beta1 0.9, beta2 0.99, epsilon 1e-8) on the relative L2 loss
 * (p - y)^2 / (p^2 + 0.01), p held fixed in the denominator. */
#define GE_CACHE_INPUTS 192
#define GE_CACHE_WIDTH 64
#define GE_CACHE_OUTPUTS 16
#define GE_CACHE_LAYERS 6
#define GE_CACHE_WEIGHTS 29696 /* 192*64 + 4*64*64 + 64*16: layer by layer, each [out][in] row-major */

/* One query, 48 bytes. position: the point in the cache's box, (p - box_min) / (box_max - box_min);
 * direction (towards the previous path vertex) and normal (the shading normal): octahedral
 * coordinates mapped to [0, 1]^2 (x, z over |x| + |y| + |z|, folded when y < 0). tag: the tracer's
 * bookkeeping (the path's bounce in a training sample), not an input. */
typedef struct ge_cache_query {
  float position[3];
  float direction[2];
  float normal[2];
  float albedo[3];
  float roughness;
  uint32_t tag;
} ge_cache_query;

/* One training sample, 64 bytes: a path vertex and the radiance its path suffix brought back to it,
 * (path radiance - radiance up to the vertex's emission) / throughput at the vertex. */
typedef struct ge_cache_sample {
  ge_cache_query query;
  float radiance[3];
  uint32_t path; /* sample * width * height + y * width + x */
} ge_cache_sample;

typedef struct ge_rt_cache ge_rt_cache;

/* A cache with He-uniform weights from `seed`, for positions in [box_min, box_max] (the path
 * tracer's frame: eye-relative, as its hit positions). */
int32_t ge_rt_cache_create(ge_rt_ctx* ctx, const float box_min[3], const float box_max[3], uint32_t seed,
                           ge_rt_cache** out);
void ge_rt_cache_destroy(ge_rt_ctx* ctx, ge_rt_cache* cache);
/* The FP32 master weights (GE_CACHE_WEIGHTS floats, host memory): `get` receives them, `set`
 * replaces them and restarts the optimizer. Either may be NULL. */
int32_t ge_rt_cache_weights(ge_rt_ctx* ctx, ge_rt_cache* cache, float* get, const float* set);
/* The FP32 weight gradient of the last optimizer step (ge_rt_cache_train, or the last step of
 * ge_rt_cache_update or a training ge_rt_path_trace_cached): d(batch mean loss)/d(weights), the
 * value Adam consumed, GE_CACHE_WEIGHTS floats into host `out`. Zero before the first step and
 * after weights are set. No counterpart in the reference, whose cache trains inside its frame: it
 * is for callers that check or log the backward pass. */
int32_t ge_rt_cache_gradient(ge_rt_ctx* ctx, ge_rt_cache* cache, float* out);
/* The network's raw outputs (n x 3 floats) for n queries. */
int32_t ge_rt_cache_infer(ge_rt_ctx* ctx, ge_rt_cache* cache, const ge_cache_query* queries, uint32_t n,
                          float* radiance, uint32_t flags);
/* One optimizer step over n samples; `loss` (host, may be NULL) receives the batch's mean loss. */
int32_t ge_rt_cache_train(ge_rt_ctx* ctx, ge_rt_cache* cache, const ge_cache_sample* samples, uint32_t n,
                          float learning_rate, double* loss, uint32_t flags);
/* Training samples from the path tracer: traces params' paths (its estimator, to max_bounces) and
 * writes one sample per path vertex that hit a surface, up to `capacity`; *count receives how many
 * vertices there were. Samples are in no particular order. */
int32_t ge_rt_cache_collect(ge_rt_ctx* ctx, const ge_rt_scene* scene, ge_rt_cache* cache,
                            const ge_path_trace_params* params, ge_cache_sample* samples, uint32_t capacity,
                            uint32_t* count, uint32_t flags);
/* The samples the last ge_rt_cache_collect, ge_rt_cache_update or training ge_rt_path_trace_cached
 * gathered, their targets filled in: up to `capacity` into `out`; *count receives how many were
 * kept. */
int32_t ge_rt_cache_samples(ge_rt_ctx* ctx, ge_rt_cache* cache, ge_cache_sample* out, uint32_t capacity,
                            uint32_t* count, uint32_t flags);
/* Online training: collect (at most `samples` samples) and `steps` optimizer steps over interleaved
 * slices of them. `loss` (host, may be NULL) receives the last step's mean loss. */
int32_t ge_rt_cache_update(ge_rt_ctx* ctx, const ge_rt_scene* scene, ge_rt_cache* cache,
                           const ge_path_trace_params* params, uint32_t samples, uint32_t steps, float learning_rate,
                           double* loss, uint32_t flags);
/* Online training inside a cached frame. Every stride-th path (phase sample_base % stride) goes on
 * past the cache's vertex: it records a training sample there, adds next-event estimation, takes
 * one more bounce and ends into the cache at the next hit (self-training); its own radiance is
 * the sample's target. The image keeps the cache's answer at that vertex for these paths too.
 * After the frame, `steps` optimizer steps run over interleaved slices of the samples. */
typedef struct ge_cache_train_params {
  uint32_t samples;    /* the most samples to gather; 0 trains nothing */
  uint32_t steps;      /* optimizer steps over them */
  float learning_rate;
  uint32_t stride;     /* 0: chosen from the last frame so that about `samples` paths reach the vertex */
} ge_cache_train_params;

/* The cached tracer: ge_rt_path_trace with each path ended at the surface its ray of bounce `depth`
 * hits, which adds its emission and throughput x max(cache prediction, 0) instead of next-event
 * estimation and further bounces. Same accum / coverage contract as ge_rt_path_trace. With `train`
 * (may be NULL) the frame also trains the cache; `loss` (host, may be NULL) receives the last
 * step's mean loss, which synchronizes. */
int32_t ge_rt_path_trace_cached(ge_rt_ctx* ctx, const ge_rt_scene* scene, ge_rt_cache* cache,
                                const ge_path_trace_params* params, uint32_t depth,
                                const ge_cache_train_params* train, float* accum, uint32_t* coverage, double* loss,
                                uint32_t flags);

/* Counters and the last kernel/build times (ReferenceFrame::trace_ms, reference.h:95). */
int32_t ge_rt_stats(ge_rt_ctx* ctx, ge_rt_counters* out);
int32_t ge_rt_reset_stats(ge_rt_ctx* ctx);

/* ---- The `.clusters` container ---------------------------------------------------------------
 * The engine's built-mesh container (domain/geometry/include/domain/geometry/cluster_file.h), read
 * from memory as geometry::read_cluster_file_memory reads it (domain/geometry/src/cluster_file.cpp:
 * 419-515, 691-800, 1184-1215): the header and section table, every section inside the file, the
 * content hash (core/hash hash_bytes over every byte after the header), the required sections at
 * their element sizes (clusters, lod, vertices, attributes, triangles, vertex_source,
 * level_cluster_counts, scalars), every cluster's runs inside the streams and every triangle's
 * corners inside its cluster, the quantized grid (the file's, or quantize_positions' from the
 * vertices), the optional materials and material map; unknown kinds are skipped. The scene it
 * gives: the finest level (lod level 0) as one mesh, one identity instance of it, and the materials
 * as GpuScene::fill_material fills their factors (systems/renderer/src/gpu_scene_textures.cpp:
 * 477-540), the alpha-mask mode as set_alpha_mask packs it. Images are not decoded: with
 * GE_CLUSTERS_IMAGES each texture slot names its image's index in the container's image list (the
 * caller uploads them as the scene's textures, in that order, and sampler 0), otherwise none.
 * Host code; no device is touched. On failure `error` holds the reason. */
#define GE_CLUSTERS_IMAGES 0x1u
typedef struct ge_clusters ge_clusters;
int32_t ge_clusters_read(const uint8_t* bytes, uint64_t size, uint32_t flags, ge_clusters** out, char* error,
                         size_t error_size);
/* The scene's arrays (pointers into the reader's copy, valid until ge_clusters_destroy). */
void ge_clusters_scene(const ge_clusters* clusters, ge_scene_desc* out);
/* The container's clusters over every level, and its levels. */
void ge_clusters_counts(const ge_clusters* clusters, uint32_t* all_clusters, uint32_t* levels);
void ge_clusters_destroy(ge_clusters* clusters);
/* The content hash the header holds (hash_bytes with the engine's seed). */
uint64_t ge_clusters_hash(const uint8_t* bytes, uint64_t size);

#ifdef __cplusplus
} /* extern "C" */

static_assert(sizeof(ge_cluster_desc) == 48, "geometry::ClusterDesc is 48 bytes");
static_assert(sizeof(ge_vertex_attributes) == 8, "geometry::VertexAttributes is 8 bytes");
static_assert(sizeof(ge_mesh_desc) == 64, "gfx::MeshDesc is 64 bytes");
static_assert(sizeof(ge_instance_desc) == 96, "gfx::InstanceDesc is 96 bytes");
static_assert(sizeof(ge_frame_eye) == 48, "gfx::FrameEye is 48 bytes");
static_assert(sizeof(ge_resolve_material) == 128, "gfx::ResolveMaterial is 128 bytes");
static_assert(sizeof(ge_resolve_light) == 64, "gfx::ResolveLight is 64 bytes");
static_assert(sizeof(ge_texture_desc) == 24, "a texture record is 24 bytes");
static_assert(sizeof(ge_sampler_desc) == 16, "a sampler record is 16 bytes");
static_assert(sizeof(ge_scene_desc) == 152, "the scene descriptor is 152 bytes");
static_assert(sizeof(ge_path_trace_params) == 368, "gfx::PathTraceParams is 368 bytes");
static_assert(sizeof(ge_ray_visibility_params) == 256, "gfx::RayVisibilityParams is 256 bytes");
static_assert(sizeof(ge_rt_counters) == 120, "counters record");
static_assert(sizeof(ge_ray) == 32, "ge_ray is 32 bytes");
static_assert(sizeof(ge_hit) == 32, "ge_hit is 32 bytes");
static_assert(sizeof(ge_bvh_node) == 64, "a BVH node is one 64-byte fetch");
static_assert(sizeof(ge_bvh_wide_node) == 64, "a wide node is one 64-byte line");
static_assert(sizeof(ge_bvh_triangle) == 48, "a BVH triangle is 48 bytes");
static_assert(sizeof(ge_shadow_receiver) == 32, "a shadow receiver is 32 bytes");
static_assert(sizeof(ge_shadow_resolve_params) == 128, "shadow resolve params are 128 bytes");
static_assert(sizeof(ge_cache_query) == 48, "a cache query is 48 bytes");
static_assert(sizeof(ge_cache_sample) == 64, "a cache sample is 64 bytes");
static_assert(sizeof(ge_cache_train_params) == 16, "cache training params are 16 bytes");
#endif

#endif /* GE_RT_H */
