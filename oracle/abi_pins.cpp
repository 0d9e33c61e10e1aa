// Compile-only check (test infrastructure): every engine record include/ge_rt.h mirrors has, field
// for field, the offset and size of the engine's own definition, and the mirrored constants their
// values. Built with -fsyntax-only against the reference's headers where they lie
// (domain/gfx/include/domain/gfx/path_trace.h:51-143, ray_visibility.h:39-61,
// cluster_cull.h:72-86,154-167,837-848,939-960, visibility_resolve.h:284-305,1012-1018,
// domain/geometry/include/domain/geometry/cluster.h:45-78); a mismatch fails the compile and names
// the field. Nothing here is linked into anything.
#include <cstddef>

#include "../include/ge_rt.h"

#include <domain/geometry/cluster.h>
#include <domain/gfx/cluster_cull.h>
#include <domain/gfx/path_trace.h>
#include <domain/gfx/ray_visibility.h>
#include <domain/gfx/visibility_resolve.h>

namespace g = engine::gfx;
namespace geo = engine::geometry;

#define GE_PIN_AS(OURS, THEIRS, OUR_FIELD, THEIR_FIELD)                                            \
  static_assert(offsetof(OURS, OUR_FIELD) == offsetof(THEIRS, THEIR_FIELD),                        \
                #OURS "::" #OUR_FIELD " is not at " #THEIRS "::" #THEIR_FIELD "'s offset");         \
  static_assert(sizeof(OURS::OUR_FIELD) == sizeof(THEIRS::THEIR_FIELD),                             \
                #OURS "::" #OUR_FIELD " is not " #THEIRS "::" #THEIR_FIELD "'s size")
#define GE_PIN(OURS, THEIRS, FIELD) GE_PIN_AS(OURS, THEIRS, FIELD, FIELD)
#define GE_SIZE(OURS, THEIRS) static_assert(sizeof(OURS) == sizeof(THEIRS), #OURS " is not " #THEIRS "'s size")

// geometry::ClusterDesc and VertexAttributes (cluster.h:45-55, 75-78).
GE_SIZE(ge_cluster_desc, geo::ClusterDesc);
GE_PIN(ge_cluster_desc, geo::ClusterDesc, vertex_offset);
GE_PIN(ge_cluster_desc, geo::ClusterDesc, triangle_offset);
GE_PIN(ge_cluster_desc, geo::ClusterDesc, vertex_count);
GE_PIN(ge_cluster_desc, geo::ClusterDesc, triangle_count);
GE_PIN(ge_cluster_desc, geo::ClusterDesc, center);
GE_PIN(ge_cluster_desc, geo::ClusterDesc, radius);
GE_PIN(ge_cluster_desc, geo::ClusterDesc, cone_apex);
GE_PIN(ge_cluster_desc, geo::ClusterDesc, cone);
GE_SIZE(ge_vertex_attributes, geo::VertexAttributes);
GE_PIN(ge_vertex_attributes, geo::VertexAttributes, normal_oct);
GE_PIN(ge_vertex_attributes, geo::VertexAttributes, uv_half2);

// gfx::FrameEye (cluster_cull.h:72-86).
GE_SIZE(ge_frame_eye, g::FrameEye);
GE_PIN(ge_frame_eye, g::FrameEye, cell);
GE_PIN(ge_frame_eye, g::FrameEye, masks);
GE_PIN(ge_frame_eye, g::FrameEye, local);
GE_PIN(ge_frame_eye, g::FrameEye, pad1);
GE_PIN(ge_frame_eye, g::FrameEye, residual);
GE_PIN(ge_frame_eye, g::FrameEye, pad2);

// gfx::MeshDesc (cluster_cull.h:837-848).
GE_SIZE(ge_mesh_desc, g::MeshDesc);
GE_PIN(ge_mesh_desc, g::MeshDesc, quant);
GE_PIN(ge_mesh_desc, g::MeshDesc, quantized);
GE_PIN(ge_mesh_desc, g::MeshDesc, first_cluster);
GE_PIN(ge_mesh_desc, g::MeshDesc, cluster_count);
GE_PIN(ge_mesh_desc, g::MeshDesc, deform_pool);
GE_PIN(ge_mesh_desc, g::MeshDesc, deform_slots);
GE_PIN(ge_mesh_desc, g::MeshDesc, templates);
GE_PIN(ge_mesh_desc, g::MeshDesc, skin);

// gfx::InstanceDesc and its flags (cluster_cull.h:679, 882, 939-960).
GE_SIZE(ge_instance_desc, g::InstanceDesc);
GE_PIN(ge_instance_desc, g::InstanceDesc, rows);
GE_PIN(ge_instance_desc, g::InstanceDesc, cell);
GE_PIN(ge_instance_desc, g::InstanceDesc, mesh);
GE_PIN(ge_instance_desc, g::InstanceDesc, material_base);
GE_PIN(ge_instance_desc, g::InstanceDesc, first_pair);
GE_PIN(ge_instance_desc, g::InstanceDesc, scale_max);
GE_PIN(ge_instance_desc, g::InstanceDesc, flags);
GE_PIN(ge_instance_desc, g::InstanceDesc, deform);
GE_PIN(ge_instance_desc, g::InstanceDesc, bounds_padding);
GE_PIN(ge_instance_desc, g::InstanceDesc, terrain);
GE_PIN(ge_instance_desc, g::InstanceDesc, band);
static_assert(GE_INSTANCE_STATIC_RT == g::k_instance_static_rt);
static_assert(GE_INSTANCE_UNTRACED == g::k_instance_untraced);
static_assert(GE_INVALID_DEFORM == g::k_invalid_deform);

// gfx::ResolveMaterial and its flags (visibility_resolve.h:38, 268, 284-305, 337-343, 499-501).
GE_SIZE(ge_resolve_material, g::ResolveMaterial);
GE_PIN(ge_resolve_material, g::ResolveMaterial, albedo);
GE_PIN(ge_resolve_material, g::ResolveMaterial, emissive);
GE_PIN(ge_resolve_material, g::ResolveMaterial, albedo_texture);
GE_PIN(ge_resolve_material, g::ResolveMaterial, sampler);
GE_PIN(ge_resolve_material, g::ResolveMaterial, uv_scale);
GE_PIN(ge_resolve_material, g::ResolveMaterial, flags);
GE_PIN(ge_resolve_material, g::ResolveMaterial, metallic_roughness_texture);
GE_PIN(ge_resolve_material, g::ResolveMaterial, normal_texture);
GE_PIN(ge_resolve_material, g::ResolveMaterial, normal_scale);
GE_PIN(ge_resolve_material, g::ResolveMaterial, occlusion_texture);
GE_PIN(ge_resolve_material, g::ResolveMaterial, emissive_texture);
GE_PIN(ge_resolve_material, g::ResolveMaterial, occlusion_strength);
GE_PIN(ge_resolve_material, g::ResolveMaterial, samplers_mr_normal);
GE_PIN(ge_resolve_material, g::ResolveMaterial, samplers_occlusion_emissive);
GE_PIN(ge_resolve_material, g::ResolveMaterial, uv_transform);
GE_PIN(ge_resolve_material, g::ResolveMaterial, uv_offset);
GE_PIN(ge_resolve_material, g::ResolveMaterial, specular);
static_assert(GE_NO_TEXTURE == g::k_no_texture);
static_assert(GE_SAME_SAMPLER == (g::k_same_samplers & 0xffffu) && GE_SAME_SAMPLER == (g::k_same_samplers >> 16));
static_assert(GE_MATERIAL_MIPPED == g::k_material_mipped);
static_assert(GE_MATERIAL_NORMAL_RG == g::k_material_normal_rg);
static_assert(GE_MATERIAL_UV_TRANSFORM == g::k_material_uv_transform);
static_assert(GE_MATERIAL_MASKED == g::k_material_alpha_mask);
static_assert(GE_MATERIAL_STOCHASTIC == g::k_material_stochastic);

// gfx::ResolveLight (visibility_resolve.h:1001-1018).
GE_SIZE(ge_resolve_light, g::ResolveLight);
GE_PIN(ge_resolve_light, g::ResolveLight, position_radius);
GE_PIN(ge_resolve_light, g::ResolveLight, color_intensity);
GE_PIN(ge_resolve_light, g::ResolveLight, direction_type);
GE_PIN(ge_resolve_light, g::ResolveLight, spot);

// gfx::PathTraceParams and its flags (path_trace.h:35-48, 51-143, 147).
GE_SIZE(ge_path_trace_params, g::PathTraceParams);
GE_PIN(ge_path_trace_params, g::PathTraceParams, clip_to_ray);
GE_PIN(ge_path_trace_params, g::PathTraceParams, camera);
GE_PIN(ge_path_trace_params, g::PathTraceParams, sky);
GE_PIN(ge_path_trace_params, g::PathTraceParams, sun);
GE_PIN(ge_path_trace_params, g::PathTraceParams, ground);
GE_PIN(ge_path_trace_params, g::PathTraceParams, accum);
GE_PIN(ge_path_trace_params, g::PathTraceParams, output);
GE_PIN(ge_path_trace_params, g::PathTraceParams, clusters);
GE_PIN(ge_path_trace_params, g::PathTraceParams, mesh);
GE_PIN(ge_path_trace_params, g::PathTraceParams, triangles);
GE_PIN(ge_path_trace_params, g::PathTraceParams, instances);
GE_PIN(ge_path_trace_params, g::PathTraceParams, attributes);
GE_PIN(ge_path_trace_params, g::PathTraceParams, materials);
GE_PIN(ge_path_trace_params, g::PathTraceParams, cluster_materials);
GE_PIN(ge_path_trace_params, g::PathTraceParams, visible);
GE_PIN(ge_path_trace_params, g::PathTraceParams, lights);
GE_PIN(ge_path_trace_params, g::PathTraceParams, width);
GE_PIN(ge_path_trace_params, g::PathTraceParams, height);
GE_PIN(ge_path_trace_params, g::PathTraceParams, light_count);
GE_PIN(ge_path_trace_params, g::PathTraceParams, scene);
GE_PIN(ge_path_trace_params, g::PathTraceParams, spp);
GE_PIN(ge_path_trace_params, g::PathTraceParams, sample_base);
GE_PIN(ge_path_trace_params, g::PathTraceParams, max_bounces);
GE_PIN(ge_path_trace_params, g::PathTraceParams, seed);
GE_PIN(ge_path_trace_params, g::PathTraceParams, flags);
GE_PIN(ge_path_trace_params, g::PathTraceParams, ray_bias);
GE_PIN(ge_path_trace_params, g::PathTraceParams, coverage);
GE_PIN(ge_path_trace_params, g::PathTraceParams, background);
GE_PIN(ge_path_trace_params, g::PathTraceParams, ray_bias_steps);
GE_PIN(ge_path_trace_params, g::PathTraceParams, ground_detail);
GE_PIN(ge_path_trace_params, g::PathTraceParams, sky_params);
GE_PIN(ge_path_trace_params, g::PathTraceParams, eye);
GE_PIN(ge_path_trace_params, g::PathTraceParams, terrain_uv_x0_mm);
GE_PIN(ge_path_trace_params, g::PathTraceParams, terrain_uv_z0_mm);
GE_PIN(ge_path_trace_params, g::PathTraceParams, terrain_uv_per_mm);
GE_PIN(ge_path_trace_params, g::PathTraceParams, extra_vertices);
GE_PIN(ge_path_trace_params, g::PathTraceParams, hit_entries);
GE_PIN(ge_path_trace_params, g::PathTraceParams, extra);
static_assert(GE_PT_UNIFORM_SKY == g::k_pt_uniform_sky);
static_assert(GE_PT_PIXEL_CENTER == g::k_pt_pixel_center);
static_assert(GE_PT_DITHER == g::k_pt_dither);
static_assert(GE_PT_ALPHA_MASKS == g::k_pt_alpha_masks);
static_assert(GE_PT_ROULETTE_DEPTH == g::k_pt_roulette_depth);

// gfx::RayVisibilityParams (ray_visibility.h:39-61), its MaskTables block (cluster_cull.h:154-167)
// taken as four addresses.
GE_SIZE(ge_ray_visibility_params, g::RayVisibilityParams);
GE_PIN(ge_ray_visibility_params, g::RayVisibilityParams, view_proj);
GE_PIN(ge_ray_visibility_params, g::RayVisibilityParams, clip_to_ray);
GE_PIN(ge_ray_visibility_params, g::RayVisibilityParams, camera);
GE_PIN(ge_ray_visibility_params, g::RayVisibilityParams, output);
GE_PIN(ge_ray_visibility_params, g::RayVisibilityParams, instance_base);
GE_PIN(ge_ray_visibility_params, g::RayVisibilityParams, width);
GE_PIN(ge_ray_visibility_params, g::RayVisibilityParams, height);
GE_PIN(ge_ray_visibility_params, g::RayVisibilityParams, scene);
GE_PIN(ge_ray_visibility_params, g::RayVisibilityParams, caster_base);
GE_PIN(ge_ray_visibility_params, g::RayVisibilityParams, visible);
GE_PIN(ge_ray_visibility_params, g::RayVisibilityParams, instances);
GE_PIN(ge_ray_visibility_params, g::RayVisibilityParams, meshes);
GE_PIN(ge_ray_visibility_params, g::RayVisibilityParams, triangles);
GE_PIN(ge_ray_visibility_params, g::RayVisibilityParams, clusters);
GE_PIN(ge_ray_visibility_params, g::RayVisibilityParams, hit_entries);
GE_PIN(ge_ray_visibility_params, g::RayVisibilityParams, masks);
static_assert(offsetof(g::MaskTables, materials) == 0 * sizeof(uint64_t));
static_assert(offsetof(g::MaskTables, cluster_materials) == 1 * sizeof(uint64_t));
static_assert(offsetof(g::MaskTables, attributes) == 2 * sizeof(uint64_t));
static_assert(offsetof(g::MaskTables, fade) == 3 * sizeof(uint64_t));

// The check's own check: a field compared with its neighbour must fail the compile.
#ifdef GE_ABI_PINS_PLANT_MISMATCH
GE_PIN_AS(ge_path_trace_params, g::PathTraceParams, width, height);
#endif

int ge_abi_pins_compiled;
