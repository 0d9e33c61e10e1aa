/*
 * ge_oracle.h — the CPU oracle's C ABI. TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker the HIP backend is compared against, and the CPU baseline the bench
 * times beside it. Only the tests, the smoke check and the bench's cpu_baseline leg load it; the
 * product (game_engine_amd/) never does.
 *
 * It restates, in scalar C++ compiled with -ffp-contract=off (the engine's CPU policy,
 * cmake/EngineFpContraction.cmake:5-53), the reference's path-tracer sources:
 *   domain/gfx/shaders/path_trace.slang, sampling.slang, brdf.slang, the material.slang and
 *   scene.slang subsets the configurations use, view_ray.slang, ray_visibility.slang, and the shadow
 *   query of visibility_resolve.slang:516-605,742-755,
 * with a software BVH (LBVH: Morton codes, stable radix sort, Karras 2012 emission, bottom-up fit)
 * standing in for the driver's acceleration structures and `RayQuery`, and a watertight
 * ray/triangle test (Woop, Benthin, Wald 2013) standing in for the hardware's.
 *
 * Two arithmetic modes: GE_ORACLE_FP32 mirrors the shader's operation order in float; GE_ORACLE_FP64
 * runs the same integrator in double (the BVH and traversal stay the float ones).
 */
#ifndef GE_ORACLE_H
#define GE_ORACLE_H

#include "../include/ge_rt.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GE_ORACLE_FP32 0u
#define GE_ORACLE_FP64 1u

typedef struct ge_oracle_scene ge_oracle_scene;

/* Builds every mesh's LBVH and the top level in the frame of `eye` (NULL: origin). */
ge_oracle_scene* ge_oracle_scene_create(const ge_scene_desc* desc, const ge_frame_eye* eye);
void ge_oracle_scene_destroy(ge_oracle_scene* scene);
int32_t ge_oracle_bvh_info(const ge_oracle_scene* scene, uint32_t mesh, uint32_t* node_count,
                           uint32_t* prim_count);
int32_t ge_oracle_bvh_download(const ge_oracle_scene* scene, uint32_t mesh, ge_bvh_node* nodes,
                               ge_bvh_triangle* triangles, uint32_t* instance_slots,
                               uint32_t* morton);
/* The 4-wide nodes of mesh `mesh` (~0: the top level) at the binary nodes' indices (node_count
 * entries), references by binary index as the backend lays them out; zero at odd depth. */
int32_t ge_oracle_bvh_download_wide(const ge_oracle_scene* scene, uint32_t mesh, ge_bvh_wide_node* wide);
/* The frame's visible list ({instance, cluster} pairs): each instance without InstanceDesc::flags
 * bit 7 gets a structure over its entries' clusters in list order, and the top level is rebuilt.
 * GE_RT_ERROR_INVALID_ARGUMENT (nothing changed) for an instance out of range, a cluster outside its
 * instance's mesh or a repeated pair. The download functions take mesh = bit 31 | instance for
 * that instance's frame structure. */
int32_t ge_oracle_set_visible(ge_oracle_scene* scene, const uint32_t* entries, uint32_t count);
/* Replaces instance records and rebuilds the top level (the per-frame placement). */
int32_t ge_oracle_update_instances(ge_oracle_scene* scene, uint32_t first, uint32_t count,
                                   const ge_instance_desc* instances, const ge_frame_eye* eye);
/* The refit of the instance structure (GE_RT_BUILD_REFIT): its topology and slots kept, its boxes
 * and wide nodes recomputed; GE_RT_ERROR_INVALID_SCENE (and nothing changed) when the instances
 * with geometry are no longer the structure's slots. */
int32_t ge_oracle_refit_instances(ge_oracle_scene* s, uint32_t first, uint32_t count, const ge_instance_desc* in,
                                  const ge_frame_eye* eye);

/* Queries. `brute` = 1 tests every triangle of every instance instead of walking the BVH. `flags`:
 * GE_RT_ALPHA_MASKS or GE_RT_CULL_MASKED (closest only), as ge_rt_trace_closest/any take them. */
void ge_oracle_trace_closest(const ge_oracle_scene* scene, const ge_ray* rays, uint32_t count,
                             uint32_t mask, ge_hit* hits, int32_t brute, int32_t threads, uint32_t flags);
void ge_oracle_trace_any(const ge_oracle_scene* scene, const ge_ray* rays, uint32_t count,
                         uint32_t mask, uint32_t* occluded, int32_t brute, int32_t threads, uint32_t flags);
void ge_oracle_ray_visibility(const ge_oracle_scene* scene, const ge_ray_visibility_params* params,
                              uint64_t* words, int32_t threads);
/* One batch of path_trace_main over rows [row0, row1) (all rows: 0, height). */
void ge_oracle_path_trace(const ge_oracle_scene* scene, const ge_path_trace_params* params,
                          float* accum, uint32_t* coverage, uint32_t mode, uint32_t row0,
                          uint32_t row1, int32_t threads);
/* Same, FP64 accumulation into a double[4] per pixel (mode FP64 only; accum64 is w*h*4). */
void ge_oracle_path_trace64(const ge_oracle_scene* scene, const ge_path_trace_params* params,
                            double* accum64, uint32_t* coverage, uint32_t row0, uint32_t row1,
                            int32_t threads);
void ge_oracle_tonemap(const ge_path_trace_params* params, const float* accum,
                       const uint32_t* coverage, uint32_t* output);
void ge_oracle_shadow_pass(const ge_oracle_scene* scene, const ge_shadow_params* params,
                           const ge_shadow_receiver* receivers, uint32_t* blocked, int32_t threads,
                           uint32_t flags);
/* The sun shadow pass over a visibility buffer (ge_rt_shadow_resolve's contract); `rays`, when not
 * null, receives each entry's ray origin and t_max (t_max -1: no ray), 4 floats an entry. */
int32_t ge_oracle_shadow_resolve(const ge_oracle_scene* s, const ge_shadow_resolve_params* p, const uint64_t* words,
                                 uint32_t* out, float* rays, int32_t threads);

/* The path-replay adjoint (ge_rt_path_adjoint's contract) in float (mode FP32, the GPU's mirror)
 * or double (FP64): per pixel the sum over its samples of the detached estimator (sampling from
 * `sampling`, or the scene's materials when null; `image_sum` in FP32 mode, `image64` in FP64, 4
 * values a pixel, the fourth the sample count) and, when `d_image` (dLoss/dI, 4 floats a pixel) is
 * given, the loss gradient over material albedo rgb and roughness (4 doubles a material) and light
 * intensities (sun, then each light). */
/* The cached tracer's view of each path of the dispatch (path i = sample * width * height + pixel):
 * radiance[i * 3]: what the path gathered before the cache's term; weight[i * 4]: the throughput
 * at the surface its ray of bounce `depth` hit and 1 when it ended there into the cache (0: it
 * missed or ended earlier); queries[i]: that query (tag = i). No counterpart in the reference. */
void ge_oracle_path_cached(const ge_oracle_scene* s, const ge_path_trace_params* p, uint32_t depth,
                           const float box_lo[3], const float box_inv[3], float* radiance, float* weight,
                           ge_cache_query* queries, int32_t threads);
/* The cached tracer with self-training: paths with i % stride == phase record a sample at the
 * surface of bounce `depth` (its query, the radiance after its emission as `radiance`, path i) and
 * go on one bounce, ending into the cache at the next hit. Per path: the image's radiance before
 * the cache term, its weight and query (as ge_oracle_path_cached); the sample and its throughput
 * (sample_weight[i * 4 + 3] = 1 when there is one); the training path's own radiance before its
 * cache term, that term's weight (train_weight[i * 4 + 3] = 1 when it ended into the cache) and
 * query. No counterpart in the reference. */
void ge_oracle_path_cached_training(const ge_oracle_scene* s, const ge_path_trace_params* p, uint32_t depth,
                                    uint32_t stride, uint32_t phase, const float box_lo[3], const float box_inv[3],
                                    float* radiance, float* weight, ge_cache_query* queries, ge_cache_sample* samples,
                                    float* sample_weight, float* train_radiance, float* train_weight,
                                    ge_cache_query* train_queries, int32_t threads);
/* The plain tracer's training samples: one per hit vertex, ordered by path then bounce (query.tag
 * = bounce, path = i), up to `capacity`; returns how many there were. */
uint32_t ge_oracle_cache_samples(const ge_oracle_scene* s, const ge_path_trace_params* p, const float box_lo[3],
                                 const float box_inv[3], ge_cache_sample* out, uint32_t capacity, int32_t threads);
int32_t ge_oracle_path_adjoint(const ge_oracle_scene* s, const ge_path_trace_params* p, const float* d_image,
                               const ge_resolve_material* sampling, uint32_t mode, float* image_sum, double* image64,
                               double* grad_materials, double* grad_lights, int32_t threads);
int32_t ge_oracle_update_materials(ge_oracle_scene* s, uint32_t first, uint32_t count, const ge_resolve_material* m);
int32_t ge_oracle_update_lights(ge_oracle_scene* s, uint32_t first, uint32_t count, const ge_resolve_light* l);

/* Traversal counters accumulated since the last reset: closest rays, any rays, then nodes,
 * triangles and instances of the closest-hit walks, then those of the any-hit walks, then the
 * masks evaluated by the closest-hit and the any-hit walks. */
void ge_oracle_counters(uint64_t out[10]);
void ge_oracle_reset_counters(void);
int32_t ge_oracle_max_threads(void);

/* ---- unit probes, for golden vectors and closed-form checks ---- */
uint32_t ge_oracle_pcg_hash(uint32_t v);
uint32_t ge_oracle_rng_seed(uint32_t x, uint32_t y, uint32_t sample_index, uint32_t seed);
/* Advances `state` and returns the draw; `state` is updated in place. */
float ge_oracle_rng_next(uint32_t* state);
uint32_t ge_oracle_morton3(float cx, float cy, float cz, const float lo[3], const float scale[3]);
void ge_oracle_sincos(float x, float* s, float* c);
/* a*b - c*d by the watertight test's accurate fallback. */
float ge_oracle_det2(float a, float b, float c, float d);
/* relative_axis (world.h:181-188): a position's axis from the frame's eye. */
float ge_oracle_relative_axis(int32_t cell, float local, int32_t eye_cell, float eye_local, float eye_residual);
/* Sampler and BSDF, in double (mode FP64) or float (FP32): `in`/`out` layouts per kind, see the
 * source. kind: 0 cosine hemisphere, 1 GGX VNDF sample, 2 ggx_vndf_pdf, 3 brdf_eval, 4 bsdf_pdf,
 * 5 specular_probability, 6 onb_from_normal, 7 ggx_g1, 8 v_smith, 9 d_ggx, 10 fresnel,
 * 11 mis_power_heuristic, 12 light_sample, 13 decode_normal_oct, 14 display_output,
 * 15 transform_normal, 16 brdf_direct, 17 brdf_eval_grad (df/d albedo per channel, df/d roughness). */
int32_t ge_oracle_probe(uint32_t kind, uint32_t mode, const double* in, double* out);
/* Level-0 texture reads: out[4 i .. 4 i + 3] = SampleLevel(texture, sampler, uv[2 i], uv[2 i + 1]). */
void ge_oracle_sample(const ge_oracle_scene* scene, uint32_t texture, uint32_t sampler, const float* uv,
                      uint32_t count, float* out);
/* The hit decode of the path tracer for committed hits (instance_id ~0: zeros): 16 floats a hit,
 * position, the shading normal after the normal map (before it is turned to the ray's side),
 * albedo, roughness, metallic, emissive, and the interpolated UV; `towards` orients the geometric
 * normal. */
void ge_oracle_shade(const ge_oracle_scene* scene, const ge_hit* hits, uint32_t count, const float towards[3],
                     float* out);
/* The structures every walk takes: 2, the binary structures as built; 4, their 4-wide collapse. */
void ge_oracle_walk_width(int32_t width);
/* f16tof32 of the low 16 bits, and the sRGB decode table's entry for an 8-bit code. */
float ge_oracle_half(uint32_t h);
float ge_oracle_srgb(uint32_t code);
/* Watertight ray/triangle test: out = {hit, t, u, v}. */
void ge_oracle_intersect(const float o[3], const float d[3], const float v0[3], const float v1[3],
                         const float v2[3], float tmin, float tmax, double out[4]);

#ifdef __cplusplus
}
#endif
#endif
