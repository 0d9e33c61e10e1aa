/*
 * scene_synth.h — seeded, deterministic synthetic scenes in the engine's GPU formats.
 *
 * The configurations the backend is measured and tested on. Every array is emitted in the engine's
 * layout (geometry::ClusterDesc, the packed i0|i1<<8|i2<<16 triangles, cluster-ordered fp32
 * vertices, the 16-bit quantized grid of geometry::quantize_positions
 * (domain/geometry/src/cluster.cpp:437-466), octahedral normals, gfx::MeshDesc, gfx::InstanceDesc,
 * gfx::ResolveMaterial, gfx::ResolveLight), so the arrays go straight into ge_rt_scene_upload.
 *
 *   config 1: 256 objects (128 icospheres of 320 triangles and 128 boxes of 192) at uniform random
 *             positions in a 100 m cube, 65,536 triangles, one mesh in clusters of <= 128
 *             triangles, one sun plus a constant sky, Lambertian albedo U[0.1, 0.9] per object.
 *   config 2: `detail` instances (default 4096) of 16 distinct 1,024-triangle meshes (8 clusters of
 *             128) under random rigid transforms, GGX roughness U[0.05, 1], one sun plus 64 point
 *             lights with a radius of influence.
 *   config 3: `detail` (default 16,384) unique cluster-group meshes of 8 clusters of 128 triangles,
 *             one instance each, plus a (n x n) fBm heightfield terrain (n = 4096 by default, 6
 *             octaves of PCG-hashed value noise, lacunarity 2, gain 0.5) in clusters of 128.
 *   config 4: 2,048 objects as config 1 (524,288 triangles) in a 200 m cube, 1,024 materials
 *             (albedo U[0.1, 0.9], roughness U[0.05, 1], one in five metal), the sun and eight
 *             point lights, 16 spp and 6 bounces: the inverse-rendering scene.
 *   config 5: `detail` trees (default 4096) of 8 meshes, each an 8-sided trunk of 96 triangles and
 *             320 to 480 alpha-tested leaf cards (two triangles each) in its crown, on a jittered
 *             6 m grid over gently rolling ground of 32,768 triangles; textures at level 0 (a 4 x 4
 *             leaf atlas whose alpha is the mask, read through a UV transform; bark colour and a
 *             normal map; ground colour and a metallic-roughness map), three samplers (linear
 *             repeat, linear clamp, nearest mirrored), the sun and sky: the alpha-masked workload.
 */
#ifndef GE_SCENE_SYNTH_H
#define GE_SCENE_SYNTH_H

#include "../../include/ge_rt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ge_synth ge_synth;

/* A scene of configuration 1 to 5 (above). `detail` = 0 takes the configuration's size;
 * `terrain` is config 3's heightfield samples per side (0: 4096). */
ge_synth* ge_synth_create(uint32_t config, uint64_t seed, uint32_t detail, uint32_t terrain);
/* One mesh, one identity instance, from raw triangles (9 floats each) in clusters of
 * `cluster_size` (<= 128) triangles, flat octahedral normals (`attributes`; 0: none), UVs from
 * `uvs` (6 floats a triangle, stored as halves; NULL: zero), one material per cluster with albedo
 * from `seed`. */
ge_synth* ge_synth_from_triangles(const float* positions, uint32_t triangle_count, uint32_t cluster_size,
                                  int32_t attributes, uint64_t seed, const float* uvs);
/* Replaces the materials (and, when not NULL, the per-cluster material indices). */
void ge_synth_set_materials(ge_synth* s, const ge_resolve_material* materials, uint32_t count,
                            const uint32_t* cluster_materials);
/* Replaces the textures, their texel pool and the samplers (copied). */
void ge_synth_set_textures(ge_synth* s, const ge_texture_desc* textures, uint32_t texture_count, const uint8_t* texels,
                           uint64_t texel_bytes, const ge_sampler_desc* samplers, uint32_t sampler_count);
void ge_synth_destroy(ge_synth* s);
/* Pointers into the synth's arrays (valid until ge_synth_destroy). */
void ge_synth_scene(const ge_synth* s, ge_scene_desc* out);
/* The configuration's camera and lighting for a width x height frame: fills clip_to_ray, camera,
 * sky, sun, ground, light_count, ray_bias(_steps), background, spp, max_bounces, seed; and the
 * visibility block's view_proj, clip_to_ray, camera, width, height. Either may be NULL. */
void ge_synth_frame(const ge_synth* s, uint32_t width, uint32_t height, ge_path_trace_params* pt,
                    ge_ray_visibility_params* rv);
/* Overrides the camera: eye, target, vertical field of view in degrees. */
void ge_synth_set_camera(ge_synth* s, const float eye[3], const float target[3], float fov_y_deg);
/* Moves instances [first, first+count) by a seeded rigid motion of up to `step` metres and `angle`
 * radians (the per-frame motion of config 3) and writes the new records into `out`. */
void ge_synth_move_instances(ge_synth* s, uint32_t first, uint32_t count, uint64_t seed, float step,
                             float angle, ge_instance_desc* out);
/* Jittered primary rays, one per pixel (x + hash, y + hash from pcg_hash of the pixel and `seed`),
 * for the shadow pass's receivers: `rays` receives width * height rays. */
void ge_synth_primary_rays(const ge_synth* s, uint32_t width, uint32_t height, uint32_t seed, ge_ray* rays);
/* splitmix64 and the unit draw it gives, exposed for tests. */
uint64_t ge_synth_splitmix64(uint64_t* state);

#ifdef __cplusplus
}
#endif
#endif
