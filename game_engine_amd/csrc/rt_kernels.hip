// rt_kernels.hip — the traced passes: closest-hit and any-hit queries, the primary visibility word,
// the reference path tracer (one path per lane, the megakernel form of path_trace_main), its
// tonemap, and the sun shadow pass.
//
// Launch shapes: 64-lane workgroups (one wave64). The path tracer's workgroup is an 8 x 8 pixel
// tile, the reference's numthreads(8, 8, 1) (path_trace.h:149-154), so one wave traces a compact
// block of coherent primaries. Counters, when asked for, are summed per wave before one atomic.
#include "rt_internal.h"
#include "rt_shade.h"

#include <algorithm>

namespace ge {
namespace {

// The per-lane traversal stack rings (dynamic LDS: k_ring entries x 64 lanes x 4 bytes).
extern __shared__ uint32_t g_lds[];
__device__ __forceinline__ uint32_t* lane_stack() { return g_lds + threadIdx.x; }

template <bool Count, bool Masks>
__global__ void __launch_bounds__(64) GE_WALK_ATTR k_trace_closest(Bvh bvh, const ge_ray* __restrict__ rays, uint32_t n,
                                                      uint32_t mask, uint32_t* head, ge_hit* __restrict__ hits,
                                                      unsigned long long* counters, uint32_t alpha, GeomMap geom) {
  Tally tally;
  uint64_t issued = 0;
  auto load = [&](uint32_t i, uint32_t& slot, Ray& ray, float& tmax, uint32_t& m) {
    const ge_ray r = rays[i];
    slot = i;
    ray = make_ray(f3{r.origin[0], r.origin[1], r.origin[2]}, f3{r.dir[0], r.dir[1], r.dir[2]}, r.tmin);
    tmax = r.tmax;
    m = mask;
    return true;
  };
  auto store = [&](uint32_t i, const Hit& h) {
    ge_hit out;
    out.pad = 0;
    if (h.found) {
      out.t = h.t;
      out.u = h.u;
      out.v = h.v;
      out.instance_id = h.inst;
      out.geometry = geom_of(geom, h.inst, h.prim >> 8);
      out.primitive = h.prim & 0xffu;
      out.key = h.key;
    } else {
      out.t = out.u = out.v = 0.0f;
      out.instance_id = out.geometry = out.primitive = GE_RT_MISS;
      out.key = 0;
    }
    hits[i] = out;
  };
  persistent_walk<false, Count, Masks>(bvh, n, head, lane_stack(), load, store, issued, tally, alpha);
  if (Count) flush_counters(counters, issued, 0, tally);
}

// Occlusion bits by atomic OR into words the host cleared: a persistent walk finishes rays in any
// order.
template <bool Count, bool Masks>
__global__ void __launch_bounds__(64) GE_WALK_ATTR k_trace_any(Bvh bvh, const ge_ray* __restrict__ rays, uint32_t n, uint32_t mask,
                                                  uint32_t* head, uint32_t* __restrict__ bits,
                                                  unsigned long long* counters, uint32_t alpha) {
  Tally tally;
  uint64_t issued = 0;
  auto load = [&](uint32_t i, uint32_t& slot, Ray& ray, float& tmax, uint32_t& m) {
    const ge_ray r = rays[i];
    slot = i;
    tmax = r.tmax;
    m = mask;
    if (!(r.tmax > 0.0f)) return false;
    ray = make_ray(f3{r.origin[0], r.origin[1], r.origin[2]}, f3{r.dir[0], r.dir[1], r.dir[2]}, r.tmin);
    return true;
  };
  auto store = [&](uint32_t i, const Hit& h) {
    if (h.found) atomicOr(&bits[i >> 5], 1u << (i & 31u));
  };
  persistent_walk<true, Count, Masks>(bvh, n, head, lane_stack(), load, store, issued, tally, alpha);
  if (Count) flush_counters(counters, 0, issued, tally);
}

// trace_main (ray_visibility.slang:106-217): the pixel centre's ray, the committed hit (whose
// (t, key) order already settles a depth tie to the larger pair << 8 | triangle, the tie scan's
// answer), and the word asuint(depth) << 32 | key.
template <bool Count, bool Masks>
__global__ void __launch_bounds__(64) GE_WALK_ATTR k_ray_visibility(Bvh bvh, ge_ray_visibility_params p, uint32_t tiles_x,
                                                                     uint32_t n, uint32_t* head,
                                                                     uint64_t* __restrict__ words,
                                                                     unsigned long long* counters, uint32_t alpha) {
  // Ray i is lane i % 64 of 8 x 8 tile i / 64 (tiles row-major), so a refill takes a tile's rays.
  auto direction = [&](uint32_t x, uint32_t y) {
    const float ndx = (float(x) + 0.5f) / float(p.width) * 2.0f - 1.0f;
    const float ndy = 1.0f - (float(y) + 0.5f) / float(p.height) * 2.0f;
    const float* m = p.clip_to_ray;
    f3 d;
    d.x = ((m[0] * ndx + m[4] * ndy) + m[8] * 0.0f) + m[12] * 1.0f;
    d.y = ((m[1] * ndx + m[5] * ndy) + m[9] * 0.0f) + m[13] * 1.0f;
    d.z = ((m[2] * ndx + m[6] * ndy) + m[10] * 0.0f) + m[14] * 1.0f;
    const float len = __builtin_sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
    const float inv = 1.0f / len;
    return f3{d.x * inv, d.y * inv, d.z * inv};
  };
  Tally tally;
  uint64_t issued = 0;
  auto load = [&](uint32_t i, uint32_t& slot, Ray& ray, float& tmax, uint32_t& mask) {
    const uint32_t tile = i >> 6, l = i & 63u;
    const uint32_t x = (tile % tiles_x) * 8 + (l & 7u), y = (tile / tiles_x) * 8 + (l >> 3);
    tmax = 1e30f;
    mask = 0xfeu;
    if (x >= p.width || y >= p.height) {
      slot = GE_RT_MISS;  // past the picture's edge: no pixel
      return false;
    }
    slot = y * p.width + x;
    ray = make_ray(f3{p.camera[0], p.camera[1], p.camera[2]}, direction(x, y), 0.0f);
    return true;
  };
  auto store = [&](uint32_t pixel, const Hit& h) {
    if (pixel == GE_RT_MISS) return;
    uint64_t value = 0;
    if (h.found) {
      const f3 o = f3{p.camera[0], p.camera[1], p.camera[2]};
      const f3 d = direction(pixel % p.width, pixel / p.width);
      const float pt[3] = {o.x + d.x * h.t, o.y + d.y * h.t, o.z + d.z * h.t};
      const float* vp = p.view_proj;
      const float cz = ((vp[2] * pt[0] + vp[6] * pt[1]) + vp[10] * pt[2]) + vp[14] * 1.0f;
      const float cw = ((vp[3] * pt[0] + vp[7] * pt[1]) + vp[11] * pt[2]) + vp[15] * 1.0f;
      const float depth = cz / cw;
      value = (uint64_t(__float_as_uint(depth)) << 32) | uint64_t(h.key);
    }
    words[pixel] = value;
  };
  persistent_walk<false, Count, Masks, Masks ? k_refill_masked : k_refill_camera, Masks ? k_batch_masked : k_batch>(
      bvh, n, head, lane_stack(), load, store, issued, tally, alpha);
  if (Count) flush_counters(counters, issued, 0, tally);
}

// ---- the path tracer ---------------------------------------------------------------------------
struct Pt {
  Bvh bvh;
  SceneView sv;
  ge_path_trace_params p;
};

// `Full` compiles textured materials and the alpha-tested candidate rule (a scene with either);
// without it every material is its factors and every triangle opaque.
template <bool Count, bool Full = false>
__device__ bool occluded(const Pt& c, f3 origin, f3 direction, float t_max, uint64_t& any, Tally& tally) {
  if (t_max <= 0.0f) return false;
  ++any;
  const Ray ray = make_ray(origin, direction, 0.0f);
  Hit h;
  const uint32_t alpha =
      (c.sv.features & k_scene_masked) && (c.p.flags & GE_PT_ALPHA_MASKS) ? k_alpha_test : k_alpha_opaque;
  return traverse<true, Count, Full>(c.bvh, ray, t_max, 0x01u, h, tally, lane_stack(), alpha);
}

// The closest-hit ray's rule for masked candidates (path_trace.slang:278-292): by the mask in a frame
// with GE_PT_ALPHA_MASKS, else left out by the camera ray and solid to every later ray.
__device__ __forceinline__ uint32_t closest_alpha(const ge_path_trace_params& p, uint32_t bounce) {
  if (p.flags & GE_PT_ALPHA_MASKS) return k_alpha_test;
  return bounce == 0 ? k_alpha_cull : k_alpha_opaque;
}

__device__ f3 environment(const ge_path_trace_params& p, f3 direction, bool primary) {
  const f3 sky = ld3(p.sky);
  if (p.flags & GE_PT_UNIFORM_SKY) return sky;
  if (primary) {
    return f3{__builtin_powf(maxr(sky.x, 0.0f), 2.2f), __builtin_powf(maxr(sky.y, 0.0f), 2.2f),
              __builtin_powf(maxr(sky.z, 0.0f), 2.2f)};
  }
  const f3 sky_r = sky * 0.35f;
  const f3 ground = ld3(p.ground) * (splat(p.sun[3] * maxr(p.sun[1], 0.0f) / k_pi) + sky_r);
  return direction.y >= 0.0f ? sky_r : ground;
}

struct HitSurface {
  f3 position, normal, geometric;
  uint32_t cluster;
  Corners corners;
  f3 b;
};

// fetch_triangle + surface_at (material.slang:244-302) at the hit's barycentrics.
__device__ HitSurface surface_of_hit(const SceneView& sv, const ge_path_trace_params& p, const Hit& hit, f3 towards,
                                     const ge_instance_desc& in) {
  const uint32_t cluster_index = sv.meshes[in.mesh].first_cluster + (hit.prim >> 8);
  const Corners c = fetch_corners(sv, p.eye, in, cluster_index, hit.prim & 0xffu);
  const Vertex v = surface_at(sv, in, c, f3{1.0f - hit.u - hit.v, hit.u, hit.v}, towards);
  HitSurface out;
  out.position = v.position;
  out.normal = v.normal;
  out.geometric = v.geometric;
  out.cluster = cluster_index;
  out.corners = c;
  out.b = f3{1.0f - hit.u - hit.v, hit.u, hit.v};
  return out;
}

template <bool Count, bool Full = false>
__device__ f3 next_event(const Pt& c, const Surface& surface, f3 position, f3 front, float bias, uint64_t& any,
                         Tally& tally) {
  const ge_path_trace_params& p = c.p;
  f3 direct = splat(0.0f);
  const f3 shadow_origin = position + front * bias;
  const f3 sun = ld3(p.sun);
  if (p.sun[3] > 0.0f && dot(surface.normal, sun) > 0.0f && !occluded<Count, Full>(c, shadow_origin, sun, 1e30f, any, tally))
    direct = direct + brdf_direct(surface, sun, splat(p.sun[3]));
  for (uint32_t i = 0; i < c.sv.light_count; ++i) {
    const ge_resolve_light light = c.sv.lights[i];
    f3 light_dir, radiance;
    float light_distance;
    if (!light_sample(light, position, light_dir, radiance, light_distance)) continue;
    if (dot(surface.normal, light_dir) <= 0.0f) continue;
    if (occluded<Count, Full>(c, shadow_origin, light_dir, light_distance - bias - light.spot[2], any, tally)) continue;
    direct = direct + brdf_direct(surface, light_dir, radiance);
  }
  return direct;
}

// A path in flight: where its next ray starts and goes, what it has gathered, and its sampler.
struct PathVars {
  f3 origin, direction, radiance, throughput;
  uint32_t rng, bounce;
  bool hit_primary;
};

// trace_path's bounce loop (path_trace.slang:257-461) from the path's current bounce to its end.
template <bool Count, bool Full = false>
__device__ void continue_path(const Pt& c, PathVars& v, uint64_t& closest, uint64_t& any, Tally& tally) {
  const ge_path_trace_params& p = c.p;
  for (;; ++v.bounce) {
    const uint32_t bounce = v.bounce;
    ++closest;
    const Ray ray = make_ray(v.origin, v.direction, 0.0f);
    Hit hit;
    traverse<false, Count, Full>(c.bvh, ray, 1e30f, bounce == 0 ? 0xfeu : 0x01u, hit, tally, lane_stack(),
                                 Full && (c.sv.features & k_scene_masked) ? closest_alpha(p, bounce) : k_alpha_opaque);
    if (!hit.found) {
      v.radiance = v.radiance + v.throughput * environment(p, v.direction, bounce == 0);
      break;
    }
    if (bounce == 0) v.hit_primary = true;
    const ge_instance_desc in = c.sv.instances[hit.inst];
    HitSurface sf = surface_of_hit(c.sv, p, hit, v.origin, in);
    const ge_resolve_material* mat = c.sv.materials + (c.sv.cluster_materials[sf.cluster] + in.material_base);
    const float4 spc = *reinterpret_cast<const float4*>(mat->specular);
    f3 albedo, emissive;
    float roughness, metallic;
    if (Full) {
      const MaterialSample ms = shade_material(c.sv, *mat, sf.corners, sf.b, sf.normal);
      albedo = ms.albedo;
      emissive = ms.emissive;
      roughness = ms.roughness;
      metallic = ms.metallic;
    } else {
      const float4 alb = *reinterpret_cast<const float4*>(mat->albedo);
      const float4 emi = *reinterpret_cast<const float4*>(mat->emissive);
      albedo = f3{alb.x, alb.y, alb.z};
      emissive = f3{emi.x, emi.y, emi.z};
      roughness = alb.w;
      metallic = emi.w;
    }
    f3 shading_normal = sf.normal;
    if (dot(shading_normal, sf.geometric) < 0.0f) shading_normal = -shading_normal;
    v.radiance = v.radiance + v.throughput * emissive;
    Surface surface;
    surface.albedo = albedo;
    surface.normal = shading_normal;
    surface.view = -v.direction;
    surface.roughness = roughness;
    surface.metallic = metallic;
    surface.spec[0] = spc.x;
    surface.spec[1] = spc.y;
    surface.spec[2] = spc.z;
    surface.spec[3] = spc.w;
    const float bias = p.ray_bias + p.ray_bias_steps * c.sv.meshes[in.mesh].quant[3] * in.scale_max;
    const f3 front = dot(sf.geometric, surface.view) >= 0.0f ? sf.geometric : -sf.geometric;
    v.radiance = v.radiance + v.throughput * next_event<Count, Full>(c, surface, sf.position, front, bias, any, tally);
    if (bounce >= p.max_bounces) break;
    const float n_dot_v = dot(surface.normal, surface.view);
    if (n_dot_v <= 0.0f) break;
    f3 tangent, bitangent;
    onb(surface.normal, tangent, bitangent);
    const f3 v_local = f3{dot(surface.view, tangent), dot(surface.view, bitangent), dot(surface.view, surface.normal)};
    const float alpha = brdf_alpha(surface.roughness);
    const f3 f = fresnel(surface, n_dot_v);
    const float p_specular = specular_probability(surface.albedo, surface.metallic, f);
    f3 l_local;
    if (rng_next(v.rng) < p_specular) {
      const float u0 = rng_next(v.rng);
      const float u1 = rng_next(v.rng);
      const f3 h = ggx_vndf(v_local, alpha, u0, u1);
      l_local = reflect3(-v_local, h);
    } else {
      const float u0 = rng_next(v.rng);
      const float u1 = rng_next(v.rng);
      l_local = cosine_hemisphere(u0, u1);
    }
    if (l_local.z <= 0.0f) break;
    const float pdf = bsdf_pdf(v_local, l_local, alpha, p_specular);
    if (!(pdf > 0.0f)) break;
    const f3 next = tangent * l_local.x + bitangent * l_local.y + surface.normal * l_local.z;
    const f3 weight = brdf_eval(surface, next) * (l_local.z / pdf);
    v.throughput = v.throughput * weight;
    if (v.throughput.x <= 0.0f && v.throughput.y <= 0.0f && v.throughput.z <= 0.0f) break;
    if (bounce + 1 >= GE_PT_ROULETTE_DEPTH) {
      const float survive = clampr(maxr(v.throughput.x, maxr(v.throughput.y, v.throughput.z)), 0.0f, 0.95f);
      if (rng_next(v.rng) >= survive) break;
      v.throughput = v.throughput / maxr(survive, 1e-6f);
    }
    v.origin = sf.position + (dot(next, front) >= 0.0f ? front : -front) * bias;
    v.direction = next;
  }
}

template <bool Count, bool Full = false>
__device__ f3 trace_path(const Pt& c, f3 origin, f3 direction, uint32_t& rng, bool& hit_primary, uint64_t& closest,
                         uint64_t& any, Tally& tally) {
  PathVars v;
  v.origin = origin;
  v.direction = direction;
  v.radiance = splat(0.0f);
  v.throughput = splat(1.0f);
  v.rng = rng;
  v.bounce = 0;
  v.hit_primary = false;
  continue_path<Count, Full>(c, v, closest, any, tally);
  rng = v.rng;
  hit_primary = v.hit_primary;
  return v.radiance;
}

// ---- the path-replay adjoint ---------------------------------------------------------------------
// The path tracer's estimator with sampling detached (directions, lobe choices, pdfs and roulette
// from the sampling materials; the BSDF values from the scene's), replayed from its counter RNG and
// differentiated: dL = sum_k T_k dC_k + sum_k T_k dw_k U_{k+1}, U_k = C_k + w_k U_{k+1}, the first
// term at each vertex's next-event sum, the second backwards over the path's recorded vertices.
// Per-path arithmetic in float; the per-parameter sums in double.
constexpr uint32_t k_adjoint_vertices = 17;  // max_bounces <= 15: the vertices and the escape

struct AdjointArgs {
  Pt c;
  const ge_resolve_material* sampling;  // null: the scene's materials
  const float4* d_image;                // dLoss/dI per pixel; null: the estimate only
  float4* image;                        // per pixel: the samples' sum, the sample count
  double* grad_materials;               // per material: albedo rgb, roughness
  double* grad_lights;                  // the sun, then each light
};

struct AdjointVertex {
  f3 C, T, w, dw_a, dw_r;
  f3 na;     // the next-event term's albedo gradient, delta T dC/dalbedo
  float nr;  // and its roughness gradient
  uint32_t mat;
  bool cont;
};

// light_acc: the workgroup's LDS sums of the sun's and each light's gradient (every lane adds to
// the same few addresses; one global atomic a workgroup at the end instead of one a vertex).
__device__ f3 replay_path(const AdjointArgs& a, f3 origin, f3 direction, uint32_t& rng, bool grad, f3 delta,
                         double* light_acc) {
  const Pt& c = a.c;
  const ge_path_trace_params& p = c.p;
  AdjointVertex vx[k_adjoint_vertices];
  uint32_t nv = 0;
  Tally tally;
  uint64_t any = 0;
  f3 radiance = splat(0.0f);
  f3 throughput = splat(1.0f);
  f3 throughput_s = splat(1.0f);
  for (uint32_t bounce = 0;; ++bounce) {
    const Ray ray = make_ray(origin, direction, 0.0f);
    Hit hit;
    traverse<false, false>(c.bvh, ray, 1e30f, bounce == 0 ? 0xfeu : 0x01u, hit, tally, lane_stack());
    AdjointVertex& v = vx[nv++];
    v.T = throughput;
    v.cont = false;
    v.mat = GE_RT_MISS;
    v.na = splat(0.0f);
    v.nr = 0.0f;
    if (!hit.found) {
      v.C = environment(p, direction, bounce == 0);
      radiance = radiance + throughput * v.C;
      if (grad && bounce != 0 && !(p.flags & GE_PT_UNIFORM_SKY) && !(direction.y >= 0.0f))
        atomicAdd(&light_acc[0], double(dot(delta, throughput * ld3(p.ground) * (maxr(p.sun[1], 0.0f) / k_pi))));
      break;
    }
    const ge_instance_desc in = c.sv.instances[hit.inst];
    const HitSurface sf = surface_of_hit(c.sv, p, hit, origin, in);
    const uint32_t mat_index = c.sv.cluster_materials[sf.cluster] + in.material_base;
    const ge_resolve_material* mat = c.sv.materials + mat_index;
    const ge_resolve_material* smat = a.sampling ? a.sampling + mat_index : mat;
    f3 shading_normal = sf.normal;
    if (dot(shading_normal, sf.geometric) < 0.0f) shading_normal = -shading_normal;
    const float4 alb = *reinterpret_cast<const float4*>(mat->albedo);
    const float4 emi = *reinterpret_cast<const float4*>(mat->emissive);
    const float4 spc = *reinterpret_cast<const float4*>(mat->specular);
    const f3 emission = f3{emi.x, emi.y, emi.z};
    radiance = radiance + throughput * emission;
    Surface surface;
    surface.albedo = f3{alb.x, alb.y, alb.z};
    surface.normal = shading_normal;
    surface.view = -direction;
    surface.roughness = alb.w;
    surface.metallic = emi.w;
    surface.spec[0] = spc.x;
    surface.spec[1] = spc.y;
    surface.spec[2] = spc.z;
    surface.spec[3] = spc.w;
    Surface ssurface = surface;
    {
      const float4 salb = *reinterpret_cast<const float4*>(smat->albedo);
      ssurface.albedo = f3{salb.x, salb.y, salb.z};
      ssurface.roughness = salb.w;
      ssurface.metallic = smat->emissive[3];
    }
    const float bias = p.ray_bias + p.ray_bias_steps * c.sv.meshes[in.mesh].quant[3] * in.scale_max;
    const f3 front = dot(sf.geometric, surface.view) >= 0.0f ? sf.geometric : -sf.geometric;
    f3 direct = splat(0.0f);
    f3 dd_a = splat(0.0f), dd_r = splat(0.0f);
    const f3 shadow_origin = sf.position + front * bias;
    const f3 sun = ld3(p.sun);
    if (p.sun[3] > 0.0f && dot(surface.normal, sun) > 0.0f && !occluded<false>(c, shadow_origin, sun, 1e30f, any, tally)) {
      direct = direct + brdf_direct(surface, sun, splat(p.sun[3]));
      if (grad) {
        f3 ga, gr;
        brdf_eval_grad(surface, sun, ga, gr);
        const float cosl = maxr(dot(surface.normal, sun), 0.0f);
        dd_a = dd_a + ga * (cosl * p.sun[3]);
        dd_r = dd_r + gr * (cosl * p.sun[3]);
        atomicAdd(&light_acc[0], double(dot(delta, throughput * (brdf_eval(surface, sun) * cosl))));
      }
    }
    for (uint32_t i = 0; i < c.sv.light_count; ++i) {
      const ge_resolve_light light = c.sv.lights[i];
      f3 light_dir, light_radiance;
      float light_distance, att;
      if (!light_sample(light, sf.position, light_dir, light_radiance, light_distance, &att)) continue;
      if (dot(surface.normal, light_dir) <= 0.0f) continue;
      if (occluded<false>(c, shadow_origin, light_dir, light_distance - bias - light.spot[2], any, tally)) continue;
      direct = direct + brdf_direct(surface, light_dir, light_radiance);
      if (grad) {
        f3 ga, gr;
        brdf_eval_grad(surface, light_dir, ga, gr);
        const float cosl = maxr(dot(surface.normal, light_dir), 0.0f);
        dd_a = dd_a + ga * cosl * light_radiance;
        dd_r = dd_r + gr * cosl * light_radiance;
        atomicAdd(&light_acc[1 + i],
                  double(dot(delta, throughput * (brdf_eval(surface, light_dir) * cosl * ld3(light.color_intensity) * att))));
      }
    }
    radiance = radiance + throughput * direct;
    v.C = emission + direct;
    v.mat = mat_index;
    if (grad) {  // added with the vertex's throughput term in the backward sweep
      v.na = delta * throughput * dd_a;
      v.nr = dot(delta, throughput * dd_r);
    }
    if (bounce >= p.max_bounces) break;
    const float n_dot_v = dot(surface.normal, surface.view);
    if (n_dot_v <= 0.0f) break;
    f3 tangent, bitangent;
    onb(surface.normal, tangent, bitangent);
    const f3 v_local = f3{dot(surface.view, tangent), dot(surface.view, bitangent), dot(surface.view, surface.normal)};
    const float alpha = brdf_alpha(ssurface.roughness);
    const f3 f = fresnel(ssurface, n_dot_v);
    const float p_specular = specular_probability(ssurface.albedo, ssurface.metallic, f);
    f3 l_local;
    if (rng_next(rng) < p_specular) {
      const float u0 = rng_next(rng);
      const float u1 = rng_next(rng);
      const f3 hh = ggx_vndf(v_local, alpha, u0, u1);
      l_local = reflect3(-v_local, hh);
    } else {
      const float u0 = rng_next(rng);
      const float u1 = rng_next(rng);
      l_local = cosine_hemisphere(u0, u1);
    }
    if (l_local.z <= 0.0f) break;
    const float pdf = bsdf_pdf(v_local, l_local, alpha, p_specular);
    if (!(pdf > 0.0f)) break;
    const f3 next = tangent * l_local.x + bitangent * l_local.y + surface.normal * l_local.z;
    const float scale = l_local.z / pdf;
    const f3 weight = brdf_eval(surface, next) * scale;
    const f3 weight_s = a.sampling ? brdf_eval(ssurface, next) * scale : weight;
    throughput = throughput * weight;
    throughput_s = throughput_s * weight_s;
    if (throughput_s.x <= 0.0f && throughput_s.y <= 0.0f && throughput_s.z <= 0.0f) break;
    float survive = 1.0f;
    if (bounce + 1 >= GE_PT_ROULETTE_DEPTH) {
      survive = clampr(maxr(throughput_s.x, maxr(throughput_s.y, throughput_s.z)), 0.0f, 0.95f);
      if (rng_next(rng) >= survive) break;
      throughput = throughput / maxr(survive, 1e-6f);
      throughput_s = throughput_s / maxr(survive, 1e-6f);
      survive = maxr(survive, 1e-6f);
    }
    if (grad) {
      f3 ga, gr;
      brdf_eval_grad(surface, next, ga, gr);
      v.w = weight / survive;
      v.dw_a = ga * (scale / survive);
      v.dw_r = gr * (scale / survive);
      v.cont = true;
    }
    if (nv >= k_adjoint_vertices - 1) break;
    origin = sf.position + (dot(next, front) >= 0.0f ? front : -front) * bias;
    direction = next;
  }
  if (grad) {
    f3 u = splat(0.0f);
    for (uint32_t k = nv; k-- > 0;) {
      const AdjointVertex& v = vx[k];
      if (v.mat != GE_RT_MISS) {
        // One update a vertex: its next-event term, plus (when the path went on) dw_k U_{k+1}.
        double* gm = a.grad_materials + size_t(v.mat) * 4;
        atomicAdd(&gm[0], double(v.na.x) + (v.cont ? double((delta * v.T * v.dw_a * u).x) : 0.0));
        atomicAdd(&gm[1], double(v.na.y) + (v.cont ? double((delta * v.T * v.dw_a * u).y) : 0.0));
        atomicAdd(&gm[2], double(v.na.z) + (v.cont ? double((delta * v.T * v.dw_a * u).z) : 0.0));
        atomicAdd(&gm[3], double(v.nr) + (v.cont ? double(dot(delta, v.T * v.dw_r * u)) : 0.0));
      }
      u = v.cont ? v.C + v.w * u : v.C;
    }
  }
  return radiance;
}

__device__ void adjoint_pixel(const AdjointArgs& a, uint32_t x, uint32_t y, bool grad, double* light_acc);

__global__ void __launch_bounds__(64) k_path_adjoint(AdjointArgs a) {
  const uint32_t x = blockIdx.x * 8 + (threadIdx.x & 7u);
  const uint32_t y = blockIdx.y * 8 + (threadIdx.x >> 3);
  const ge_path_trace_params& p = a.c.p;
  const bool grad = a.d_image != nullptr;
  double* light_acc = reinterpret_cast<double*>(g_lds + k_ring * k_lanes);
  const uint32_t nl = a.c.sv.light_count + 1;
  if (grad) {
    for (uint32_t i = threadIdx.x; i < nl; i += 64) light_acc[i] = 0.0;
    __syncthreads();
  }
  if (x < p.width && y < p.height) adjoint_pixel(a, x, y, grad, light_acc);
  if (grad) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nl; i += 64) {
      const double v = light_acc[i];
      if (v != 0.0) atomicAdd(&a.grad_lights[i], v);
    }
  }
}

__device__ void adjoint_pixel(const AdjointArgs& a, uint32_t x, uint32_t y, bool grad, double* light_acc) {
  const ge_path_trace_params& p = a.c.p;
  const size_t pixel = size_t(y) * p.width + x;
  f3 delta = splat(0.0f);
  if (grad) {
    const float4 d = a.d_image[pixel];
    delta = f3{d.x, d.y, d.z} / float(p.spp);
  }
  const float sx = float(p.width), sy = float(p.height);
  const bool centered = (p.flags & GE_PT_PIXEL_CENTER) != 0;
  f3 sum = splat(0.0f);
  for (uint32_t s = 0; s < p.spp; ++s) {
    uint32_t rng = rng_seed(x, y, p.sample_base + s, p.seed);
    float jx = 0.5f, jy = 0.5f;
    if (!centered) {
      jx = rng_next(rng);
      jy = rng_next(rng);
    }
    const float ndx = (float(x) + jx) / sx * 2.0f - 1.0f;
    const float ndy = 1.0f - (float(y) + jy) / sy * 2.0f;
    const float* m = p.clip_to_ray;
    f3 dir;
    dir.x = ((m[0] * ndx + m[4] * ndy) + m[8] * 0.0f) + m[12] * 1.0f;
    dir.y = ((m[1] * ndx + m[5] * ndy) + m[9] * 0.0f) + m[13] * 1.0f;
    dir.z = ((m[2] * ndx + m[6] * ndy) + m[10] * 0.0f) + m[14] * 1.0f;
    dir = normalize(dir);
    sum = sum + replay_path(a, ld3(p.camera), dir, rng, grad, delta, light_acc);
  }
  if (a.image) a.image[pixel] = make_float4(sum.x, sum.y, sum.z, float(p.spp));
}

template <bool Count, bool Full>
__global__ void __launch_bounds__(64) k_path_trace(Pt c, float4* __restrict__ accum, uint32_t* __restrict__ coverage,
                                                   unsigned long long* counters) {
  const uint32_t x = blockIdx.x * 8 + (threadIdx.x & 7u);
  const uint32_t y = blockIdx.y * 8 + (threadIdx.x >> 3);
  const ge_path_trace_params& p = c.p;
  Tally tally;
  uint64_t closest = 0, any = 0;
  if (x < p.width && y < p.height) {
    const size_t pixel = size_t(y) * p.width + x;
    const float sx = float(p.width), sy = float(p.height);
    const bool centered = (p.flags & GE_PT_PIXEL_CENTER) != 0;
    f3 sum = splat(0.0f);
    bool covered = false;
    for (uint32_t s = 0; s < p.spp; ++s) {
      uint32_t rng = rng_seed(x, y, p.sample_base + s, p.seed);
      float jx = 0.5f, jy = 0.5f;
      if (!centered) {
        jx = rng_next(rng);
        jy = rng_next(rng);
      }
      const float ndx = (float(x) + jx) / sx * 2.0f - 1.0f;
      const float ndy = 1.0f - (float(y) + jy) / sy * 2.0f;
      const float* m = p.clip_to_ray;
      f3 dir;
      dir.x = ((m[0] * ndx + m[4] * ndy) + m[8] * 0.0f) + m[12] * 1.0f;
      dir.y = ((m[1] * ndx + m[5] * ndy) + m[9] * 0.0f) + m[13] * 1.0f;
      dir.z = ((m[2] * ndx + m[6] * ndy) + m[10] * 0.0f) + m[14] * 1.0f;
      dir = normalize(dir);
      bool hit_primary = false;
      sum = sum + trace_path<Count, Full>(c, ld3(p.camera), dir, rng, hit_primary, closest, any, tally);
      covered = covered || hit_primary;
    }
    float4 prev = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (p.sample_base != 0) prev = accum[pixel];
    accum[pixel] = make_float4(prev.x + sum.x, prev.y + sum.y, prev.z + sum.z, prev.w + float(p.spp));
    const uint32_t before = p.sample_base == 0 ? 0u : coverage[pixel];
    coverage[pixel] = before | (covered ? 1u : 0u);
  }
  if (Count) flush_counters(counters, closest, any, tally);
}

// tonemap_main (path_trace.slang:511-543) with display_output (display.slang).
__global__ void __launch_bounds__(64) k_tonemap(ge_path_trace_params p, const float4* __restrict__ accum,
                                                const uint32_t* __restrict__ coverage, uint32_t* __restrict__ out) {
  const uint32_t x = blockIdx.x * 8 + (threadIdx.x & 7u);
  const uint32_t y = blockIdx.y * 8 + (threadIdx.x >> 3);
  if (x >= p.width || y >= p.height) return;
  const size_t pixel = size_t(y) * p.width + x;
  const float4 v = accum[pixel];
  float mean[3] = {0.0f, 0.0f, 0.0f};
  if (v.w > 0.0f) {
    mean[0] = v.x / v.w;
    mean[1] = v.y / v.w;
    mean[2] = v.z / v.w;
  }
  if (coverage[pixel] == 0 && (p.flags & GE_PT_UNIFORM_SKY) == 0) {
    out[pixel] = p.background;
    return;
  }
  float e[3];
  for (int k = 0; k < 3; ++k) e[k] = __builtin_powf(__builtin_fmaxf(mean[k], 0.0f), 1.0f / 2.2f);
  if (p.flags & GE_PT_DITHER) {
    const uint32_t inner = x * 288237660u + y * 25070368u;
    const uint32_t outer = uint32_t((uint64_t(inner) * uint64_t(3472289u)) >> 16);
    const float u = float(outer >> 8) * (1.0f / 16777216.0f);
    const float dither = u < 0.5f ? __builtin_sqrtf(2.0f * u) - 1.0f : 1.0f - __builtin_sqrtf(2.0f * (1.0f - u));
    const float s = 255.0f;
    for (int k = 0; k < 3; ++k) {
      const float vv = __builtin_fminf(__builtin_fmaxf(e[k], 0.0f), 1.0f);
      const float room = __builtin_fminf(__builtin_fminf(vv * s, (1.0f - vv) * s), 1.0f);
      e[k] = vv + dither * room / s;
    }
  }
  uint32_t rgb[3];
  for (int k = 0; k < 3; ++k) rgb[k] = uint32_t(__builtin_fminf(__builtin_fmaxf(e[k], 0.0f), 1.0f) * 255.0f + 0.5f);
  out[pixel] = rgb[0] | (rgb[1] << 8) | (rgb[2] << 16) | (255u << 24);
}

// The resolve's sun query (shadow_visibility + shadow_ray_blocked, visibility_resolve.slang:516-605,
// 742-755): origin = position + geometric * ray_offset, one terminate-on-first-hit ray with the caster
// mask, a ballot of the answers per 32 receivers.
template <bool Count, bool Masks>
__global__ void __launch_bounds__(64) k_shadow(Bvh bvh, SceneView sv, ge_shadow_params p,
                                               const ge_shadow_receiver* __restrict__ rec, uint32_t* __restrict__ bits,
                                               unsigned long long* counters, uint32_t alpha) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  Tally tally;
  uint64_t issued = 0;
  bool blocked = false;
  if (i < p.count) {
    const ge_shadow_receiver r = rec[i];
    if (r.instance_id != GE_RT_MISS && p.t_max > 0.0f) {
      const ge_instance_desc in = sv.instances[r.instance_id];
      const float offset = p.offset + p.offset_steps * sv.meshes[in.mesh].quant[3] * in.scale_max;
      const f3 o = f3{r.position[0] + r.geometric[0] * offset, r.position[1] + r.geometric[1] * offset,
                      r.position[2] + r.geometric[2] * offset};
      issued = 1;
      const Ray ray = make_ray(o, f3{p.sun[0], p.sun[1], p.sun[2]}, 0.0f);
      Hit h;
      blocked = traverse<true, Count, Masks>(bvh, ray, p.t_max, 0x01u, h, tally, lane_stack(), alpha);
    }
  }
  const uint64_t ballot = __ballot(blocked);
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t word0 = (blockIdx.x * blockDim.x + (threadIdx.x & ~63u)) >> 5;
  if (lane == 0 && word0 * 32 < p.count) bits[word0] = uint32_t(ballot);
  if (lane == 32 && (word0 + 1) * 32 < p.count) bits[word0 + 1] = uint32_t(ballot >> 32);
  if (Count) flush_counters(counters, 0, issued, tally);
}

inline uint32_t blocks(uint64_t n, uint32_t t) { return uint32_t((n + t - 1) / t); }
inline size_t lds(const Bvh&) { return size_t(k_ring) * k_lanes * 4; }

}  // namespace

hipError_t launch_trace_closest(hipStream_t stream, const Bvh& bvh, const ge_ray* rays, uint32_t n, uint32_t mask,
                                ge_hit* hits, unsigned long long* counters, uint32_t* head, uint32_t max_blocks,
                                uint32_t alpha, GeomMap geom) {
  if (n == 0) return hipSuccess;
  hipError_t e = hipMemsetAsync(head, 0, 4, stream);
  if (e != hipSuccess) return e;
  const uint32_t grid = std::min(blocks(n, 64), max_blocks);
  if (alpha != k_alpha_opaque) {
    if (counters) k_trace_closest<true, true><<<grid, 64, lds(bvh), stream>>>(bvh, rays, n, mask, head, hits, counters, alpha, geom);
    else k_trace_closest<false, true><<<grid, 64, lds(bvh), stream>>>(bvh, rays, n, mask, head, hits, counters, alpha, geom);
  } else {
    if (counters) k_trace_closest<true, false><<<grid, 64, lds(bvh), stream>>>(bvh, rays, n, mask, head, hits, counters, 0u, geom);
    else k_trace_closest<false, false><<<grid, 64, lds(bvh), stream>>>(bvh, rays, n, mask, head, hits, counters, 0u, geom);
  }
  return hipGetLastError();
}

hipError_t launch_trace_any(hipStream_t stream, const Bvh& bvh, const ge_ray* rays, uint32_t n, uint32_t mask,
                            uint32_t* bits, unsigned long long* counters, uint32_t* head, uint32_t max_blocks,
                            uint32_t alpha) {
  if (n == 0) return hipSuccess;
  hipError_t e = hipMemsetAsync(head, 0, 4, stream);
  if (e == hipSuccess) e = hipMemsetAsync(bits, 0, size_t((n + 31) / 32) * 4, stream);
  if (e != hipSuccess) return e;
  const uint32_t grid = std::min(blocks(n, 64), max_blocks);
  if (alpha != k_alpha_opaque) {
    if (counters) k_trace_any<true, true><<<grid, 64, lds(bvh), stream>>>(bvh, rays, n, mask, head, bits, counters, alpha);
    else k_trace_any<false, true><<<grid, 64, lds(bvh), stream>>>(bvh, rays, n, mask, head, bits, counters, alpha);
  } else {
    if (counters) k_trace_any<true, false><<<grid, 64, lds(bvh), stream>>>(bvh, rays, n, mask, head, bits, counters, 0u);
    else k_trace_any<false, false><<<grid, 64, lds(bvh), stream>>>(bvh, rays, n, mask, head, bits, counters, 0u);
  }
  return hipGetLastError();
}

hipError_t launch_ray_visibility(hipStream_t stream, const Bvh& bvh, const ge_ray_visibility_params& p, uint64_t* words,
                                 unsigned long long* counters, uint32_t* head, uint32_t max_blocks, uint32_t alpha) {
  const uint32_t tiles_x = blocks(p.width, 8), tiles_y = blocks(p.height, 8);
  const uint64_t n = uint64_t(tiles_x) * tiles_y * 64;
  if (n == 0) return hipSuccess;
  if (n >= (1ull << 32)) return hipErrorInvalidValue;
  hipError_t e = hipMemsetAsync(head, 0, 4, stream);
  if (e != hipSuccess) return e;
  const uint32_t grid = std::min(blocks(n, 64), max_blocks);
  const uint32_t m = uint32_t(n);
  if (alpha != k_alpha_opaque) {
    if (counters) k_ray_visibility<true, true><<<grid, 64, lds(bvh), stream>>>(bvh, p, tiles_x, m, head, words, counters, alpha);
    else k_ray_visibility<false, true><<<grid, 64, lds(bvh), stream>>>(bvh, p, tiles_x, m, head, words, counters, alpha);
  } else {
    if (counters) k_ray_visibility<true, false><<<grid, 64, lds(bvh), stream>>>(bvh, p, tiles_x, m, head, words, counters, 0u);
    else k_ray_visibility<false, false><<<grid, 64, lds(bvh), stream>>>(bvh, p, tiles_x, m, head, words, counters, 0u);
  }
  return hipGetLastError();
}

hipError_t launch_path_adjoint(hipStream_t stream, const Bvh& bvh, const SceneView& sv, const ge_path_trace_params& p,
                               const ge_resolve_material* sampling, const float4* d_image, float4* image,
                               double* grad_materials, double* grad_lights) {
  AdjointArgs a;
  a.c.bvh = bvh;
  a.c.sv = sv;
  a.c.p = p;
  a.sampling = sampling;
  a.d_image = d_image;
  a.image = image;
  a.grad_materials = grad_materials;
  a.grad_lights = grad_lights;
  const dim3 grid(blocks(p.width, 8), blocks(p.height, 8));
  k_path_adjoint<<<grid, 64, lds(bvh) + (size_t(sv.light_count) + 1) * 8, stream>>>(a);
  return hipGetLastError();
}

hipError_t launch_path_trace(hipStream_t stream, const Bvh& bvh, const SceneView& sv, const ge_path_trace_params& p,
                             float4* accum, uint32_t* coverage, unsigned long long* counters, bool full) {
  const dim3 grid(blocks(p.width, 8), blocks(p.height, 8));
  Pt c{bvh, sv, p};
  if (full) {
    if (counters) k_path_trace<true, true><<<grid, 64, lds(bvh), stream>>>(c, accum, coverage, counters);
    else k_path_trace<false, true><<<grid, 64, lds(bvh), stream>>>(c, accum, coverage, counters);
  } else {
    if (counters) k_path_trace<true, false><<<grid, 64, lds(bvh), stream>>>(c, accum, coverage, counters);
    else k_path_trace<false, false><<<grid, 64, lds(bvh), stream>>>(c, accum, coverage, counters);
  }
  return hipGetLastError();
}

hipError_t launch_tonemap(hipStream_t stream, const ge_path_trace_params& p, const float4* accum,
                          const uint32_t* coverage, uint32_t* out) {
  const dim3 grid(blocks(p.width, 8), blocks(p.height, 8));
  k_tonemap<<<grid, 64, 0, stream>>>(p, accum, coverage, out);
  return hipGetLastError();
}

hipError_t launch_shadow_pass(hipStream_t stream, const Bvh& bvh, const SceneView& sv, const ge_shadow_params& p,
                              const ge_shadow_receiver* rec, uint32_t* bits, unsigned long long* counters,
                              uint32_t alpha) {
  if (p.count == 0) return hipSuccess;
  const uint32_t g = blocks(p.count, 64);
  if (alpha != k_alpha_opaque) {
    if (counters) k_shadow<true, true><<<g, 64, lds(bvh), stream>>>(bvh, sv, p, rec, bits, counters, alpha);
    else k_shadow<false, true><<<g, 64, lds(bvh), stream>>>(bvh, sv, p, rec, bits, counters, alpha);
  } else {
    if (counters) k_shadow<true, false><<<g, 64, lds(bvh), stream>>>(bvh, sv, p, rec, bits, counters, 0u);
    else k_shadow<false, false><<<g, 64, lds(bvh), stream>>>(bvh, sv, p, rec, bits, counters, 0u);
  }
  return hipGetLastError();
}

}  // namespace ge
