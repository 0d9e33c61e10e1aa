// rt_wavefront.hip — the reference path tracer as a wavefront: one kernel per stage, paths in
// queues compacted with workgroup prefix sums, so each traversal kernel runs one kind of ray with a small
// register footprint instead of the whole integrator in every lane.
//
// Per bounce, over the queue of live paths:
//   closest  — the closest hit of each path's ray (camera mask on bounce 0, caster mask after);
//   shade    — a miss adds the environment and ends the path; a hit decodes the surface, adds its
//              emission, writes one shadow-ray record per light next-event estimation asks about
//              (sun first, then the lights in order, as next_event does, path_trace.slang:215-252),
//              samples the BSDF, plays roulette and appends the path to the next queue;
//   any      — one terminate-on-first-hit query per shadow record;
//   resolve  — each path sums its unblocked lights in record order and adds throughput x that sum.
// The sums are formed in the shader's order, so every path's radiance is the same float as the
// one-path-per-lane integrator's (and the oracle's).
//
// Path state is double-buffered (read from one, write the other), which makes a shade launch
// idempotent: if its shadow records overflow their buffer, the host grows it and runs the launch
// again.
#include "rt_internal.h"
#include "rt_mlp.h"
#include "rt_shade.h"

namespace ge {
namespace {

extern __shared__ uint32_t g_lds[];
__device__ __forceinline__ uint32_t* lane_stack() { return g_lds + threadIdx.x; }

#ifndef GE_SHADE_THREADS
#define GE_SHADE_THREADS 256
#endif
// Shade workgroups: one queue-append atomic per workgroup (same-address atomics serialize).
constexpr uint32_t k_shade_threads = GE_SHADE_THREADS;
static_assert(k_shade_threads % 64 == 0 && k_shade_threads <= 1024, "whole waves, at most 1024 threads");

// Appends `count` entries per lane, one atomic per workgroup of k_shade_threads; returns each lane's
// first slot (lanes in order within a wave, waves in order within the workgroup). `part` is a
// workgroup-shared array of k_shade_threads / 64 + 1 words, one per call site, so two appends in one
// kernel never share it. Every thread of the workgroup must call it.
__device__ __forceinline__ uint32_t block_append(uint32_t* counter, uint32_t count, uint32_t* part) {
  constexpr uint32_t waves = k_shade_threads / 64;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t incl = count;  // inclusive prefix over the wave by shuffles
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t v = uint32_t(__shfl_up(int(incl), off));
    if (lane >= uint32_t(off)) incl += v;
  }
  if (lane == 63) part[wave] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t total = 0;
    for (uint32_t k = 0; k < waves; ++k) {
      const uint32_t c = part[k];
      part[k] = total;
      total += c;
    }
    part[waves] = total ? atomicAdd(counter, total) : 0u;
  }
  __syncthreads();
  return part[waves] + part[wave] + incl - count;
}

// Camera rays in 8 x 8 pixel tiles, samples outermost: path id = sample * pixels + pixel. The first
// queue lists the paths tile by tile (tiles row-major, pixels row-major within a tile), so a wave
// of the first closest-hit walk takes one tile; each path's queue slot is that order's closed form.
__global__ void __launch_bounds__(64) k_generate(ge_path_trace_params p, uint32_t tiles_x, uint32_t tiles,
                                                 PathState* __restrict__ state, uint32_t* __restrict__ queue) {
  const uint32_t tile = blockIdx.x % tiles, sample = blockIdx.x / tiles;
  const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
  const uint32_t lx = threadIdx.x & 7u, ly = threadIdx.x >> 3;
  const uint32_t x = tx * 8 + lx, y = ty * 8 + ly;
  if (x >= p.width || y >= p.height) return;
  const uint32_t pixels = p.width * p.height;
  const uint32_t rows = std::min(8u, p.height - ty * 8), cols = std::min(8u, p.width - tx * 8);
  const uint32_t slot = sample * pixels + ty * 8 * p.width + tx * 8 * rows + ly * cols + lx;
  const uint32_t pid = sample * pixels + y * p.width + x;
  uint32_t rng = rng_seed(x, y, p.sample_base + sample, p.seed);
  float jx = 0.5f, jy = 0.5f;
  if (!(p.flags & GE_PT_PIXEL_CENTER)) {
    jx = rng_next(rng);
    jy = rng_next(rng);
  }
  const float ndx = (float(x) + jx) / float(p.width) * 2.0f - 1.0f;
  const float ndy = 1.0f - (float(y) + jy) / float(p.height) * 2.0f;
  const float* m = p.clip_to_ray;
  f3 dir;
  dir.x = ((m[0] * ndx + m[4] * ndy) + m[8] * 0.0f) + m[12] * 1.0f;
  dir.y = ((m[1] * ndx + m[5] * ndy) + m[9] * 0.0f) + m[13] * 1.0f;
  dir.z = ((m[2] * ndx + m[6] * ndy) + m[10] * 0.0f) + m[14] * 1.0f;
  dir = normalize(dir);
  PathState st;
  st.o[0] = p.camera[0];
  st.o[1] = p.camera[1];
  st.o[2] = p.camera[2];
  st.d[0] = dir.x;
  st.d[1] = dir.y;
  st.d[2] = dir.z;
  st.rng = rng;
  st.T[0] = st.T[1] = st.T[2] = 1.0f;
  st.L[0] = st.L[1] = st.L[2] = 0.0f;
  st.bounce = 0;
  st.flags = 0;
  st.pad = 0;
  state[pid] = st;
  queue[slot] = pid;
}

// One bounce's closest-hit rays: path `queue[i]`'s ray from its state, its hit into hits[pid].
struct ClosestJob {
  const PathState* state;
  const uint32_t* queue;
  uint32_t n;
  uint32_t* head;
  PathHit* hits;
  uint32_t alpha;  // the masked-candidate rule of the bounce
  const uint32_t* count = nullptr;  // set: the rays are the first min(n, *count), read on the device
};

template <bool Count, bool Masks>
__device__ void closest_walks(const Bvh& bvh, const ClosestJob& job, uint64_t& issued, Tally& tally) {
  const PathState* __restrict__ state = job.state;
  const uint32_t* __restrict__ queue = job.queue;
  PathHit* __restrict__ hits = job.hits;
  auto load = [&](uint32_t i, uint32_t& slot, Ray& ray, float& tmax, uint32_t& mask) {
    const uint32_t pid = queue[i];
    const PathState& st = state[pid];
    slot = pid;
    ray = make_ray(f3{st.o[0], st.o[1], st.o[2]}, f3{st.d[0], st.d[1], st.d[2]}, 0.0f);
    tmax = 1e30f;
    mask = st.bounce == 0 ? 0xfeu : 0x01u;
    return true;
  };
  auto store = [&](uint32_t pid, const Hit& h) {
    PathHit out;
    out.t = h.t;
    out.u = h.u;
    out.v = h.v;
    out.inst = h.found ? h.inst : GE_RT_MISS;
    out.prim = h.prim;
    hits[pid] = out;
  };
  persistent_walk<false, Count, Masks, k_refill_paths>(bvh, job.n, job.head, lane_stack(), load, store, issued, tally,
                                                       job.alpha);
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

// fetch_triangle + surface_at (material.slang:244-302) at the hit's barycentrics.
__device__ Vertex vertex_of(const SceneView& sv, const ge_path_trace_params& p, const PathHit& hit, f3 towards,
                            const ge_instance_desc& in, uint32_t& cluster_index, Corners* corners = nullptr) {
  cluster_index = sv.meshes[in.mesh].first_cluster + (hit.prim >> 8);
  const Corners c = fetch_corners(sv, p.eye, in, cluster_index, hit.prim & 0xffu);
  if (corners) *corners = c;
  return surface_at(sv, in, c, f3{1.0f - hit.u - hit.v, hit.u, hit.v}, towards);
}

struct ShadeArgs {
  SceneView sv;
  ge_path_trace_params p;
  const PathState* in;
  PathState* out;
  const uint32_t* queue;
  uint32_t n;
  uint32_t shadow_capacity;
  const PathHit* hits;
  NeeRec* nee;
  ShadowRec* shadow;
  uint32_t* next_queue;
  uint32_t* counts;  // [0] next queue, [1] shadow records; [2] [3] the walk heads, cleared for the next walk
  // The radiance cache (kShadeCached, kShadeRecord). Queries (with their throughput in
  // query_weights) are appended at counts[6] while they fit in query_capacity; training samples
  // (with theirs in sample_weights) at counts[7] while they fit in sample_capacity.
  ge_cache_query* queries;   // cached: where a path ends into the cache (tag = the result slot)
  float4* query_weights;
  ge_cache_sample* samples;  // radiance = L after the vertex's emission, until the targets pass
  float4* sample_weights;
  uint32_t* cache_counts;
  uint32_t query_capacity, sample_capacity;
  uint32_t cache_depth;      // cached: the bounce whose hit ends the path into the cache
  uint32_t train_stride;     // cached: paths with pid % stride == phase train (0: none do)
  uint32_t train_phase;
  uint32_t npaths;           // a training path's own radiance goes to result[npaths + pid]
  float4* result;            // cached: the image's radiance of a path that goes on to train
  CacheBox box;
  // The adjoint (kShadeAdjoint): sampling table, per-path sampling throughput, and with grad the
  // per-record and per-vertex gradient terms.
  const ge_resolve_material* sampling;  // null: the scene's materials
  const float4* ts_in;  // the sampling throughput, double-buffered as the path state is
  float4* ts_out;
  bool grad;
  const float4* d_image;
  uint32_t pixels;
  float spp;
  ShadowGrad* sgrad;
  float4* adj_a;
  float4* adj_r;
  double* grad_lights;
  // A launch run again after its shadow records overflowed: its one non-idempotent effect, the
  // workgroup's ground term added to grad_lights[0], was made by the first run and is not repeated.
  bool rerun;
};

enum : int {
  kShadePlain = 0,   // the path tracer
  kShadeCached = 1,  // paths end into the cache at the hit of bounce cache_depth; the training
                     // paths among them record a sample there and go on one more bounce (next-event
                     // estimation, then the cache at the next hit: self-training)
  kShadeRecord = 2,  // the path tracer, recording a training sample at every hit vertex
  kShadeAdjoint = 3, // the detached estimator (sampling from its own table), and with grad its gradient terms
};
constexpr uint32_t kTrainingPath = 4u;  // PathState::flags: the path goes on past the cache to train

// Octahedral coordinates of a unit vector in [0, 1]^2: (x, z) / (|x| + |y| + |z|), folded over
// the diagonals when y < 0.
__device__ __forceinline__ void octahedral(f3 v, float& u, float& w) {
  const float s = (fabsf(v.x) + fabsf(v.y)) + fabsf(v.z);
  float px = v.x / s, pz = v.z / s;
  if (v.y < 0.0f) {
    const float fx = (1.0f - fabsf(pz)) * (px >= 0.0f ? 1.0f : -1.0f);
    const float fz = (1.0f - fabsf(px)) * (pz >= 0.0f ? 1.0f : -1.0f);
    px = fx;
    pz = fz;
  }
  u = px * 0.5f + 0.5f;
  w = pz * 0.5f + 0.5f;
}

// The cache's query at a surface point, looking back along `s.view`.
__device__ __forceinline__ ge_cache_query cache_query(const CacheBox& box, f3 position, const Surface& s, uint32_t tag) {
  ge_cache_query q;
  q.position[0] = (position.x - box.lo[0]) * box.inv[0];
  q.position[1] = (position.y - box.lo[1]) * box.inv[1];
  q.position[2] = (position.z - box.lo[2]) * box.inv[2];
  octahedral(s.view, q.direction[0], q.direction[1]);
  octahedral(s.normal, q.normal[0], q.normal[1]);
  q.albedo[0] = s.albedo.x;
  q.albedo[1] = s.albedo.y;
  q.albedo[2] = s.albedo.z;
  q.roughness = s.roughness;
  q.tag = tag;
  return q;
}

// Which lights next-event estimation traces from this surface, in next_event's order: the sun, then
// each light whose sample is not zero and that is in front of the shading normal.
__device__ __forceinline__ bool sun_asked(const ge_path_trace_params& p, const Surface& s) {
  return p.sun[3] > 0.0f && dot(s.normal, ld3(p.sun)) > 0.0f;
}

// `Full`: textured materials (sample_material and the normal map at level 0); without it every
// material is its factors.
// Occupancy asked of the shading, in waves per SIMD: 5 holds the plain, cached and recording
// shadings at 96 VGPRs without spills and runs the 4M-triangle frame 1% faster than the compiler's
// 98 VGPRs (4 waves); 6 spills and runs 3% slower. The adjoint's shading (137 VGPRs) keeps 3.
#ifndef GE_SHADE_WAVES_PER_EU
#define GE_SHADE_WAVES_PER_EU 5
#endif
#ifndef GE_SHADE_ADJOINT_WAVES_PER_EU
#define GE_SHADE_ADJOINT_WAVES_PER_EU 3
#endif
template <int Mode, bool Full = false>
__global__ void __launch_bounds__(k_shade_threads)
    __attribute__((amdgpu_waves_per_eu(Mode == kShadeAdjoint ? GE_SHADE_ADJOINT_WAVES_PER_EU : GE_SHADE_WAVES_PER_EU)))
    k_shade(ShadeArgs a) {
  __shared__ uint32_t parts[Mode == kShadePlain ? 2 : 4][k_shade_threads / 64 + 1];
  __shared__ double ground_acc;  // kShadeAdjoint: the workgroup's sum of the sun's ground-radiance term
  const bool grad = Mode == kShadeAdjoint && a.grad;
  if (grad) {
    if (threadIdx.x == 0) ground_acc = 0.0;
    __syncthreads();
  }
  // The walk heads: this bounce's walk is over, the next starts after this launch.
  if (blockIdx.x == 0 && threadIdx.x < 2) a.counts[2 + threadIdx.x] = 0u;
  const uint32_t g = blockIdx.x * k_shade_threads + threadIdx.x;
  const bool valid = g < a.n;
  const uint32_t pid = valid ? a.queue[g] : 0u;
  PathState st;
  bool alive = false;
  uint32_t nshadow = 0;
  uint64_t lit = 0;  // lights 0..63 that give this vertex a shadow record
  Surface surface, ssurface;
  Vertex vx;
  float bias = 0.0f;
  f3 front = splat(0.0f);
  const ge_path_trace_params& p = a.p;
  bool hit_found = false, cache_here = false, train_here = false;
  f3 Ts = splat(1.0f), delta = splat(0.0f);
  uint32_t mat_index = GE_RT_MISS;
  if (valid) {
    st = a.in[pid];
    const PathHit hit = a.hits[pid];
    const f3 dir = f3{st.d[0], st.d[1], st.d[2]};
    f3 L = f3{st.L[0], st.L[1], st.L[2]};
    const f3 T = f3{st.T[0], st.T[1], st.T[2]};
    if (Mode == kShadeAdjoint && st.bounce != 0) {
      const float4 t4 = a.ts_in[pid];
      Ts = f3{t4.x, t4.y, t4.z};
    }
    if (grad) {
      const float4 d = a.d_image[pid % a.pixels];
      delta = f3{d.x, d.y, d.z} / a.spp;
    }
    if (hit.inst == GE_RT_MISS) {
      L = L + T * environment(p, dir, st.bounce == 0);
      // An escaped ray below the horizon sees the sun-lit ground (path_trace.slang:166-176).
      if (grad && st.bounce != 0 && !(p.flags & GE_PT_UNIFORM_SKY) && !(dir.y >= 0.0f))
        atomicAdd(&ground_acc, double(dot(delta, T * ld3(p.ground) * (maxr(p.sun[1], 0.0f) / k_pi))));
      if (grad) a.adj_a[pid] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(GE_RT_MISS));
      st.L[0] = L.x;
      st.L[1] = L.y;
      st.L[2] = L.z;
      NeeRec nr;
      nr.T[0] = nr.T[1] = nr.T[2] = 0.0f;
      nr.first = 0;
      nr.count = GE_RT_MISS;
      a.nee[pid] = nr;
    } else {
      hit_found = true;
      if (st.bounce == 0) st.flags |= 1u;
      const ge_instance_desc in = a.sv.instances[hit.inst];
      const f3 origin = f3{st.o[0], st.o[1], st.o[2]};
      uint32_t cluster_index;
      Corners corners;
      vx = vertex_of(a.sv, p, hit, origin, in, cluster_index, Full ? &corners : nullptr);
      mat_index = a.sv.cluster_materials[cluster_index] + in.material_base;
      const ge_resolve_material* mat = a.sv.materials + mat_index;
      const float4 spc = *reinterpret_cast<const float4*>(mat->specular);
      f3 albedo, emissive;
      float roughness, metallic;
      if (Full) {
        const MaterialSample ms = shade_material(a.sv, *mat, corners, f3{1.0f - hit.u - hit.v, hit.u, hit.v}, vx.normal);
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
      f3 shading_normal = vx.normal;
      if (dot(shading_normal, vx.geometric) < 0.0f) shading_normal = -shading_normal;
      L = L + T * emissive;
      st.L[0] = L.x;
      st.L[1] = L.y;
      st.L[2] = L.z;
      surface.albedo = albedo;
      surface.normal = shading_normal;
      surface.view = -dir;
      surface.roughness = roughness;
      surface.metallic = metallic;
      surface.spec[0] = spc.x;
      surface.spec[1] = spc.y;
      surface.spec[2] = spc.z;
      surface.spec[3] = spc.w;
      ssurface = surface;
      if (Mode == kShadeAdjoint && a.sampling) {
        const ge_resolve_material* smat = a.sampling + mat_index;
        const float4 salb = *reinterpret_cast<const float4*>(smat->albedo);
        ssurface.albedo = f3{salb.x, salb.y, salb.z};
        ssurface.roughness = salb.w;
        ssurface.metallic = smat->emissive[3];
      }
      bias = p.ray_bias + p.ray_bias_steps * a.sv.meshes[in.mesh].quant[3] * in.scale_max;
      front = dot(vx.geometric, surface.view) >= 0.0f ? vx.geometric : -vx.geometric;
      if (Mode == kShadeCached) {
        const bool training = (st.flags & kTrainingPath) != 0;
        train_here = !training && st.bounce == a.cache_depth && a.train_stride != 0 &&
                     pid % a.train_stride == a.train_phase;
        cache_here = training ? st.bounce == a.cache_depth + 1 : (st.bounce == a.cache_depth && !train_here);
      }
      // Count the records first; the same tests run again when they are written.
      if (!cache_here) {
        nshadow = sun_asked(p, surface) ? 1u : 0u;
        // Lights 0..63 through the light grid's cell mask (a light not in the cell's mask is out
        // of range there), the rest in full.
        const uint32_t low = a.sv.light_count < 64u ? a.sv.light_count : 64u;
        uint64_t cand = low == 64u ? ~0ull : (1ull << low) - 1ull;
        if (a.sv.light_cells) {
          const float gx = (vx.position.x - a.sv.light_lo[0]) * a.sv.light_inv[0];
          const float gy = (vx.position.y - a.sv.light_lo[1]) * a.sv.light_inv[1];
          const float gz = (vx.position.z - a.sv.light_lo[2]) * a.sv.light_inv[2];
          constexpr float G = float(k_light_grid);
          if (gx >= 0.0f && gx < G && gy >= 0.0f && gy < G && gz >= 0.0f && gz < G) {
            const uint32_t cell = (uint32_t(gz) * k_light_grid + uint32_t(gy)) * k_light_grid + uint32_t(gx);
            cand &= a.sv.light_cells[cell];
          } else {
            cand &= a.sv.light_far;
          }
        }
        while (cand) {
          const uint32_t i = uint32_t(__builtin_ctzll(cand));
          cand &= cand - 1ull;
          f3 ld, rad;
          float dist;
          if (!light_sample(a.sv.lights[i], vx.position, ld, rad, dist)) continue;
          if (dot(surface.normal, ld) <= 0.0f) continue;
          ++nshadow;
          lit |= 1ull << i;
        }
        for (uint32_t i = 64u; i < a.sv.light_count; ++i) {
          f3 ld, rad;
          float dist;
          if (!light_sample(a.sv.lights[i], vx.position, ld, rad, dist)) continue;
          if (dot(surface.normal, ld) <= 0.0f) continue;
          ++nshadow;
        }
      }
    }
  }
  const uint32_t first = block_append(&a.counts[1], nshadow, parts[0]);
  if (Mode != kShadePlain) {
    // A query where the path ends into the cache, or where a training path leaves the image's.
    const bool query = Mode == kShadeCached && (cache_here || train_here);
    const bool sample = hit_found && (Mode == kShadeRecord || train_here);
    const uint32_t qslot = block_append(&a.cache_counts[0], query ? 1u : 0u, parts[2]);
    const uint32_t sslot = block_append(&a.cache_counts[1], sample ? 1u : 0u, parts[3]);
    if (query || sample) {
      const ge_cache_query q = cache_query(a.box, vx.position, surface, st.bounce);
      const float4 T = make_float4(st.T[0], st.T[1], st.T[2], 0.0f);
      if (query && qslot < a.query_capacity) {
        ge_cache_query qq = q;
        qq.tag = (st.flags & kTrainingPath) ? a.npaths + pid : pid;
        a.queries[qslot] = qq;
        a.query_weights[qslot] = T;
      }
      if (sample && sslot < a.sample_capacity) {
        ge_cache_sample smp;
        smp.query = q;
        smp.radiance[0] = st.L[0];
        smp.radiance[1] = st.L[1];
        smp.radiance[2] = st.L[2];
        smp.path = pid;
        a.samples[sslot] = smp;
        a.sample_weights[sslot] = T;
      }
      if (train_here) {
        // The image's radiance for this path stops here (the cache's answer is added to it); the
        // path goes on with its own, into result[npaths + pid].
        a.result[pid] = make_float4(st.L[0], st.L[1], st.L[2], __uint_as_float(st.flags & 1u));
        st.flags |= kTrainingPath;
      }
    }
  }
  if (valid && hit_found) {
    const bool fits = first + nshadow <= a.shadow_capacity;
    NeeRec nr;
    nr.T[0] = st.T[0];
    nr.T[1] = st.T[1];
    nr.T[2] = st.T[2];
    nr.first = first;
    nr.count = nshadow;
    a.nee[pid] = nr;
    if (fits && nshadow) {
      const f3 shadow_origin = vx.position + front * bias;
      uint32_t k = first;
      const f3 T = f3{st.T[0], st.T[1], st.T[2]};
      // kShadeAdjoint with grad: what the record adds to the gradient if nothing blocks it. `unit`
      // is the light's radiance per unit intensity.
      auto emit_grad = [&](uint32_t k, uint32_t light, f3 ld, f3 radiance, f3 unit) {
        f3 ga, gr;
        brdf_eval_grad(surface, ld, ga, gr);
        const float cosl = maxr(dot(surface.normal, ld), 0.0f);
        ShadowGrad q;
        const f3 ta = delta * T * (ga * cosl * radiance);
        q.ga[0] = ta.x;
        q.ga[1] = ta.y;
        q.ga[2] = ta.z;
        q.gr = dot(delta, T * (gr * cosl * radiance));
        q.gl = dot(delta, T * (brdf_eval(surface, ld) * cosl * unit));
        q.light = light;
        q.pad[0] = q.pad[1] = 0;
        a.sgrad[k] = q;
      };
      auto emit = [&](f3 ld, float tmax, f3 c) {
        ShadowRec r;
        r.o[0] = shadow_origin.x;
        r.o[1] = shadow_origin.y;
        r.o[2] = shadow_origin.z;
        r.tmax = tmax;
        r.d[0] = ld.x;
        r.d[1] = ld.y;
        r.d[2] = ld.z;
        r.c[0] = c.x;
        r.c[1] = c.y;
        r.c[2] = c.z;
        r.pad0 = r.pad1 = 0;
        a.shadow[k++] = r;
      };
      if (sun_asked(p, surface)) {
        if (grad) emit_grad(k, 0, ld3(p.sun), splat(p.sun[3]), splat(1.0f));
        emit(ld3(p.sun), 1e30f, brdf_direct(surface, ld3(p.sun), splat(p.sun[3])));
      }
      auto emit_light = [&](uint32_t i) {
        const ge_resolve_light light = a.sv.lights[i];
        f3 ld, rad;
        float dist, att;
        if (!light_sample(light, vx.position, ld, rad, dist, &att)) return;
        if (dot(surface.normal, ld) <= 0.0f) return;
        if (grad) emit_grad(k, 1 + i, ld, rad, ld3(light.color_intensity) * att);
        emit(ld, dist - bias - light.spot[2], brdf_direct(surface, ld, rad));
      };
      // The counting pass's lights again, in the same order: the first 64 from its mask, the rest
      // tested in full.
      for (uint64_t m = lit; m != 0u; m &= m - 1u) emit_light(uint32_t(__builtin_ctzll(m)));
      for (uint32_t i = 64; i < a.sv.light_count; ++i) emit_light(i);
    }
    f3 ratio_a = splat(0.0f), ratio_r = splat(0.0f);  // kShadeAdjoint: (df/dparameter) / f along the next direction
    // ---- the next direction (path_trace.slang:397-455) ----
    alive = !cache_here && st.bounce < p.max_bounces;
    if (alive) {
      const float n_dot_v = dot(surface.normal, surface.view);
      alive = n_dot_v > 0.0f;
      if (alive) {
        f3 tangent, bitangent;
        onb(surface.normal, tangent, bitangent);
        const f3 v_local = f3{dot(surface.view, tangent), dot(surface.view, bitangent), dot(surface.view, surface.normal)};
        // The sampling decisions come from the sampling surface (the scene's own outside the adjoint).
        const float alpha = brdf_alpha(ssurface.roughness);
        const f3 f = fresnel(ssurface, n_dot_v);
        const float p_specular = specular_probability(ssurface.albedo, ssurface.metallic, f);
        uint32_t rng = st.rng;
        f3 l_local;
        if (rng_next(rng) < p_specular) {
          const float u0 = rng_next(rng);
          const float u1 = rng_next(rng);
          const f3 h = ggx_vndf(v_local, alpha, u0, u1);
          l_local = reflect3(-v_local, h);
        } else {
          const float u0 = rng_next(rng);
          const float u1 = rng_next(rng);
          l_local = cosine_hemisphere(u0, u1);
        }
        alive = l_local.z > 0.0f;
        float pdf = 0.0f;
        if (alive) {
          pdf = bsdf_pdf(v_local, l_local, alpha, p_specular);
          alive = pdf > 0.0f;
        }
        if (alive) {
          const f3 next = tangent * l_local.x + bitangent * l_local.y + surface.normal * l_local.z;
          const float scale = l_local.z / pdf;
          const f3 fe = brdf_eval(surface, next);
          const f3 weight = fe * scale;
          f3 T = f3{st.T[0], st.T[1], st.T[2]} * weight;
          if (Mode == kShadeAdjoint) {
            // Roulette and the end test follow the sampling throughput (the scene's: T itself).
            Ts = Ts * (a.sampling ? brdf_eval(ssurface, next) * scale : weight);
            alive = !(Ts.x <= 0.0f && Ts.y <= 0.0f && Ts.z <= 0.0f);
            if (alive && st.bounce + 1 >= GE_PT_ROULETTE_DEPTH) {
              const float survive = clampr(maxr(Ts.x, maxr(Ts.y, Ts.z)), 0.0f, 0.95f);
              if (rng_next(rng) >= survive) {
                alive = false;
              } else {
                T = T / maxr(survive, 1e-6f);
                Ts = Ts / maxr(survive, 1e-6f);
              }
            }
            if (alive && grad) {
              f3 ga, gr;
              brdf_eval_grad(surface, next, ga, gr);
              ratio_a = f3{fe.x > 0.0f ? ga.x / fe.x : 0.0f, fe.y > 0.0f ? ga.y / fe.y : 0.0f,
                           fe.z > 0.0f ? ga.z / fe.z : 0.0f};
              ratio_r = f3{fe.x > 0.0f ? gr.x / fe.x : 0.0f, fe.y > 0.0f ? gr.y / fe.y : 0.0f,
                           fe.z > 0.0f ? gr.z / fe.z : 0.0f};
            }
          } else {
            alive = !(T.x <= 0.0f && T.y <= 0.0f && T.z <= 0.0f);
            if (alive && st.bounce + 1 >= GE_PT_ROULETTE_DEPTH) {
              const float survive = clampr(maxr(T.x, maxr(T.y, T.z)), 0.0f, 0.95f);
              if (rng_next(rng) >= survive) alive = false;
              else T = T / maxr(survive, 1e-6f);
            }
          }
          if (alive) {
            const f3 o = vx.position + (dot(next, front) >= 0.0f ? front : -front) * bias;
            st.o[0] = o.x;
            st.o[1] = o.y;
            st.o[2] = o.z;
            st.d[0] = next.x;
            st.d[1] = next.y;
            st.d[2] = next.z;
            st.T[0] = T.x;
            st.T[1] = T.y;
            st.T[2] = T.z;
            st.bounce += 1;
          }
        }
        st.rng = rng;
      }
    }
    if (grad) {
      a.adj_a[pid] = make_float4(ratio_a.x, ratio_a.y, ratio_a.z, __uint_as_float(mat_index));
      a.adj_r[pid] = make_float4(ratio_r.x, ratio_r.y, ratio_r.z, alive ? 1.0f : 0.0f);
    }
  }
  if (valid) {
    if (!alive) st.flags |= 2u;
    a.out[pid] = st;
    if (Mode == kShadeAdjoint && alive) a.ts_out[pid] = make_float4(Ts.x, Ts.y, Ts.z, 0.0f);
  }
  const uint32_t slot = block_append(&a.counts[0], alive ? 1u : 0u, parts[1]);
  if (alive) a.next_queue[slot] = pid;
  if (grad) {
    __syncthreads();
    if (threadIdx.x == 0 && ground_acc != 0.0 && !a.rerun) atomicAdd(&a.grad_lights[0], ground_acc);
  }
}

// One bounce's shadow rays: record j's ray, its answer into blocked[j].
struct AnyJob {
  const ShadowRec* shadow;
  uint32_t n;
  uint32_t* head;
  uint8_t* blocked;
  uint32_t alpha;  // the masked-candidate rule of light rays
  const uint32_t* count = nullptr;  // set: the records are the first min(n, *count), read on the device
};

template <bool Count, bool Masks>
__device__ void any_walks(const Bvh& bvh, const AnyJob& job, uint64_t& issued, Tally& tally) {
  const ShadowRec* __restrict__ shadow = job.shadow;
  uint8_t* __restrict__ blocked = job.blocked;
  auto load = [&](uint32_t j, uint32_t& slot, Ray& ray, float& tmax, uint32_t& mask) {
    const ShadowRec& r = shadow[j];
    slot = j;
    tmax = r.tmax;
    mask = 0x01u;
    if (!(r.tmax > 0.0f)) return false;  // the bias ate the distance: never occluded (occluded(), :181)
    ray = make_ray(f3{r.o[0], r.o[1], r.o[2]}, f3{r.d[0], r.d[1], r.d[2]}, 0.0f);
    return true;
  };
  auto store = [&](uint32_t j, const Hit& h) { blocked[j] = h.found ? 1 : 0; };
  persistent_walk<true, Count, Masks, k_refill_paths>(bvh, job.n, job.head, lane_stack(), load, store, issued, tally,
                                                      job.alpha);
}

// The walks of one launch: a bounce's closest-hit rays, then the previous bounce's shadow rays (the
// two do not depend on each other: both were written by the previous shading launch). Each wave
// takes closest-hit rays until their queue is drained and then shadow rays, so the waves that run
// out of the first kind fill the other's tail instead of idling until the launch ends.
template <bool Count, bool Masks>
__global__ void __launch_bounds__(64) GE_WALK_ATTR k_walks(Bvh bvh, ClosestJob closest, AnyJob any,
                                                          unsigned long long* counters) {
  Tally tally;
  uint64_t closest_issued = 0, any_issued = 0;
  if (closest.count) closest.n = min(closest.n, *closest.count);
  if (any.count) any.n = min(any.n, *any.count);
  // A wave past the rays' count would find the counters spent: it leaves without contending for them.
  if (uint64_t(blockIdx.x) * 64u >= uint64_t(closest.n) + any.n) return;
  if (closest.n) closest_walks<Count, Masks>(bvh, closest, closest_issued, tally);
  if (any.n) any_walks<Count, Masks>(bvh, any, any_issued, tally);
  if (Count) flush_counters(counters, closest_issued, any_issued, tally);
}

// radiance + throughput * next_event(...): the unblocked lights summed in record order. A path that
// ended this bounce (a miss, or shade's break) leaves its radiance in `result`.
struct ResolveAdjoint {
  const ShadowGrad* sgrad;
  const float4* adj_a;
  const float4* adj_r;
  const float4* lpath;
  const float4* d_image;
  uint32_t pixels;
  float spp;
  uint32_t lights;  // the sun and the lights: 1 + light_count
  double* grad_materials;
  double* grad_lights;
};

// With Adjoint, also each vertex's gradient: its unblocked records' terms, and (df/f) x (L - prefix),
// what the path gathered after this vertex, when the path went on from it. The lights' terms are
// summed in LDS (every lane adds to the same few addresses) and flushed once a workgroup.
// Resolve workgroups: 256 threads, so the adjoint's light sums go to memory once a 256 paths (its
// resolve 3.5% faster than at 64; the path tracer's alike).
#ifndef GE_RESOLVE_THREADS
#define GE_RESOLVE_THREADS 256
#endif
constexpr uint32_t k_resolve_threads = GE_RESOLVE_THREADS;

template <bool Adjoint>
__global__ void __launch_bounds__(k_resolve_threads) k_resolve(PathState* __restrict__ state, const uint32_t* __restrict__ queue,
                                                uint32_t n, const NeeRec* __restrict__ nee,
                                                const ShadowRec* __restrict__ shadow,
                                                const uint8_t* __restrict__ blocked, float4* __restrict__ result,
                                                uint32_t npaths, ResolveAdjoint adj, uint32_t* clear) {
  // The next shading's queue and shadow-record counts (the walk that read them has ended).
  if (blockIdx.x == 0 && threadIdx.x < 2) clear[threadIdx.x] = 0u;
  extern __shared__ double light_acc[];
  if (Adjoint) {
    for (uint32_t i = threadIdx.x; i < adj.lights; i += k_resolve_threads) light_acc[i] = 0.0;
    __syncthreads();
  }
  const uint32_t g = blockIdx.x * k_resolve_threads + threadIdx.x;
  if (g < n) {
    const uint32_t pid = queue[g];
    const NeeRec nr = nee[pid];
    PathState& st = state[pid];
    f3 L = f3{st.L[0], st.L[1], st.L[2]};
    if (nr.count != GE_RT_MISS) {
      f3 direct = splat(0.0f);
      f3 ga = splat(0.0f);
      float gr = 0.0f;
      for (uint32_t k = 0; k < nr.count; ++k) {
        const uint32_t j = nr.first + k;
        if (blocked[j]) continue;
        direct = direct + f3{shadow[j].c[0], shadow[j].c[1], shadow[j].c[2]};
        if (Adjoint) {
          const ShadowGrad q = adj.sgrad[j];
          ga = ga + f3{q.ga[0], q.ga[1], q.ga[2]};
          gr += q.gr;
          atomicAdd(&light_acc[q.light], double(q.gl));
        }
      }
      L = L + f3{nr.T[0], nr.T[1], nr.T[2]} * direct;
      st.L[0] = L.x;
      st.L[1] = L.y;
      st.L[2] = L.z;
      if (Adjoint) {
        const float4 ra = adj.adj_a[pid], rr = adj.adj_r[pid];
        const uint32_t mat = __float_as_uint(ra.w);
        if (mat != GE_RT_MISS) {
          if (rr.w != 0.0f) {
            const float4 lp = adj.lpath[pid];
            const f3 after = f3{lp.x, lp.y, lp.z} - L;
            const float4 d = adj.d_image[pid % adj.pixels];
            const f3 delta = f3{d.x, d.y, d.z} / adj.spp;
            ga = ga + delta * f3{ra.x, ra.y, ra.z} * after;
            gr += dot(delta, f3{rr.x, rr.y, rr.z} * after);
          }
          double* gm = adj.grad_materials + size_t(mat) * 4;
          if (ga.x != 0.0f) atomicAdd(&gm[0], double(ga.x));
          if (ga.y != 0.0f) atomicAdd(&gm[1], double(ga.y));
          if (ga.z != 0.0f) atomicAdd(&gm[2], double(ga.z));
          if (gr != 0.0f) atomicAdd(&gm[3], double(gr));
        }
      }
    }
    if (st.flags & 2u)
      result[(st.flags & kTrainingPath) ? npaths + pid : pid] = make_float4(L.x, L.y, L.z, __uint_as_float(st.flags & 1u));
  }
  if (Adjoint) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < adj.lights; i += k_resolve_threads) {
      const double v = light_acc[i];
      if (v != 0.0) atomicAdd(&adj.grad_lights[i], v);
    }
  }
}

// accum[p] += (sum over samples in order, spp); coverage |= any primary hit (path_trace.slang:490-494).
__global__ void __launch_bounds__(64) k_accumulate(ge_path_trace_params p, const float4* __restrict__ result,
                                                   float4* __restrict__ accum, uint32_t* __restrict__ coverage) {
  const uint32_t pixel = blockIdx.x * 64 + threadIdx.x;
  const uint32_t pixels = p.width * p.height;
  if (pixel >= pixels) return;
  f3 sum = splat(0.0f);
  bool covered = false;
  for (uint32_t s = 0; s < p.spp; ++s) {
    const float4 r = result[size_t(s) * pixels + pixel];
    sum = sum + f3{r.x, r.y, r.z};
    covered = covered || (__float_as_uint(r.w) & 1u);
  }
  float4 prev = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (p.sample_base != 0) prev = accum[pixel];
  accum[pixel] = make_float4(prev.x + sum.x, prev.y + sum.y, prev.z + sum.z, prev.w + float(p.spp));
  const uint32_t before = p.sample_base == 0 ? 0u : coverage[pixel];
  coverage[pixel] = before | (covered ? 1u : 0u);
}

// A training sample's radiance: what the path gathered after the vertex's emission, over the
// throughput that reached the vertex (zero in a channel the throughput has none of).
__global__ void __launch_bounds__(256) k_cache_targets(ge_cache_sample* __restrict__ samples,
                                                       const float4* __restrict__ weights, const uint32_t* count,
                                                       uint32_t capacity, const float4* __restrict__ result,
                                                       uint32_t offset) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= min(*count, capacity)) return;
  ge_cache_sample& s = samples[i];
  const float4 R = result[offset + s.path], T = weights[i];
  s.radiance[0] = T.x > 0.0f ? (R.x - s.radiance[0]) / T.x : 0.0f;
  s.radiance[1] = T.y > 0.0f ? (R.y - s.radiance[1]) / T.y : 0.0f;
  s.radiance[2] = T.z > 0.0f ? (R.z - s.radiance[2]) / T.z : 0.0f;
}

inline uint32_t blocks(uint64_t n, uint32_t t) { return uint32_t((n + t - 1) / t); }
constexpr size_t k_lds = size_t(k_ring) * k_lanes * 4;


}  // namespace

hipError_t stage_mark(ge_rt_ctx* ctx, int kind, int end) {
  const uint32_t i = ctx->stage_n[kind];
  if (i >= GE_RT_STAGE_TIMED) return hipSuccess;
  const hipError_t e = hipEventRecord(ctx->stage_ev[kind][2 * i + end], ctx->stream);
  if (end) ctx->stage_n[kind] = i + 1;
  return e;
}

hipError_t wavefront_path_trace(ge_rt_ctx* ctx, const Bvh& bvh, const SceneView& sv, const ge_path_trace_params& p,
                                float4* accum, uint32_t* coverage, unsigned long long* counters,
                                const CacheHook* hook, const AdjointHook* adjoint) {
  Wavefront& w = ctx->wave;
  const hipStream_t st = ctx->stream;
  const uint32_t pixels = p.width * p.height;
  const uint64_t npaths64 = uint64_t(pixels) * p.spp;
  if (npaths64 == 0) return hipSuccess;
  if (npaths64 >= (1ull << 31)) return hipErrorInvalidValue;
  const uint32_t npaths = uint32_t(npaths64);
  hipError_t e;
  if ((e = w.state[0].reserve(size_t(npaths) * sizeof(PathState))) != hipSuccess) return e;
  if ((e = w.state[1].reserve(size_t(npaths) * sizeof(PathState))) != hipSuccess) return e;
  if ((e = w.hits.reserve(size_t(npaths) * sizeof(PathHit))) != hipSuccess) return e;
  if ((e = w.nee.reserve(size_t(npaths) * sizeof(NeeRec))) != hipSuccess) return e;
  const bool self_train = hook && hook->mode == CacheHook::kCached && hook->train_stride != 0;
  if ((e = w.result.reserve(size_t(npaths) * (self_train ? 32 : 16))) != hipSuccess) return e;
  if ((e = w.queue[0].reserve(size_t(npaths) * 4)) != hipSuccess) return e;
  if ((e = w.queue[1].reserve(size_t(npaths) * 4)) != hipSuccess) return e;
  if ((e = w.counts.reserve(256)) != hipSuccess) return e;
  if (!w.host_counts && (e = hipHostMalloc(reinterpret_cast<void**>(&w.host_counts), 64)) != hipSuccess) return e;
  if (!w.counted && (e = hipEventCreateWithFlags(&w.counted, hipEventDisableTiming)) != hipSuccess) return e;
  const bool grad = adjoint && adjoint->grad;
  if (adjoint) {
    if ((e = w.ts.reserve(size_t(npaths) * 32)) != hipSuccess) return e;  // two buffers, as the state
    if (grad) {
      if ((e = w.adj_a.reserve(size_t(npaths) * 16)) != hipSuccess) return e;
      if ((e = w.adj_r.reserve(size_t(npaths) * 16)) != hipSuccess) return e;
    }
  }
  auto grow_shadow = [&](uint64_t capacity) -> hipError_t {
    w.shadow_capacity = capacity;
    w.shadow.release();
    w.blocked.release();
    w.sgrad.release();
    hipError_t r = w.shadow.reserve(w.shadow_capacity * sizeof(ShadowRec));
    if (r == hipSuccess) r = w.blocked.reserve(w.shadow_capacity);
    return r;
  };
  if (w.shadow_capacity < 2ull * npaths && (e = grow_shadow(2ull * npaths)) != hipSuccess) return e;
  if (grad && (e = w.sgrad.reserve(w.shadow_capacity * sizeof(ShadowGrad))) != hipSuccess) return e;
  // [0] first queue; [2] next queue, [3] shadow records, [4] [5] walk heads; [6] cache records
  uint32_t* counts = w.counts.as<uint32_t>();
  if ((e = hipMemsetAsync(counts, 0, 32, st)) != hipSuccess) return e;
  const uint32_t tiles_x = (p.width + 7) / 8, tiles = tiles_x * ((p.height + 7) / 8);
  k_generate<<<tiles * p.spp, 64, 0, st>>>(p, tiles_x, tiles, w.state[0].as<PathState>(), w.queue[0].as<uint32_t>());
  if ((e = hipGetLastError()) != hipSuccess) return e;
  int cur = 0;
  uint32_t n = npaths;  // every path starts live
  uint32_t cache_before[2] = {0u, 0u};  // the cache's query and sample counts before this bounce's shade
  uint32_t* c2 = counts + 2;
  // Every live path is at the same bounce: the loop's iteration. The walks' rule for masked
  // candidates follows it (path_trace.slang:278-292), and the shading's textured form the scene.
  const bool full = sv.features != 0u;
  const bool masked = (sv.features & k_scene_masked) != 0u;
  const uint32_t any_alpha = masked && (p.flags & GE_PT_ALPHA_MASKS) ? k_alpha_test : k_alpha_opaque;
  // A bounce's resolve waits for its shadow rays, which are walked in the same launch as the next
  // bounce's closest-hit rays: walks(b + 1 closest, b shadow), resolve(b), shade(b + 1). Each path's
  // radiance still takes its additions in the order of the single-bounce loop (bounce b's lights
  // before bounce b + 1's emission or environment).
  struct Pending {
    bool live = false;        // a shaded bounce whose resolve has not run
    uint32_t n = 0;           // its paths
    uint32_t shadows = 0;     // its shadow records
    PathState* state = nullptr;
    const uint32_t* queue = nullptr;
  } pending;
  auto alpha_of = [&](uint32_t bounce) {
    return !masked ? k_alpha_opaque
                   : ((p.flags & GE_PT_ALPHA_MASKS) ? k_alpha_test : (bounce == 0 ? k_alpha_cull : k_alpha_opaque));
  };
  // One walk launch: closest-hit rays of `cj`, shadow records of `aj`. Its heads (counts[4], [5])
  // are zero: the frame's first clear, then each shading's.
  auto launch_walks = [&](const ClosestJob& cj, const AnyJob& aj, bool mask_walk, uint32_t grid) -> hipError_t {
    hipError_t r;
    if ((r = stage_mark(ctx, 0, 0)) != hipSuccess) return r;
    if (mask_walk) {
      if (counters) k_walks<true, true><<<grid, 64, k_lds, st>>>(bvh, cj, aj, counters);
      else k_walks<false, true><<<grid, 64, k_lds, st>>>(bvh, cj, aj, counters);
    } else {
      if (counters) k_walks<true, false><<<grid, 64, k_lds, st>>>(bvh, cj, aj, counters);
      else k_walks<false, false><<<grid, 64, k_lds, st>>>(bvh, cj, aj, counters);
    }
    if ((r = hipGetLastError()) != hipSuccess) return r;
    return stage_mark(ctx, 0, 1);
  };
  // Plain tracing enqueues the next bounce's walks before it waits for the shading's counts (they
  // read the counts on the device), so the host's round trip hides behind the walk; a shading that
  // overflows its shadow records runs again, and so does that walk. Counted walks, the cache's
  // recording and the adjoint's passes wait first (a walk run twice would count twice; the
  // recording and the adjoint keep their plain order).
  const bool ahead = !counters && !adjoint && (!hook || hook->mode == CacheHook::kCached);
  bool walked = false;        // this bounce's walks were enqueued ahead
  uint32_t last_shadows = 0;  // the previous bounce's shadow records: the size of the walk ahead's grid
  for (uint32_t bounce = 0; n > 0 || pending.live; ++bounce) {
    const uint32_t alpha = alpha_of(bounce);
    const uint64_t walk_rays = uint64_t(n) + pending.shadows;
    if (!walked && walk_rays > 0) {
      ClosestJob cj{w.state[cur].as<PathState>(), w.queue[0].as<uint32_t>(), n, counts + 4, w.hits.as<PathHit>(), alpha};
      AnyJob aj{w.shadow.as<ShadowRec>(), pending.shadows, counts + 5, w.blocked.as<uint8_t>(), any_alpha};
      const bool mask_walk = (n > 0 && alpha != k_alpha_opaque) || (pending.shadows > 0 && any_alpha != k_alpha_opaque);
      if ((e = launch_walks(cj, aj, mask_walk, std::min(blocks(walk_rays, 64), ctx->persistent_blocks))) != hipSuccess)
        return e;
    }
    walked = false;
    if (pending.live) {
      if (grad) {
        ResolveAdjoint ra;
        ra.sgrad = w.sgrad.as<ShadowGrad>();
        ra.adj_a = w.adj_a.as<float4>();
        ra.adj_r = w.adj_r.as<float4>();
        ra.lpath = adjoint->lpath;
        ra.d_image = adjoint->d_image;
        ra.pixels = pixels;
        ra.spp = float(p.spp);
        ra.lights = sv.light_count + 1;
        ra.grad_materials = adjoint->grad_materials;
        ra.grad_lights = adjoint->grad_lights;
        k_resolve<true><<<blocks(pending.n, k_resolve_threads), k_resolve_threads, size_t(ra.lights) * 8, st>>>(
            pending.state, pending.queue, pending.n, w.nee.as<NeeRec>(), w.shadow.as<ShadowRec>(),
            w.blocked.as<uint8_t>(), w.result.as<float4>(), npaths, ra, c2);
      } else {
        k_resolve<false><<<blocks(pending.n, k_resolve_threads), k_resolve_threads, 0, st>>>(pending.state, pending.queue, pending.n,
                                                              w.nee.as<NeeRec>(), w.shadow.as<ShadowRec>(),
                                                              w.blocked.as<uint8_t>(), w.result.as<float4>(), npaths,
                                                              ResolveAdjoint{}, c2);
      }
      if ((e = hipGetLastError()) != hipSuccess) return e;
      pending.live = false;
      pending.shadows = 0;
    }
    if (n == 0) break;
    ShadeArgs a{};
    a.sv = sv;
    a.p = p;
    a.in = w.state[cur].as<PathState>();
    a.out = w.state[cur ^ 1].as<PathState>();
    a.queue = w.queue[0].as<uint32_t>();
    a.n = n;
    a.hits = w.hits.as<PathHit>();
    a.nee = w.nee.as<NeeRec>();
    a.next_queue = w.queue[1].as<uint32_t>();
    a.counts = c2;
    a.cache_counts = counts + 6;
    a.npaths = npaths;
    a.result = w.result.as<float4>();
    if (adjoint) {
      a.sampling = adjoint->sampling;
      a.ts_in = w.ts.as<float4>() + size_t(cur) * npaths;
      a.ts_out = w.ts.as<float4>() + size_t(cur ^ 1) * npaths;
      a.grad = grad;
      a.d_image = adjoint->d_image;
      a.pixels = pixels;
      a.spp = float(p.spp);
      a.sgrad = w.sgrad.as<ShadowGrad>();
      a.adj_a = w.adj_a.as<float4>();
      a.adj_r = w.adj_r.as<float4>();
      a.grad_lights = adjoint->grad_lights;
    }
    if (hook) {
      a.queries = hook->queries;
      a.query_weights = hook->query_weights;
      a.query_capacity = hook->query_capacity;
      a.samples = hook->samples;
      a.sample_weights = hook->sample_weights;
      a.sample_capacity = hook->sample_capacity;
      a.cache_depth = hook->depth;
      a.train_stride = hook->train_stride;
      a.train_phase = hook->train_phase;
      a.box = hook->box;
    }
    // The next queue's and the shadow records' counts are zero: the frame's first clear, then the
    // previous bounce's resolve's (every bounce after the first resolves one before it shades).
    uint32_t nshadow = 0, next = 0;
    a.rerun = false;
    for (;;) {
      a.shadow = w.shadow.as<ShadowRec>();
      a.shadow_capacity = uint32_t(std::min<uint64_t>(w.shadow_capacity, 0xffffffffu));
      a.sgrad = w.sgrad.as<ShadowGrad>();
      const uint32_t sg = blocks(n, k_shade_threads);
      if (adjoint)
        k_shade<kShadeAdjoint><<<sg, k_shade_threads, 0, st>>>(a);
      else if (!hook)
        full ? k_shade<kShadePlain, true><<<sg, k_shade_threads, 0, st>>>(a)
             : k_shade<kShadePlain><<<sg, k_shade_threads, 0, st>>>(a);
      else if (hook->mode == CacheHook::kCached)
        full ? k_shade<kShadeCached, true><<<sg, k_shade_threads, 0, st>>>(a)
             : k_shade<kShadeCached><<<sg, k_shade_threads, 0, st>>>(a);
      else
        full ? k_shade<kShadeRecord, true><<<sg, k_shade_threads, 0, st>>>(a)
             : k_shade<kShadeRecord><<<sg, k_shade_threads, 0, st>>>(a);
      if ((e = hipGetLastError()) != hipSuccess) return e;
      // The next queue's and the shadow records' counts, and the cache's query and sample counts.
      if ((e = hipMemcpyAsync(w.host_counts, c2, 24, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
      if (ahead) {
        // The next bounce's closest-hit rays (at most this bounce's n) and this bounce's shadow
        // records (at most the buffer's capacity), counted on the device. The host waits for the
        // counts only, while the walk runs. The grid is sized by the previous bounce's shadow
        // records (the first bounce: the capacity): a persistent walk covers every ray whatever
        // its grid, but spare waves contend for the ray counter.
        if ((e = hipEventRecord(w.counted, st)) != hipSuccess) return e;
        const uint32_t alpha1 = alpha_of(bounce + 1);
        ClosestJob cj{w.state[cur ^ 1].as<PathState>(), w.queue[1].as<uint32_t>(), n, counts + 4, w.hits.as<PathHit>(),
                      alpha1, c2};
        AnyJob aj{w.shadow.as<ShadowRec>(), a.shadow_capacity, counts + 5, w.blocked.as<uint8_t>(), any_alpha, c2 + 1};
        const bool mask_walk = alpha1 != k_alpha_opaque || any_alpha != k_alpha_opaque;
        const uint64_t guess = uint64_t(n) + (bounce == 0 ? a.shadow_capacity : last_shadows);
        if ((e = launch_walks(cj, aj, mask_walk, std::min(blocks(guess, 64), ctx->persistent_blocks))) != hipSuccess) return e;
        walked = true;
        if ((e = hipEventSynchronize(w.counted)) != hipSuccess) return e;
      } else if ((e = hipStreamSynchronize(st)) != hipSuccess) {
        return e;
      }
      next = w.host_counts[0];
      nshadow = w.host_counts[1];
      if (nshadow <= a.shadow_capacity) {
        cache_before[0] = w.host_counts[4];
        cache_before[1] = w.host_counts[5];
        break;
      }
      // The records did not fit: grow the buffer and shade this bounce again (idempotent: the
      // launch reads one state buffer and writes the other), from the counts it started with: the
      // queue and shadow counts at zero and the cache's counts as the previous bounce left them.
      // A walk enqueued ahead reads the buffers being replaced: it ends first.
      if (ahead && (e = hipStreamSynchronize(st)) != hipSuccess) return e;
      if ((e = grow_shadow(uint64_t(nshadow) + nshadow / 4)) != hipSuccess) return e;
      if (grad && (e = w.sgrad.reserve(w.shadow_capacity * sizeof(ShadowGrad))) != hipSuccess) return e;
      if ((e = hipMemsetAsync(c2, 0, 8, st)) != hipSuccess) return e;
      w.host_counts[6] = cache_before[0];
      w.host_counts[7] = cache_before[1];
      if ((e = hipMemcpyAsync(counts + 6, w.host_counts + 6, 8, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
      a.rerun = true;
    }
    last_shadows = nshadow;
    pending.live = true;
    pending.n = n;
    pending.shadows = nshadow;
    pending.state = w.state[cur ^ 1].as<PathState>();
    pending.queue = w.queue[0].as<uint32_t>();  // swapped to [1] below; the next shade writes it after the resolve
    cur ^= 1;
    std::swap(w.queue[0], w.queue[1]);
    n = next;
  }
  if (hook && hook->mode == CacheHook::kRecord) {
    k_cache_targets<<<blocks(hook->sample_capacity, 256), 256, 0, st>>>(
        hook->samples, hook->sample_weights, counts + 7, hook->sample_capacity, w.result.as<float4>(), 0);
    return hipGetLastError();
  }
  if (hook) {
    // Each query's result slot += throughput x max(prediction, 0): the image's paths, and the
    // training paths' own radiance; then the training samples' targets from the latter.
    if ((e = hook->cache->apply(st, hook->queries, hook->query_weights, hook->query_capacity, counts + 6,
                                w.result.as<float4>())) != hipSuccess)
      return e;
    if (self_train) {
      k_cache_targets<<<blocks(hook->sample_capacity, 256), 256, 0, st>>>(
          hook->samples, hook->sample_weights, counts + 7, hook->sample_capacity, w.result.as<float4>(), npaths);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
  }
  if (grad || !accum) return hipSuccess;  // the gradient pass, or a caller that wants only `result`
  k_accumulate<<<blocks(pixels, 64), 64, 0, st>>>(p, w.result.as<float4>(), accum, coverage);
  return hipGetLastError();
}

hipError_t read_cache_counts(ge_rt_ctx* ctx, uint32_t* queries, uint32_t* samples) {
  Wavefront& w = ctx->wave;
  hipError_t e = hipMemcpyAsync(w.host_counts + 8, w.counts.as<uint32_t>() + 6, 8, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) {
    if (queries) *queries = w.host_counts[8];
    if (samples) *samples = w.host_counts[9];
  }
  return e;
}

}  // namespace ge
