// rt_shadow.hip — the sun shadow pass over a visibility buffer.
//
// Two launches. k_receivers turns each visibility word (or each 2x2 block of one pair, at half
// rate) into a receiver as the resolve reconstructs it: the pixel's triangle fetched into the
// frame, the barycentrics where the pixel's ray meets its plane (reconstruct_screen,
// material.slang:386-458), the geometric normal turned to the camera, and the sun ray's origin
// offset along it by ray_offset (shadow_visibility, visibility_resolve.slang:742-755). It writes
// the entry's answer without the blocked bit, and a 16-byte ray record (origin, t_max; t_max <= 0
// for no ray). k_sun_walk is a persistent any-hit walk over those records along the one sun
// direction that sets the blocked bit.
#include "rt_internal.h"
#include "rt_shade.h"

namespace ge {
namespace {

extern __shared__ uint32_t g_lds[];
__device__ __forceinline__ uint32_t* lane_stack() { return g_lds + threadIdx.x; }

struct ReceiverArgs {
  SceneView sv;
  ge_frame_eye eye;
  ge_shadow_resolve_params p;
  const uint64_t* words;
  const uint2* pairs;  // pair -> (instance, cluster within its mesh); x = ~0 where unclaimed
  uint32_t pair_count;
  uint32_t cols, n;    // entries a row, entries
  float4* rays;
  uint32_t* out;
  uint32_t* head;      // the walk's ray counter, cleared here for the launch after this one
};

// The receiver of pixel (x, y) showing visibility word `word`; false when it asks no ray.
__device__ bool receiver(const ReceiverArgs& a, uint32_t x, uint32_t y, uint64_t word, uint32_t& pair, float4& ray) {
  if (word == 0) return false;
  const uint32_t key = uint32_t(word);
  pair = key >> 8;
  if (pair >= a.pair_count) return false;
  const uint2 entry = a.pairs[pair];
  if (entry.x == GE_RT_MISS) return false;
  const ge_instance_desc in = a.sv.instances[entry.x];
  const uint32_t cluster_index = a.sv.meshes[in.mesh].first_cluster + entry.y;
  const Corners c = fetch_corners(a.sv, a.eye, in, cluster_index, key & 0xffu);
  float px = float(x) + 0.5f, py = float(y) + 0.5f;
  if (a.p.flags & GE_SHADOW_JITTER) {
    uint32_t rng = rng_seed(x, y, 0u, a.p.seed);
    px = float(x) + rng_next(rng);
    py = float(y) + rng_next(rng);
  }
  const f3 cam = f3{a.p.camera[0], a.p.camera[1], a.p.camera[2]};
  const f3 b = screen_barycentrics(c, px, py, a.p.view_proj, cam, a.p.width, a.p.height);
  const Vertex s = surface_at(a.sv, in, c, b, cam);
  const f3 from_eye = b.x * (c.p[0] - cam) + b.y * (c.p[1] - cam) + b.z * (c.p[2] - cam);
  const f3 position = cam + from_eye;
  const f3 sun = f3{a.p.sun[0], a.p.sun[1], a.p.sun[2]};
  // Facing away by both normals: the resolve asks no sun there (resolve_half_sun.slang:80-83).
  if (dot(s.normal, sun) <= 0.0f && dot(s.geometric, sun) <= 0.0f) return false;
  const float bias = a.p.shadow_bias + a.p.shadow_bias_steps * a.sv.meshes[in.mesh].quant[3] * in.scale_max;
  const f3 o = position + s.geometric * bias;
  ray = make_float4(o.x, o.y, o.z, 1e30f);
  return true;
}

template <bool Half>
__global__ void __launch_bounds__(256) k_receivers(ReceiverArgs a) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i == 0) *a.head = 0u;  // a kernel's store, not a memset: the walk after this launch starts from 0
  if (i >= a.n) return;
  const uint32_t bx = i % a.cols, by = i / a.cols;
  const uint32_t W = a.p.width, H = a.p.height;
  uint32_t pair = 0;
  float4 ray = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
  bool taken = false;
  if (!Half) {
    taken = receiver(a, bx, by, a.words[size_t(by) * W + bx], pair, ray);
  } else {
    // half_sun_block (resolve_half_sun.slang:35-96): four non-empty words of one pair, answered by
    // the top-left pixel's ray.
    const uint32_t x = bx * 2, y = by * 2;
    if (x + 1 < W && y + 1 < H) {
      const size_t row = size_t(y) * W + x;
      const uint64_t v00 = a.words[row], v10 = a.words[row + 1], v01 = a.words[row + W], v11 = a.words[row + W + 1];
      const uint32_t pr = uint32_t(v00) >> 8;
      if (v00 && v10 && v01 && v11 && (uint32_t(v10) >> 8) == pr && (uint32_t(v01) >> 8) == pr &&
          (uint32_t(v11) >> 8) == pr)
        taken = receiver(a, x, y, v00, pair, ray);
    }
  }
  if (!taken) ray.w = -1.0f;
  a.out[i] = taken ? pair << 1 : GE_SHADOW_NOT_TAKEN;
  a.rays[i] = ray;
}

template <bool Count, bool Masks>
__global__ void __launch_bounds__(64) GE_WALK_ATTR k_sun_walk(Bvh bvh, const float4* __restrict__ rays, float sx,
                                                               float sy, float sz, uint32_t n, uint32_t* head,
                                                               uint32_t* __restrict__ out,
                                                               unsigned long long* counters, uint32_t alpha) {
  Tally tally;
  uint64_t issued = 0;
  auto load = [&](uint32_t i, uint32_t& slot, Ray& ray, float& tmax, uint32_t& mask) {
    const float4 r = rays[i];
    slot = i;
    tmax = r.w;
    mask = 0x01u;
    if (!(r.w > 0.0f)) return false;
    ray = make_ray(f3{r.x, r.y, r.z}, f3{sx, sy, sz}, 0.0f);
    return true;
  };
  auto store = [&](uint32_t i, const Hit& h) {
    if (h.found) out[i] |= 1u;
  };
  persistent_walk<true, Count, Masks, Masks ? k_refill_masked : k_refill_sun, Masks ? k_batch_masked : k_batch_sun>(
      bvh, n, head, lane_stack(), load, store, issued, tally, alpha);
  if (Count) flush_counters(counters, 0, issued, tally);
}

inline uint32_t blocks(uint64_t n, uint32_t t) { return uint32_t((n + t - 1) / t); }

}  // namespace

hipError_t launch_shadow_resolve(ge_rt_ctx* ctx, const Bvh& bvh, const SceneView& sv, const ge_frame_eye& eye,
                                 const ge_shadow_resolve_params& p, const uint64_t* words, const uint2* pairs,
                                 uint32_t pair_count, float4* rays, uint32_t* out, unsigned long long* counters,
                                 uint32_t* head, uint32_t max_blocks, uint32_t alpha) {
  const hipStream_t stream = ctx->stream;
  const bool half = (p.flags & GE_SHADOW_HALF_RATE) != 0;
  ReceiverArgs a;
  a.sv = sv;
  a.eye = eye;
  a.p = p;
  a.words = words;
  a.pairs = pairs;
  a.pair_count = pair_count;
  a.cols = half ? (p.width + 1) / 2 : p.width;
  const uint64_t n = uint64_t(a.cols) * (half ? (p.height + 1) / 2 : p.height);
  if (n == 0) return hipSuccess;
  if (n >= (1ull << 31)) return hipErrorInvalidValue;
  a.n = uint32_t(n);
  a.rays = rays;
  a.out = out;
  a.head = head;
  if (half) k_receivers<true><<<blocks(n, 256), 256, 0, stream>>>(a);
  else k_receivers<false><<<blocks(n, 256), 256, 0, stream>>>(a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const uint32_t grid = std::min(blocks(n, 64), max_blocks);
  const size_t lds = size_t(k_ring) * k_lanes * 4;
  if ((e = stage_mark(ctx, 1, 0)) != hipSuccess) return e;
  const float sx = p.sun[0], sy = p.sun[1], sz = p.sun[2];
  if (alpha != k_alpha_opaque) {
    if (counters) k_sun_walk<true, true><<<grid, 64, lds, stream>>>(bvh, rays, sx, sy, sz, a.n, head, out, counters, alpha);
    else k_sun_walk<false, true><<<grid, 64, lds, stream>>>(bvh, rays, sx, sy, sz, a.n, head, out, counters, alpha);
  } else {
    if (counters) k_sun_walk<true, false><<<grid, 64, lds, stream>>>(bvh, rays, sx, sy, sz, a.n, head, out, counters, 0u);
    else k_sun_walk<false, false><<<grid, 64, lds, stream>>>(bvh, rays, sx, sy, sz, a.n, head, out, counters, 0u);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return stage_mark(ctx, 1, 1);
}

}  // namespace ge
