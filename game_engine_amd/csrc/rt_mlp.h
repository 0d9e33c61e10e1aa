// rt_mlp.h — the radiance cache's network: sizes, the encoding of a query, and the host-side state
// (FP32 master weights, Adam moments, BF16 copies as MFMA fragments, training activations).
#pragma once

#include "rt_internal.h"

namespace ge {

constexpr int kCacheFeatures = 11;   // position 3, direction 2, normal 2, albedo 3, roughness 1
constexpr int kCacheBands = 12;      // sin and cos of 2^b pi x, b = 0..11, for the 7 geometric values
constexpr int kCacheInputs = GE_CACHE_INPUTS;    // 168 encoded + 4 material + 1 constant, padded to 192
constexpr int kCacheWidth = GE_CACHE_WIDTH;      // 64
constexpr int kCacheOutPad = GE_CACHE_OUTPUTS;   // 16 output rows, the first 3 RGB
constexpr int kCacheLayers = GE_CACHE_LAYERS;    // 6 weight matrices: 5 hidden ReLU layers
constexpr uint32_t kCacheWeights = GE_CACHE_WEIGHTS;
static_assert(kCacheInputs * kCacheWidth + 4 * kCacheWidth * kCacheWidth + kCacheWidth * kCacheOutPad ==
                  GE_CACHE_WEIGHTS,
              "weight count");
static_assert(sizeof(ge_cache_query) == 48 && sizeof(ge_cache_sample) == 64, "cache record sizes");

// Layer l maps cache_dim(l) inputs to cache_dim(l + 1) outputs; its weights are [out][in] from
// cache_offset(l) in every flat layout.
__host__ __device__ constexpr uint32_t cache_dim(int l) {
  return l == 0 ? uint32_t(kCacheInputs) : (l == kCacheLayers ? uint32_t(kCacheOutPad) : uint32_t(kCacheWidth));
}
__host__ __device__ constexpr uint32_t cache_offset(int l) {
  uint32_t at = 0;
  for (int i = 0; i < l; ++i) at += cache_dim(i) * cache_dim(i + 1);
  return at;
}

// The encoded query's inputs, in the order the fused kernels compute them: input index
// idx = 32 kb + 8 g + j with group G = 6 g + kb; groups 0..20 hold sin, cos pairs of bands
// 4 (G % 3) .. + 3 of geometric value G / 3 (even j sin, odd j cos); group 21 holds albedo rgb,
// roughness, then 1 (the bias input) and zeros; groups 22 and 23 are zero. The order lets lane
// group g of a wave encode values 2 g and 2 g + 1 with compile-time band numbers.
//
// One group's four bands of x: sin and cos of 2^b pi x for b = 4 q .. 4 q + 3. The first pair is
// sin(2 pi frac(2^(4q-1) x)) from the hardware (exact argument reduction for any x), the next three
// by doubling the angle (sin 2a = 2 sin a cos a, cos 2a = 1 - 2 sin a sin a).
__device__ __forceinline__ void cache_bands4(float x, int q, float v[8]) {
  const float r = __builtin_amdgcn_fractf(x * __builtin_ldexpf(1.0f, 4 * q - 1));
  float sn = __builtin_amdgcn_sinf(r), cs = __builtin_amdgcn_cosf(r);  // arguments in revolutions
  v[0] = sn;
  v[1] = cs;
#pragma unroll
  for (int i = 1; i < 4; ++i) {
    const float twice = sn + sn;
    const float s2 = twice * cs, c2 = 1.0f - twice * sn;
    sn = s2;
    cs = c2;
    v[2 * i] = sn;
    v[2 * i + 1] = cs;
  }
}

// The cache's normalisation: position -> (p - lo) * inv.
struct CacheBox {
  float lo[3], inv[3];
};

struct RadianceCache {
  CacheBox box{{0, 0, 0}, {1, 1, 1}};
  DevBuf w, m, v, grad;  // FP32 master weights, Adam moments, gradient (kCacheWeights each)
  DevBuf frag, frag_t;   // BF16 weights as the forward pass's MFMA fragments, and W^T as the backward pass's
  DevBuf partial;        // the weight gradients' per-split partial sums
  DevBuf loss;           // double
  DevBuf x, act, dz;     // a training batch's encoded inputs, activations and deltas (bf16)
  // The tracer's records: queries and training samples, each with its float4 throughput.
  DevBuf queries, query_weights, samples, sample_weights;
  uint32_t capacity = 0;    // rows the training buffers hold
  uint32_t step = 0;        // Adam steps taken
  float hit_rate = 0.5f;    // the share of training paths that reached the cache's vertex, last frame
  uint32_t kept = 0;        // samples with targets in `samples`, from the last frame that gathered them
  double* host_loss = nullptr;  // pinned

  hipError_t init(hipStream_t st, uint32_t seed);
  hipError_t set_weights(hipStream_t st, const float* host);
  hipError_t get_weights(hipStream_t st, float* host);
  hipError_t get_gradient(hipStream_t st, float* host);  // the last step's FP32 weight gradient
  // Raw network outputs for n queries (stride floats apart, the first 11 the features): out n x 3.
  // With n_dev (device count, clamped to n) the launch reads the count from the device.
  hipError_t infer(hipStream_t st, const float* queries, uint32_t stride, uint32_t n, const uint32_t* n_dev, float* out);
  // The cached tracer's apply: result[tag] += w.xyz * max(prediction, 0) for each query (tag = the
  // path id), the queries counted on the device.
  hipError_t apply(hipStream_t st, const ge_cache_query* queries, const float4* w, uint32_t capacity,
                   const uint32_t* n_dev, float4* result);
  // One Adam step on the relative L2 loss over n samples `stride` records apart from `samples`;
  // the batch's mean loss lands in `loss` (device double).
  hipError_t train(hipStream_t st, const ge_cache_sample* samples, uint32_t stride, uint32_t n, float lr);
  void release();

 private:
  hipError_t reserve_rows(uint32_t n);
  hipError_t publish(hipStream_t st, float lr, float c1, float c2, bool update);
};

// How the wavefront tracer feeds the cache (wavefront_path_trace's `hook`).
struct CacheHook {
  enum Mode { kCached = 1, kRecord = 2 } mode;
  uint32_t depth;              // kCached: the bounce whose hit ends each path into the cache
  uint32_t train_stride;       // kCached: paths with pid % stride == phase go on to train (0: none)
  uint32_t train_phase;
  CacheBox box;
  RadianceCache* cache;        // kCached: the network the ended paths query
  ge_cache_query* queries;     // kCached: query_capacity records (at most two a path)
  float4* query_weights;
  uint32_t query_capacity;
  ge_cache_sample* samples;    // kRecord, or kCached with training: sample_capacity records
  float4* sample_weights;
  uint32_t sample_capacity;
};

// The cache queries and training samples the last wavefront_path_trace with a hook appended (all of
// them, including any past the capacities). Synchronizes the context's stream.
hipError_t read_cache_counts(ge_rt_ctx* ctx, uint32_t* queries, uint32_t* samples);

}  // namespace ge
