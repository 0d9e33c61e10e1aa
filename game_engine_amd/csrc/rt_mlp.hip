// rt_mlp.hip — the radiance cache's network on CDNA4 matrix cores (v_mfma_f32_16x16x32_bf16: BF16
// operands, FP32 accumulation).
//
// Inference is one fused kernel: a workgroup stages every weight in LDS once (58 KB, already laid
// out as MFMA fragments), and each wave pushes 32 queries at a time through all six layers with
// the activations in registers. It computes H^T = W X^T, so a layer's accumulator (lane l: neurons
// 16 t + 4 (l / 16) + i of query l % 16) becomes the next layer's B operand in the same lane once
// the next layer's weights are stored with their inputs in that order; nothing moves between lanes.
//
// Training runs layer by layer through one strided GEMM kernel (the activations of 2^16 samples are
// a few MB and stay in L2): the forward products, the deltas through ReLU's derivative, and the
// weight gradients split over the batch, then Adam on the FP32 master weights, which also refreshes
// the three BF16 copies (the GEMM's [out][in], the backward pass's [in][out], and the fragments).
#include <cmath>
#include <vector>

#include "rt_internal.h"
#include "rt_mlp.h"

namespace ge {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

// Round to nearest even (v_cvt_pk_bf16_f32, two values an instruction).
__device__ __forceinline__ __bf16 to_bf16(float f) { return static_cast<__bf16>(f); }
__device__ __forceinline__ uint16_t bf16_bits(float f) { return __builtin_bit_cast(uint16_t, to_bf16(f)); }

// Which fragment slot holds weight (row r, column c) of layer l: rows in 16-row tiles t, columns in
// 32-wide k blocks kb, lane = 16 g + r % 16, element j. Layer 0 takes its inputs in order
// (c % 32 = 8 g + j); later layers in the order the previous layer's accumulators sit in a lane
// (c % 32 = 4 g + j for j < 4, 16 + 4 g + j - 4 after).
__host__ __device__ inline uint32_t frag_index(int l, uint32_t r, uint32_t c) {
  const uint32_t t = r / 16, kb = c / 32, w = c % 32;
  uint32_t g, j;
  if (l == 0) {
    g = w / 8;
    j = w % 8;
  } else if (w < 16) {
    g = w / 4;
    j = w % 4;
  } else {
    g = (w - 16) / 4;
    j = 4 + (w - 16) % 4;
  }
  const uint32_t lane = 16 * g + r % 16;
  return cache_offset(l) + ((t * (cache_dim(l) / 32) + kb) * 64 + lane) * 8 + j;
}

// ---- fused inference -----------------------------------------------------------------------------
struct InferArgs {
  const float* queries;  // the first 11 floats of each record are the features, the 12th the tag
  uint32_t stride;       // floats between records (a multiple of 4)
  uint32_t n;
  const uint32_t* n_dev;  // optional device count (clamped to n)
  const uint4* frag;      // kCacheWeights bf16 in fragment order
  float* out;             // n x 3 (infer)
  const float4* w;        // per query: throughput (apply)
  float4* result;         // per path: radiance (apply)
};

__device__ __forceinline__ bf16x8 pack8(const float* v) {
  bf16x8 r;
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = to_bf16(v[j]);
  return r;
}

__device__ __forceinline__ uint2 pack4(const float* v) {
  return make_uint2(uint32_t(bf16_bits(v[0])) | (uint32_t(bf16_bits(v[1])) << 16),
                    uint32_t(bf16_bits(v[2])) | (uint32_t(bf16_bits(v[3])) << 16));
}
__device__ __forceinline__ bf16x8 join8(uint2 lo, uint2 hi) {
  const uint4 u = make_uint4(lo.x, lo.y, hi.x, hi.y);
  return __builtin_bit_cast(bf16x8, u);
}

typedef short __attribute__((ext_vector_type(2))) short2v;

// max(x, 0) of the two bf16 in a word, as one packed int16 max: a negative bf16 (-0 included) is a
// negative int16, a positive one a positive int16. Rounding first and clamping after gives the
// same bits as clamping in float and rounding.
__device__ __forceinline__ uint32_t relu2(uint32_t w) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(short2v, w), short2v{0, 0}));
}

// One accumulator tile, ReLU'd, as four bf16 (the hidden activations neurons 16 t + 4 g + i).
__device__ __forceinline__ uint2 relu4(const f32x4& a) {
  const float v[4] = {a[0], a[1], a[2], a[3]};
  const uint2 p = pack4(v);
  return make_uint2(relu2(p.x), relu2(p.y));
}

// ReLU, then the two B operands of the next layer from this layer's four accumulator tiles.
__device__ __forceinline__ void relu_pack(const f32x4 acc[4], bf16x8 out[2]) {
  out[0] = join8(relu4(acc[0]), relu4(acc[1]));
  out[1] = join8(relu4(acc[2]), relu4(acc[3]));
}

#ifndef GE_CACHE_WAVES
#define GE_CACHE_WAVES 16
#endif
// One workgroup a CU shares one LDS copy of the weights: 16 waves, 4 a SIMD.
#ifndef GE_CACHE_TILES
#define GE_CACHE_TILES 2
#endif
constexpr uint32_t kInferWaves = GE_CACHE_WAVES;
constexpr int kTiles = GE_CACHE_TILES;  // 16-query tiles a wave pushes through the network per pass
constexpr uint32_t kInferQueries = 16 * kTiles;

template <bool Apply>
__global__ void __launch_bounds__(kInferWaves * 64) k_cache_infer(InferArgs a) {
  __shared__ uint4 frag[kCacheWeights / 8];
  for (uint32_t i = threadIdx.x; i < kCacheWeights / 8; i += blockDim.x) frag[i] = a.frag[i];
  __syncthreads();
  const uint32_t n = a.n_dev ? min(*a.n_dev, a.n) : a.n;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, g = lane >> 4, r = lane & 15u;
  for (uint32_t base = (blockIdx.x * kInferWaves + wave) * kInferQueries; base < n;
       base += gridDim.x * kInferWaves * kInferQueries) {
    // The weights are the same every pass; an offset the compiler cannot see through keeps it
    // from hoisting all 58 KB of fragment reads into registers ahead of the loop.
    uint32_t lane_at = lane;
    asm volatile("" : "+v"(lane_at));
    const bf16x8* fr = reinterpret_cast<const bf16x8*>(frag) + lane_at;
    // ---- the two queries' features this lane encodes ----
    // Lane group g encodes geometric values 2g (k blocks 0..2) and 2g + 1 (3..5) of each query;
    // group 3's second half is the material tail (albedo, roughness, the constant 1).
    float xa[kTiles], xb[kTiles], tail[kTiles][4];
#pragma unroll
    for (int h = 0; h < kTiles; ++h) {
      const uint32_t q = base + 16 * h + r;
      float f[12];
      if (q < n) {
        const float4* rec = reinterpret_cast<const float4*>(a.queries + size_t(q) * a.stride);
        const float4 f0 = rec[0], f1 = rec[1], f2 = rec[2];
        f[0] = f0.x; f[1] = f0.y; f[2] = f0.z; f[3] = f0.w;
        f[4] = f1.x; f[5] = f1.y; f[6] = f1.z; f[7] = f1.w;
        f[8] = f2.x; f[9] = f2.y; f[10] = f2.z; f[11] = 0.0f;
      } else {
#pragma unroll
        for (int k = 0; k < 12; ++k) f[k] = 0.0f;
      }
      xa[h] = g == 0 ? f[0] : (g == 1 ? f[2] : (g == 2 ? f[4] : f[6]));
      xb[h] = g == 0 ? f[1] : (g == 1 ? f[3] : f[5]);
#pragma unroll
      for (int k = 0; k < 4; ++k) tail[h][k] = f[7 + k];
    }
    f32x4 acc[kTiles][4];
    // ---- layer 0: 192 -> 64, each k block encoded just before its products ----
#pragma unroll
    for (int h = 0; h < kTiles; ++h)
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[h][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 6; ++kb) {
      bf16x8 x[kTiles];
#pragma unroll
      for (int h = 0; h < kTiles; ++h) {
        float v[8];
        if (kb < 3 || g < 3) {
          cache_bands4(kb < 3 ? xa[h] : xb[h], kb % 3, v);
        } else if (kb == 3) {
          v[0] = tail[h][0]; v[1] = tail[h][1]; v[2] = tail[h][2]; v[3] = tail[h][3];
          v[4] = 1.0f; v[5] = 0.0f; v[6] = 0.0f; v[7] = 0.0f;
        } else {
#pragma unroll
          for (int k = 0; k < 8; ++k) v[k] = 0.0f;
        }
        x[h] = pack8(v);
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const bf16x8 wf = fr[(t * 6 + kb) * 64];
#pragma unroll
        for (int h = 0; h < kTiles; ++h)
          acc[h][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, x[h], acc[h][t], 0, 0, 0);
      }
    }
    // ---- hidden layers 1..4: 64 -> 64 ----
#pragma unroll
    for (int l = 1; l < kCacheLayers - 1; ++l) {
      bf16x8 hin[kTiles][2];
#pragma unroll
      for (int h = 0; h < kTiles; ++h) relu_pack(acc[h], hin[h]);
      const bf16x8* lw = fr + cache_offset(l) / 8;
#pragma unroll
      for (int h = 0; h < kTiles; ++h)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[h][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const bf16x8 wf = lw[(t * 2 + kb) * 64];
#pragma unroll
          for (int h = 0; h < kTiles; ++h)
            acc[h][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, hin[h][kb], acc[h][t], 0, 0, 0);
        }
      }
    }
    // ---- output layer: 64 -> 16 (rows 0..2 used) ----
    bf16x8 hin[kTiles][2];
#pragma unroll
    for (int h = 0; h < kTiles; ++h) relu_pack(acc[h], hin[h]);
    const bf16x8* lw = fr + cache_offset(kCacheLayers - 1) / 8;
    f32x4 o[kTiles];
#pragma unroll
    for (int h = 0; h < kTiles; ++h) o[h] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
      const bf16x8 wf = lw[kb * 64];
#pragma unroll
      for (int h = 0; h < kTiles; ++h) o[h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, hin[h][kb], o[h], 0, 0, 0);
    }
    // Lanes 0..15 hold rows 0..3 of query r.
    if (g == 0) {
#pragma unroll
      for (int h = 0; h < kTiles; ++h) {
        const uint32_t q = base + 16 * h + r;
        if (q >= n) continue;
        if (Apply) {
          const uint32_t pid = __float_as_uint(a.queries[size_t(q) * a.stride + 11]);
          const float4 w = a.w[q];
          float4 res = a.result[pid];
          res.x += w.x * (o[h][0] > 0.0f ? o[h][0] : 0.0f);
          res.y += w.y * (o[h][1] > 0.0f ? o[h][1] : 0.0f);
          res.z += w.z * (o[h][2] > 0.0f ? o[h][2] : 0.0f);
          a.result[pid] = res;
        } else {
          a.out[size_t(q) * 3 + 0] = o[h][0];
          a.out[size_t(q) * 3 + 1] = o[h][1];
          a.out[size_t(q) * 3 + 2] = o[h][2];
        }
      }
    }
  }
}

// ---- training --------------------------------------------------------------------------------------
// The backward pass's fragments: layer l >= 1 as W_l^T, rows = its 64 inputs in 16-row tiles t,
// k = its outputs in 32-wide blocks kb (the output layer's 16 padded with zeros to one block), in
// the order a layer's accumulators sit in a lane (the forward order of the hidden layers).
__host__ __device__ constexpr uint32_t frag_t_blocks(int l) { return (cache_dim(l + 1) + 31) / 32; }
__host__ __device__ constexpr uint32_t frag_t_offset(int l) {
  uint32_t at = 0;
  for (int i = 1; i < l; ++i) at += cache_dim(i) * 32 * frag_t_blocks(i);
  return at;
}
constexpr uint32_t kFragT = frag_t_offset(kCacheLayers);  // 18432 bf16

__host__ __device__ inline uint32_t frag_t_index(int l, uint32_t o, uint32_t c) {
  const uint32_t t = c / 16, kb = o / 32, w = o % 32;
  const uint32_t g = w < 16 ? w / 4 : (w - 16) / 4, j = w < 16 ? w % 4 : 4 + (w - 16) % 4;
  return frag_t_offset(l) + ((t * frag_t_blocks(l) + kb) * 64 + 16 * g + c % 16) * 8 + j;
}

struct TrainArgs {
  const ge_cache_sample* samples;
  uint32_t stride;  // records between samples
  uint32_t n;
  const uint4* frag;      // forward fragments (staged in LDS)
  const bf16x8* frag_t;   // backward fragments
  __bf16* x;              // n x 192: the encoded inputs
  __bf16* act;            // layer l's activations at act + l n 64 (n x 64), l = 0..4
  __bf16* dz;             // dLoss/dZ_l at dz + l n 64 (n x 64), l = 0..4; the output's at dz + 5 n 64 (n x 16)
  double* loss;
};


// The relative L2 loss's forward and backward pass, 32 samples a wave: the encoding, six products
// with the activations kept (bf16, in accumulator order) for ReLU's derivative, the loss and its
// gradient, and five products back through W^T. Writes the encoded inputs, every activation and
// every delta for the weight gradients; adds the batch's mean loss into `loss`.
constexpr uint32_t kTrainWaves = 4;
constexpr uint32_t kTrainQueries = 32;  // two 16-sample tiles a wave pass

__global__ void __launch_bounds__(kTrainWaves * 64, 2) k_cache_train_pass(TrainArgs a) {
  __shared__ uint4 frag[kCacheWeights / 8];
  for (uint32_t i = threadIdx.x; i < kCacheWeights / 8; i += blockDim.x) frag[i] = a.frag[i];
  __syncthreads();
  const uint32_t n = a.n;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, g = lane >> 4, r = lane & 15u;
  const float inv_n = 1.0f / float(n);
  double loss_sum = 0.0;
  for (uint32_t base = (blockIdx.x * kTrainWaves + wave) * kTrainQueries; base < n;
       base += gridDim.x * kTrainWaves * kTrainQueries) {
    uint32_t lane_at = lane;
    asm volatile("" : "+v"(lane_at));
    const bf16x8* fr = reinterpret_cast<const bf16x8*>(frag) + lane_at;
    const bf16x8* ft = a.frag_t + lane_at;
    uint32_t q[2];
    bool live[2];
    float xa[2], xb[2], tail[2][4], y[2][3];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      q[h] = base + 16 * h + r;
      live[h] = q[h] < n;
      float f[12];
      y[h][0] = y[h][1] = y[h][2] = 0.0f;
      if (live[h]) {
        const ge_cache_sample& smp = a.samples[size_t(q[h]) * a.stride];
        const float4* rec = reinterpret_cast<const float4*>(&smp);
        const float4 f0 = rec[0], f1 = rec[1], f2 = rec[2], f3 = rec[3];
        f[0] = f0.x; f[1] = f0.y; f[2] = f0.z; f[3] = f0.w;
        f[4] = f1.x; f[5] = f1.y; f[6] = f1.z; f[7] = f1.w;
        f[8] = f2.x; f[9] = f2.y; f[10] = f2.z; f[11] = 0.0f;
        y[h][0] = f3.x; y[h][1] = f3.y; y[h][2] = f3.z;
      } else {
#pragma unroll
        for (int k = 0; k < 12; ++k) f[k] = 0.0f;
      }
      xa[h] = g == 0 ? f[0] : (g == 1 ? f[2] : (g == 2 ? f[4] : f[6]));
      xb[h] = g == 0 ? f[1] : (g == 1 ? f[3] : f[5]);
#pragma unroll
      for (int k = 0; k < 4; ++k) tail[h][k] = f[7 + k];
    }
    // ---- forward ----
    f32x4 acc[2][4];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[h][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 6; ++kb) {
      bf16x8 xv[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        float v[8];
        if (kb < 3 || g < 3) {
          cache_bands4(kb < 3 ? xa[h] : xb[h], kb % 3, v);
        } else if (kb == 3) {
          v[0] = tail[h][0]; v[1] = tail[h][1]; v[2] = tail[h][2]; v[3] = tail[h][3];
          v[4] = 1.0f; v[5] = 0.0f; v[6] = 0.0f; v[7] = 0.0f;
        } else {
#pragma unroll
          for (int k = 0; k < 8; ++k) v[k] = 0.0f;
        }
        xv[h] = pack8(v);
        if (live[h]) *reinterpret_cast<bf16x8*>(a.x + size_t(q[h]) * kCacheInputs + 32 * kb + 8 * g) = xv[h];
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const bf16x8 wf = fr[(t * 6 + kb) * 64];
        acc[0][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, xv[0], acc[0][t], 0, 0, 0);
        acc[1][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, xv[1], acc[1][t], 0, 0, 0);
      }
    }
    uint2 act[kCacheLayers - 1][2][4];  // bf16 activations of hidden layer l, tile t: neurons 16 t + 4 g + i
#pragma unroll
    for (int l = 0; l < kCacheLayers - 1; ++l) {
      bf16x8 hin[2][2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          act[l][h][t] = relu4(acc[h][t]);
          if (live[h])
            *reinterpret_cast<uint2*>(a.act + (size_t(l) * n + q[h]) * kCacheWidth + 16 * t + 4 * g) = act[l][h][t];
        }
        hin[h][0] = join8(act[l][h][0], act[l][h][1]);
        hin[h][1] = join8(act[l][h][2], act[l][h][3]);
      }
      const int nl = l + 1;  // the next layer's product
      const bf16x8* lw = fr + cache_offset(nl) / 8;
      const int tiles = int(cache_dim(nl + 1) / 16);
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[h][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          if (t >= tiles) continue;
          const bf16x8 wf = lw[(t * 2 + kb) * 64];
          acc[0][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, hin[0][kb], acc[0][t], 0, 0, 0);
          acc[1][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, hin[1][kb], acc[1][t], 0, 0, 0);
        }
      }
    }
    // ---- the loss: acc[h][0] holds output rows 4 g + i of sample r ----
    bf16x8 dzb[2][2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      float d[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      if (g == 0 && live[h]) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float p = acc[h][0][c], yv = y[h][c];
          const float denom = p * p + 0.01f;
          d[c] = 2.0f * (p - yv) / denom * inv_n;
          loss_sum += double((p - yv) * (p - yv) / denom);
        }
      }
      const uint2 db = pack4(d);
      if (live[h]) *reinterpret_cast<uint2*>(a.dz + size_t(kCacheLayers - 1) * n * kCacheWidth + size_t(q[h]) * kCacheOutPad + 4 * g) = db;
      dzb[h][0] = join8(db, make_uint2(0u, 0u));
      dzb[h][1] = join8(make_uint2(0u, 0u), make_uint2(0u, 0u));
    }
    // ---- backward: dA_{l-1} = W_l^T dZ_l, dZ_{l-1} = dA_{l-1} where A_{l-1} > 0 ----
#pragma unroll
    for (int l = kCacheLayers - 1; l >= 1; --l) {
      const bf16x8* lt = ft + frag_t_offset(l) / 8;
      const int kbs = int(frag_t_blocks(l));
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[h][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) {
        if (kb >= kbs) continue;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const bf16x8 wf = lt[(t * kbs + kb) * 64];
          acc[0][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, dzb[0][kb], acc[0][t], 0, 0, 0);
          acc[1][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, dzb[1][kb], acc[1][t], 0, 0, 0);
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        uint2 db[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const uint2 m = act[l - 1][h][t];
          const uint32_t bits[4] = {m.x & 0xffffu, m.x >> 16, m.y & 0xffffu, m.y >> 16};
          float v[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) v[i] = bits[i] != 0u ? acc[h][t][i] : 0.0f;
          db[t] = pack4(v);
          if (live[h])
            *reinterpret_cast<uint2*>(a.dz + (size_t(l - 1) * n + q[h]) * kCacheWidth + 16 * t + 4 * g) = db[t];
        }
        dzb[h][0] = join8(db[0], db[1]);
        dzb[h][1] = join8(db[2], db[3]);
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) loss_sum += __shfl_xor(loss_sum, off);
  if (lane == 0 && loss_sum != 0.0) atomicAdd(a.loss, loss_sum * double(inv_n));
}

// ---- the weight gradients: dW_l = dZ_l^T A_{l-1}, every layer in one launch, split over samples ----
constexpr int kTile = 64;  // output tile: 64 x 64, four waves of 16 rows
constexpr int kStep = 32;  // k per MFMA step

struct GemmArgs {
  const __bf16* a;  // A(m, k) = a[m * a_m + k * a_k]
  const __bf16* b;  // B(n, k) = b[n * b_n + k * b_k]
  float* c;         // C(m, n) of split z at c[z * c_z + m * c_m + n]
  uint32_t m, n, k;
  uint32_t a_m, a_k, b_n, b_k, c_m, c_z;
  uint32_t k_split;  // k a split covers (a multiple of kStep)
};

// One 64 x 64 tile (m0, n0) of split z of C = A B^T, A and B staged in LDS 32 k at a time with the
// threads running along whichever of the operand's axes is contiguous in memory.
__device__ void gemm_tile(const GemmArgs& g, uint32_t m0, uint32_t n0, uint32_t z) {
  __shared__ __attribute__((aligned(16))) __bf16 sa[kTile][kStep + 8];
  __shared__ __attribute__((aligned(16))) __bf16 sb[kTile][kStep + 8];
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const uint32_t k_begin = z * g.k_split;
  const uint32_t k_end = min(k_begin + g.k_split, g.k);
  const bool a_rows = g.a_m == 1, b_rows = g.b_n == 1;
  f32x4 acc[4];
  for (int j = 0; j < 4; ++j) acc[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (uint32_t k0 = k_begin; k0 < k_end; k0 += kStep) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const uint32_t idx = uint32_t(e) * 256u + t;
      {
        const uint32_t rr = a_rows ? idx % kTile : idx / kStep, kk = a_rows ? idx / kTile : idx % kStep;
        const uint32_t gm = m0 + rr, gk = k0 + kk;
        sa[rr][kk] = (gm < g.m && gk < k_end) ? g.a[size_t(gm) * g.a_m + size_t(gk) * g.a_k] : __bf16(0.0f);
      }
      {
        const uint32_t rr = b_rows ? idx % kTile : idx / kStep, kk = b_rows ? idx / kTile : idx % kStep;
        const uint32_t gn = n0 + rr, gk = k0 + kk;
        sb[rr][kk] = (gn < g.n && gk < k_end) ? g.b[size_t(gn) * g.b_n + size_t(gk) * g.b_k] : __bf16(0.0f);
      }
    }
    __syncthreads();
    const bf16x8 af = *reinterpret_cast<const bf16x8*>(&sa[wave * 16 + (lane & 15u)][8 * (lane >> 4)]);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const bf16x8 bf = *reinterpret_cast<const bf16x8*>(&sb[j * 16 + (lane & 15u)][8 * (lane >> 4)]);
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bf, acc[j], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t gm = m0 + wave * 16 + 4 * (lane >> 4) + uint32_t(i);
      const uint32_t gn = n0 + uint32_t(j) * 16 + (lane & 15u);
      if (gm < g.m && gn < g.n) g.c[size_t(z) * g.c_z + size_t(gm) * g.c_m + gn] = acc[j][i];
    }
  }
}

// Tiles of every layer's weight gradient: layer 0 is 64 x 192 (three tiles), the hidden layers one
// tile each, the output layer 16 x 64 (one partial tile).
constexpr int kGradTiles = 3 + (kCacheLayers - 2) + 1;
struct GradArgs {
  GemmArgs layer[kCacheLayers];
};

__global__ void __launch_bounds__(256) k_weight_grads(GradArgs a) {
  const int tile = int(blockIdx.x);
  const int l = tile < 3 ? 0 : tile - 2;
  const uint32_t n0 = tile < 3 ? uint32_t(tile) * kTile : 0u;
  gemm_tile(a.layer[l], 0, n0, blockIdx.y);
}

__global__ void k_reduce_splits(uint32_t count, uint32_t splits, const float* __restrict__ partial,
                                float* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float s = 0.0f;
  for (uint32_t z = 0; z < splits; ++z) s += partial[size_t(z) * count + i];
  out[i] = s;
}

// Adam on the master weights (when `update`), then the BF16 copies: the forward fragments, and
// the backward ones of every layer after the first.
__global__ void k_publish(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                          const float* __restrict__ grad, __bf16* __restrict__ frag, __bf16* __restrict__ frag_t,
                          float lr, float beta1, float beta2, float eps, float c1, float c2, int update) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kCacheWeights) return;
  float wi = w[i];
  if (update) {
    const float gr = grad[i];
    const float mi = beta1 * m[i] + (1.0f - beta1) * gr;
    const float vi = beta2 * v[i] + (1.0f - beta2) * gr * gr;
    m[i] = mi;
    v[i] = vi;
    wi = wi - lr * (mi / c1) / (__builtin_sqrtf(vi / c2) + eps);
    w[i] = wi;
  }
  int l = 0;
  while (l + 1 < kCacheLayers && i >= cache_offset(l + 1)) ++l;
  const uint32_t off = cache_offset(l), cols = cache_dim(l);
  const uint32_t o = (i - off) / cols, c = (i - off) % cols;
  const __bf16 b = to_bf16(wi);
  frag[frag_index(l, o, c)] = b;
  if (l >= 1) frag_t[frag_t_index(l, o, c)] = b;
}

inline uint32_t blocks(uint64_t n, uint32_t t) { return uint32_t((n + t - 1) / t); }

// Splits of k: each a multiple of kStep, the last possibly short. Returns how many there are.
uint32_t split_k(uint32_t k, uint32_t want, uint32_t& per) {
  if (want < 1) want = 1;
  per = ((k + want - 1) / want + kStep - 1) / kStep * kStep;
  return (k + per - 1) / per;
}

constexpr uint32_t kMaxSplits = 256;

}  // namespace

// ---- host side -----------------------------------------------------------------------------------
hipError_t RadianceCache::init(hipStream_t st, uint32_t seed) {
  hipError_t e;
  for (DevBuf* b : {&w, &m, &v, &grad})
    if ((e = b->reserve(size_t(kCacheWeights) * 4)) != hipSuccess) return e;
  if ((e = frag.reserve(size_t(kCacheWeights) * 2)) != hipSuccess) return e;
  if ((e = frag_t.reserve(size_t(kFragT) * 2)) != hipSuccess) return e;
  // The output layer's backward fragments are padded with zeros that nothing writes again.
  if ((e = hipMemsetAsync(frag_t.ptr, 0, size_t(kFragT) * 2, st)) != hipSuccess) return e;
  if ((e = partial.reserve(size_t(kMaxSplits) * kCacheWeights * 4)) != hipSuccess) return e;
  if ((e = loss.reserve(8)) != hipSuccess) return e;
  if (!host_loss && (e = hipHostMalloc(reinterpret_cast<void**>(&host_loss), 8)) != hipSuccess) return e;
  // He-uniform weights from a splitmix64 sequence; the unused output rows 3.. start at zero.
  std::vector<float> h(kCacheWeights);
  uint64_t state = 0x9e3779b97f4a7c15ull * (uint64_t(seed) + 1);
  for (int l = 0; l < kCacheLayers; ++l) {
    const uint32_t rows = cache_dim(l + 1), cols = cache_dim(l), off = cache_offset(l);
    const float bound = std::sqrt(6.0f / float(cols));
    for (uint32_t i = 0; i < rows * cols; ++i) {
      state += 0x9e3779b97f4a7c15ull;
      uint64_t z = state;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      z ^= z >> 31;
      const float u = float(z >> 40) * (1.0f / 16777216.0f);
      const bool unused = l == kCacheLayers - 1 && i / cols >= 3;
      h[off + i] = unused ? 0.0f : (2.0f * u - 1.0f) * bound;
    }
  }
  return set_weights(st, h.data());
}

hipError_t RadianceCache::set_weights(hipStream_t st, const float* host) {
  const size_t bytes = size_t(kCacheWeights) * 4;
  hipError_t e = hipMemcpyAsync(w.ptr, host, bytes, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(m.ptr, 0, bytes, st);
  if (e == hipSuccess) e = hipMemsetAsync(v.ptr, 0, bytes, st);
  if (e == hipSuccess) e = hipMemsetAsync(grad.ptr, 0, bytes, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return e;
  step = 0;
  return publish(st, 0.0f, 1.0f, 1.0f, false);
}

hipError_t RadianceCache::get_weights(hipStream_t st, float* host) {
  hipError_t e = hipMemcpyAsync(host, w.ptr, size_t(kCacheWeights) * 4, hipMemcpyDeviceToHost, st);
  return e == hipSuccess ? hipStreamSynchronize(st) : e;
}

hipError_t RadianceCache::get_gradient(hipStream_t st, float* host) {
  hipError_t e = hipMemcpyAsync(host, grad.ptr, size_t(kCacheWeights) * 4, hipMemcpyDeviceToHost, st);
  return e == hipSuccess ? hipStreamSynchronize(st) : e;
}

hipError_t RadianceCache::publish(hipStream_t st, float lr, float c1, float c2, bool update) {
  k_publish<<<blocks(kCacheWeights, 256), 256, 0, st>>>(w.as<float>(), m.as<float>(), v.as<float>(), grad.as<float>(),
                                                       frag.as<__bf16>(), frag_t.as<__bf16>(), lr, 0.9f, 0.99f, 1e-8f,
                                                       c1, c2, update ? 1 : 0);
  return hipGetLastError();
}

hipError_t RadianceCache::reserve_rows(uint32_t n) {
  if (n <= capacity) return hipSuccess;
  hipError_t e;
  if ((e = x.reserve(size_t(n) * kCacheInputs * 2)) != hipSuccess) return e;
  if ((e = act.reserve(size_t(n) * (kCacheLayers - 1) * kCacheWidth * 2)) != hipSuccess) return e;
  if ((e = dz.reserve(size_t(n) * ((kCacheLayers - 1) * kCacheWidth + kCacheOutPad) * 2)) != hipSuccess) return e;
  capacity = n;
  return hipSuccess;
}

// Persistent grids: 256 CUs, one 16-wave inference workgroup a CU (its LDS copy of the weights
// serves 4 waves a SIMD), or two 4-wave training workgroups (their 224 VGPRs allow 2 waves a SIMD).
static uint32_t persistent_grid(uint32_t n, uint32_t waves, uint32_t per_wave) {
  const uint32_t per_cu = waves >= 16 ? 1u : 2u;
  return std::max(1u, std::min(blocks(n, waves * per_wave), 256u * per_cu));
}

hipError_t RadianceCache::infer(hipStream_t st, const float* queries, uint32_t stride, uint32_t n,
                                const uint32_t* n_dev, float* dst) {
  if (n == 0) return hipSuccess;
  if (stride % 4 != 0 || stride < 12) return hipErrorInvalidValue;
  InferArgs a{};
  a.queries = queries;
  a.stride = stride;
  a.n = n;
  a.n_dev = n_dev;
  a.frag = frag.as<uint4>();
  a.out = dst;
  k_cache_infer<false><<<persistent_grid(n, kInferWaves, kInferQueries), kInferWaves * 64, 0, st>>>(a);
  return hipGetLastError();
}

hipError_t RadianceCache::apply(hipStream_t st, const ge_cache_query* queries, const float4* wts, uint32_t cap,
                                const uint32_t* n_dev, float4* result) {
  if (cap == 0) return hipSuccess;
  InferArgs a{};
  a.queries = reinterpret_cast<const float*>(queries);
  a.stride = sizeof(ge_cache_query) / 4;
  a.n = cap;
  a.n_dev = n_dev;
  a.frag = frag.as<uint4>();
  a.w = wts;
  a.result = result;
  k_cache_infer<true><<<persistent_grid(cap, kInferWaves, kInferQueries), kInferWaves * 64, 0, st>>>(a);
  return hipGetLastError();
}

hipError_t RadianceCache::train(hipStream_t st, const ge_cache_sample* smp, uint32_t stride, uint32_t n, float lr) {
  if (n == 0) return hipSuccess;
  hipError_t e;
  if ((e = reserve_rows(n)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(loss.ptr, 0, 8, st)) != hipSuccess) return e;
  TrainArgs ta{};
  ta.samples = smp;
  ta.stride = stride;
  ta.n = n;
  ta.frag = frag.as<uint4>();
  ta.frag_t = frag_t.as<bf16x8>();
  ta.x = x.as<__bf16>();
  ta.act = act.as<__bf16>();
  ta.dz = dz.as<__bf16>();
  ta.loss = loss.as<double>();
  k_cache_train_pass<<<persistent_grid(n, kTrainWaves, kTrainQueries), kTrainWaves * 64, 0, st>>>(ta);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // dW_l = dZ_l^T A_{l-1} for every layer in one launch, split over the samples; then the splits summed.
  uint32_t per;
  const uint32_t splits = split_k(n, std::max(1u, std::min(kMaxSplits, n / 128u)), per);
  GradArgs ga{};
  for (int l = 0; l < kCacheLayers; ++l) {
    GemmArgs& g = ga.layer[l];
    const uint32_t dout = cache_dim(l + 1), din = cache_dim(l);
    g.a = dz.as<__bf16>() + size_t(l) * n * kCacheWidth;
    g.a_m = 1;
    g.a_k = dout;
    g.b = l == 0 ? x.as<__bf16>() : act.as<__bf16>() + size_t(l - 1) * n * kCacheWidth;
    g.b_n = 1;
    g.b_k = din;
    g.c = partial.as<float>() + cache_offset(l);
    g.c_m = din;
    g.c_z = kCacheWeights;
    g.m = dout;
    g.n = din;
    g.k = n;
    g.k_split = per;
  }
  k_weight_grads<<<dim3(kGradTiles, splits), 256, 0, st>>>(ga);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  k_reduce_splits<<<blocks(kCacheWeights, 256), 256, 0, st>>>(kCacheWeights, splits, partial.as<float>(),
                                                               grad.as<float>());
  if ((e = hipGetLastError()) != hipSuccess) return e;
  ++step;
  const float c1 = 1.0f - std::pow(0.9f, float(step)), c2 = 1.0f - std::pow(0.99f, float(step));
  return publish(st, lr, c1, c2, true);
}

void RadianceCache::release() {
  for (DevBuf* b : {&w, &m, &v, &grad, &frag, &frag_t, &partial, &loss, &x, &act, &dz, &queries, &query_weights,
                    &samples, &sample_weights})
    b->release();
  if (host_loss) (void)hipHostFree(host_loss);
  host_loss = nullptr;
  capacity = 0;
}

}  // namespace ge
