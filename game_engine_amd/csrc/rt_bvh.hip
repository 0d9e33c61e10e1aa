// rt_bvh.hip — the LBVH build on the GPU: segmented over many meshes in one set of launches.
//
// Per primitive: a box and the 30-bit Morton code of its centroid, quantized on its segment's
// centroid bounds (10 bits an axis). A stable LSD radix sort on (segment << 30 | code) orders each
// segment by code with ties in primitive order. Karras 2012 emits each segment's binary radix tree
// from the sorted codes (one thread per internal node, the index breaking equal codes). The boxes
// are fitted bottom-up tile by tile (a node's range of sorted primitives holds its own index, so the
// nodes inside a run of primitives are one workgroup's work), the nodes crossing a tile's edge by
// the next level's wider tiles, then the depths and the 4-wide collapse come back down the levels.
// Kernel boundaries and workgroup-local counters are the only synchronization: no device-scope
// atomics or fences, which on a part whose L2 is split over eight XCDs cost a cache write-back
// each. Everything is integer work or exact float min/max, so the structure is bit-identical to the
// oracle's sequential build.
//
// Memory: keys 8 B + values 4 B per primitive, double-buffered, a 256 x tiles histogram, and the
// nodes themselves (64 B each). Bytes per triangle of the whole build are stated beside the
// measured build time.
#include "rt_internal.h"

namespace ge {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

// Float <-> order-preserving uint, for min/max by integer atomics.
__device__ __forceinline__ uint32_t ordered(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float unordered(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__device__ __forceinline__ uint32_t expand10(uint32_t v) {
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__device__ __forceinline__ uint32_t quantize10(float c, float lo, float scale) {
  const float q = __builtin_fminf(__builtin_fmaxf((c - lo) * scale, 0.0f), 1023.0f);
  return uint32_t(q);
}

__global__ void k_init_bounds(uint32_t nseg, uint32_t* cbox) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  for (int a = 0; a < 3; ++a) {
    cbox[s * 6 + a] = 0xffffffffu;  // min accumulator
    cbox[s * 6 + 3 + a] = 0u;       // max accumulator
  }
}

// Centroid bounds per segment. Each wave takes a contiguous run of `steps` x 64 primitives and keeps
// per-lane running bounds while its lanes stay in one segment (segments are contiguous), so a
// segment costs one wave reduction and six atomics per run instead of per primitive: same-address
// atomics serialize, and one segment can hold tens of millions of primitives. `steps` grows with
// the batch (bounds_steps) so that small batches still spread over many waves. With Partials the
// items are a batch's clusters, each carrying the lowest and highest ordered centroids of its
// triangles (reduced by the records pass): the same bounds from a 1/128th of the reads.
constexpr uint32_t kBoundsMaxSteps = 64;

template <bool Partials>
__global__ void __launch_bounds__(256) k_centroid_bounds(uint32_t n, uint32_t steps, const float* __restrict__ boxes,
                                                         const uint32_t* __restrict__ part,
                                                         const uint32_t* __restrict__ seg_of, uint32_t* cbox) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t base = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 64u * steps;
  uint32_t cur = 0xffffffffu;  // the segment the running bounds belong to (the same in every lane)
  uint32_t lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
  auto flush = [&]() {
    for (int a = 0; a < 3; ++a) {
      for (int off = 32; off > 0; off >>= 1) {
        lo[a] = min(lo[a], uint32_t(__shfl_xor(int(lo[a]), off)));
        hi[a] = max(hi[a], uint32_t(__shfl_xor(int(hi[a]), off)));
      }
    }
    if (lane == 0 && cur != 0xffffffffu) {
      for (int a = 0; a < 3; ++a) {
        atomicMin(&cbox[cur * 6 + a], lo[a]);
        atomicMax(&cbox[cur * 6 + 3 + a], hi[a]);
      }
    }
    for (int a = 0; a < 3; ++a) {
      lo[a] = 0xffffffffu;
      hi[a] = 0u;
    }
  };
  for (uint32_t k = 0; k < steps; ++k) {
    const uint32_t i = base + k * 64u + lane;
    const bool valid = i < n;
    if (!__any(valid)) break;
    const uint32_t s = valid ? seg_of[i] : 0xffffffffu;
    // An item's (lowest, highest) ordered centroid: a primitive's own, or a cluster's partial bounds.
    uint32_t clo[3], chi[3];
    for (int a = 0; a < 3; ++a) {
      if (Partials) {
        clo[a] = valid ? part[size_t(i) * 6 + a] : 0xffffffffu;
        chi[a] = valid ? part[size_t(i) * 6 + 3 + a] : 0u;
      } else {
        clo[a] = chi[a] = valid ? ordered((boxes[size_t(i) * 6 + a] + boxes[size_t(i) * 6 + 3 + a]) * 0.5f) : 0u;
      }
    }
    const uint32_t s0 = uint32_t(__shfl(int(s), 0));
    if (__all(!valid || s == s0)) {
      if (s0 != cur) {
        flush();
        cur = s0;
      }
      if (valid) {
        for (int a = 0; a < 3; ++a) {
          lo[a] = min(lo[a], clo[a]);
          hi[a] = max(hi[a], chi[a]);
        }
      }
    } else {
      // A segment boundary inside this step: close the run and add these lanes one by one.
      flush();
      cur = 0xffffffffu;
      if (valid) {
        for (int a = 0; a < 3; ++a) {
          atomicMin(&cbox[s * 6 + a], clo[a]);
          atomicMax(&cbox[s * 6 + 3 + a], chi[a]);
        }
      }
    }
  }
  flush();
}

__global__ void k_segment_scale(uint32_t nseg, const uint32_t* cbox, float* lo_scale) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  for (int a = 0; a < 3; ++a) {
    const float lo = unordered(cbox[s * 6 + a]);
    const float hi = unordered(cbox[s * 6 + 3 + a]);
    const float ext = hi - lo;
    lo_scale[s * 6 + a] = lo;
    lo_scale[s * 6 + 3 + a] = ext > 0.0f ? 1024.0f / ext : 0.0f;
  }
}

__global__ void k_morton(uint32_t n, const float* __restrict__ boxes, const uint32_t* __restrict__ seg_of,
                         const float* __restrict__ lo_scale, uint64_t* keys, uint32_t* vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = seg_of[i];
  uint32_t q[3];
  for (int a = 0; a < 3; ++a) {
    const float c = (boxes[size_t(i) * 6 + a] + boxes[size_t(i) * 6 + 3 + a]) * 0.5f;
    q[a] = quantize10(c, lo_scale[s * 6 + a], lo_scale[s * 6 + 3 + a]);
  }
  const uint32_t code = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
  keys[i] = (uint64_t(s) << 30) | code;
  vals[i] = i;
}

// ---- one segment of up to kSmallMax primitives: bounds, codes and the sort in one workgroup ----
// The instance structure of a scene (tens of thousands of instances) and mid-sized meshes take
// this path instead of the device-wide passes, whose ~25 launches (bounds, codes, then a
// histogram, a scan and a scatter a digit) cost more than their work at this size. The keys
// (30-bit codes) live in LDS between the four 8-bit LSD passes. A pass: each wave counts its
// keys' digits (LDS atomics), the block scans the (digit, wave) counts in digit-major order into
// places, and each wave scatters its keys a slot at a time, a key's place being its (digit, wave)
// place plus the keys of that digit in the wave's earlier slots and lower lanes. The lanes that
// share a digit (the peers) come from a 64-bit mask a digit in LDS that each lane ORs its bit
// into: a few LDS operations a key, where ballots over the digit's 8 bits cost ~50 vector
// instructions on the four SIMDs one workgroup has. The element order inside a wave is
// slot-major (element s * 64 + lane of the wave's run), so places follow positions and every pass
// is stable: the result is the device-wide sort's, key for key and index for index.
constexpr uint32_t kSmallThreads = 1024;
constexpr uint32_t kSmallItems = 17;  // keys a lane holds
constexpr uint32_t kSmallWaves = kSmallThreads / 64;
constexpr uint32_t kSmallMax = kSmallThreads * kSmallItems;
static_assert(kSmallMax <= 0x10000u, "16-bit indices in LDS");

__global__ void __launch_bounds__(kSmallThreads) k_sort_small(uint32_t n, const float* __restrict__ boxes,
                                                               uint64_t* __restrict__ keys_out,
                                                               uint32_t* __restrict__ vals_out,
                                                               uint32_t* __restrict__ perm_out,
                                                               uint32_t* __restrict__ codes_out) {
  __shared__ uint32_t lkey[kSmallMax];
  __shared__ uint16_t lval[kSmallMax];
  __shared__ uint32_t cnt[kSmallWaves][256];            // a wave's count of each digit, then its place
  __shared__ unsigned long long peer[kSmallWaves][256];  // the lanes holding each digit in this slot
  __shared__ uint32_t wsum[kSmallWaves];
  __shared__ uint32_t red[6];
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  // Slots in use: a wave's run is 64 x slots elements (padded past n), so a smaller segment takes
  // fewer passes of each loop below (a uniform branch a slot).
  const uint32_t slots = (n + kSmallThreads - 1) / kSmallThreads;
  const uint32_t padded = slots * kSmallThreads;
  const uint32_t run = wave * 64u * slots;  // the wave's first element
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const unsigned long long me = 1ull << lane;

  // Centroid bounds (the same order-preserving min and max as the device-wide passes).
  if (t < 6) red[t] = t < 3 ? 0xffffffffu : 0u;
  for (uint32_t d = lane; d < 256; d += 64) peer[wave][d] = 0;
  __syncthreads();
  {
    uint32_t lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
    for (uint32_t i = t; i < n; i += kSmallThreads)
      for (int a = 0; a < 3; ++a) {
        const uint32_t c = ordered((boxes[size_t(i) * 6 + a] + boxes[size_t(i) * 6 + 3 + a]) * 0.5f);
        lo[a] = min(lo[a], c);
        hi[a] = max(hi[a], c);
      }
    for (int a = 0; a < 3; ++a)
      for (int off = 32; off > 0; off >>= 1) {
        lo[a] = min(lo[a], uint32_t(__shfl_xor(lo[a], off)));
        hi[a] = max(hi[a], uint32_t(__shfl_xor(hi[a], off)));
      }
    if (lane == 0)
      for (int a = 0; a < 3; ++a) {
        atomicMin(&red[a], lo[a]);
        atomicMax(&red[3 + a], hi[a]);
      }
  }
  __syncthreads();
  float clo[3], cscale[3];
  for (int a = 0; a < 3; ++a) {
    clo[a] = unordered(red[a]);
    const float ext = unordered(red[3 + a]) - clo[a];
    cscale[a] = ext > 0.0f ? 1024.0f / ext : 0.0f;
  }
  // Codes, through LDS into the lanes' slots; past n, ~0 (above every 30-bit code: sorts last).
  for (uint32_t i = t; i < padded; i += kSmallThreads) {
    uint32_t code = 0xffffffffu;
    if (i < n) {
      uint32_t q[3];
      for (int a = 0; a < 3; ++a) {
        const float c = (boxes[size_t(i) * 6 + a] + boxes[size_t(i) * 6 + 3 + a]) * 0.5f;
        q[a] = quantize10(c, clo[a], cscale[a]);
      }
      code = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
    }
    lkey[i] = code;
  }
  __syncthreads();
  uint32_t key[kSmallItems], val[kSmallItems];
#pragma unroll
  for (uint32_t s = 0; s < kSmallItems; ++s) {
    val[s] = run + s * 64u + lane;
    key[s] = s < slots ? lkey[val[s]] : 0u;
  }
  __syncthreads();  // lkey is rewritten by the first scatter
#pragma unroll  // four passes, unrolled: the arrays are not carried through a loop as register tuples
  for (uint32_t shift = 0; shift < 32; shift += 8) {
    for (uint32_t d = lane; d < 256; d += 64) cnt[wave][d] = 0;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (uint32_t s = 0; s < kSmallItems; ++s)
      if (s < slots) atomicAdd(&cnt[wave][(key[s] >> shift) & 255u], 1u);
    __syncthreads();
    // Places: the exclusive scan of the counts in (digit, wave) order; thread t holds digit t / 4,
    // waves 4 (t % 4) to 4 (t % 4) + 3.
    {
      const uint32_t d = t >> 2, w0 = (t & 3u) * 4u;
      uint32_t c[4], sum = 0;
      for (uint32_t j = 0; j < 4; ++j) {
        c[j] = cnt[w0 + j][d];
        sum += c[j];
      }
      uint32_t incl = sum;
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = uint32_t(__shfl_up(incl, off));
        if (lane >= uint32_t(off)) incl += y;
      }
      if (lane == 63) wsum[wave] = incl;
      __syncthreads();
      uint32_t base = 0;
      for (uint32_t w = 0; w < wave; ++w) base += wsum[w];
      uint32_t at = base + incl - sum;
      for (uint32_t j = 0; j < 4; ++j) {
        cnt[w0 + j][d] = at;
        at += c[j];
      }
    }
    __syncthreads();
    // The scatter, a slot at a time. LDS operations of one wave complete in order: every lane's OR
    // lands before any lane reads the mask back, and the reads before the mask is cleared.
#pragma unroll
    for (uint32_t s = 0; s < kSmallItems; ++s) {
      if (s >= slots) break;
      const uint32_t k = key[s];
      const uint32_t d = (k >> shift) & 255u;
      atomicOr(&peer[wave][d], me);
      __builtin_amdgcn_wave_barrier();
      const uint64_t peers = peer[wave][d];
      const uint32_t at = cnt[wave][d];
      __builtin_amdgcn_wave_barrier();
      peer[wave][d] = 0;
      const uint32_t p = at + uint32_t(__popcll(peers & lt));
      lkey[p] = k;
      lval[p] = uint16_t(val[s]);
      if ((peers & lt) == 0) cnt[wave][d] = at + uint32_t(__popcll(peers));
      __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    if (shift + 8 < 32) {
#pragma unroll
      for (uint32_t s = 0; s < kSmallItems; ++s) {
        if (s >= slots) break;
        key[s] = lkey[run + s * 64u + lane];
        val[s] = lval[run + s * 64u + lane];
      }
      __syncthreads();  // cnt is cleared and lkey rewritten by the next pass
    }
  }
  for (uint32_t i = t; i < n; i += kSmallThreads) {
    const uint32_t k = lkey[i], v = lval[i];
    keys_out[i] = k;
    vals_out[i] = v;
    perm_out[i] = v;
    codes_out[i] = k & 0x3fffffffu;
  }
}

// ---- stable LSD radix sort, 8-bit digits -------------------------------------------------------
__global__ void __launch_bounds__(kSortThreads) k_sort_hist(const uint64_t* __restrict__ keys, uint32_t n, int shift,
                                                            uint32_t* hist, uint32_t tiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kSortTile;
  for (int k = 0; k < kSortItems; ++k) {
    const uint32_t i = base + uint32_t(k) * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[size_t(threadIdx.x) * tiles + blockIdx.x] = h[threadIdx.x];
}

// A tile's keys in place, stable: wave w holds the tile's elements [1024 w, 1024 (w + 1)), slot s
// of a lane element 1024 w + 64 s + lane, so ranks follow positions. Each wave counts its digits
// (LDS atomics); a digit's place for a wave within the tile is the tile's earlier digits plus the
// earlier waves' counts of it. The keys go to their places in the tile in LDS, a slot at a time (a
// key's place is its wave's place for the digit plus the keys of that digit in the wave's earlier
// slots and lower lanes, the lanes sharing a digit from a 64-bit LDS mask a digit that each lane
// ORs its bit into), and then out in tile order, each to the tile's place for its digit (the scan
// of the histograms) plus its rank in the digit's run: neighbouring lanes write neighbouring words.
template <bool Keys>
__global__ void __launch_bounds__(kSortThreads) k_sort_scatter(const uint64_t* __restrict__ kin,
                                                               const uint32_t* __restrict__ vin, uint64_t* kout,
                                                               uint32_t* vout, uint32_t n, int shift,
                                                               const uint32_t* __restrict__ offsets, uint32_t tiles) {
  constexpr uint32_t kWaves = kSortThreads / 64;
  static_assert(kSortItems * 64 * kWaves == kSortTile, "a wave holds a contiguous quarter of the tile");
  static_assert(kSortThreads == 256, "a thread a digit");
  __shared__ uint32_t cnt[kWaves][256];             // a wave's count of each digit, then its place in the tile
  __shared__ unsigned long long peer[kWaves][256];  // the lanes holding each digit in this slot
  __shared__ uint32_t start[256];                   // a digit's first place in the tile
  __shared__ uint32_t dest[256];                    // a digit's first place in the output
  __shared__ uint32_t wsum[kWaves];
  __shared__ uint64_t lkey[kSortTile];
  __shared__ uint32_t lval[kSortTile];
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const unsigned long long me = 1ull << lane;
  for (uint32_t w = 0; w < kWaves; ++w) {
    cnt[w][t] = 0;
    peer[w][t] = 0;
  }
  const uint32_t base = blockIdx.x * kSortTile;
  const uint32_t count = min(uint32_t(kSortTile), n - base);  // the tile's keys
  const uint32_t first = wave * 64u * kSortItems;             // the wave's first element in the tile
  uint64_t key[kSortItems];
  uint32_t val[kSortItems];
#pragma unroll
  for (uint32_t s = 0; s < kSortItems; ++s) {
    const uint32_t i = first + s * 64u + lane;
    key[s] = i < count ? kin[base + i] : 0ull;
    val[s] = i < count ? vin[base + i] : 0u;
  }
  __syncthreads();
#pragma unroll
  for (uint32_t s = 0; s < kSortItems; ++s)
    if (first + s * 64u + lane < count) atomicAdd(&cnt[wave][uint32_t(key[s] >> shift) & 255u], 1u);
  __syncthreads();
  {
    // Digit t: its count in the tile, the exclusive scan over digits (a wave scan, then the waves'
    // sums), and each wave's place for it.
    uint32_t c[kWaves], sum = 0;
    for (uint32_t w = 0; w < kWaves; ++w) {
      c[w] = cnt[w][t];
      sum += c[w];
    }
    uint32_t incl = sum;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = uint32_t(__shfl_up(incl, off));
      if (lane >= uint32_t(off)) incl += y;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t at = incl - sum;
    for (uint32_t w = 0; w < wave; ++w) at += wsum[w];
    start[t] = at;
    dest[t] = offsets[size_t(t) * tiles + blockIdx.x];
    for (uint32_t w = 0; w < kWaves; ++w) {
      cnt[w][t] = at;
      at += c[w];
    }
  }
  __syncthreads();
#pragma unroll
  for (uint32_t s = 0; s < kSortItems; ++s) {
    if (first + s * 64u >= count) break;  // the wave's remaining slots are past the keys
    const bool valid = first + s * 64u + lane < count;
    uint64_t k = key[s];
    asm volatile("" : "+v"(k));  // recomputed, not kept from the count above: registers
    const uint32_t d = uint32_t(k >> shift) & 255u;
    if (valid) atomicOr(&peer[wave][d], me);
    __builtin_amdgcn_wave_barrier();
    const uint64_t peers = valid ? peer[wave][d] : 0ull;
    const uint32_t at = valid ? cnt[wave][d] : 0u;
    __builtin_amdgcn_wave_barrier();
    if (valid) {
      peer[wave][d] = 0;
      const uint32_t p = at + uint32_t(__popcll(peers & lt));
      lkey[p] = k;
      lval[p] = val[s];
      if ((peers & lt) == 0) cnt[wave][d] = at + uint32_t(__popcll(peers));
    }
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  for (uint32_t j = t; j < count; j += kSortThreads) {
    const uint64_t k = lkey[j];
    const uint32_t d = uint32_t(k >> shift) & 255u;
    const uint32_t p = dest[d] + (j - start[d]);
    if (Keys) kout[p] = k;
    vout[p] = lval[j];
  }
}

// Exclusive scan of u32, in tiles of kScanTile with a recursive pass over the tile sums.
__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(uint32_t* data, uint32_t n, uint32_t* sums) {
  __shared__ uint32_t s[kScanThreads];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t local = 0;
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < n) ? data[base + k] : 0u;
    local += v[k];
  }
  s[threadIdx.x] = local;
  __syncthreads();
  for (uint32_t off = 1; off < kScanThreads; off <<= 1) {
    const uint32_t add = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += add;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - local;
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) data[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == kScanThreads - 1 && sums) sums[blockIdx.x] = s[threadIdx.x];
}

__global__ void k_scan_add(uint32_t* data, uint32_t n, const uint32_t* sums) {
  const uint32_t i = blockIdx.x * kScanThreads * kScanItems + threadIdx.x;
  const uint32_t add = sums[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k) {
    const uint32_t j = i + uint32_t(k) * kScanThreads;
    if (j < n) data[j] += add;
  }
}

hipError_t exclusive_scan(uint32_t* data, uint32_t n, DevBuf& tmp, size_t tmp_offset, hipStream_t stream) {
  const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles <= 1) {
    k_scan_tiles<<<1, kScanThreads, 0, stream>>>(data, n, nullptr);
    return hipGetLastError();
  }
  uint32_t* sums = reinterpret_cast<uint32_t*>(static_cast<char*>(tmp.ptr) + tmp_offset);
  k_scan_tiles<<<tiles, kScanThreads, 0, stream>>>(data, n, sums);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = exclusive_scan(sums, tiles, tmp, tmp_offset + ((size_t(tiles) * 4 + 255) & ~size_t(255)), stream);
  if (e != hipSuccess) return e;
  k_scan_add<<<tiles, kScanThreads, 0, stream>>>(data, n, sums);
  return hipGetLastError();
}

// ---- the fit and the collapse, tile by tile ------------------------------------------------
// Karras' internal node i covers a contiguous range [lo, hi] of sorted primitives with i = lo or
// i = hi, so a node whose range lies inside a run of primitives has its index inside that run too.
// A node's level is the first in a ladder of tile widths (2^10 primitives, then 64 times wider at
// each level) at which one tile holds its whole range; Karras' kernel, which knows the range,
// drops every node above level 1 (a few in a hundred) into its tile's list. Then:
// - the levels above 1, top first, give their nodes their depths (a walk up the tile to its top
//   node, plus the depth of that node's parent, which the level above wrote);
// - one launch takes every level-1 tile (99% of the nodes) in one workgroup from LDS: the leaf
//   boxes and child links loaded once, the nodes fitted bottom-up (a node whose children are leaves
//   starts; the last child to finish a node moves on to it, by LDS counters), the tile's top nodes'
//   boxes handed up, each node's depth, and the collapse of the even-depth ones into 4-wide nodes;
// - the levels above 1, bottom first, fit their nodes the same way from the boxes below, and one
//   launch collapses their even-depth nodes.
// The boxes are exact min/max, so the order of the unions does not change them, and every node is
// fitted once. No pointer jumping, no depth sort, no level-per-launch fit and no host round trip.
constexpr uint32_t kTileShift = 10;     // a level-1 tile: 1,024 primitives
constexpr uint32_t kTileMax = 1u << kTileShift;
constexpr uint32_t kTileThreads = 512;
constexpr uint32_t kFanShift = 6;       // 64 tiles of a level make one of the next
constexpr uint32_t kSparseCap = 4096;   // nodes a tile above level 1 holds
constexpr uint32_t kSparseThreads = 512;
constexpr uint32_t kFitMaxDepth = 254;  // deeper structures are refused (the flag)
constexpr uint32_t kMaxLevels = 6;
constexpr uint16_t kNoSlot = 0xffffu;
constexpr uint16_t kLeafRef = 0x8000u;

struct TileArgs {
  uint32_t nseg, levels;
  uint32_t bucket_base[kMaxLevels + 1];  // level l's lists start at bucket_base[l] (l >= 1)
  const uint32_t* seg_first;
  const uint32_t* seg_count;
  const uint32_t* node_base;
  const uint32_t* root_index;
  const float* boxes;
  const uint32_t* perm;
  ge_bvh_node* nodes;
  const uint4* links;  // per batch entry first + i: internal node i's child[0], child[1], -, primitive count
  const uint32_t* node_parent;
  float* node_box;  // 8 floats a node above level 1 or at a level-1 tile's top: box, height bits
  float* roots;
  uint32_t* heights;
  uint32_t* list;   // kSparseCap entries a tile above level 1
  uint32_t* count;
  uint32_t* depth;  // per batch entry: the depth of the node of that index above level 1
  WideNode* wide;
  uint32_t* flag;   // set on a tile over capacity or a structure deeper than kFitMaxDepth
};

__device__ __forceinline__ uint32_t level_shift(uint32_t l) { return kTileShift + kFanShift * l; }
__device__ __forceinline__ uint32_t tiles_of(uint32_t n, uint32_t shift) {
  return uint32_t((uint64_t(n) + (1ull << shift) - 1) >> shift);
}
// Tile t of a level as (segment, tile within it), and back.
__device__ void tile_place(const TileArgs& a, uint32_t shift, uint32_t t, uint32_t& s, uint32_t& k) {
  for (s = 0; s + 1 < a.nseg; ++s) {
    const uint32_t m = tiles_of(a.seg_count[s], shift);
    if (t < m) break;
    t -= m;
  }
  k = t;
}
__device__ uint32_t tile_index(const TileArgs& a, uint32_t shift, uint32_t s, uint32_t k) {
  uint32_t base = 0;
  for (uint32_t q = 0; q < s; ++q) base += tiles_of(a.seg_count[q], shift);
  return base + k;
}
// The level of a node of range [lo, hi]; a node above level 1 goes into its tile's list.
__device__ void file_node(const TileArgs& a, uint32_t s, uint32_t i, uint32_t lo, uint32_t hi) {
  uint32_t l = 0;
  while (l + 1 < a.levels && (lo >> level_shift(l)) != (hi >> level_shift(l))) ++l;
  if (l == 0) return;
  const uint32_t b = a.bucket_base[l] + tile_index(a, level_shift(l), s, i >> level_shift(l));
  const uint32_t pos = atomicAdd(&a.count[b], 1u);
  if (pos < kSparseCap) a.list[size_t(b) * kSparseCap + pos] = i;
  else atomicOr(a.flag, 1u);
}

// Stores of whole records as 16-byte vectors (records and wide nodes are 16-byte aligned).
__device__ __forceinline__ void store_record(ge_bvh_node* out, const float* lo0, const float* hi0, const float* lo1,
                                             const float* hi1, uint32_t c0, uint32_t c1, uint32_t parent,
                                             uint32_t count) {
  uint4* o = reinterpret_cast<uint4*>(out);
  o[0] = make_uint4(__float_as_uint(lo0[0]), __float_as_uint(lo0[1]), __float_as_uint(lo0[2]), __float_as_uint(hi0[0]));
  o[1] = make_uint4(__float_as_uint(hi0[1]), __float_as_uint(hi0[2]), __float_as_uint(lo1[0]), __float_as_uint(lo1[1]));
  o[2] = make_uint4(__float_as_uint(lo1[2]), __float_as_uint(hi1[0]), __float_as_uint(hi1[1]), __float_as_uint(hi1[2]));
  o[3] = make_uint4(c0, c1, parent, count);
}

// Fits internal node i of segment s from global memory: its record's child boxes and parent, its
// box and height (node_box), and the root's box and height.
__device__ void fit_node(const TileArgs& a, uint32_t s, uint32_t i) {
  const uint32_t first = a.seg_first[s], nb = a.node_base[s];
  ge_bvh_node* nd = a.nodes + nb + i;
  const uint4 link = a.links[first + i];
  const uint32_t child[2] = {link.x, link.y};
  float b[2][6];
  uint32_t h[2];
  for (int c = 0; c < 2; ++c) {
    const uint32_t ch = child[c];
    if (ch & GE_BVH_LEAF) {
      const uint32_t p = a.perm[first + (ch & ~GE_BVH_LEAF)];
      for (int q = 0; q < 6; ++q) b[c][q] = a.boxes[size_t(p) * 6 + q];
      h[c] = 0;
    } else {
      const float* cb = a.node_box + size_t(nb + ch) * 8;
      for (int q = 0; q < 6; ++q) b[c][q] = cb[q];
      h[c] = __float_as_uint(cb[6]);
    }
  }
  store_record(nd, b[0], b[0] + 3, b[1], b[1] + 3, child[0], child[1], i == 0 ? GE_BVH_NONE : a.node_parent[nb + i],
               link.w);
  float u[6];
  for (int q = 0; q < 3; ++q) {
    u[q] = __builtin_fminf(b[0][q], b[1][q]);
    u[3 + q] = __builtin_fmaxf(b[0][3 + q], b[1][3 + q]);
  }
  const uint32_t ht = 1 + (h[0] > h[1] ? h[0] : h[1]);
  float* o = a.node_box + size_t(nb + i) * 8;
  for (int q = 0; q < 6; ++q) o[q] = u[q];
  o[6] = __uint_as_float(ht);
  if (i == 0) {
    const uint32_t out = a.root_index ? a.root_index[s] : s;
    for (int q = 0; q < 6; ++q) a.roots[out * 6 + q] = u[q];
    a.heights[out] = ht;
  }
}

// The unbiased exponent of a finite x >= 0 (-127 for zero and subnormals).
__device__ __forceinline__ int exponent_of(float x) { return int((__float_as_uint(x) >> 23) & 0xffu) - 127; }
// 2^e for -126 <= e <= 127, exactly.
__device__ __forceinline__ float pow2(int e) { return __uint_as_float(uint32_t(e + 127) << 23); }

// The grid of one axis of a wide node whose children's low planes reach down to L and high planes
// up to H: the smallest exponent e >= -100 at which L / 2^e and H / 2^e are below 2^23 in magnitude
// (so the floors and ceilings below are exact in float) and ceil(H / 2^e) - floor(L / 2^e) <= 253
// (so every index fits a byte with the extra step on each side); k0 = floor(L / 2^e). Every
// operation is exact, so the grid is a function of the box planes alone. (Planes beyond 2^100 or
// not finite end the search at e = 126 with indices that mean nothing; boxes of finite geometry
// never get there.)
__device__ void axis_grid(float L, float H, int& e, float& k0) {
  const float m = __builtin_fmaxf(__builtin_fabsf(L), __builtin_fabsf(H));
  int g = exponent_of(m) - 22;
  const int span = exponent_of(H - L) - 7;
  g = g > span ? g : span;
  g = g > -100 ? g : -100;
  float a = 0.0f;
  for (; g < 126; ++g) {
    const float down = pow2(-g);
    a = __builtin_floorf(L * down);
    const float b = __builtin_ceilf(H * down);
    if (b - a <= 253.0f && __builtin_fabsf(a) < 8388608.0f && __builtin_fabsf(b) < 8388608.0f) break;
  }
  e = g;
  k0 = a;
}

// A plane x, in grid steps, rounded outward: down (up) to the grid, and one step further when the
// rounding left less than 2^-10 of a step (x - floor(x) and ceil(x) - x are exact), so that a
// quantized box holds its child's box with at least 2^-10 of a step to spare on every side.
__device__ __forceinline__ float grid_down(float x) {
  const float f = __builtin_floorf(x);
  return x - f < 0.0009765625f ? f - 1.0f : f;
}
__device__ __forceinline__ float grid_up(float x) {
  const float c = __builtin_ceilf(x);
  return c - x < 0.0009765625f ? c + 1.0f : c;
}

// A wide node from four (reference, box) entries, of which the first `valid` are real; the rest
// are empty (GE_BVH_NONE, no occupied bit, zero planes); a leaf reference sets its slot's leaf bit
// (bit 28 + slot of meta, beside the occupied bit at 24 + slot). Per axis, the grid of axis_grid with its
// origin at k0 - 1 steps, and each child's planes at grid_down(lo / 2^e) - (k0 - 1) and
// grid_up(hi / 2^e) - (k0 - 1), within [0, 255].
__device__ __forceinline__ void store_wide(WideNode* out, const uint32_t ref[4], const float* box[4], int valid) {
  uint32_t meta = ((1u << valid) - 1u) << 24, lo[3] = {0u, 0u, 0u}, hi[3] = {0u, 0u, 0u};
  float origin[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float L = box[0][a], H = box[0][3 + a];
    for (int q = 1; q < valid; ++q) {
      L = __builtin_fminf(L, box[q][a]);
      H = __builtin_fmaxf(H, box[q][3 + a]);
    }
    int e;
    float k0;
    axis_grid(L, H, e, k0);
    const float down = pow2(-e), base = k0 - 1.0f;
    origin[a] = base * pow2(e);
    meta |= uint32_t(e + 127) << (8 * a);
    for (int q = 0; q < valid; ++q) {
      lo[a] |= uint32_t(grid_down(box[q][a] * down) - base) << (8 * q);
      hi[a] |= uint32_t(grid_up(box[q][3 + a] * down) - base) << (8 * q);
    }
  }
  uint32_t c[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    c[q] = q < valid ? ref[q] : GE_BVH_NONE;
    if (q < valid && (ref[q] & GE_BVH_LEAF)) meta |= 1u << (28 + q);
  }
  uint4* o = reinterpret_cast<uint4*>(out);
  o[0] = make_uint4(__float_as_uint(origin[0]), __float_as_uint(origin[1]), __float_as_uint(origin[2]), meta);
  o[1] = make_uint4(c[0], c[1], c[2], c[3]);
  o[2] = make_uint4(lo[0], lo[1], lo[2], hi[0]);
  o[3] = make_uint4(hi[1], hi[2], 0u, 0u);
}

// The 4-wide node of binary node i from the records in global memory: its children's children (a
// leaf child taken as it is), in child order, with their boxes, at the node's own binary index.
__device__ void collapse_node(const TileArgs& a, uint32_t nb, uint32_t i) {
  const ge_bvh_node nd = a.nodes[nb + i];
  const bool lx = nd.child[0] & GE_BVH_LEAF, ly = nd.child[1] & GE_BVH_LEAF;
  const ge_bvh_node mx = a.nodes[nb + (lx ? i : nd.child[0])];
  const ge_bvh_node my = a.nodes[nb + (ly ? i : nd.child[1])];
  const uint32_t ref[4] = {lx ? nd.child[0] : mx.child[0], lx ? (ly ? nd.child[1] : my.child[0]) : mx.child[1],
                           lx ? my.child[1] : (ly ? nd.child[1] : my.child[0]), my.child[1]};
  const float* lo[4] = {lx ? nd.lo0 : mx.lo0, lx ? (ly ? nd.lo1 : my.lo0) : mx.lo1, lx ? my.lo1 : (ly ? nd.lo1 : my.lo0),
                        my.lo1};
  const float* hi[4] = {lx ? nd.hi0 : mx.hi0, lx ? (ly ? nd.hi1 : my.hi0) : mx.hi1, lx ? my.hi1 : (ly ? nd.hi1 : my.hi0),
                        my.hi1};
  float boxes[4][6];
#pragma unroll
  for (int e = 0; e < 4; ++e)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      boxes[e][q] = lo[e][q];
      boxes[e][3 + q] = hi[e][q];
    }
  const float* box[4] = {boxes[0], boxes[1], boxes[2], boxes[3]};
  store_wide(a.wide + nb + i, ref, box, (lx ? 1 : 2) + (ly ? 1 : 2));
}

// ---- Karras 2012 --------------------------------------------------------------------------------
__device__ __forceinline__ int delta(const uint64_t* __restrict__ keys, int n, int i, int j) {
  if (j < 0 || j > n - 1) return -1;
  const uint32_t a = uint32_t(keys[i]) & 0x3fffffffu, b = uint32_t(keys[j]) & 0x3fffffffu;
  if (a != b) return __clz(a ^ b);
  return 32 + __clz(uint32_t(i) ^ uint32_t(j));
}

// Karras' emission; with more than one tile level, each node above level 1 is also filed into its
// tile's list.
__global__ void k_karras(uint32_t n_total, const uint64_t* __restrict__ sorted_keys, const uint32_t* __restrict__ seg_of,
                         const uint32_t* __restrict__ seg_first, const uint32_t* __restrict__ seg_count,
                         const uint32_t* __restrict__ node_base, ge_bvh_node* nodes, uint32_t* node_parent,
                         uint4* __restrict__ links, TileArgs ta) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_total) return;
  const uint32_t s = seg_of[g];
  const uint32_t first_g = seg_first[s];
  const int n = int(seg_count[s]);
  const int i = int(g - first_g);
  const uint64_t* keys = sorted_keys + first_g;
  ge_bvh_node* seg_nodes = nodes + node_base[s];
  if (n == 1) {
    ge_bvh_node node;
    node.child[0] = GE_BVH_LEAF;
    node.child[1] = GE_BVH_NONE;
    node.parent = GE_BVH_NONE;
    node.count = 1;
    for (int a = 0; a < 3; ++a) {
      node.lo0[a] = node.hi0[a] = 0.0f;  // written by the fit
      node.lo1[a] = __uint_as_float(0x7f800000u);
      node.hi1[a] = __uint_as_float(0xff800000u);
    }
    seg_nodes[0] = node;
    return;
  }
  if (i >= n - 1) return;
  const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) > 0 ? 1 : -1;
  const int dmin = delta(keys, n, i, i - d);
  int lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = delta(keys, n, i, j);
  int sp = 0;
  int t = l;
  do {
    t = (t + 1) >> 1;
    if (delta(keys, n, i, i + (sp + t) * d) > dnode) sp += t;
  } while (t > 1);
  const int gamma = i + sp * d + (d < 0 ? d : 0);
  const int lo = i < j ? i : j, hi = i < j ? j : i;
  const uint32_t c0 = (lo == gamma) ? (GE_BVH_LEAF | uint32_t(gamma)) : uint32_t(gamma);
  const uint32_t c1 = (hi == gamma + 1) ? (GE_BVH_LEAF | uint32_t(gamma + 1)) : uint32_t(gamma + 1);
  links[g] = make_uint4(c0, c1, 0u, uint32_t(hi - lo + 1));  // the fit writes the whole record
  if (!(c0 & GE_BVH_LEAF)) node_parent[node_base[s] + gamma] = uint32_t(i);
  if (!(c1 & GE_BVH_LEAF)) node_parent[node_base[s] + gamma + 1] = uint32_t(i);
  if (ta.levels > 1) file_node(ta, s, uint32_t(i), uint32_t(lo), uint32_t(hi));
}

// ---- levels above 1: a tile's list in LDS, sorted, each node's parent slot inside the tile ----
struct SparseTile {
  uint32_t list[kSparseCap];
  uint32_t arrived[kSparseCap];
  uint16_t pslot[kSparseCap];
  uint8_t pending[kSparseCap];
};

__device__ uint32_t sparse_setup(const TileArgs& a, SparseTile& sh, uint32_t b, uint32_t s, bool fit) {
  const uint32_t t = threadIdx.x;
  const uint32_t nb = a.node_base[s];
  uint32_t cnt = a.count[b];
  if (cnt > kSparseCap) cnt = kSparseCap;  // flagged where it was filled
  uint32_t p2 = 1;
  while (p2 < cnt) p2 <<= 1;
  const uint32_t* src = a.list + size_t(b) * kSparseCap;
  for (uint32_t j = t; j < p2; j += kSparseThreads) sh.list[j] = j < cnt ? src[j] : 0xffffffffu;
  __syncthreads();
  for (uint32_t w = 2; w <= p2; w <<= 1)
    for (uint32_t v = w >> 1; v > 0; v >>= 1) {
      for (uint32_t i = t; i < p2; i += kSparseThreads) {
        const uint32_t l = i ^ v;
        if (l > i) {
          const uint32_t x = sh.list[i], y = sh.list[l];
          if ((x > y) == ((i & w) == 0)) {
            sh.list[i] = y;
            sh.list[l] = x;
          }
        }
      }
      __syncthreads();
    }
  auto slot_of = [&](uint32_t i) -> uint32_t {
    uint32_t lo = 0, hi = cnt;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (sh.list[mid] < i) lo = mid + 1;
      else hi = mid;
    }
    return (lo < cnt && sh.list[lo] == i) ? lo : 0xffffffffu;
  };
  for (uint32_t j = t; j < cnt; j += kSparseThreads) {
    const uint32_t i = sh.list[j];
    const uint32_t ps = i == 0 ? 0xffffffffu : slot_of(a.node_parent[nb + i]);
    sh.pslot[j] = ps == 0xffffffffu ? kNoSlot : uint16_t(ps);
    sh.arrived[j] = 0;
    if (fit) {
      const uint4 link = a.links[a.seg_first[s] + i];
      uint32_t pend = 0;
      for (int c = 0; c < 2; ++c) {
        const uint32_t ch = c == 0 ? link.x : link.y;
        if (!(ch & GE_BVH_LEAF) && slot_of(ch) != 0xffffffffu) ++pend;
      }
      sh.pending[j] = uint8_t(pend);
    }
  }
  __syncthreads();
  return cnt;
}

// Depths of a level's nodes, top level first.
__global__ void __launch_bounds__(kSparseThreads) k_depth_sparse(TileArgs a, uint32_t l) {
  __shared__ SparseTile sh;
  uint32_t s, k;
  tile_place(a, level_shift(l), blockIdx.x, s, k);
  const uint32_t b = a.bucket_base[l] + blockIdx.x;
  const uint32_t cnt = sparse_setup(a, sh, b, s, false);
  const uint32_t first = a.seg_first[s], nb = a.node_base[s];
  for (uint32_t j = threadIdx.x; j < cnt; j += kSparseThreads) {
    uint32_t q = j, steps = 0;
    while (sh.pslot[q] != kNoSlot) {
      q = sh.pslot[q];
      ++steps;
    }
    const uint32_t top = sh.list[q];
    const uint32_t d = (top == 0 ? 0u : a.depth[first + a.node_parent[nb + top]] + 1u) + steps;
    a.depth[first + sh.list[j]] = d;
    if (d > kFitMaxDepth) atomicOr(a.flag, 1u);
  }
}

// The fit of a level's nodes, level 2 first: children below the level are read from node_box.
__global__ void __launch_bounds__(kSparseThreads) k_fit_sparse(TileArgs a, uint32_t l) {
  __shared__ SparseTile sh;
  uint32_t s, k;
  tile_place(a, level_shift(l), blockIdx.x, s, k);
  const uint32_t cnt = sparse_setup(a, sh, a.bucket_base[l] + blockIdx.x, s, true);
  for (uint32_t j = threadIdx.x; j < cnt; j += kSparseThreads) {
    if (sh.pending[j] != 0) continue;
    uint32_t q = j;
    for (;;) {
      fit_node(a, s, sh.list[q]);
      const uint16_t p = sh.pslot[q];
      if (p == kNoSlot) break;
      __threadfence_block();
      if (atomicAdd(&sh.arrived[p], 1u) + 1u < sh.pending[p]) break;
      __threadfence_block();
      q = p;
    }
  }
}

// The collapse of the even-depth nodes above level 1, every such tile in one launch.
__global__ void __launch_bounds__(kSparseThreads) k_collapse_sparse(TileArgs a) {
  const uint32_t b = blockIdx.x;
  uint32_t l = 1;
  while (l + 1 < a.levels && b >= a.bucket_base[l + 1]) ++l;
  uint32_t s, k;
  tile_place(a, level_shift(l), b - a.bucket_base[l], s, k);
  const uint32_t first = a.seg_first[s], nb = a.node_base[s];
  const uint32_t cnt = min(a.count[b], kSparseCap);
  for (uint32_t j = threadIdx.x; j < cnt; j += kSparseThreads) {
    const uint32_t i = a.list[size_t(b) * kSparseCap + j];
    if (!(a.depth[first + i] & 1u)) collapse_node(a, nb, i);
  }
}

// ---- level 1: a tile of up to 1,024 primitives in one workgroup, from LDS ----
__global__ void __launch_bounds__(kTileThreads) k_tile_build(TileArgs a) {
  __shared__ float lbox[kTileMax][6];   // the tile's leaves, sorted order
  __shared__ float nbox[kTileMax][6];   // its internal nodes
  __shared__ uint16_t ch[kTileMax][2];  // kLeafRef | leaf slot, or internal slot; kNoSlot: not this tile's
  __shared__ uint16_t pslot[kTileMax];
  __shared__ uint16_t span[kTileMax];   // primitives below, minus one
  __shared__ uint32_t arrived[kTileMax];
  __shared__ uint8_t pending[kTileMax];
  __shared__ uint8_t height[kTileMax];
  const uint32_t t = threadIdx.x;
  uint32_t s, k;
  tile_place(a, kTileShift, blockIdx.x, s, k);
  const uint32_t n = a.seg_count[s], first = a.seg_first[s], nb = a.node_base[s];
  const uint32_t A = k << kTileShift;
  const uint32_t B = min(n, A + kTileMax);
  const uint32_t nleaf = B - A;
  const uint32_t cnt = min(B, n - 1) > A ? min(B, n - 1) - A : 0u;  // internal node indices [A, A + cnt)
  for (uint32_t x = t; x < nleaf; x += kTileThreads) {
    const uint32_t p = a.perm[first + A + x];
    for (int q = 0; q < 6; ++q) lbox[x][q] = a.boxes[size_t(p) * 6 + q];
  }
  // Child links; a node whose range leaves the tile is not this tile's (ch[j][0] = kNoSlot).
  for (uint32_t j = t; j < cnt; j += kTileThreads) {
    const uint32_t i = A + j;
    const uint4 tail = a.links[first + i];  // child[2], -, count
    const uint32_t c0 = tail.x, c1 = tail.y, cn = tail.w;
    const uint32_t gamma = c0 & ~GE_BVH_LEAF;
    const uint32_t lo = i <= gamma ? i : i + 1 - cn, hi = i <= gamma ? i + cn - 1 : i;
    arrived[j] = 0;
    if (lo < A || hi >= B) {
      ch[j][0] = ch[j][1] = kNoSlot;
      continue;
    }
    const bool l0 = c0 & GE_BVH_LEAF, l1 = c1 & GE_BVH_LEAF;
    ch[j][0] = uint16_t(l0 ? (kLeafRef | ((c0 & ~GE_BVH_LEAF) - A)) : (c0 - A));
    ch[j][1] = uint16_t(l1 ? (kLeafRef | ((c1 & ~GE_BVH_LEAF) - A)) : (c1 - A));
    pending[j] = uint8_t((l0 ? 0 : 1) + (l1 ? 0 : 1));
    span[j] = uint16_t(cn - 1);
  }
  __syncthreads();
  for (uint32_t j = t; j < cnt; j += kTileThreads) {
    if (ch[j][0] == kNoSlot) continue;
    const uint32_t i = A + j;
    uint32_t ps = 0xffffffffu;
    if (i != 0) {
      const uint32_t p = a.node_parent[nb + i];
      if (p >= A && p < A + cnt && ch[p - A][0] != kNoSlot) ps = p - A;
    }
    pslot[j] = ps == 0xffffffffu ? kNoSlot : uint16_t(ps);
  }
  __syncthreads();
  // The fit, bottom-up in LDS alone (the records go out afterwards).
  auto box_of = [&](uint16_t c) -> const float* { return (c & kLeafRef) ? lbox[c & (kTileMax - 1)] : nbox[c & (kTileMax - 1)]; };
  for (uint32_t j0 = t; j0 < cnt; j0 += kTileThreads) {
    if (ch[j0][0] == kNoSlot || pending[j0] != 0) continue;
    uint32_t j = j0;
    for (;;) {
      const uint16_t x = ch[j][0], y = ch[j][1];
      const float* b0 = box_of(x);
      const float* b1 = box_of(y);
      const uint32_t h0 = (x & kLeafRef) ? 0u : height[x], h1 = (y & kLeafRef) ? 0u : height[y];
      for (int q = 0; q < 3; ++q) {
        nbox[j][q] = __builtin_fminf(b0[q], b1[q]);
        nbox[j][3 + q] = __builtin_fmaxf(b0[3 + q], b1[3 + q]);
      }
      const uint32_t ht = 1 + (h0 > h1 ? h0 : h1);
      height[j] = uint8_t(ht < 255u ? ht : 255u);
      const uint16_t p = pslot[j];
      if (p == kNoSlot) break;
      __threadfence_block();
      if (atomicAdd(&arrived[p], 1u) + 1u < pending[p]) break;
      __threadfence_block();
      j = p;
    }
  }
  __syncthreads();
  // The records, the tile's top boxes, depths and the collapse.
  for (uint32_t j = t; j < cnt; j += kTileThreads) {
    const uint16_t x = ch[j][0], y = ch[j][1];
    if (x == kNoSlot) continue;
    const uint32_t i = A + j;
    const bool lx = x & kLeafRef, ly = y & kLeafRef;
    const float* bx = box_of(x);
    const float* by = box_of(y);
    const uint32_t gx = lx ? (GE_BVH_LEAF | (A + (x & (kTileMax - 1)))) : (A + x);
    const uint32_t gy = ly ? (GE_BVH_LEAF | (A + (y & (kTileMax - 1)))) : (A + y);
    const uint16_t ps = pslot[j];
    const uint32_t parent = i == 0 ? GE_BVH_NONE : a.node_parent[nb + i];
    store_record(a.nodes + nb + i, bx, bx + 3, by, by + 3, gx, gy, parent, uint32_t(span[j]) + 1);
    uint32_t steps = 0, q = j;
    while (pslot[q] != kNoSlot) {
      q = pslot[q];
      ++steps;
    }
    const uint32_t top = A + q;
    if (ps == kNoSlot) {  // the tile's top: its box and height for the level above, or the root's
      float* o = a.node_box + size_t(nb + i) * 8;
      for (int f = 0; f < 6; ++f) o[f] = nbox[j][f];
      o[6] = __uint_as_float(height[j]);
      if (i == 0) {
        const uint32_t out = a.root_index ? a.root_index[s] : s;
        for (int f = 0; f < 6; ++f) a.roots[out * 6 + f] = nbox[j][f];
        a.heights[out] = height[j];
      }
    }
    const uint32_t d = (top == 0 ? 0u : a.depth[first + a.node_parent[nb + top]] + 1u) + steps;
    if (d > kFitMaxDepth) atomicOr(a.flag, 1u);
    if ((d & 1u) || !a.wide) continue;
    // Entries: a leaf child as it is, an inner child's two children, in child order.
    const uint16_t x0 = ch[x & (kTileMax - 1)][0], x1 = ch[x & (kTileMax - 1)][1];
    const uint16_t y0 = ch[y & (kTileMax - 1)][0], y1 = ch[y & (kTileMax - 1)][1];
    const uint16_t r[4] = {lx ? x : x0, lx ? (ly ? y : y0) : x1, lx ? y1 : (ly ? y : y0), y1};
    uint32_t ref[4];
    const float* box[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint16_t c = r[e];
      ref[e] = (c & kLeafRef) ? (GE_BVH_LEAF | (A + (c & (kTileMax - 1)))) : (A + (c & (kTileMax - 1)));
      box[e] = box_of(c);
    }
    store_wide(a.wide + nb + i, ref, box, (lx ? 1 : 2) + (ly ? 1 : 2));
  }
}

// Triangle t of cluster cl: its three corners from the packed 8-bit vertex indices.
__device__ __forceinline__ void fetch_corners(const ge_cluster_desc& cl, const float* __restrict__ vertices,
                                              const uint32_t* __restrict__ triangles, uint32_t t, ge_bvh_triangle& tri) {
  const uint32_t packed = triangles[cl.triangle_offset + t];
  float* v[3] = {tri.v0, tri.v1, tri.v2};
  for (int q = 0; q < 3; ++q) {
    const uint32_t vi = cl.vertex_offset + ((packed >> (8 * q)) & 0xffu);
    v[q][0] = vertices[size_t(vi) * 3];
    v[q][1] = vertices[size_t(vi) * 3 + 1];
    v[q][2] = vertices[size_t(vi) * 3 + 2];
  }
}

// The device-wide batch's clusters: segment s owns items [pre[s], pre[s + 1]) of the batch's cluster
// sequence, in order.
__device__ __forceinline__ uint32_t item_segment(const uint32_t* __restrict__ pre, uint32_t nseg, uint32_t b) {
  uint32_t lo = 0, hi = nseg - 1;  // the last segment whose first item is at or before b
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
__device__ __forceinline__ uint32_t item_cluster(const GlobalSeg& g, const uint32_t* __restrict__ lists, uint32_t j) {
  return g.list ? lists[g.first + j] : g.first + j;
}

__global__ void k_cluster_counts(uint32_t ncl, uint32_t nseg, const uint32_t* __restrict__ pre,
                                 const GlobalSeg* __restrict__ segs, const uint32_t* __restrict__ lists,
                                 const ge_cluster_desc* __restrict__ clusters, uint32_t* counts) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= ncl) return;
  const uint32_t s = item_segment(pre, nseg, b);
  counts[b] = clusters[item_cluster(segs[s], lists, b - pre[s])].triangle_count;
}

// Item b's triangles (a workgroup a cluster) at batch index out[b] on (the scan of the counts): each
// one's record, box and segment, and the cluster's lowest and highest ordered centroids.
__global__ void __launch_bounds__(128) k_triangle_records(uint32_t nseg, const uint32_t* __restrict__ pre,
                                                          const GlobalSeg* __restrict__ segs,
                                                          const uint32_t* __restrict__ lists,
                                                          const ge_cluster_desc* __restrict__ clusters,
                                                          const float* __restrict__ vertices,
                                                          const uint32_t* __restrict__ triangles,
                                                          const uint32_t* __restrict__ out, ge_bvh_triangle* tris,
                                                          float* boxes, uint32_t* seg_of, uint32_t* part,
                                                          uint32_t* item_seg) {
  __shared__ uint32_t red[2][6];
  const uint32_t b = blockIdx.x;
  const uint32_t s = item_segment(pre, nseg, b);
  const GlobalSeg g = segs[s];
  const uint32_t gc = item_cluster(g, lists, b - pre[s]);
  const ge_cluster_desc cl = clusters[gc];
  const uint32_t out0 = out[b];
  uint32_t lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
  for (uint32_t t = threadIdx.x; t < cl.triangle_count; t += blockDim.x) {
    ge_bvh_triangle tri;
    fetch_corners(cl, vertices, triangles, t, tri);
    const uint32_t o = out0 + t;
    tri.prim = ((gc - g.cbase) << 8) | t;
    tri.index = o;  // rewritten to the mesh-local index by the gather
    tri.pad = 0;
    tris[o] = tri;
    for (int a = 0; a < 3; ++a) {
      const float blo = __builtin_fminf(__builtin_fminf(tri.v0[a], tri.v1[a]), tri.v2[a]);
      const float bhi = __builtin_fmaxf(__builtin_fmaxf(tri.v0[a], tri.v1[a]), tri.v2[a]);
      boxes[size_t(o) * 6 + a] = blo;
      boxes[size_t(o) * 6 + 3 + a] = bhi;
      const uint32_t c = ordered((blo + bhi) * 0.5f);
      lo[a] = min(lo[a], c);
      hi[a] = max(hi[a], c);
    }
    seg_of[o] = s;
  }
  for (int a = 0; a < 3; ++a)
    for (int off = 32; off > 0; off >>= 1) {
      lo[a] = min(lo[a], uint32_t(__shfl_xor(lo[a], off)));
      hi[a] = max(hi[a], uint32_t(__shfl_xor(hi[a], off)));
    }
  const uint32_t w = threadIdx.x >> 6;
  if ((threadIdx.x & 63u) == 0)
    for (int a = 0; a < 3; ++a) {
      red[w][a] = lo[a];
      red[w][3 + a] = hi[a];
    }
  __syncthreads();
  if (threadIdx.x < 6) {
    const uint32_t x = red[0][threadIdx.x], y = red[1][threadIdx.x];
    part[size_t(b) * 6 + threadIdx.x] = threadIdx.x < 3 ? min(x, y) : max(x, y);
  }
  if (threadIdx.x == 0) item_seg[b] = s;
}

__global__ void k_gather_triangles(uint32_t n, const uint32_t* __restrict__ perm, const uint32_t* __restrict__ seg_of,
                                   const uint32_t* __restrict__ seg_first, const uint32_t* __restrict__ seg_out,
                                   const ge_bvh_triangle* __restrict__ in, ge_bvh_triangle* out,
                                   const uint32_t* __restrict__ morton_in, uint32_t* morton_out) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const uint32_t s = seg_of[g];
  const uint32_t o = seg_out[s] + (g - seg_first[s]);
  ge_bvh_triangle tri = in[perm[g]];
  tri.index = perm[g] - seg_first[s];
  out[o] = tri;
  morton_out[o] = morton_in[g];
}

__global__ void k_codes(uint32_t n, const uint64_t* keys, uint32_t* codes) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) codes[g] = uint32_t(keys[g]) & 0x3fffffffu;
}

// Slot k's box: the 8 corners of its instance's structure's root box (its mesh's, or the frame's
// structure's: root_of[id]) through the instance's 3x4, padded by 2^-20 of its largest coordinate.
__global__ void k_instance_boxes(uint32_t n, const uint32_t* __restrict__ live, const ge_instance_desc* __restrict__ inst,
                                 const float* __restrict__ T, const float* __restrict__ roots,
                                 const uint32_t* __restrict__ root_of, float* boxes, uint32_t* seg_of) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t id = live[k];
  const ge_instance_desc in = inst[id];
  const float* b = roots + size_t(root_of[id]) * 6;
  float lo[3] = {__uint_as_float(0x7f800000u), __uint_as_float(0x7f800000u), __uint_as_float(0x7f800000u)};
  float hi[3] = {__uint_as_float(0xff800000u), __uint_as_float(0xff800000u), __uint_as_float(0xff800000u)};
  for (int i = 0; i < 8; ++i) {
    const float p[3] = {(i & 1) ? b[3] : b[0], (i & 2) ? b[4] : b[1], (i & 4) ? b[5] : b[2]};
    for (int a = 0; a < 3; ++a) {
      const float v = (in.rows[a][0] * p[0] + in.rows[a][1] * p[1] + in.rows[a][2] * p[2]) + T[size_t(id) * 3 + a];
      lo[a] = __builtin_fminf(lo[a], v);
      hi[a] = __builtin_fmaxf(hi[a], v);
    }
  }
  float m = 0.0f;
  for (int a = 0; a < 3; ++a) m = __builtin_fmaxf(m, __builtin_fmaxf(__builtin_fabsf(lo[a]), __builtin_fabsf(hi[a])));
  const float pad = m * 9.5367431640625e-7f + 1e-30f;
  for (int a = 0; a < 3; ++a) {
    boxes[size_t(k) * 6 + a] = lo[a] - pad;
    boxes[size_t(k) * 6 + 3 + a] = hi[a] + pad;
  }
  seg_of[k] = 0;
}

__global__ void k_gather_slots(uint32_t n, const uint32_t* perm, const uint32_t* live, uint32_t* slots) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) slots[g] = live[perm[g]];
}

// ---- a whole mesh's structure in one workgroup --------------------------------------------------
// A segment of at most kBlockMax primitives is built by one workgroup from start to end in LDS: its
// centroid bounds, Morton codes, a bitonic sort of (code, index) (the stable order of the global
// sort), Karras' emission (with each node's primitive count, its range's length), each node's
// depth by pointer jumping, the fit level by level from the deepest (a barrier a level; a node's
// internal children sit one level deeper), the binary nodes and the 4-wide collapse. One launch
// builds every small mesh of a batch, where the device-wide passes cost a launch a level; the
// results are the device-wide build's to the bit.
//
// Over triangles (Tri), a workgroup also reads its segment's clusters itself (a mesh's own, or a
// frame structure's list) and writes the sorted triangle records and codes straight to the
// structure's place: no batch records, boxes, order or gather pass go through memory.
constexpr uint32_t kBlockMax = k_block_build_max;
constexpr uint32_t kBlockThreads = 512;
constexpr uint16_t kBlockLeaf = 0x8000u;
static_assert(kBlockMax == 2 * kBlockThreads && kBlockMax <= 0x8000u && (kBlockMax & (kBlockMax - 1)) == 0,
              "a thread holds two keys of a power-of-two sort; 15-bit indices");

struct BlockArgs {
  // Boxes given (Tri false): segment seg_base + blockIdx.x of the batch tables, its primitives'
  // boxes at boxes[6 seg_first ..], its sorted order (batch indices) to perm[seg_first ..] and its
  // codes to morton[seg_first ..].
  uint32_t seg_base;
  const float* boxes;
  const uint32_t* seg_first;
  const uint32_t* seg_count;
  const uint32_t* node_base;
  const uint32_t* root_index;
  uint32_t* perm;
  // Triangles (Tri): segment blockIdx.x of `segs`; its sorted records to tris[out ..] and codes to
  // morton[out ..].
  const BlockSeg* segs;
  const uint32_t* lists;  // frame structures' cluster lists
  const ge_cluster_desc* clusters;
  const float* vertices;
  const uint32_t* triangles;
  ge_bvh_triangle* tris;
  // Both.
  ge_bvh_node* nodes;
  uint32_t* morton;
  float* roots;
  uint32_t* heights;
  WideNode* wide;
};

__device__ __forceinline__ uint32_t seg_cluster(const BlockSeg& g, const uint32_t* lists, uint32_t j) {
  return g.list ? lists[g.first + j] : g.first + j;
}

template <bool Tri>
__global__ void __launch_bounds__(kBlockThreads) k_block_build(BlockArgs ba) {
  // The centroid bounds, then the sort keys, then each internal node's ancestor link, depth and
  // height (the fit's level order) share one array.
  __shared__ union {
    uint32_t red[6];
    uint32_t scan[kBlockThreads + 1 + kBlockThreads / 64];  // the triangles' cluster prefix
    uint64_t keys[kBlockMax];  // (code << 32 | local index), sorted in place
    struct {
      uint16_t anc[kBlockMax];  // pointer jumping: an ancestor 2^r levels up (0xffff past the root)
      uint16_t dep[kBlockMax];  // the distance to it, then the node's depth
      uint16_t hgt[kBlockMax];  // the node's height (a leaf's is 0)
      uint32_t max_depth;
    } f;
  } u;
  __shared__ float lbox[kBlockMax][6];     // the leaves' boxes, sorted order
  __shared__ float nbox[kBlockMax][6];     // internal nodes' boxes (first, over triangles, the leaves' in batch order)
  __shared__ __align__(16) uint16_t child[kBlockMax][2];  // internal children; kBlockLeaf | sorted position for a leaf
  __shared__ uint16_t parent[kBlockMax];   // internal node -> its parent
  __shared__ uint16_t count[kBlockMax];    // internal node -> the primitives below it
  const uint32_t t = threadIdx.x;
  uint32_t n, nb, out, first = 0;
  BlockSeg g{};
  if (Tri) {
    g = ba.segs[blockIdx.x];
    n = g.count;
    nb = g.node;
    out = g.root;
  } else {
    const uint32_t s = ba.seg_base + blockIdx.x;
    first = ba.seg_first[s];
    n = ba.seg_count[s];
    nb = ba.node_base[s];
    out = ba.root_index ? ba.root_index[s] : s;
  }
  // A triangle's source, (its cluster's place in the segment's list << 8 | triangle in the
  // cluster), kept where the child links go later.
  uint32_t* src = reinterpret_cast<uint32_t*>(child);
  // Primitive i's box: from memory, or from the triangle phase's boxes in nbox.
  auto pbox = [&](uint32_t i, int a) -> float { return Tri ? nbox[i][a] : ba.boxes[size_t(first + i) * 6 + a]; };

  if (Tri) {
    // The segment's triangles in batch order: its clusters in list order, a chunk of
    // kBlockThreads clusters at a time (a scan of their triangle counts), each triangle finding
    // its cluster by a binary search of the chunk's prefix. Boxes to nbox, sources to src.
    uint32_t* pre = u.scan;
    uint32_t* wsum = u.scan + kBlockThreads + 1;
    const uint32_t lane = t & 63u, wave = t >> 6;
    uint32_t done = 0;
    for (uint32_t c0 = 0; c0 < g.ncl; c0 += kBlockThreads) {
      const uint32_t j = c0 + t;
      const uint32_t tc = j < g.ncl ? ba.clusters[seg_cluster(g, ba.lists, j)].triangle_count : 0u;
      uint32_t incl = tc;
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = uint32_t(__shfl_up(incl, off));
        if (lane >= uint32_t(off)) incl += y;
      }
      if (lane == 63) wsum[wave] = incl;
      __syncthreads();
      uint32_t base = 0, total = 0;
      for (uint32_t w = 0; w < kBlockThreads / 64; ++w) {
        base += w < wave ? wsum[w] : 0u;
        total += wsum[w];
      }
      pre[t] = base + incl - tc;
      __syncthreads();
      const uint32_t m = min(kBlockThreads, g.ncl - c0);
      for (uint32_t x = t; x < total; x += kBlockThreads) {
        uint32_t lo = 0, hi = m - 1;  // the last cluster of the chunk that starts at or before x
        while (lo < hi) {
          const uint32_t mid = (lo + hi + 1) >> 1;
          if (pre[mid] <= x) lo = mid;
          else hi = mid - 1;
        }
        const uint32_t tri_in = x - pre[lo];
        const uint32_t o = done + x;
        ge_bvh_triangle tri;
        fetch_corners(ba.clusters[seg_cluster(g, ba.lists, c0 + lo)], ba.vertices, ba.triangles, tri_in, tri);
        for (int a = 0; a < 3; ++a) {
          nbox[o][a] = __builtin_fminf(__builtin_fminf(tri.v0[a], tri.v1[a]), tri.v2[a]);
          nbox[o][3 + a] = __builtin_fmaxf(__builtin_fmaxf(tri.v0[a], tri.v1[a]), tri.v2[a]);
        }
        src[o] = ((c0 + lo) << 8) | tri_in;
      }
      done += total;
      __syncthreads();
    }
  }

  // Centroid bounds, as order-preserving integers (min and max are the float ones).
  if (t < 6) u.red[t] = t < 3 ? 0xffffffffu : 0u;
  __syncthreads();
  {
    uint32_t lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
    for (uint32_t i = t; i < n; i += kBlockThreads)
      for (int a = 0; a < 3; ++a) {
        const uint32_t c = ordered((pbox(i, a) + pbox(i, 3 + a)) * 0.5f);
        lo[a] = min(lo[a], c);
        hi[a] = max(hi[a], c);
      }
    // The wave's bounds first: one LDS atomic a wave and axis, not one a lane on six addresses.
    for (int a = 0; a < 3; ++a)
      for (int off = 32; off > 0; off >>= 1) {
        lo[a] = min(lo[a], uint32_t(__shfl_xor(lo[a], off)));
        hi[a] = max(hi[a], uint32_t(__shfl_xor(hi[a], off)));
      }
    if ((t & 63u) == 0)
      for (int a = 0; a < 3; ++a) {
        atomicMin(&u.red[a], lo[a]);
        atomicMax(&u.red[3 + a], hi[a]);
      }
  }
  __syncthreads();
  float clo[3], cscale[3];
  for (int a = 0; a < 3; ++a) {
    clo[a] = unordered(u.red[a]);
    const float ext = unordered(u.red[3 + a]) - clo[a];
    cscale[a] = ext > 0.0f ? 1024.0f / ext : 0.0f;
  }
  __syncthreads();  // red is read; the keys take its place
  // Codes: thread t holds keys 2t and 2t + 1, (code << 32 | index); past n, ~0, which sorts last.
  uint64_t kv[2];
  for (uint32_t e = 0; e < 2; ++e) {
    const uint32_t i = 2 * t + e;
    if (i < n) {
      uint32_t q[3];
      for (int a = 0; a < 3; ++a) {
        const float c = (pbox(i, a) + pbox(i, 3 + a)) * 0.5f;
        q[a] = quantize10(c, clo[a], cscale[a]);
      }
      const uint32_t code = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
      kv[e] = (uint64_t(code) << 32) | i;
    } else {
      kv[e] = ~0ull;
    }
  }
  // Batcher's bitonic sort of the 1,024 keys (unique, so the order is the stable sort's): element i
  // meets i ^ j, keeping the smaller where i is the lower index of an ascending block (i & k == 0)
  // or the higher of a descending one. Partners 1 apart are in the thread, 2 to 64 apart in the
  // wave (lane t ^ j/2, a shuffle), 128 and more apart in other waves (through LDS, a barrier each):
  // 6 of the 55 stages wait at a barrier.
  for (uint32_t k = 2; k <= kBlockMax; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      if (j == 1) {
        const bool up = ((2 * t) & k) == 0;
        const uint64_t lo = kv[0] < kv[1] ? kv[0] : kv[1], hi = kv[0] < kv[1] ? kv[1] : kv[0];
        kv[0] = up ? lo : hi;
        kv[1] = up ? hi : lo;
        continue;
      }
      uint64_t other[2];
      if (j < 128) {
        for (uint32_t e = 0; e < 2; ++e) {
          const uint32_t m = j >> 1;
          const uint32_t lo32 = uint32_t(__shfl_xor(uint32_t(kv[e]), int(m)));
          const uint32_t hi32 = uint32_t(__shfl_xor(uint32_t(kv[e] >> 32), int(m)));
          other[e] = (uint64_t(hi32) << 32) | lo32;
        }
      } else {
        u.keys[2 * t] = kv[0];
        u.keys[2 * t + 1] = kv[1];
        __syncthreads();
        other[0] = u.keys[(2 * t) ^ j];
        other[1] = u.keys[(2 * t + 1) ^ j];
        __syncthreads();
      }
      for (uint32_t e = 0; e < 2; ++e) {
        const uint32_t i = 2 * t + e;
        const bool keep_small = ((i & j) == 0) == ((i & k) == 0);
        const uint64_t y = other[e];
        kv[e] = keep_small ? (kv[e] < y ? kv[e] : y) : (kv[e] < y ? y : kv[e]);
      }
    }
  }
  u.keys[2 * t] = kv[0];
  u.keys[2 * t + 1] = kv[1];
  __syncthreads();
  for (uint32_t i = t; i < n; i += kBlockThreads) {
    const uint32_t o = uint32_t(u.keys[i]);
    const uint32_t code = uint32_t(u.keys[i] >> 32);
    for (int a = 0; a < 6; ++a) lbox[i][a] = pbox(o, a);
    if (Tri) {
      // The record of the triangle sorted to place i, its corners fetched again (L2-resident).
      const uint32_t from = src[o], j = from >> 8, tri_in = from & 0xffu;
      const uint32_t gc = seg_cluster(g, ba.lists, j);
      ge_bvh_triangle tri;
      fetch_corners(ba.clusters[gc], ba.vertices, ba.triangles, tri_in, tri);
      tri.prim = ((gc - g.cbase) << 8) | tri_in;
      tri.index = o;
      tri.pad = 0;
      ba.tris[g.out + i] = tri;
      ba.morton[g.out + i] = code;
    } else {
      ba.perm[first + i] = first + o;
      ba.morton[first + i] = code;
    }
  }
  if (n == 1) {
    __syncthreads();
    if (t == 0) {
      const float* lo = lbox[0];
      const float* hi = lbox[0] + 3;
      const float empty_lo[3] = {__uint_as_float(0x7f800000u), __uint_as_float(0x7f800000u), __uint_as_float(0x7f800000u)};
      const float empty_hi[3] = {__uint_as_float(0xff800000u), __uint_as_float(0xff800000u), __uint_as_float(0xff800000u)};
      store_record(ba.nodes + nb, lo, hi, empty_lo, empty_hi, GE_BVH_LEAF, GE_BVH_NONE, GE_BVH_NONE, 1u);
      for (int a = 0; a < 6; ++a) ba.roots[out * 6 + a] = lbox[0][a];
      ba.heights[out] = 1;
      const uint32_t ref[4] = {GE_BVH_LEAF, 0u, 0u, 0u};
      const float* box[4] = {lbox[0], lbox[0], lbox[0], lbox[0]};
      store_wide(ba.wide + nb, ref, box, 1);
    }
    return;
  }
  // Karras 2012 over the sorted codes, equal codes split by position.
  const uint32_t* code32 = reinterpret_cast<const uint32_t*>(u.keys);  // a key's code is its high word
  auto delta = [&](int i, int j) -> int {
    if (j < 0 || j > int(n) - 1) return -1;
    const uint32_t a = code32[2 * i + 1], b = code32[2 * j + 1];
    if (a != b) return __clz(a ^ b);
    return 32 + __clz(uint32_t(i) ^ uint32_t(j));
  };
  __syncthreads();
  if (t == 0) parent[0] = 0xffffu;
  for (uint32_t ii = t; ii + 1 < n; ii += kBlockThreads) {
    const int i = int(ii);
    const int d = (delta(i, i + 1) - delta(i, i - 1)) > 0 ? 1 : -1;
    const int dmin = delta(i, i - d);
    int lmax = 2;
    while (delta(i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int st = lmax / 2; st >= 1; st /= 2)
      if (delta(i, i + (l + st) * d) > dmin) l += st;
    const int j = i + l * d;
    const int dnode = delta(i, j);
    int sp = 0, st = l;
    do {
      st = (st + 1) >> 1;
      if (delta(i, i + (sp + st) * d) > dnode) sp += st;
    } while (st > 1);
    const int gamma = i + sp * d + (d < 0 ? d : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const bool l0 = lo == gamma, l1 = hi == gamma + 1;
    child[i][0] = l0 ? uint16_t(kBlockLeaf | gamma) : uint16_t(gamma);
    child[i][1] = l1 ? uint16_t(kBlockLeaf | (gamma + 1)) : uint16_t(gamma + 1);
    if (!l0) parent[gamma] = uint16_t(i);
    if (!l1) parent[gamma + 1] = uint16_t(i);
    count[i] = uint16_t(hi - lo + 1);
  }
  __syncthreads();  // the keys are done with; the depths take their place
  // Each internal node's depth by pointer jumping: after round r a node's link is its ancestor
  // 2^(r+1) levels up (or none) and `dep` the levels to it; ceil(log2 n) rounds reach the root.
  constexpr uint32_t kPer = kBlockMax / kBlockThreads;  // internal nodes a thread holds
  const uint32_t ni = n - 1;
  for (uint32_t k = 0; k < kPer; ++k) {
    const uint32_t i = t + k * kBlockThreads;
    if (i < ni) {
      u.f.anc[i] = i == 0 ? 0xffffu : parent[i];
      u.f.dep[i] = i == 0 ? 0u : 1u;
    }
  }
  if (t == 0) u.f.max_depth = 0;
  __syncthreads();
  for (uint32_t span = 1; span < n; span <<= 1) {
    uint16_t na[kPer], nd[kPer];
    for (uint32_t k = 0; k < kPer; ++k) {
      const uint32_t i = t + k * kBlockThreads;
      if (i >= ni) continue;
      const uint16_t a = u.f.anc[i];
      na[k] = a == 0xffffu ? a : u.f.anc[a];
      nd[k] = a == 0xffffu ? u.f.dep[i] : uint16_t(u.f.dep[i] + u.f.dep[a]);
    }
    __syncthreads();
    for (uint32_t k = 0; k < kPer; ++k) {
      const uint32_t i = t + k * kBlockThreads;
      if (i >= ni) continue;
      u.f.anc[i] = na[k];
      u.f.dep[i] = nd[k];
    }
    __syncthreads();
  }
  {
    uint32_t m = 0;
    for (uint32_t k = 0; k < kPer; ++k) {
      const uint32_t i = t + k * kBlockThreads;
      if (i < ni) m = u.f.dep[i] > m ? u.f.dep[i] : m;
    }
    atomicMax(&u.f.max_depth, m);
  }
  __syncthreads();
  auto box_of = [&](uint16_t c) -> const float* {
    return (c & kBlockLeaf) ? lbox[c & (kBlockMax - 1)] : nbox[c & (kBlockMax - 1)];
  };
  auto height_of = [&](uint16_t c) -> uint32_t { return (c & kBlockLeaf) ? 0u : u.f.hgt[c]; };
  // The fit, level by level from the deepest: a node's internal children are one level deeper, so
  // they are fitted a barrier earlier. Its box is the union of theirs (exact min and max), its
  // height one more than the taller one's.
  for (int level = int(u.f.max_depth); level >= 0; --level) {
    for (uint32_t k = 0; k < kPer; ++k) {
      const uint32_t i = t + k * kBlockThreads;
      if (i >= ni || u.f.dep[i] != uint32_t(level)) continue;
      const uint16_t x = child[i][0], y = child[i][1];
      const float* b0 = box_of(x);
      const float* b1 = box_of(y);
      float a[6], b[6];
      for (int q = 0; q < 6; ++q) {
        a[q] = b0[q];
        b[q] = b1[q];
      }
      const uint32_t h0 = height_of(x), h1 = height_of(y);
      for (int q = 0; q < 3; ++q) {
        nbox[i][q] = __builtin_fminf(a[q], b[q]);
        nbox[i][3 + q] = __builtin_fmaxf(a[3 + q], b[3 + q]);
      }
      u.f.hgt[i] = uint16_t(1u + (h0 > h1 ? h0 : h1));
    }
    __syncthreads();
  }
  // The binary nodes, the root, each node's depth and the collapse.
  for (uint32_t i = t; i + 1 < n; i += kBlockThreads) {
    const uint16_t x = child[i][0], y = child[i][1];
    const bool lx = x & kBlockLeaf, ly = y & kBlockLeaf;
    const float* bx = box_of(x);
    const float* by = box_of(y);
    store_record(ba.nodes + nb + i, bx, bx + 3, by, by + 3, lx ? (GE_BVH_LEAF | uint32_t(x & ~kBlockLeaf)) : uint32_t(x),
                 ly ? (GE_BVH_LEAF | uint32_t(y & ~kBlockLeaf)) : uint32_t(y), i == 0 ? GE_BVH_NONE : uint32_t(parent[i]),
                 count[i]);
    if (u.f.dep[i] & 1u) continue;
    const uint16_t x0 = child[x & (kBlockMax - 1)][0], x1 = child[x & (kBlockMax - 1)][1];
    const uint16_t y0 = child[y & (kBlockMax - 1)][0], y1 = child[y & (kBlockMax - 1)][1];
    const uint16_t r[4] = {lx ? x : x0, lx ? (ly ? y : y0) : x1, lx ? y1 : (ly ? y : y0), y1};
    uint32_t ref[4];
    const float* box[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      ref[e] = (r[e] & kBlockLeaf) ? (GE_BVH_LEAF | uint32_t(r[e] & ~kBlockLeaf)) : uint32_t(r[e]);
      box[e] = box_of(r[e]);
    }
    store_wide(ba.wide + nb + i, ref, box, (lx ? 1 : 2) + (ly ? 1 : 2));
  }
  if (t == 0) {
    for (int a = 0; a < 6; ++a) ba.roots[out * 6 + a] = nbox[0][a];
    ba.heights[out] = u.f.hgt[0];
  }
}

// For a refit: the parent links and the tile lists from the records a build left, and the
// identity as the sorted order (the leaves' boxes are given in slot order).
__global__ void k_parents_from_records(uint32_t n, const ge_bvh_node* __restrict__ nodes, uint32_t* node_parent,
                                       uint4* links) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 >= n) return;
  const ge_bvh_node nd = nodes[i];
  node_parent[i] = nd.parent;
  links[i] = make_uint4(nd.child[0], nd.child[1], 0u, nd.count);
}
__global__ void k_file_records(uint32_t n, const ge_bvh_node* __restrict__ nodes, TileArgs ta) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 >= n) return;
  const uint32_t gamma = nodes[i].child[0] & ~GE_BVH_LEAF, span = nodes[i].count;
  const uint32_t lo = i <= gamma ? i : i + 1 - span, hi = i <= gamma ? i + span - 1 : i;
  file_node(ta, 0, i, lo, hi);
}
__global__ void k_identity(uint32_t n, uint32_t* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = i;
}

inline uint32_t blocks(uint64_t n, uint32_t t) { return uint32_t((n + t - 1) / t); }

// The tile ladder of a device-wide build: how many levels, each level's tiles, the lists above
// level 1 (cleared).
struct TileLadder {
  TileArgs args{};
  uint32_t levels = 1;
  uint32_t tiles[kMaxLevels] = {};
};

hipError_t tile_ladder(BuildScratch& sc, hipStream_t stream, uint32_t nseg, const uint32_t* h_seg_count, TileLadder& lad) {
  uint32_t levels = 1;
  for (uint32_t q = 0; q < nseg; ++q)
    while (levels < kMaxLevels && (uint64_t(1) << (kTileShift + kFanShift * (levels - 1))) < h_seg_count[q]) ++levels;
  lad.levels = levels;
  lad.args.nseg = nseg;
  lad.args.levels = levels;
  for (uint32_t l = 0; l < levels; ++l) {
    const uint32_t shift = kTileShift + kFanShift * l;
    uint64_t t = 0;
    for (uint32_t q = 0; q < nseg; ++q) t += (uint64_t(h_seg_count[q]) + (1ull << shift) - 1) >> shift;
    lad.tiles[l] = uint32_t(t);
    lad.args.bucket_base[l + 1] = lad.args.bucket_base[l] + (l == 0 ? 0u : lad.tiles[l]);  // level 1 keeps no list
  }
  const uint32_t lists = lad.args.bucket_base[levels];
  hipError_t e;
  if ((e = sc.buckets.reserve(size_t(std::max(lists, 1u)) * kSparseCap * 4)) != hipSuccess) return e;
  if ((e = sc.bucket_count.reserve(size_t(std::max(lists, 1u)) * 4)) != hipSuccess) return e;
  if (lists && (e = hipMemsetAsync(sc.bucket_count.ptr, 0, size_t(lists) * 4, stream)) != hipSuccess) return e;
  lad.args.list = sc.buckets.as<uint32_t>();
  lad.args.count = sc.bucket_count.as<uint32_t>();
  lad.args.depth = sc.depth.as<uint32_t>();
  return hipSuccess;
}

// Depths above level 1 (top down), level 1 whole, the fit above level 1 (bottom up), the collapse
// above level 1.
hipError_t tile_fit(hipStream_t stream, const TileLadder& lad) {
  const TileArgs& ta = lad.args;
  for (uint32_t l = lad.levels; l-- > 1;) k_depth_sparse<<<lad.tiles[l], kSparseThreads, 0, stream>>>(ta, l);
  k_tile_build<<<lad.tiles[0], kTileThreads, 0, stream>>>(ta);
  for (uint32_t l = 1; l < lad.levels; ++l) k_fit_sparse<<<lad.tiles[l], kSparseThreads, 0, stream>>>(ta, l);
  const uint32_t lists = ta.bucket_base[lad.levels];
  if (lists && ta.wide) k_collapse_sparse<<<lists, kSparseThreads, 0, stream>>>(ta);
  return hipGetLastError();
}

}  // namespace

hipError_t build_segments(BuildScratch& sc, hipStream_t stream, uint32_t n, const float* boxes, uint32_t nseg,
                          const uint32_t* d_seg_of, const uint32_t* d_seg_first, const uint32_t* d_seg_count,
                          const uint32_t* d_node_base, ge_bvh_node* nodes, uint32_t* perm, uint32_t* morton,
                          float* roots, uint32_t* heights, const uint32_t* d_root_index, uint32_t node_pool,
                          WideNode* wide, uint32_t nseg_global, uint32_t n_global, const uint32_t* h_seg_count,
                          uint32_t* flag, uint32_t nparts) {
  if (n == 0) return hipSuccess;
  hipError_t e;
  // Segments [nseg_global, nseg) fit a workgroup's LDS: one launch builds them all.
  if (nseg_global < nseg) {
    BlockArgs ba{};
    ba.seg_base = nseg_global;
    ba.boxes = boxes;
    ba.seg_first = d_seg_first;
    ba.seg_count = d_seg_count;
    ba.node_base = d_node_base;
    ba.root_index = d_root_index;
    ba.perm = perm;
    ba.nodes = nodes;
    ba.morton = morton;
    ba.roots = roots;
    ba.heights = heights;
    ba.wide = wide;
    k_block_build<false><<<nseg - nseg_global, kBlockThreads, 0, stream>>>(ba);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (nseg_global == 0) return hipSuccess;
  n = n_global;
  nseg = nseg_global;
  const uint32_t tiles = blocks(n, kSortTile);
  if ((e = sc.keys[0].reserve(size_t(n) * 8)) != hipSuccess) return e;
  if ((e = sc.keys[1].reserve(size_t(n) * 8)) != hipSuccess) return e;
  if ((e = sc.vals[0].reserve(size_t(n) * 4)) != hipSuccess) return e;
  if ((e = sc.vals[1].reserve(size_t(n) * 4)) != hipSuccess) return e;
  if ((e = sc.hist.reserve(size_t(tiles) * 256 * 4)) != hipSuccess) return e;
  if ((e = sc.scan_tmp.reserve(size_t(tiles) * 256 * 4 / 256 + 65536)) != hipSuccess) return e;
  if ((e = sc.node_box.reserve(size_t(node_pool) * 32)) != hipSuccess) return e;
  if ((e = sc.node_parent.reserve(size_t(node_pool) * 4)) != hipSuccess) return e;
  if ((e = sc.depth.reserve(size_t(n) * 4)) != hipSuccess) return e;
  if ((e = sc.seg_cbox.reserve(size_t(nseg) * 24)) != hipSuccess) return e;
  if ((e = sc.seg_scale.reserve(size_t(nseg) * 24)) != hipSuccess) return e;

  // One segment small enough for one workgroup's LDS: bounds, codes, the sort, the codes and the
  // order written out, in one launch.
  const bool small = nseg == 1 && n <= kSmallMax;
  int cur = 0;
  if (small) {
    k_sort_small<<<1, kSmallThreads, 0, stream>>>(n, boxes, sc.keys[0].as<uint64_t>(), sc.vals[0].as<uint32_t>(), perm,
                                                  morton);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  } else {
    k_init_bounds<<<blocks(nseg, 256), 256, 0, stream>>>(nseg, sc.seg_cbox.as<uint32_t>());
    // Runs of 64 x steps items a wave: at least 2,048 waves, at most 64 steps a run.
    const uint32_t items = nparts ? nparts : n;
    const uint32_t steps = std::max(1u, std::min(kBoundsMaxSteps, items / (64u * 2048u)));
    const uint32_t bound_grid = blocks(uint64_t(blocks(items, 64u * steps)) * 64, 256);
    if (nparts)
      k_centroid_bounds<true><<<bound_grid, 256, 0, stream>>>(nparts, steps, nullptr, sc.cl_part.as<uint32_t>(),
                                                              sc.cl_seg.as<uint32_t>(), sc.seg_cbox.as<uint32_t>());
    else
      k_centroid_bounds<false><<<bound_grid, 256, 0, stream>>>(n, steps, boxes, nullptr, d_seg_of,
                                                               sc.seg_cbox.as<uint32_t>());
    k_segment_scale<<<blocks(nseg, 256), 256, 0, stream>>>(nseg, sc.seg_cbox.as<uint32_t>(), sc.seg_scale.as<float>());
    k_morton<<<blocks(n, 256), 256, 0, stream>>>(n, boxes, d_seg_of, sc.seg_scale.as<float>(), sc.keys[0].as<uint64_t>(),
                                                 sc.vals[0].as<uint32_t>());
    if ((e = hipGetLastError()) != hipSuccess) return e;
    uint32_t seg_bits = 0;
    while ((1u << seg_bits) < nseg) ++seg_bits;
    const int key_bits = 30 + int(seg_bits);
    for (int shift = 0; shift < key_bits; shift += 8) {
      k_sort_hist<<<tiles, kSortThreads, 0, stream>>>(sc.keys[cur].as<uint64_t>(), n, shift, sc.hist.as<uint32_t>(), tiles);
      if ((e = hipGetLastError()) != hipSuccess) return e;
      if ((e = exclusive_scan(sc.hist.as<uint32_t>(), tiles * 256, sc.scan_tmp, 0, stream)) != hipSuccess) return e;
      k_sort_scatter<true><<<tiles, kSortThreads, 0, stream>>>(sc.keys[cur].as<uint64_t>(), sc.vals[cur].as<uint32_t>(),
                                                         sc.keys[cur ^ 1].as<uint64_t>(), sc.vals[cur ^ 1].as<uint32_t>(),
                                                         n, shift, sc.hist.as<uint32_t>(), tiles);
      if ((e = hipGetLastError()) != hipSuccess) return e;
      cur ^= 1;
    }
  }
  const uint64_t* sorted_keys = sc.keys[cur].as<uint64_t>();
  const uint32_t* sorted_vals = sc.vals[cur].as<uint32_t>();
  TileLadder lad;
  if ((e = tile_ladder(sc, stream, nseg, h_seg_count, lad)) != hipSuccess) return e;
  TileArgs& ta = lad.args;
  ta.seg_first = d_seg_first;
  ta.seg_count = d_seg_count;
  ta.node_base = d_node_base;
  ta.root_index = d_root_index;
  ta.boxes = boxes;
  ta.nodes = nodes;
  ta.node_parent = sc.node_parent.as<uint32_t>();
  ta.node_box = sc.node_box.as<float>();
  ta.roots = roots;
  ta.heights = heights;
  ta.wide = wide;
  ta.flag = flag;
  if ((e = sc.links.reserve(size_t(n) * 16)) != hipSuccess) return e;
  ta.links = sc.links.as<uint4>();
  k_karras<<<blocks(n, 256), 256, 0, stream>>>(n, sorted_keys, d_seg_of, d_seg_first, d_seg_count, d_node_base, nodes,
                                               sc.node_parent.as<uint32_t>(), sc.links.as<uint4>(), ta);
  if (!small) {
    k_codes<<<blocks(n, 256), 256, 0, stream>>>(n, sorted_keys, morton);
    if ((e = hipMemcpyAsync(perm, sorted_vals, size_t(n) * 4, hipMemcpyDeviceToDevice, stream)) != hipSuccess) return e;
  }
  ta.perm = sorted_vals;
  return tile_fit(stream, lad);
}

hipError_t refit_segment(BuildScratch& sc, hipStream_t stream, uint32_t n, const float* slot_boxes,
                         const uint32_t* d_seg_tables, ge_bvh_node* nodes, WideNode* wide, float* root, uint32_t* height,
                         uint32_t* flag) {
  if (n < 2) return hipErrorInvalidValue;
  hipError_t e;
  if ((e = sc.node_box.reserve(size_t(n) * 32)) != hipSuccess) return e;
  if ((e = sc.node_parent.reserve(size_t(n) * 4)) != hipSuccess) return e;
  if ((e = sc.depth.reserve(size_t(n) * 4)) != hipSuccess) return e;
  if ((e = sc.vals[0].reserve(size_t(n) * 4)) != hipSuccess) return e;
  TileLadder lad;
  if ((e = tile_ladder(sc, stream, 1, &n, lad)) != hipSuccess) return e;
  TileArgs& ta = lad.args;
  ta.seg_first = d_seg_tables;
  ta.seg_count = d_seg_tables + 1;
  ta.node_base = d_seg_tables + 2;
  ta.root_index = nullptr;
  ta.boxes = slot_boxes;
  ta.nodes = nodes;
  ta.node_parent = sc.node_parent.as<uint32_t>();
  ta.node_box = sc.node_box.as<float>();
  ta.roots = root;
  ta.heights = height;
  ta.wide = wide;
  ta.flag = flag;
  k_identity<<<blocks(n, 256), 256, 0, stream>>>(n, sc.vals[0].as<uint32_t>());
  ta.perm = sc.vals[0].as<uint32_t>();
  if ((e = sc.links.reserve(size_t(n) * 16)) != hipSuccess) return e;
  ta.links = sc.links.as<uint4>();
  k_parents_from_records<<<blocks(n, 256), 256, 0, stream>>>(n, nodes, sc.node_parent.as<uint32_t>(), sc.links.as<uint4>());
  if (lad.levels > 1) k_file_records<<<blocks(n, 256), 256, 0, stream>>>(n, nodes, ta);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return tile_fit(stream, lad);
}

hipError_t launch_triangle_records(hipStream_t stream, const SceneDev& s, BuildScratch& sc, uint32_t nseg,
                                   uint32_t ncl, const uint32_t* pre, const GlobalSeg* segs, const uint32_t* lists,
                                   ge_bvh_triangle* tris, float* boxes, uint32_t* seg_of) {
  if (ncl == 0) return hipSuccess;
  hipError_t e;
  if ((e = sc.cl_out.reserve(size_t(ncl) * 4)) != hipSuccess) return e;
  if ((e = sc.cl_part.reserve(size_t(ncl) * 24)) != hipSuccess) return e;
  if ((e = sc.cl_seg.reserve(size_t(ncl) * 4)) != hipSuccess) return e;
  if ((e = sc.scan_tmp.reserve(size_t(ncl) / 256 + 65536)) != hipSuccess) return e;
  const ge_cluster_desc* clusters = s.clusters.as<ge_cluster_desc>();
  k_cluster_counts<<<blocks(ncl, 256), 256, 0, stream>>>(ncl, nseg, pre, segs, lists, clusters, sc.cl_out.as<uint32_t>());
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = exclusive_scan(sc.cl_out.as<uint32_t>(), ncl, sc.scan_tmp, 0, stream)) != hipSuccess) return e;
  k_triangle_records<<<ncl, 128, 0, stream>>>(nseg, pre, segs, lists, clusters, s.vertices.as<float>(),
                                              s.triangles.as<uint32_t>(), sc.cl_out.as<uint32_t>(), tris, boxes, seg_of,
                                              sc.cl_part.as<uint32_t>(), sc.cl_seg.as<uint32_t>());
  return hipGetLastError();
}

hipError_t launch_block_meshes(hipStream_t stream, const SceneDev& s, uint32_t nseg, const BlockSeg* segs,
                               const uint32_t* lists, ge_bvh_node* nodes, ge_bvh_triangle* tris, uint32_t* morton,
                               float* roots, uint32_t* heights, WideNode* wide) {
  if (nseg == 0) return hipSuccess;
  BlockArgs ba{};
  ba.segs = segs;
  ba.lists = lists;
  ba.clusters = s.clusters.as<ge_cluster_desc>();
  ba.vertices = s.vertices.as<float>();
  ba.triangles = s.triangles.as<uint32_t>();
  ba.tris = tris;
  ba.nodes = nodes;
  ba.morton = morton;
  ba.roots = roots;
  ba.heights = heights;
  ba.wide = wide;
  k_block_build<true><<<nseg, kBlockThreads, 0, stream>>>(ba);
  return hipGetLastError();
}

hipError_t launch_gather_triangles(hipStream_t stream, uint32_t n, const uint32_t* perm, const uint32_t* seg_of,
                                   const uint32_t* seg_first, const uint32_t* seg_out, const ge_bvh_triangle* in,
                                   ge_bvh_triangle* out, const uint32_t* morton_in, uint32_t* morton_out) {
  if (n == 0) return hipSuccess;
  k_gather_triangles<<<blocks(n, 256), 256, 0, stream>>>(n, perm, seg_of, seg_first, seg_out, in, out, morton_in,
                                                         morton_out);
  return hipGetLastError();
}

hipError_t launch_instance_boxes(hipStream_t stream, uint32_t n, const uint32_t* live, const ge_instance_desc* inst,
                                 const float* T, const float* roots, const uint32_t* root_of, float* boxes,
                                 uint32_t* seg_of) {
  if (n == 0) return hipSuccess;
  k_instance_boxes<<<blocks(n, 256), 256, 0, stream>>>(n, live, inst, T, roots, root_of, boxes, seg_of);
  return hipGetLastError();
}

hipError_t launch_gather_slots(hipStream_t stream, uint32_t n, const uint32_t* perm, const uint32_t* live,
                               uint32_t* slots) {
  if (n == 0) return hipSuccess;
  k_gather_slots<<<blocks(n, 256), 256, 0, stream>>>(n, perm, live, slots);
  return hipGetLastError();
}

}  // namespace ge
