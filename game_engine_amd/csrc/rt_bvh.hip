// rt_bvh.hip — the LBVH build on the GPU: segmented over many meshes in one set of launches.
//
// Per primitive: a box and the 30-bit Morton code of its centroid, quantized on its segment's
// centroid bounds (10 bits an axis). A stable LSD radix sort on (segment << 30 | code) orders each
// segment by code with ties in primitive order. Karras 2012 emits each segment's binary radix tree
// from the sorted codes (one thread per internal node, the index breaking equal codes). The boxes
// are fitted level by level from the deepest: each node's depth comes from pointer jumping over the
// parent links, one pass of the radix sort groups the nodes by depth, and one launch per level
// unions each node's two children, whose boxes the previous launch wrote. The kernel boundary is the
// only synchronization: no per-node atomics or device-scope fences, which on a part whose L2 is
// split over eight XCDs cost a cache write-back each. Everything is integer work or exact float
// min/max, so the structure is bit-identical to the oracle's sequential build.
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
// the batch (bounds_steps) so that small batches still spread over many waves.
constexpr uint32_t kBoundsMaxSteps = 64;

__global__ void __launch_bounds__(256) k_centroid_bounds(uint32_t n, uint32_t steps, const float* __restrict__ boxes,
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
    uint32_t c[3];
    for (int a = 0; a < 3; ++a)
      c[a] = valid ? ordered((boxes[size_t(i) * 6 + a] + boxes[size_t(i) * 6 + 3 + a]) * 0.5f) : 0u;
    const uint32_t s0 = uint32_t(__shfl(int(s), 0));
    if (__all(!valid || s == s0)) {
      if (s0 != cur) {
        flush();
        cur = s0;
      }
      if (valid) {
        for (int a = 0; a < 3; ++a) {
          lo[a] = min(lo[a], c[a]);
          hi[a] = max(hi[a], c[a]);
        }
      }
    } else {
      // A segment boundary inside this step: close the run and add these lanes one by one.
      flush();
      cur = 0xffffffffu;
      if (valid) {
        for (int a = 0; a < 3; ++a) {
          atomicMin(&cbox[s * 6 + a], c[a]);
          atomicMax(&cbox[s * 6 + 3 + a], c[a]);
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

template <bool Keys>
__global__ void __launch_bounds__(kSortThreads) k_sort_scatter(const uint64_t* __restrict__ kin,
                                                               const uint32_t* __restrict__ vin, uint64_t* kout,
                                                               uint32_t* vout, uint32_t n, int shift,
                                                               const uint32_t* __restrict__ offsets, uint32_t tiles) {
  __shared__ uint32_t running[256];
  __shared__ uint32_t wave_counts[kSortThreads / 64][256];
  const uint32_t t = threadIdx.x;
  const uint32_t lane = t & 63u, wave = t >> 6;
  running[t] = offsets[size_t(t) * tiles + blockIdx.x];
  for (int w = 0; w < kSortThreads / 64; ++w) wave_counts[w][t] = 0;
  __syncthreads();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint32_t base = blockIdx.x * kSortTile;
  for (int k = 0; k < kSortItems; ++k) {
    const uint32_t i = base + uint32_t(k) * kSortThreads + t;
    const bool valid = i < n;
    const uint64_t key = valid ? kin[i] : 0ull;
    const uint32_t d = uint32_t(key >> shift) & 255u;
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t bits = __ballot((d >> b) & 1u);
      match &= ((d >> b) & 1u) ? bits : ~bits;
    }
    const uint32_t below = uint32_t(__popcll(match & lt));
    if (valid && below == 0) wave_counts[wave][d] = uint32_t(__popcll(match));
    __syncthreads();
    if (valid) {
      uint32_t pos = running[d] + below;
      for (uint32_t w = 0; w < wave; ++w) pos += wave_counts[w][d];
      if (Keys) kout[pos] = key;
      vout[pos] = vin[i];
    }
    __syncthreads();
    uint32_t sum = 0;
    for (int w = 0; w < kSortThreads / 64; ++w) {
      sum += wave_counts[w][t];
      wave_counts[w][t] = 0;
    }
    running[t] += sum;
    __syncthreads();
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

// ---- Karras 2012 --------------------------------------------------------------------------------
__device__ __forceinline__ int delta(const uint64_t* __restrict__ keys, int n, int i, int j) {
  if (j < 0 || j > n - 1) return -1;
  const uint32_t a = uint32_t(keys[i]) & 0x3fffffffu, b = uint32_t(keys[j]) & 0x3fffffffu;
  if (a != b) return __clz(a ^ b);
  return 32 + __clz(uint32_t(i) ^ uint32_t(j));
}

__global__ void k_karras(uint32_t n_total, const uint64_t* __restrict__ sorted_keys, const uint32_t* __restrict__ seg_of,
                         const uint32_t* __restrict__ seg_first, const uint32_t* __restrict__ seg_count,
                         const uint32_t* __restrict__ node_base, ge_bvh_node* nodes, uint32_t* node_parent) {
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
  ge_bvh_node& node = seg_nodes[i];
  const uint32_t c0 = (lo == gamma) ? (GE_BVH_LEAF | uint32_t(gamma)) : uint32_t(gamma);
  const uint32_t c1 = (hi == gamma + 1) ? (GE_BVH_LEAF | uint32_t(gamma + 1)) : uint32_t(gamma + 1);
  node.child[0] = c0;
  node.child[1] = c1;
  node.count = uint32_t(hi - lo + 1);
  if (i == 0) node.parent = GE_BVH_NONE;
  if (!(c0 & GE_BVH_LEAF)) node_parent[node_base[s] + gamma] = uint32_t(i);
  if (!(c1 & GE_BVH_LEAF)) node_parent[node_base[s] + gamma + 1] = uint32_t(i);
}


// ---- the fit, level by level ----------------------------------------------------------------
// A node is named by the batch index of its segment's primitive of the same index: internal node i
// of segment s is batch entry seg_first[s] + i (i < n - 1, or i = 0 for a one-primitive segment),
// so a batch of n primitives has at most n nodes and every array below is n long.
constexpr uint32_t kNoNode = 0xffffffffu;
constexpr uint32_t kNoDepth = 0xffffffffu;  // the distance of a batch entry that names no node
constexpr uint32_t kNotANode = 255u;        // its depth key
constexpr int kJumps = 8;                   // pointer-jumping rounds: depths up to 2^8 - 1

__global__ void k_depth_init(uint32_t n_total, const uint32_t* __restrict__ seg_of,
                             const uint32_t* __restrict__ seg_first, const uint32_t* __restrict__ seg_count,
                             const uint32_t* __restrict__ node_base, const uint32_t* __restrict__ node_parent,
                             uint32_t* anc, uint32_t* dist) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_total) return;
  const uint32_t s = seg_of[g];
  const uint32_t first_g = seg_first[s], n = seg_count[s];
  const uint32_t i = g - first_g;
  const bool node = n == 1 ? i == 0 : i < n - 1;
  uint32_t a = kNoNode;
  if (node && i != 0) a = first_g + node_parent[node_base[s] + i];
  anc[g] = a;
  dist[g] = node ? (a == kNoNode ? 0u : 1u) : kNoDepth;
}

// One round of pointer jumping: the distance to the ancestor doubles its reach.
__global__ void k_depth_jump(uint32_t n_total, const uint32_t* __restrict__ anc, const uint32_t* __restrict__ dist,
                             uint32_t* anc_out, uint32_t* dist_out) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_total) return;
  const uint32_t a = anc[g];
  uint32_t d = dist[g];
  uint32_t up = kNoNode;
  if (a != kNoNode) {
    d += dist[a];
    up = anc[a];
  }
  anc_out[g] = up;
  dist_out[g] = d;
}

// The sort keys: a node's depth (kNotANode for the rest); a node short of its root after kJumps
// rounds, or as deep as kNotANode, is deeper than any structure the traversal accepts and raises
// the error flag.
__global__ void k_depth_keys(uint32_t n_total, const uint32_t* __restrict__ anc, const uint32_t* __restrict__ dist,
                             uint64_t* keys, uint32_t* vals, uint32_t* error) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_total) return;
  uint32_t d = dist[g];
  if (d == kNoDepth) {
    d = kNotANode;
  } else if (anc[g] != kNoNode || d >= kNotANode) {
    atomicOr(error, 1u);
    d = kNotANode - 1;
  }
  keys[g] = d;
  vals[g] = g;
}

// Each level's first position in the depth-sorted order: the exclusive-scanned histogram's entry
// for (digit d, tile 0). `out` gets 256 starts, then the error flag.
__global__ void k_level_starts(const uint32_t* __restrict__ scanned, uint32_t tiles, const uint32_t* error,
                               uint32_t* out) {
  const uint32_t d = threadIdx.x;
  out[d] = scanned[size_t(d) * tiles];
  if (d == 0) out[256] = *error;
}

__device__ __forceinline__ void leaf_box(const float* boxes, const uint32_t* perm, uint32_t g, float lo[3], float hi[3]) {
  const uint32_t p = perm[g];
  for (int a = 0; a < 3; ++a) {
    lo[a] = boxes[size_t(p) * 6 + a];
    hi[a] = boxes[size_t(p) * 6 + 3 + a];
  }
}

// One level of the fit: each node unions its children (a leaf's box, or the box and height the
// previous launch wrote for an internal child) into its record, its box and its height.
__global__ void k_fit_level(const uint32_t* __restrict__ level, uint32_t count, const float* __restrict__ boxes,
                            const uint32_t* __restrict__ perm, const uint32_t* __restrict__ seg_of,
                            const uint32_t* __restrict__ seg_first, const uint32_t* __restrict__ seg_count,
                            const uint32_t* __restrict__ node_base, const uint32_t* __restrict__ root_index,
                            ge_bvh_node* nodes, const uint32_t* __restrict__ node_parent, float* node_box,
                            float* roots, uint32_t* heights) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const uint32_t g = level[k];
  const uint32_t s = seg_of[g];
  const uint32_t first_g = seg_first[s];
  const uint32_t nb = node_base[s];
  const uint32_t i = g - first_g;
  const uint32_t out = root_index ? root_index[s] : s;
  ge_bvh_node& node = nodes[nb + i];
  if (seg_count[s] == 1) {
    float lo[3], hi[3];
    leaf_box(boxes, perm, first_g, lo, hi);
    for (int a = 0; a < 3; ++a) {
      node.lo0[a] = lo[a];
      node.hi0[a] = hi[a];
      roots[out * 6 + a] = lo[a];
      roots[out * 6 + 3 + a] = hi[a];
    }
    heights[out] = 1;
    return;
  }
  float clo[2][3], chi[2][3];
  uint32_t ch_height[2];
  for (int c = 0; c < 2; ++c) {
    const uint32_t ch = node.child[c];
    if (ch & GE_BVH_LEAF) {
      leaf_box(boxes, perm, first_g + (ch & ~GE_BVH_LEAF), clo[c], chi[c]);
      ch_height[c] = 0;
    } else {
      const float* b = node_box + size_t(nb + ch) * 8;
      for (int a = 0; a < 3; ++a) {
        clo[c][a] = b[a];
        chi[c][a] = b[3 + a];
      }
      ch_height[c] = __float_as_uint(b[6]);
    }
  }
  for (int a = 0; a < 3; ++a) {
    node.lo0[a] = clo[0][a];
    node.hi0[a] = chi[0][a];
    node.lo1[a] = clo[1][a];
    node.hi1[a] = chi[1][a];
  }
  const uint32_t h = 1 + (ch_height[0] > ch_height[1] ? ch_height[0] : ch_height[1]);
  float* b = node_box + size_t(nb + i) * 8;
  for (int a = 0; a < 3; ++a) {
    b[a] = __builtin_fminf(clo[0][a], clo[1][a]);
    b[3 + a] = __builtin_fmaxf(chi[0][a], chi[1][a]);
  }
  b[6] = __uint_as_float(h);
  if (i == 0) {
    node.parent = GE_BVH_NONE;
    for (int a = 0; a < 6; ++a) roots[out * 6 + a] = b[a];
    heights[out] = h;
    return;
  }
  node.parent = node_parent[nb + i];
}

__global__ void k_triangle_records(const ge_cluster_desc* __restrict__ clusters, const float* __restrict__ vertices,
                                   const uint32_t* __restrict__ triangles, const uint32_t* __restrict__ cluster_list,
                                   const uint32_t* __restrict__ cluster_out, const uint32_t* __restrict__ cluster_seg,
                                   const uint32_t* __restrict__ cluster_local, ge_bvh_triangle* tris, float* boxes,
                                   uint32_t* seg_of) {
  const uint32_t k = blockIdx.x;
  const uint32_t c = cluster_list[k];
  const ge_cluster_desc cl = clusters[c];
  const uint32_t out0 = cluster_out[k];
  const uint32_t local_cluster = cluster_local[k];
  const uint32_t s = cluster_seg[k];
  for (uint32_t t = threadIdx.x; t < cl.triangle_count; t += blockDim.x) {
    const uint32_t packed = triangles[cl.triangle_offset + t];
    ge_bvh_triangle tri;
    float* v[3] = {tri.v0, tri.v1, tri.v2};
    for (int q = 0; q < 3; ++q) {
      const uint32_t vi = cl.vertex_offset + ((packed >> (8 * q)) & 0xffu);
      v[q][0] = vertices[size_t(vi) * 3];
      v[q][1] = vertices[size_t(vi) * 3 + 1];
      v[q][2] = vertices[size_t(vi) * 3 + 2];
    }
    const uint32_t o = out0 + t;
    tri.prim = (local_cluster << 8) | t;
    tri.index = o;  // rewritten to the mesh-local index by the gather
    tri.pad = 0;
    tris[o] = tri;
    for (int a = 0; a < 3; ++a) {
      boxes[size_t(o) * 6 + a] = __builtin_fminf(__builtin_fminf(tri.v0[a], tri.v1[a]), tri.v2[a]);
      boxes[size_t(o) * 6 + 3 + a] = __builtin_fmaxf(__builtin_fmaxf(tri.v0[a], tri.v1[a]), tri.v2[a]);
    }
    seg_of[o] = s;
  }
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

__global__ void k_instance_boxes(uint32_t n, const uint32_t* __restrict__ live, const ge_instance_desc* __restrict__ inst,
                                 const float* __restrict__ T, const float* __restrict__ roots, float* boxes,
                                 uint32_t* seg_of) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t id = live[k];
  const ge_instance_desc in = inst[id];
  const float* b = roots + size_t(in.mesh) * 6;
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

// The 4-wide collapse: each node at even depth gathers its children's children (a leaf child taken
// as it is), in child order, with their boxes, into the wide node at its own binary index.
__global__ void k_collapse(uint32_t n_total, const uint32_t* __restrict__ seg_of, const uint32_t* __restrict__ seg_first,
                           const uint32_t* __restrict__ node_base, const uint32_t* __restrict__ dist,
                           const ge_bvh_node* __restrict__ nodes, WideNode* __restrict__ wide) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_total) return;
  const uint32_t d = dist[g];
  if (d == kNoDepth || (d & 1u)) return;
  const uint32_t s = seg_of[g];
  const uint32_t nb = node_base[s];
  const uint32_t i = g - seg_first[s];
  const ge_bvh_node a = nodes[nb + i];
  WideNode w;
  int k = 0;
  auto put = [&](uint32_t ref, const float* lo, const float* hi) {
    w.child[k] = ref;
    w.lox[k] = lo[0];
    w.loy[k] = lo[1];
    w.loz[k] = lo[2];
    w.hix[k] = hi[0];
    w.hiy[k] = hi[1];
    w.hiz[k] = hi[2];
    ++k;
  };
  for (int c = 0; c < 2; ++c) {
    const uint32_t ch = a.child[c];
    if (ch == GE_BVH_NONE) continue;
    if (ch & GE_BVH_LEAF) {
      put(ch, c == 0 ? a.lo0 : a.lo1, c == 0 ? a.hi0 : a.hi1);
      continue;
    }
    const ge_bvh_node m = nodes[nb + ch];
    for (int e = 0; e < 2; ++e)
      if (m.child[e] != GE_BVH_NONE) put(m.child[e], e == 0 ? m.lo0 : m.lo1, e == 0 ? m.hi0 : m.hi1);
  }
  // An empty slot's box sits at +infinity on every axis: every slab test against it fails (its near
  // distance is +inf, or both its distances -inf), so the walk needs no test of the reference.
  for (; k < 4; ++k) {
    w.child[k] = GE_BVH_NONE;
    w.lox[k] = w.loy[k] = w.loz[k] = w.hix[k] = w.hiy[k] = w.hiz[k] = __builtin_huge_valf();
  }
  for (int q = 0; q < 4; ++q) w.pad[q] = 0;
  wide[nb + i] = w;
}

__global__ void k_gather_slots(uint32_t n, const uint32_t* perm, const uint32_t* live, uint32_t* slots) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) slots[g] = live[perm[g]];
}

// ---- a whole mesh's structure in one workgroup --------------------------------------------------
// A segment of at most kBlockMax primitives is built by one workgroup from start to end in LDS: its
// centroid bounds, Morton codes, a bitonic sort of (code, index) (the stable order of the global
// sort), Karras' emission, the fit bottom-up (the second child to finish a node fits it: LDS
// counters, a workgroup-scope fence between a box's store and the counter), the binary nodes, each
// node's depth and the 4-wide collapse. One launch builds every small mesh of a batch, where the
// device-wide passes cost a launch a level; the results are the device-wide build's to the bit.
constexpr uint32_t kBlockMax = k_block_build_max;
constexpr uint32_t kBlockThreads = 256;
constexpr uint16_t kBlockLeaf = 0x8000u;

__global__ void __launch_bounds__(kBlockThreads) k_block_build(
    uint32_t seg_base, const float* __restrict__ boxes, const uint32_t* __restrict__ seg_first,
    const uint32_t* __restrict__ seg_count, const uint32_t* __restrict__ node_base,
    const uint32_t* __restrict__ root_index, ge_bvh_node* __restrict__ nodes, uint32_t* __restrict__ perm,
    uint32_t* __restrict__ morton, float* __restrict__ roots, uint32_t* __restrict__ heights,
    WideNode* __restrict__ wide) {
  __shared__ uint64_t keys[kBlockMax];     // (code << 32 | local index), sorted in place
  __shared__ float nbox[kBlockMax][6];     // internal nodes' boxes
  __shared__ uint16_t child[kBlockMax][2]; // internal children; kBlockLeaf | sorted position for a leaf
  __shared__ uint16_t parent[kBlockMax];   // internal node -> its parent
  __shared__ uint16_t leaf_parent[kBlockMax];
  __shared__ uint32_t arrived[kBlockMax];  // fit counters
  __shared__ uint8_t height[kBlockMax];
  __shared__ uint32_t red[6];
  const uint32_t t = threadIdx.x;
  const uint32_t s = seg_base + blockIdx.x;
  const uint32_t first = seg_first[s], n = seg_count[s], nb = node_base[s];
  const uint32_t out = root_index ? root_index[s] : s;
  const float* sb = boxes + size_t(first) * 6;
  auto leaf_box = [&](uint32_t sorted, float lo[3], float hi[3]) {
    const uint32_t o = uint32_t(keys[sorted]);
    for (int a = 0; a < 3; ++a) {
      lo[a] = sb[size_t(o) * 6 + a];
      hi[a] = sb[size_t(o) * 6 + 3 + a];
    }
  };

  // Centroid bounds, as order-preserving integers (min and max are the float ones).
  if (t < 6) red[t] = t < 3 ? 0xffffffffu : 0u;
  __syncthreads();
  {
    uint32_t lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
    for (uint32_t i = t; i < n; i += kBlockThreads)
      for (int a = 0; a < 3; ++a) {
        const uint32_t c = ordered((sb[size_t(i) * 6 + a] + sb[size_t(i) * 6 + 3 + a]) * 0.5f);
        lo[a] = min(lo[a], c);
        hi[a] = max(hi[a], c);
      }
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
  // Codes, and the bitonic sort over the next power of two (padding keys sort last).
  uint32_t p2 = 1;
  while (p2 < n) p2 <<= 1;
  for (uint32_t i = t; i < p2; i += kBlockThreads) {
    if (i < n) {
      uint32_t q[3];
      for (int a = 0; a < 3; ++a) {
        const float c = (sb[size_t(i) * 6 + a] + sb[size_t(i) * 6 + 3 + a]) * 0.5f;
        q[a] = quantize10(c, clo[a], cscale[a]);
      }
      const uint32_t code = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
      keys[i] = (uint64_t(code) << 32) | i;
    } else {
      keys[i] = ~0ull;
    }
  }
  __syncthreads();
  for (uint32_t k = 2; k <= p2; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = t; i < p2; i += kBlockThreads) {
        const uint32_t l = i ^ j;
        if (l > i) {
          const uint64_t a = keys[i], b = keys[l];
          const bool up = (i & k) == 0;
          if ((a > b) == up) {
            keys[i] = b;
            keys[l] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (uint32_t i = t; i < n; i += kBlockThreads) {
    perm[first + i] = first + uint32_t(keys[i]);
    morton[first + i] = uint32_t(keys[i] >> 32);
    arrived[i] = 0;
  }
  if (n == 1) {
    if (t == 0) {
      float lo[3], hi[3];
      leaf_box(0, lo, hi);
      ge_bvh_node node;
      node.child[0] = GE_BVH_LEAF;
      node.child[1] = GE_BVH_NONE;
      node.parent = GE_BVH_NONE;
      node.count = 1;
      WideNode w;
      for (int a = 0; a < 3; ++a) {
        node.lo0[a] = lo[a];
        node.hi0[a] = hi[a];
        node.lo1[a] = __uint_as_float(0x7f800000u);
        node.hi1[a] = __uint_as_float(0xff800000u);
        roots[out * 6 + a] = lo[a];
        roots[out * 6 + 3 + a] = hi[a];
      }
      nodes[nb] = node;
      heights[out] = 1;
      for (int k = 0; k < 4; ++k) {
        w.child[k] = k == 0 ? GE_BVH_LEAF : GE_BVH_NONE;
        w.lox[k] = k == 0 ? lo[0] : __builtin_huge_valf();
        w.loy[k] = k == 0 ? lo[1] : __builtin_huge_valf();
        w.loz[k] = k == 0 ? lo[2] : __builtin_huge_valf();
        w.hix[k] = k == 0 ? hi[0] : __builtin_huge_valf();
        w.hiy[k] = k == 0 ? hi[1] : __builtin_huge_valf();
        w.hiz[k] = k == 0 ? hi[2] : __builtin_huge_valf();
        w.pad[k] = 0;
      }
      wide[nb] = w;
    }
    return;
  }
  // Karras 2012 over the sorted codes, equal codes split by position.
  auto delta = [&](int i, int j) -> int {
    if (j < 0 || j > int(n) - 1) return -1;
    const uint32_t a = uint32_t(keys[i] >> 32), b = uint32_t(keys[j] >> 32);
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
    if (l0) leaf_parent[gamma] = uint16_t(i); else parent[gamma] = uint16_t(i);
    if (l1) leaf_parent[gamma + 1] = uint16_t(i); else parent[gamma + 1] = uint16_t(i);
  }
  __syncthreads();
  // The fit: every leaf walks up; the second arrival at a node fits it and goes on.
  for (uint32_t leaf = t; leaf < n; leaf += kBlockThreads) {
    uint32_t p = leaf_parent[leaf];
    for (;;) {
      __threadfence_block();
      if (atomicAdd(&arrived[p], 1u) == 0u) break;
      float b[2][6];
      uint32_t h[2];
      for (int c = 0; c < 2; ++c) {
        const uint16_t ch = child[p][c];
        if (ch & kBlockLeaf) {
          leaf_box(ch & ~kBlockLeaf, &b[c][0], &b[c][3]);
          h[c] = 0;
        } else {
          for (int a = 0; a < 6; ++a) b[c][a] = nbox[ch][a];
          h[c] = height[ch];
        }
      }
      for (int a = 0; a < 3; ++a) {
        nbox[p][a] = __builtin_fminf(b[0][a], b[1][a]);
        nbox[p][3 + a] = __builtin_fmaxf(b[0][3 + a], b[1][3 + a]);
      }
      height[p] = uint8_t(1 + (h[0] > h[1] ? h[0] : h[1]));
      if (p == 0) break;
      p = parent[p];
    }
  }
  __syncthreads();
  // The binary nodes, the root, and each node's depth for the collapse.
  for (uint32_t i = t; i + 1 < n; i += kBlockThreads) {
    ge_bvh_node node;
    for (int c = 0; c < 2; ++c) {
      const uint16_t ch = child[i][c];
      float lo[3], hi[3];
      if (ch & kBlockLeaf) {
        leaf_box(ch & ~kBlockLeaf, lo, hi);
        node.child[c] = GE_BVH_LEAF | uint32_t(ch & ~kBlockLeaf);
      } else {
        for (int a = 0; a < 3; ++a) {
          lo[a] = nbox[ch][a];
          hi[a] = nbox[ch][3 + a];
        }
        node.child[c] = ch;
      }
      float* dlo = c == 0 ? node.lo0 : node.lo1;
      float* dhi = c == 0 ? node.hi0 : node.hi1;
      for (int a = 0; a < 3; ++a) {
        dlo[a] = lo[a];
        dhi[a] = hi[a];
      }
    }
    // The primitives below: the node's key range, from its children.
    uint32_t lo_i = 0, hi_i = 0;
    {
      // Leftmost and rightmost leaves: descend the first and the last child.
      uint16_t c = child[i][0];
      while (!(c & kBlockLeaf)) c = child[c][0];
      lo_i = c & ~kBlockLeaf;
      c = child[i][1];
      while (!(c & kBlockLeaf)) c = child[c][1];
      hi_i = c & ~kBlockLeaf;
    }
    node.count = hi_i - lo_i + 1;
    node.parent = i == 0 ? GE_BVH_NONE : uint32_t(parent[i]);
    nodes[nb + i] = node;
  }
  if (t == 0) {
    for (int a = 0; a < 6; ++a) roots[out * 6 + a] = nbox[0][a];
    heights[out] = height[0];
  }
  for (uint32_t i = t; i + 1 < n; i += kBlockThreads) {
    uint32_t d = 0;
    for (uint32_t q = i; q != 0; q = parent[q]) ++d;
    if (d & 1u) continue;
    WideNode w;
    int k = 0;
    auto put = [&](uint32_t ref, const float* lo, const float* hi) {
      w.child[k] = ref;
      w.lox[k] = lo[0];
      w.loy[k] = lo[1];
      w.loz[k] = lo[2];
      w.hix[k] = hi[0];
      w.hiy[k] = hi[1];
      w.hiz[k] = hi[2];
      ++k;
    };
    auto take = [&](uint16_t ch) {
      float lo[3], hi[3];
      if (ch & kBlockLeaf) {
        leaf_box(ch & ~kBlockLeaf, lo, hi);
        put(GE_BVH_LEAF | uint32_t(ch & ~kBlockLeaf), lo, hi);
      } else {
        put(ch, &nbox[ch][0], &nbox[ch][3]);
      }
    };
    for (int c = 0; c < 2; ++c) {
      const uint16_t ch = child[i][c];
      if (ch & kBlockLeaf) {
        take(ch);
        continue;
      }
      take(child[ch][0]);
      take(child[ch][1]);
    }
    for (; k < 4; ++k) {
      w.child[k] = GE_BVH_NONE;
      w.lox[k] = w.loy[k] = w.loz[k] = w.hix[k] = w.hiy[k] = w.hiz[k] = __builtin_huge_valf();
    }
    for (int q = 0; q < 4; ++q) w.pad[q] = 0;
    wide[nb + i] = w;
  }
}

inline uint32_t blocks(uint64_t n, uint32_t t) { return uint32_t((n + t - 1) / t); }

}  // namespace

hipError_t build_segments(BuildScratch& sc, hipStream_t stream, uint32_t n, const float* boxes, uint32_t nseg,
                          const uint32_t* d_seg_of, const uint32_t* d_seg_first, const uint32_t* d_seg_count,
                          const uint32_t* d_node_base, ge_bvh_node* nodes, uint32_t* perm, uint32_t* morton,
                          float* roots, uint32_t* heights, const uint32_t* d_root_index, uint32_t node_pool,
                          WideNode* wide, uint32_t nseg_global, uint32_t n_global) {
  if (n == 0) return hipSuccess;
  hipError_t e;
  // Segments [nseg_global, nseg) fit a workgroup's LDS: one launch builds them all.
  if (nseg_global < nseg) {
    k_block_build<<<nseg - nseg_global, kBlockThreads, 0, stream>>>(nseg_global, boxes, d_seg_first, d_seg_count,
                                                                    d_node_base, d_root_index, nodes, perm, morton,
                                                                    roots, heights, wide);
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
  for (int k = 0; k < 2; ++k) {
    if ((e = sc.anc[k].reserve(size_t(n) * 4)) != hipSuccess) return e;
    if ((e = sc.dist[k].reserve(size_t(n) * 4)) != hipSuccess) return e;
  }
  if ((e = sc.levels.reserve(260 * 4)) != hipSuccess) return e;
  if (!sc.host_levels && (e = hipHostMalloc(reinterpret_cast<void**>(&sc.host_levels), 260 * 4)) != hipSuccess) return e;
  if ((e = sc.seg_cbox.reserve(size_t(nseg) * 24)) != hipSuccess) return e;
  if ((e = sc.seg_scale.reserve(size_t(nseg) * 24)) != hipSuccess) return e;

  k_init_bounds<<<blocks(nseg, 256), 256, 0, stream>>>(nseg, sc.seg_cbox.as<uint32_t>());
  // Runs of 64 x steps primitives a wave: at least 2,048 waves, at most 64 steps a run.
  const uint32_t steps = std::max(1u, std::min(kBoundsMaxSteps, n / (64u * 2048u)));
  const uint32_t bound_waves = blocks(n, 64u * steps);
  k_centroid_bounds<<<blocks(uint64_t(bound_waves) * 64, 256), 256, 0, stream>>>(n, steps, boxes, d_seg_of,
                                                                                  sc.seg_cbox.as<uint32_t>());
  k_segment_scale<<<blocks(nseg, 256), 256, 0, stream>>>(nseg, sc.seg_cbox.as<uint32_t>(), sc.seg_scale.as<float>());
  k_morton<<<blocks(n, 256), 256, 0, stream>>>(n, boxes, d_seg_of, sc.seg_scale.as<float>(), sc.keys[0].as<uint64_t>(),
                                               sc.vals[0].as<uint32_t>());
  if ((e = hipGetLastError()) != hipSuccess) return e;

  uint32_t seg_bits = 0;
  while ((1u << seg_bits) < nseg) ++seg_bits;
  const int key_bits = 30 + int(seg_bits);
  int cur = 0;
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
  const uint64_t* sorted_keys = sc.keys[cur].as<uint64_t>();
  const uint32_t* sorted_vals = sc.vals[cur].as<uint32_t>();
  k_karras<<<blocks(n, 256), 256, 0, stream>>>(n, sorted_keys, d_seg_of, d_seg_first, d_seg_count, d_node_base, nodes,
                                               sc.node_parent.as<uint32_t>());
  k_codes<<<blocks(n, 256), 256, 0, stream>>>(n, sorted_keys, morton);
  if ((e = hipMemcpyAsync(perm, sorted_vals, size_t(n) * 4, hipMemcpyDeviceToDevice, stream)) != hipSuccess) return e;

  // ---- the fit: depths by pointer jumping, nodes grouped by depth, one launch per level ----
  uint32_t* error = sc.levels.as<uint32_t>() + 257;
  if ((e = hipMemsetAsync(error, 0, 4, stream)) != hipSuccess) return e;
  k_depth_init<<<blocks(n, 256), 256, 0, stream>>>(n, d_seg_of, d_seg_first, d_seg_count, d_node_base,
                                                   sc.node_parent.as<uint32_t>(), sc.anc[0].as<uint32_t>(),
                                                   sc.dist[0].as<uint32_t>());
  int jc = 0;
  for (int r = 0; r < kJumps; ++r, jc ^= 1)
    k_depth_jump<<<blocks(n, 256), 256, 0, stream>>>(n, sc.anc[jc].as<uint32_t>(), sc.dist[jc].as<uint32_t>(),
                                                     sc.anc[jc ^ 1].as<uint32_t>(), sc.dist[jc ^ 1].as<uint32_t>());
  // The Morton keys and order stay in keys[cur] / vals[cur]; the depth sort uses the other pair
  // and its own value buffer.
  uint64_t* dkeys = sc.keys[cur ^ 1].as<uint64_t>();
  uint32_t* dvals = sc.anc[jc ^ 1].as<uint32_t>();
  uint32_t* dsorted = sc.vals[cur ^ 1].as<uint32_t>();
  k_depth_keys<<<blocks(n, 256), 256, 0, stream>>>(n, sc.anc[jc].as<uint32_t>(), sc.dist[jc].as<uint32_t>(), dkeys,
                                                   dvals, error);
  k_sort_hist<<<tiles, kSortThreads, 0, stream>>>(dkeys, n, 0, sc.hist.as<uint32_t>(), tiles);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = exclusive_scan(sc.hist.as<uint32_t>(), tiles * 256, sc.scan_tmp, 0, stream)) != hipSuccess) return e;
  // One 8-bit pass orders the nodes by depth (the sorted keys themselves are not kept).
  k_sort_scatter<false><<<tiles, kSortThreads, 0, stream>>>(dkeys, dvals, nullptr, dsorted, n, 0,
                                                            sc.hist.as<uint32_t>(), tiles);
  k_level_starts<<<1, 256, 0, stream>>>(sc.hist.as<uint32_t>(), tiles, error, sc.levels.as<uint32_t>());
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(sc.host_levels, sc.levels.ptr, 257 * 4, hipMemcpyDeviceToHost, stream)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
  if (sc.host_levels[256]) return hipErrorNotSupported;  // deeper than 2^8 - 2 levels
  for (int d = int(kNotANode) - 1; d >= 0; --d) {
    const uint32_t start = sc.host_levels[d];
    const uint32_t end = sc.host_levels[d + 1];
    if (end <= start) continue;
    k_fit_level<<<blocks(end - start, 256), 256, 0, stream>>>(dsorted + start, end - start, boxes, sorted_vals,
                                                              d_seg_of, d_seg_first, d_seg_count, d_node_base,
                                                              d_root_index, nodes, sc.node_parent.as<uint32_t>(),
                                                              sc.node_box.as<float>(), roots, heights);
  }
  // The depths are still those of the pointer jumping.
  if (wide)
    k_collapse<<<blocks(n, 256), 256, 0, stream>>>(n, d_seg_of, d_seg_first, d_node_base, sc.dist[jc].as<uint32_t>(),
                                                   nodes, wide);
  return hipGetLastError();
}

hipError_t launch_triangle_records(hipStream_t stream, const SceneDev& s, uint32_t ncl, const uint32_t* cluster_list,
                                   const uint32_t* cluster_out, const uint32_t* cluster_seg, const uint32_t* cluster_local,
                                   ge_bvh_triangle* tris, float* boxes, uint32_t* seg_of) {
  if (ncl == 0) return hipSuccess;
  k_triangle_records<<<ncl, 128, 0, stream>>>(s.clusters.as<ge_cluster_desc>(), s.vertices.as<float>(),
                                               s.triangles.as<uint32_t>(), cluster_list, cluster_out, cluster_seg,
                                               cluster_local, tris, boxes, seg_of);
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
                                 const float* T, const float* roots, float* boxes, uint32_t* seg_of) {
  if (n == 0) return hipSuccess;
  k_instance_boxes<<<blocks(n, 256), 256, 0, stream>>>(n, live, inst, T, roots, boxes, seg_of);
  return hipGetLastError();
}

hipError_t launch_gather_slots(hipStream_t stream, uint32_t n, const uint32_t* perm, const uint32_t* live,
                               uint32_t* slots) {
  if (n == 0) return hipSuccess;
  k_gather_slots<<<blocks(n, 256), 256, 0, stream>>>(n, perm, live, slots);
  return hipGetLastError();
}

}  // namespace ge
