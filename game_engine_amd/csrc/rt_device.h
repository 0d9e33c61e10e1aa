// rt_device.h — device code shared by the backend's kernels: vector arithmetic in the shaders'
// operation order, the random stream, the robust box test, the watertight triangle test and the
// two-level traversal.
//
// Floating point: every kernel is compiled with -ffp-contract=off and correctly rounded division
// and square root (hipcc's default for HIP), so the arithmetic below is IEEE single precision
// operation for operation and a hit, a Morton code or a radiance sum is the same float on every
// launch. Nothing here is fused or reassociated.
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ge_rt.h"

#define GE_DEV __device__ __forceinline__

namespace ge {

struct f3 {
  float x, y, z;
};
GE_DEV f3 mk(float x, float y, float z) { return f3{x, y, z}; }
GE_DEV f3 splat(float s) { return f3{s, s, s}; }
GE_DEV f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
GE_DEV f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
GE_DEV f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
GE_DEV f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
GE_DEV f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
GE_DEV f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
GE_DEV f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
GE_DEV float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
GE_DEV f3 cross(f3 a, f3 b) { return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
GE_DEV f3 normalize(f3 v) { return v * (1.0f / __builtin_sqrtf(dot(v, v))); }
GE_DEV float rsqrt_cr(float x) { return 1.0f / __builtin_sqrtf(x); }
// HLSL max/min/clamp on non-NaN operands.
GE_DEV float maxr(float a, float b) { return a > b ? a : b; }
GE_DEV float minr(float a, float b) { return a < b ? a : b; }
GE_DEV float clampr(float v, float lo, float hi) {
  const float m = v < lo ? lo : v;
  return hi < m ? hi : m;
}
GE_DEV float lerp1(float a, float b, float t) { return a + (b - a) * t; }
GE_DEV f3 lerp3(f3 a, f3 b, float t) { return f3{lerp1(a.x, b.x, t), lerp1(a.y, b.y, t), lerp1(a.z, b.z, t)}; }
GE_DEV f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }

constexpr float k_pi = 3.14159265358979323846f;

// sin and cos in float arithmetic alone, accurate to within one ulp of the exact value for
// |x| <= 64 (callers pass angles in [0, 2 pi]).
// Reduction (Cody and Waite 1980): q = rint(x 2/pi), r = x - q pi/2 with pi/2 split into
// C1 + C2 + C3, C1 and C2 of 16 significant bits so q*C1 and q*C2 are exact for |q| < 2^8; x - q*C1
// is exact (Sterbenz), the subtraction of q*C2 keeps its rounding error (Knuth's TwoSum), and
// r = r_hi + r_lo to about 2^-48 relative.
// Polynomials on [-pi/4, pi/4], z = r_hi^2: sin r = r + r^3 S(z) and cos r = 1 - z/2 + z^2 K(z),
// S and K of degree 3, minimax in relative error (a Remez exchange, game_engine_amd/
// derive_sincos.py; 3.7e-9 and 7.0e-10 with these float coefficients). r_lo enters to first order
// (sin: r_lo (1 - z/2); cos: -r_hi r_lo), and cos's 1 - z/2 is summed with its rounding error
// (Kahan's trick) and z's own. The result's one final rounding dominates the error.
GE_DEV void sincos_f32(float x, float* s, float* c) {
  const float q = __builtin_rintf(x * 0.636619747f);
  const float r1 = __builtin_fmaf(-q, 1.57080078125f, x);
  const float p2 = q * -4.45439946e-06f;
  const float rh = r1 - p2;
  const float bb = rh - r1;
  const float e2 = (r1 - (rh - bb)) - (p2 + bb);
  const float rl = __builtin_fmaf(-q, -5.56443156e-11f, e2);
  const float z = rh * rh;
  const float ze = __builtin_fmaf(rh, rh, -z);
  const float sp = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(2.71812155e-06f, z, -1.98393129e-04f), z,
                                                 8.33332911e-03f), z, -1.66666672e-01f);
  const float sn = rh + __builtin_fmaf(rh * z, sp, __builtin_fmaf(-0.5f * z, rl, rl));
  const float kp = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(-2.72057548e-07f, z, 2.47994594e-05f), z,
                                                 -1.38888834e-03f), z, 4.16666679e-02f);
  const float hz = 0.5f * z;
  const float w = 1.0f - hz;
  const float cs = w + __builtin_fmaf(z * z, kp, ((1.0f - w) - hz) - __builtin_fmaf(rh, rl, 0.5f * ze));
  const int quadrant = int(q) & 3;
  *s = quadrant == 0 ? sn : (quadrant == 1 ? cs : (quadrant == 2 ? -sn : -cs));
  *c = quadrant == 0 ? cs : (quadrant == 1 ? -sn : (quadrant == 2 ? -cs : sn));
}

// ---- the random stream: PCG hash over (pixel, sample, seed) ------------------------------------
GE_DEV uint32_t pcg_hash(uint32_t v) {
  const uint32_t state = v * 747796405u + 2891336453u;
  const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}
GE_DEV uint32_t rng_seed(uint32_t px, uint32_t py, uint32_t sample_index, uint32_t seed) {
  const uint32_t mixed = pcg_hash(px * 73856093u ^ py * 19349663u ^ seed * 83492791u);
  return pcg_hash(mixed + sample_index * 2654435761u);
}
GE_DEV float rng_next(uint32_t& state) {
  state = state * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  word = (word >> 22u) ^ word;
  return float(word) * (1.0f / 4294967296.0f);
}

// ---- the built structures, as the kernels read them ---------------------------------------------
// A top-level leaf's record: the frame-to-mesh 3x4 and where its mesh's structure lives. 64 bytes.
struct InstRec {
  float w2o[3][4];
  uint32_t node_base;  // the mesh's root in the node pool
  uint32_t tri_base;   // the mesh's first triangle in the triangle pool
  uint32_t first_pair;
  uint32_t mask;  // the instance's 8-bit ray mask: a ray enters it only when (ray mask & mask) != 0
};
static_assert(sizeof(InstRec) == 64, "one instance record is one 64-byte fetch");

// The top-level record's mask from the instance's flags (instance_ray_mask, scene.slang:480-492): an
// instance flagged untraced (bit 6, never met by a light's ray) has every bit but the caster bit,
// every other instance all eight. Camera rays trace with 0xfe and meet both; bounce and shadow rays
// trace with the caster bit 0x01 alone (path_trace.slang:196-203, 289-298) and pass untraced
// instances by.
__host__ __device__ inline uint32_t instance_ray_mask(uint32_t flags) {
  return (flags & GE_INSTANCE_UNTRACED) != 0u ? 0xfeu : 0xffu;
}

// ---- textures at level 0 and the alpha test ------------------------------------------------------
// The scene's bindless textures (RGBA8, level 0), samplers and the sRGB decode of each 8-bit code.
struct TexView {
  const ge_texture_desc* textures;
  const ge_sampler_desc* samplers;
  const uint8_t* texels;
  const float* srgb;  // 256 entries
};

// What a walk needs to decide a masked candidate: per instance its mesh's first cluster and its
// material base with bit 31 set when one of its clusters' materials is alpha-tested, then the
// tables pt_candidate_keeps reads (path_trace.slang:117-135).
struct MaskView {
  const uint2* inst;
  const uint32_t* cluster_materials;
  const ge_resolve_material* materials;
  const ge_cluster_desc* clusters;
  const uint32_t* triangles;
  const ge_vertex_attributes* attributes;  // null: every UV is zero
  TexView tex;
};

// How a walk treats a masked candidate: as opaque (RAY_FLAG_FORCE_OPAQUE), left out
// (RAY_FLAG_CULL_NON_OPAQUE), or by its mask at level 0.
constexpr uint32_t k_alpha_opaque = 0, k_alpha_cull = 1, k_alpha_test = 2;

// f16tof32 (decode_half2, material.slang:217): exact.
GE_DEV float half_to_float(uint32_t h) { return float(__builtin_bit_cast(_Float16, uint16_t(h))); }

GE_DEV float texel_space(float x) {
  if (!(x == x)) return 0.0f;
  return x > 16777216.0f ? 16777216.0f : (x < -16777216.0f ? -16777216.0f : x);
}
// The Vulkan wrap of an integer texel coordinate; false for a texel of the border.
GE_DEV bool wrap_texel(int i, int n, uint32_t mode, int& out) {
  if (mode == GE_ADDRESS_REPEAT) {
    out = ((i % n) + n) % n;
    return true;
  }
  if (mode == GE_ADDRESS_MIRRORED_REPEAT) {
    const int t = ((i % (2 * n)) + 2 * n) % (2 * n) - n;
    out = (n - 1) - (t >= 0 ? t : -(1 + t));
    return true;
  }
  if (mode == GE_ADDRESS_CLAMP_TO_EDGE) {
    out = i < 0 ? 0 : (i >= n ? n - 1 : i);
    return true;
  }
  const bool inside = i >= 0 && i < n;
  out = inside ? i : 0;
  return inside;
}
GE_DEV float4 fetch_texel(const TexView& tv, const ge_texture_desc& t, int x, int y) {
  const uint32_t w = *reinterpret_cast<const uint32_t*>(tv.texels + t.offset + (uint64_t(uint32_t(y)) * t.width + uint32_t(x)) * 4);
  float4 c;
  if (t.format == GE_FORMAT_R8G8B8A8_SRGB) {
    c.x = tv.srgb[w & 0xffu];
    c.y = tv.srgb[(w >> 8) & 0xffu];
    c.z = tv.srgb[(w >> 16) & 0xffu];
  } else {
    c.x = float(w & 0xffu) / 255.0f;
    c.y = float((w >> 8) & 0xffu) / 255.0f;
    c.z = float((w >> 16) & 0xffu) / 255.0f;
  }
  c.w = float(w >> 24) / 255.0f;
  return c;
}
// SampleLevel(sampler, uv, 0): level 0 through the magnification filter; linear weighs the four
// texels around (u W - 1/2, v H - 1/2) by (1-a)(1-b), a(1-b), (1-a)b, ab in that order (the Vulkan
// filtering equation); a texel of the border is transparent black.
GE_DEV float4 sample_level0(const TexView& tv, uint32_t texture, uint32_t sampler, float u, float v) {
  const ge_texture_desc t = tv.textures[texture];
  const ge_sampler_desc sm = tv.samplers[sampler];
  const float fw = float(t.width), fh = float(t.height);
  const int W = int(t.width), H = int(t.height);
  if (sm.mag == GE_FILTER_NEAREST) {
    const float x = texel_space(u * fw), y = texel_space(v * fh);
    int i, j;
    const bool in_u = wrap_texel(int(__builtin_floorf(x)), W, sm.address_u, i);
    const bool in_v = wrap_texel(int(__builtin_floorf(y)), H, sm.address_v, j);
    return in_u && in_v ? fetch_texel(tv, t, i, j) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  const float x = texel_space(u * fw - 0.5f), y = texel_space(v * fh - 0.5f);
  const float x0 = __builtin_floorf(x), y0 = __builtin_floorf(y);
  const float a = x - x0, b = y - y0;
  int i[2], j[2];
  bool in_i[2], in_j[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    in_i[k] = wrap_texel(int(x0) + k, W, sm.address_u, i[k]);
    in_j[k] = wrap_texel(int(y0) + k, H, sm.address_v, j[k]);
  }
  const float w[4] = {(1.0f - a) * (1.0f - b), a * (1.0f - b), (1.0f - a) * b, a * b};
  float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int di = k & 1, dj = k >> 1;
    const float4 c = in_i[di] && in_j[dj] ? fetch_texel(tv, t, i[di], j[dj]) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    out.x = out.x + w[k] * c.x;
    out.y = out.y + w[k] * c.y;
    out.z = out.z + w[k] * c.z;
    out.w = out.w + w[k] * c.w;
  }
  return out;
}
// slot_sampler (material.slang:94-97).
GE_DEV uint32_t slot_sampler(const ge_resolve_material& m, uint32_t word, uint32_t shift) {
  const uint32_t slot = (word >> shift) & 0xffffu;
  return slot == GE_SAME_SAMPLER ? m.sampler : slot;
}
// material_uv (material.slang:136-141).
GE_DEV void material_uv(uint32_t flags, const float tr[4], const float off[4], float u, float v, float& su, float& sv) {
  if (!(flags & GE_MATERIAL_UV_TRANSFORM)) {
    su = u;
    sv = v;
    return;
  }
  su = (tr[0] * u + tr[1] * v) + off[0];
  sv = (tr[2] * u + tr[3] * v) + off[1];
}

// pt_candidate_keeps for one candidate of a cluster (scene-wide index) with material `material`:
// true for an opaque material; a masked one is left out under k_alpha_cull and otherwise kept where
// mask_keeps_level0 (alpha_mask.slang:66-75) keeps its hit_uv (:79-88) at barycentrics u, v.
// `evaluated` is set when the mask was read (a masked candidate under k_alpha_test).
GE_DEV bool candidate_kept(const MaskView& mv, uint32_t cluster, uint32_t material, uint32_t triangle, float u, float v,
                           uint32_t alpha, bool& evaluated) {
  const ge_resolve_material* m = mv.materials + material;
  const uint32_t flags = m->flags;
  if (!(flags & GE_MATERIAL_MASKED)) return true;
  if (alpha == k_alpha_cull) return false;
  evaluated = true;
  float uv[3][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}, {0.0f, 0.0f}};
  if (mv.attributes) {
    const ge_cluster_desc* cl = mv.clusters + cluster;
    const uint32_t packed = mv.triangles[cl->triangle_offset + triangle];
    const uint32_t base = cl->vertex_offset;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const uint32_t h = mv.attributes[base + ((packed >> (8 * k)) & 0xffu)].uv_half2;
      uv[k][0] = half_to_float(h & 0xffffu);
      uv[k][1] = half_to_float(h >> 16);
    }
  }
  const float hu = (uv[0][0] + u * (uv[1][0] - uv[0][0])) + v * (uv[2][0] - uv[0][0]);
  const float hv = (uv[0][1] + u * (uv[1][1] - uv[0][1])) + v * (uv[2][1] - uv[0][1]);
  const float4 offset = *reinterpret_cast<const float4*>(m->uv_offset);
  float alpha_value = offset.w;
  const uint32_t texture = m->albedo_texture;
  if (texture != GE_NO_TEXTURE) {
    const float4 tr4 = *reinterpret_cast<const float4*>(m->uv_transform);
    const float tr[4] = {tr4.x, tr4.y, tr4.z, tr4.w}, off[4] = {offset.x, offset.y, offset.z, offset.w};
    float su, sv;
    material_uv(flags, tr, off, hu, hv, su, sv);
    alpha_value *= sample_level0(mv.tex, texture, m->sampler, su, sv).w;
  }
  return alpha_value >= offset.z;
}

// A 4-wide node (ge_bvh_wide_node): two binary levels collapsed, its four references and its four
// child boxes on a per-axis grid of 8-bit plane indices, each plane rounded outward (one step
// further where the rounding left less than 2^-10 of a step). 64 bytes, one line, loaded as four
// 16-byte words: origin and exponents (byte 3 of word 3: bit c for an occupied slot c, bit 4 + c
// for a leaf slot c), references, the low planes (byte c of lo[a] is slot c's) and the high planes.
using WideNode = ge_bvh_wide_node;
static_assert(sizeof(WideNode) == 64, "a wide node is one 64-byte line");

struct Bvh {
  const ge_bvh_node* tlas;       // the top level's nodes (node 0 the root)
  const uint32_t* tlas_slots;    // sorted slot -> instance id
  const InstRec* inst;           // per instance id
  const ge_bvh_node* nodes;      // every mesh's nodes, pooled
  const ge_bvh_triangle* tris;   // every mesh's sorted triangles, pooled
  const WideNode* wtlas;         // the top level's wide nodes, at the binary indices
  const WideNode* wnodes;        // every mesh's wide nodes, at the binary indices of the pool
  uint32_t tlas_node_count;      // 0: an empty scene
  uint32_t stack_cap;            // entries of the per-lane stack the structures' depths need
  MaskView masks;                // read only by the walks that decide masked candidates
};

// A lane's traversal counts, [0] for closest-hit walks and [1] for any-hit walks.
struct Tally {
  uint32_t nodes[2] = {0, 0}, tris[2] = {0, 0}, inst[2] = {0, 0}, masks[2] = {0, 0};
};

// Adds a wave's counts to the context's counters (closest rays, any rays, then nodes, triangles and
// instances of the closest-hit walks and of the any-hit walks, then past the walks' head at [8] the
// mask evaluations of each): a wave reduction, then one lane adds.
GE_DEV void flush_counters(unsigned long long* counters, uint64_t closest, uint64_t any, const Tally& t) {
  uint64_t v[10] = {closest, any, t.nodes[0], t.tris[0], t.inst[0], t.nodes[1], t.tris[1], t.inst[1],
                    t.masks[0], t.masks[1]};
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    uint64_t x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    v[k] = x;
  }
  if ((threadIdx.x & 63u) == 0) {
#pragma unroll
    for (int k = 0; k < 10; ++k)
      if (v[k]) atomicAdd(&counters[k < 8 ? k : k + 1], (unsigned long long)v[k]);
  }
}

constexpr float k_box_robust = 1.00000095367431640625f;  // 1 + 2^-20 on the far distance (Ize 2013)
constexpr uint32_t k_sentinel = 0xfffffffeu;
// The most entries a lane's traversal stack may need: the build refuses deeper structures.
constexpr int k_stack = 160;
// Each kernel's workgroup is one wave of 64 lanes. The top k_ring entries of a lane's stack live in
// an LDS ring (entry k of lane l at word 64 (k mod k_ring) + l, so a wave's pushes and pops hit 64
// different banks); deeper entries spill to the lane's private memory and come back as the stack
// shrinks, so LDS stays at 4 KiB a wave whatever the structures' depth.
constexpr int k_lanes = 64;
#ifndef GE_WALK_RING
#define GE_WALK_RING 16
#endif
constexpr int k_ring = GE_WALK_RING;
static_assert(k_ring >= 4 && k_ring <= 64 && (k_ring & (k_ring - 1)) == 0, "the ring is indexed by a mask");

struct Ray {
  float o[3], d[3], inv[3];
  float tmin;
  float teps;  // absolute slack on the box test, in t: 2^-19 of the origin's largest coordinate
  int kx, ky, kz;
  float Sx, Sy, Sz;
};

GE_DEV float sel3(const float v[3], int k) {
  // k is 0, 1 or 2. Bit masks rather than selects: the compiler turns chained selects on a lane's k
  // into branches.
  const uint32_t m0 = 0u - ((uint32_t(k) - 1u) >> 31);      // all ones where k == 0
  const uint32_t m1 = 0u - ((uint32_t(k ^ 1) - 1u) >> 31);  // all ones where k == 1
  const uint32_t r12 = (__float_as_uint(v[1]) & m1) | (__float_as_uint(v[2]) & ~m1);  // v_bfi_b32
  return __uint_as_float((__float_as_uint(v[0]) & m0) | (r12 & ~m0));
}

GE_DEV void prep(Ray& r) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float d = __builtin_fabsf(r.d[a]) < 1e-20f ? __builtin_copysignf(1e-20f, r.d[a]) : r.d[a];
    r.inv[a] = 1.0f / d;
  }
  int kz = 0;
  if (__builtin_fabsf(r.d[1]) > __builtin_fabsf(r.d[kz])) kz = 1;
  if (__builtin_fabsf(r.d[2]) > __builtin_fabsf(r.d[kz])) kz = 2;
  int kx = kz == 2 ? 0 : kz + 1;
  int ky = kx == 2 ? 0 : kx + 1;
  if (sel3(r.d, kz) < 0.0f) {
    const int t = kx;
    kx = ky;
    ky = t;
  }
  r.kx = kx;
  r.ky = ky;
  r.kz = kz;
  // The shear shares the slab test's reciprocal of the dominant component (never clamped: it is
  // the largest of a non-zero direction), so a ray costs three divisions.
  r.Sz = sel3(r.inv, kz);
  r.Sx = sel3(r.d, kx) * r.Sz;
  r.Sy = sel3(r.d, ky) * r.Sz;
  r.teps = maxr(maxr(__builtin_fabsf(r.o[0]), __builtin_fabsf(r.o[1])), __builtin_fabsf(r.o[2])) * 1.9073486328125e-6f *
           __builtin_fabsf(r.Sz);
}

GE_DEV Ray make_ray(f3 o, f3 d, float tmin) {
  Ray r;
  r.o[0] = o.x;
  r.o[1] = o.y;
  r.o[2] = o.z;
  r.d[0] = d.x;
  r.d[1] = d.y;
  r.d[2] = d.z;
  r.tmin = tmin;
  prep(r);
  return r;
}

// The box test of a quantized box (Ize 2013, "Robust BVH Ray Traversal"): the slab interval, its
// far end widened by 1 + 2^-20, against [tmin, tmax], both widened by the ray's absolute slack
// `teps`, so that a triangle a few coordinate ulps short of its box's computed entry (a hit very
// near an origin far from the world's) is never culled with its box. Each axis's near and far
// distances come in (nt, ft), computed as fma(q, inv 2^(e - 127), (origin - o) inv) from the
// plane's grid index q. Against the exact distance of that plane, the computed one errs by at most
// 3 2^-24 of itself plus 2^-15 of a grid step's length along the ray (the node's origin lies at
// most 255 steps beyond the plane), so the 2^-10 of a step or more to spare on every side of a
// quantized box covers the second part and the far end's widening the first: the test passes wherever the fp32 slab
// test, (plane - o) inv on the child's own box, would. fma is monotone in q, so the near distance
// is the smaller of the axis's two. No operand is NaN (a direction's reciprocal is finite and
// non-zero), so the hardware's min and max (v_min3/v_max3) give the compare-and-select forms'
// values up to the sign of a zero, which no comparison below can see.
GE_DEV bool slab_q(const float nt[3], const float ft[3], const Ray& r, float tmax, float& tnear) {
  const float tn = __builtin_fmaxf(__builtin_fmaxf(nt[0], nt[1]), __builtin_fmaxf(nt[2], r.tmin));
  const float tf = __builtin_fminf(__builtin_fminf(__builtin_fminf(ft[0], ft[1]), ft[2]) * k_box_robust, tmax);
  tnear = tn;
  return tn <= tf + r.teps;
}

// Byte c of a plane word as a float (v_cvt_f32_ubyte0..3).
GE_DEV float plane_byte(uint32_t w, int c) { return float((w >> (8 * c)) & 0xffu); }

// a*b - c*d to within 2 ulp, exact in sign and exactly zero only when the value is (Kahan's
// algorithm with one fused multiply-add; Jeannerod, Louvet and Muller 2013, Math. Comp. 82).
GE_DEV float det2(float a, float b, float c, float d) {
  const float w = c * d;
  const float e = __builtin_fmaf(-c, d, w);
  return __builtin_fmaf(a, b, -w) + e;
}

// Watertight ray/triangle intersection (Woop, Benthin, Wald 2013): an edge function that is
// exactly zero in float is recomputed accurately (det2), the paper's higher-precision fallback.
// No face culling.
GE_DEV bool intersect(const Ray& r, const float* v0, const float* v1, const float* v2, float& t, float& u,
                      float& v) {
  const float A[3] = {v0[0] - r.o[0], v0[1] - r.o[1], v0[2] - r.o[2]};
  const float B[3] = {v1[0] - r.o[0], v1[1] - r.o[1], v1[2] - r.o[2]};
  const float C[3] = {v2[0] - r.o[0], v2[1] - r.o[1], v2[2] - r.o[2]};
  const float Akz = sel3(A, r.kz), Bkz = sel3(B, r.kz), Ckz = sel3(C, r.kz);
  const float Ax = sel3(A, r.kx) - r.Sx * Akz, Ay = sel3(A, r.ky) - r.Sy * Akz;
  const float Bx = sel3(B, r.kx) - r.Sx * Bkz, By = sel3(B, r.ky) - r.Sy * Bkz;
  const float Cx = sel3(C, r.kx) - r.Sx * Ckz, Cy = sel3(C, r.ky) - r.Sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = det2(Cx, By, Cy, Bx);
    V = det2(Ax, Cy, Ay, Cx);
    W = det2(Bx, Ay, By, Ax);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  const float det = U + V + W;
  if (det == 0.0f) return false;
  const float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
  const float T = U * Az + V * Bz + W * Cz;
  const float rcp = 1.0f / det;  // one division, the paper's scaling of T, V and W
  t = T * rcp;
  u = V * rcp;
  v = W * rcp;
  return true;
}

struct Hit {
  float t, u, v;
  uint32_t inst, prim, key;
  bool found;
};

GE_DEV void to_object(const InstRec& rec, const Ray& w, Ray& o) {
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    o.o[r] = ((rec.w2o[r][0] * w.o[0] + rec.w2o[r][1] * w.o[1]) + rec.w2o[r][2] * w.o[2]) + rec.w2o[r][3];
    o.d[r] = (rec.w2o[r][0] * w.d[0] + rec.w2o[r][1] * w.d[1]) + rec.w2o[r][2] * w.d[2];
  }
  o.tmin = w.tmin;
  prep(o);
}

// The two-level walk over the 4-wide nodes in its canonical order: a node's four boxes against the
// distance at entry; then a mesh's leaves among them in child order, each skipped if its box's entry
// lies beyond the distance so far; then the inner children whose entry is within the distance after
// the leaves, the nearest next (ties: the first) and the others pushed so that they come back in
// child order; a top-level leaf enters its instance (the ray goes into mesh space and a sentinel
// marks the way back). Closest hits are ordered by (t, then the larger pair << 8 | triangle), a
// total order, so the answer does not depend on the walk; the counters do, and match the oracle's
// walk.
//
// The walk is a state machine of three kinds of work, so that a wave can run each kind for the
// lanes that need it together instead of one lane at a time:
//   kNode  — test the four boxes of node `cur`; a mesh node with leaves among them goes to kLeaf
//            with the node's entries kept, any other goes straight on to its inner children;
//   kLeaf  — the kept leaves' triangle tests in child order, then the inner children;
//   kEnter — `cur` is a top-level leaf: the ray goes into its instance's mesh space (or passes it
//            by when the instance's mask shares no bit with the ray's);
// and kDone. The inner children's step chooses the next node (the nearest, the rest pushed) or pops
// one, which sets the next state. Every lane runs the same sequence of operations whenever the
// wave schedules them, so hits and counters do not depend on the schedule.
//
// `Masks` compiles the candidate rule for alpha-tested materials (`alpha`, one of k_alpha_*): a
// triangle that passes the intersection and distance tests is committed only where candidate_kept
// keeps it, so the answer is the nearest kept triangle (the same order, the same tie rule) and the
// walk and its counters are the opaque walk's but for the candidates left out.
constexpr uint32_t k_walk_node = 0, k_walk_leaf = 1, k_walk_enter = 2, k_walk_done = 3;

template <bool AnyHit, bool Count, bool Masks = false>
struct Walker {
  Ray world, ray;
  Hit hit;
  int sp;
  uint32_t cur, node_base, tri_base, first_pair, inst_id, ray_mask;
  uint32_t alpha, mask_cluster, mask_material;  // Masks: the rule, the instance's first cluster and base
  bool in_blas, mask_live;                      // Masks: the instance has masked clusters
  uint32_t state;
  // The node whose boxes were tested last: its references and entry distances, the leaves among
  // them still to test and the inner children that passed at entry.
  uint32_t nref[4];
  float ntn[4];
  uint32_t leaf_mask, hit_mask;

  // False when there is nothing to walk (mask 0 or an empty scene): the hit is already the answer.
  GE_DEV bool start(const Bvh& bvh, const Ray& w, float tmax, uint32_t mask, uint32_t alpha_rule = k_alpha_opaque) {
    alpha = alpha_rule;
    mask_cluster = mask_material = 0;
    mask_live = false;
    hit.found = false;
    hit.t = tmax;
    hit.key = 0;
    hit.u = hit.v = 0.0f;
    hit.inst = hit.prim = GE_RT_MISS;
    world = w;
    ray = w;
    sp = 0;
    cur = 0;
    node_base = tri_base = first_pair = inst_id = 0;
    ray_mask = mask;
    in_blas = false;
    leaf_mask = hit_mask = 0;
    const bool go = mask != 0 && bvh.tlas_node_count != 0;
    state = go ? k_walk_node : k_walk_done;
    return go;
  }

  GE_DEV void push(uint32_t* __restrict__ stack, uint32_t* __restrict__ spill, int cap, uint32_t x) {
    if (sp >= cap) return;
    uint32_t& slot = stack[k_lanes * (sp & (k_ring - 1))];
    if (sp >= k_ring) spill[sp - k_ring] = slot;
    slot = x;
    ++sp;
  }
  GE_DEV uint32_t pop(uint32_t* __restrict__ stack, const uint32_t* __restrict__ spill) {
    --sp;
    uint32_t& slot = stack[k_lanes * (sp & (k_ring - 1))];
    const uint32_t x = slot;
    if (sp >= k_ring) slot = spill[sp - k_ring];
    return x;
  }
  // The state that goes with node reference `cur`: a top-level leaf is entered, any other node visited.
  GE_DEV void go_to(uint32_t node) {
    cur = node;
    state = (!in_blas && (node & GE_BVH_LEAF)) ? k_walk_enter : k_walk_node;
  }
  // Takes the next node off the stack (a sentinel leads back to the top level), or ends the walk.
  GE_DEV void next(const Bvh& bvh, uint32_t* __restrict__ stack, const uint32_t* __restrict__ spill) {
    for (;;) {
      if (sp == 0) {
        state = k_walk_done;
        return;
      }
      const uint32_t e = pop(stack, spill);
      if (e == k_sentinel) {
        in_blas = false;
        ray = world;
        continue;
      }
      go_to(e);
      return;
    }
  }

  // kEnter: into the instance of top-level leaf `cur`, or past it.
  GE_DEV void enter(const Bvh& bvh, uint32_t* __restrict__ stack, uint32_t* __restrict__ spill, Tally& tally) {
    inst_id = bvh.tlas_slots[cur & ~GE_BVH_LEAF];
    const InstRec rec = bvh.inst[inst_id];
    if ((rec.mask & ray_mask) == 0u) {
      next(bvh, stack, spill);
      return;
    }
    node_base = rec.node_base;
    tri_base = rec.tri_base;
    first_pair = rec.first_pair;
    if (Masks && alpha != k_alpha_opaque) {
      const uint2 mi = bvh.masks.inst[inst_id];
      mask_cluster = mi.x;
      mask_material = mi.y & 0x7fffffffu;
      mask_live = (mi.y >> 31) != 0u;
    }
    to_object(rec, world, ray);
    push(stack, spill, int(bvh.stack_cap), k_sentinel);
    in_blas = true;
    if (Count) ++tally.inst[AnyHit];
    cur = 0;
    state = k_walk_node;
  }

  // kNode: the four boxes of node `cur` against the distance at entry (one pass of box tests for
  // the whole wave, branch-free, so that the node's four 16-byte loads stay whole). A mesh node
  // with leaves among its passing children goes to kLeaf; any other node goes on to its children.
  GE_DEV void visit(const Bvh& bvh, uint32_t* __restrict__ stack, uint32_t* __restrict__ spill, Tally& tally) {
    if (Count) ++tally.nodes[AnyHit];
    const uint4* np = reinterpret_cast<const uint4*>(in_blas ? bvh.wnodes + node_base + cur : bvh.wtlas + cur);
    const uint4 w0 = np[0], refs = np[1], w2 = np[2], w3 = np[3];
    const uint32_t lo[3] = {w2.x, w2.y, w2.z}, hi[3] = {w2.w, w3.x, w3.y};
    const float origin[3] = {__uint_as_float(w0.x), __uint_as_float(w0.y), __uint_as_float(w0.z)};
    // Per axis: the grid step times the reciprocal (exact: a power of two), the origin's distance,
    // and the near and far plane words by the direction's sign (bit masks: v_bfi_b32).
    float A[3], B[3];
    uint32_t qn[3], qf[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      A[a] = ray.inv[a] * __uint_as_float(((w0.w >> (8 * a)) & 0xffu) << 23);
      B[a] = (origin[a] - ray.o[a]) * ray.inv[a];
      const uint32_t flip = uint32_t(int32_t(__float_as_uint(ray.inv[a])) >> 31);
      qn[a] = (hi[a] & flip) | (lo[a] & ~flip);
      qf[a] = (lo[a] & flip) | (hi[a] & ~flip);
    }
    nref[0] = refs.x;
    nref[1] = refs.y;
    nref[2] = refs.z;
    nref[3] = refs.w;
    uint32_t in = 0;  // the slots whose boxes pass
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      float nt[3], ft[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        nt[a] = __builtin_fmaf(plane_byte(qn[a], c), A[a], B[a]);
        ft[a] = __builtin_fmaf(plane_byte(qf[a], c), A[a], B[a]);
      }
      in |= uint32_t(slab_q(nt, ft, ray, hit.t, ntn[c])) << c;
    }
    in &= w0.w >> 24;                                      // an empty slot's planes are not a box
    const uint32_t lm = in & (in_blas ? w0.w >> 28 : 0u);  // a mesh's leaves, to test
    leaf_mask = lm;
    hit_mask = in & ~lm;                                   // inner children (and instances) to go on to
    if (lm) {
      state = k_walk_leaf;
      return;
    }
    children<false>(bvh, stack, spill);
  }

  // kLeaf: the kept leaves in child order, one triangle test a trip for every lane that has one
  // left, so lanes whose leaves sit in different slots share the trips (at most four, unrolled). A
  // box that passed at entry is still within the distance so far exactly when tn <= t + teps (the
  // slab test's other bound held at entry). Then the inner children.
  GE_DEV void leaves(const Bvh& bvh, uint32_t* __restrict__ stack, uint32_t* __restrict__ spill, Tally& tally) {
    uint32_t lm = leaf_mask;
    constexpr int unroll = Masks ? 1 : 4;
#pragma unroll unroll
    for (int trip = 0; trip < 4; ++trip) {
      if (!lm) break;
      const int c = __builtin_ctz(lm);
      lm &= lm - 1u;
      const float tc = c == 0 ? ntn[0] : c == 1 ? ntn[1] : c == 2 ? ntn[2] : ntn[3];
      if (!(tc <= hit.t + ray.teps)) continue;
      const uint32_t rc = c == 0 ? nref[0] : c == 1 ? nref[1] : c == 2 ? nref[2] : nref[3];
      const ge_bvh_triangle* tri = bvh.tris + tri_base + (rc & ~GE_BVH_LEAF);
      const float4* tp = reinterpret_cast<const float4*>(tri);
      const float4 q0 = tp[0], q1 = tp[1], q2 = tp[2];
      const float v0[3] = {q0.x, q0.y, q0.z}, v1[3] = {q0.w, q1.x, q1.y}, v2[3] = {q1.z, q1.w, q2.x};
      const uint32_t prim = __float_as_uint(q2.y);
      if (Count) ++tally.tris[AnyHit];
      float t, u, v;
      if (!intersect(ray, v0, v1, v2, t, u, v)) continue;
      if (!(t >= ray.tmin && t <= hit.t)) continue;
      if (Masks && mask_live) {
        const uint32_t cluster = mask_cluster + (prim >> 8);
        const uint32_t material = bvh.masks.cluster_materials[cluster] + mask_material;
        bool evaluated = false;
        const bool kept = candidate_kept(bvh.masks, cluster, material, prim & 0xffu, u, v, alpha, evaluated);
        if (Count && evaluated) ++tally.masks[AnyHit];
        if (!kept) continue;
      }
      const uint32_t key_hit = (first_pair << 8) + prim;
      if (AnyHit || t < hit.t || !hit.found || key_hit > hit.key) {
        hit.found = true;
        hit.t = t;
        hit.u = u;
        hit.v = v;
        hit.inst = inst_id;
        hit.prim = prim;
        hit.key = key_hit;
        if (AnyHit) {
          state = k_walk_done;
          return;
        }
      }
    }
    leaf_mask = 0;
    children<true>(bvh, stack, spill);
  }

  // The inner children of the node just tested that are still within the distance: the nearest
  // next (ties: the first), the others pushed last child first, so that they come back in child
  // order; none: the next node off the stack. Selects, not branches.
  // `Refilter`: the leaves have been tested since the boxes were, so the distance may have shrunk;
  // straight after the box tests every inner child that passed is within it (its entry is at most
  // t + teps), so the re-filter would keep them all.
  template <bool Refilter>
  GE_DEV void children(const Bvh& bvh, uint32_t* __restrict__ stack, uint32_t* __restrict__ spill) {
    const int cap = int(bvh.stack_cap);
    if (hit_mask) {
      const float reach = hit.t + ray.teps;
      uint32_t keep = 0;
      int best = 0;
      float best_t = 0.0f;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const bool in = ((hit_mask >> c) & 1u) != 0u && (!Refilter || ntn[c] <= reach);
        const bool nearer = in && (keep == 0u || ntn[c] < best_t);
        keep |= uint32_t(in) << c;
        best = nearer ? c : best;
        best_t = nearer ? ntn[c] : best_t;
      }
      if (keep) {
        const uint32_t rest = keep & ~(1u << best);
        if (sp + 4 <= k_ring && sp + 3 <= cap) {
          // The ring has room below its first wrap: each child is written to the free slot at the
          // top whether or not it is pushed (one that is not is overwritten by the next).
#pragma unroll
          for (int c = 3; c >= 0; --c) {
            stack[k_lanes * sp] = nref[c];
            sp += int((rest >> c) & 1u);
          }
        } else {
#pragma unroll
          for (int c = 3; c >= 0; --c)
            if ((rest >> c) & 1u) push(stack, spill, cap, nref[c]);
        }
        go_to(best == 0 ? nref[0] : best == 1 ? nref[1] : best == 2 ? nref[2] : nref[3]);
        return;
      }
    }
    next(bvh, stack, spill);
  }

  // Everything up to the next node test, in order; true when the walk has ended.
  GE_DEV bool step(const Bvh& bvh, uint32_t* __restrict__ stack, uint32_t* __restrict__ spill, Tally& tally) {
    if (state == k_walk_enter) enter(bvh, stack, spill, tally);
    if (state == k_walk_node) visit(bvh, stack, spill, tally);
    if (state == k_walk_leaf) leaves(bvh, stack, spill, tally);
    return state == k_walk_done;
  }
};

template <bool AnyHit, bool Count, bool Masks = false>
GE_DEV bool traverse(const Bvh& bvh, const Ray& world, float tmax, uint32_t mask, Hit& hit, Tally& tally,
                     uint32_t* __restrict__ stack, uint32_t alpha = k_alpha_opaque) {
  Walker<AnyHit, Count, Masks> w;
  uint32_t spill[k_stack - k_ring];
  if (w.start(bvh, world, tmax, mask, alpha)) {
    while (!w.step(bvh, stack, spill, tally)) {
    }
  }
  hit = w.hit;
  return w.hit.found;
}

// A persistent walk: the wave takes rays from a shared counter, and whenever k_refill or more of its
// lanes have finished their rays it fetches new ones for them (one atomic for the wave), so lanes
// are not left idle until the wave's longest ray ends. `load(i, slot, ray, tmax, mask)` fetches ray
// i (false: nothing to trace, the answer is a miss) and `store(slot, hit)` writes an answer. Every
// wave leaves the loop once the counter has passed n and its lanes are idle.
//
// Each pass of the loop runs the box tests of every lane whose walk is at a node. A lane with
// triangle tests to make, or an instance to enter, waits for them until k_batch lanes of the wave
// wait for the same kind of work (or until fewer than k_batch lanes are left at a node), and those
// lanes then run it together: a triangle test or an instance entry run for a few lanes at a time
// costs the wave as much as for all of them.
// Occupancy asked of the walk kernels, in waves per SIMD: 5 gives the 4-wide walk 96 VGPRs with a
// few spilled; 6 (80 VGPRs) spills more of it and runs 8% slower, 4 runs 1-3% slower on the
// 4M-triangle scene; 0 leaves it to the compiler.
#ifndef GE_WALK_WAVES_PER_EU
#define GE_WALK_WAVES_PER_EU 5
#endif
#if GE_WALK_WAVES_PER_EU > 0
#define GE_WALK_ATTR __attribute__((amdgpu_waves_per_eu(GE_WALK_WAVES_PER_EU)))
#else
#define GE_WALK_ATTR
#endif
// Refill threshold: 32 idle lanes before a wave fetches rays. A refill (the atomic, the dependent
// ray loads, three divisions and a shear per ray) stalls the wave and runs as divergent code beside
// the walk, so batching it pays: on the 4M-triangle scene 48 runs 0.8% slower, 24 1.6%, 16 14%.
#ifndef GE_WALK_REFILL
#define GE_WALK_REFILL 32
#endif
constexpr uint32_t k_refill = GE_WALK_REFILL;
static_assert(k_refill >= 1 && k_refill <= 64, "a refill needs at most a wave of idle lanes");
// Lanes that wait for triangle tests or an instance entry before the wave runs them: on the
// 4M-triangle scene 16 is 5% faster than running them as they come; 12 and 20 measure alike, 8 and
// 24 0.5% slower, 32 3% slower.
#ifndef GE_WALK_BATCH
#define GE_WALK_BATCH 16
#endif
constexpr uint32_t k_batch = GE_WALK_BATCH;
static_assert(k_batch >= 1 && k_batch <= 64, "a batch is at most a wave");
// Per kind of ray, where it measured better than the defaults above (the 20-wave-a-CU grid):
// - the path tracer's bounces refill at 40 idle lanes (0.3% faster than 32, 48 0.5% slower);
// - the visibility words' camera rays at 56 (their walk 8.6% faster than at 32, 48 8%): rays of
//   one 8 x 8 tile end close together, so a wave waits for most of them before its next tile;
// - the sun rays of the shadow resolve batch 24 lanes (3% faster than 16; refills at 48 1%).
// The walks that decide masked candidates keep the defaults: through the alpha-tested forest the
// camera rays' walk ran 7% slower refilling at 56, the sun rays' 1.4% slower batching 24, and
// refills at 24 or 40 or batches of 8 or 24 measured no better for either.
#ifndef GE_WALK_REFILL_PATHS
#define GE_WALK_REFILL_PATHS 40
#endif
#ifndef GE_WALK_REFILL_CAMERA
#define GE_WALK_REFILL_CAMERA 56
#endif
#ifndef GE_WALK_REFILL_SUN
#define GE_WALK_REFILL_SUN 32
#endif
#ifndef GE_WALK_BATCH_SUN
#define GE_WALK_BATCH_SUN 24
#endif
constexpr uint32_t k_refill_paths = GE_WALK_REFILL_PATHS;    // the path tracer's bounces
constexpr uint32_t k_refill_camera = GE_WALK_REFILL_CAMERA;  // the visibility words' camera rays
constexpr uint32_t k_refill_sun = GE_WALK_REFILL_SUN;        // the shadow resolve's sun rays
constexpr uint32_t k_batch_sun = GE_WALK_BATCH_SUN;
#ifndef GE_WALK_REFILL_MASKED
#define GE_WALK_REFILL_MASKED GE_WALK_REFILL
#endif
#ifndef GE_WALK_BATCH_MASKED
#define GE_WALK_BATCH_MASKED GE_WALK_BATCH
#endif
constexpr uint32_t k_refill_masked = GE_WALK_REFILL_MASKED;  // walks that decide masked candidates
constexpr uint32_t k_batch_masked = GE_WALK_BATCH_MASKED;

GE_DEV uint64_t lanes_below() {
  const uint32_t lane = threadIdx.x & 63u;
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

template <bool AnyHit, bool Count, bool Masks = false, uint32_t Refill = k_refill, uint32_t Batch = k_batch,
          class Load, class Store>
GE_DEV void persistent_walk(const Bvh& bvh, uint32_t n, uint32_t* head, uint32_t* __restrict__ stack, Load load,
                            Store store, uint64_t& issued, Tally& tally, uint32_t alpha = k_alpha_opaque) {
  Walker<AnyHit, Count, Masks> w;
  uint32_t spill[k_stack - k_ring];
  bool active = false, exhausted = false;
  uint32_t slot = 0;
  for (;;) {
    const uint64_t idle = __ballot(!active);
    const uint32_t nidle = uint32_t(__popcll(idle));
    if (!exhausted && nidle >= Refill) {
      uint32_t base = 0;
      if ((threadIdx.x & 63u) == 0) base = atomicAdd(head, nidle);
      base = __builtin_amdgcn_readfirstlane(base);  // lane 0 took part: one value for the wave
      if (uint64_t(base) + nidle >= n) exhausted = true;
      if (!active) {
        const uint32_t i = base + uint32_t(__popcll(idle & lanes_below()));  // rank among the idle lanes
        if (i < n) {
          Ray ray;
          float tmax;
          uint32_t mask;
          if (load(i, slot, ray, tmax, mask)) {
            ++issued;
            active = w.start(bvh, ray, tmax, mask, alpha);
            if (!active) store(slot, w.hit);
          } else {
            Hit miss;
            miss.found = false;
            miss.t = miss.u = miss.v = 0.0f;
            miss.inst = miss.prim = GE_RT_MISS;
            miss.key = 0;
            store(slot, miss);
          }
        }
      }
    }
    if (__ballot(active) == 0) {
      if (exhausted) break;
      continue;
    }
    // What each kind of work waits for: the wave runs a kind when Batch lanes want it, or when
    // fewer than Batch lanes are left at a node (so that a wave never waits on itself).
    const uint32_t at_node = uint32_t(__popcll(__ballot(active && w.state == k_walk_node)));
    const uint32_t at_leaf = uint32_t(__popcll(__ballot(active && w.state == k_walk_leaf)));
    const uint32_t at_enter = uint32_t(__popcll(__ballot(active && w.state == k_walk_enter)));
    const bool few = at_node < Batch;
    if (at_enter && (few || at_enter >= Batch) && active && w.state == k_walk_enter) w.enter(bvh, stack, spill, tally);
    if (at_leaf && (few || at_leaf >= Batch) && active && w.state == k_walk_leaf) w.leaves(bvh, stack, spill, tally);
    if (active && w.state == k_walk_node) w.visit(bvh, stack, spill, tally);
    if (active && w.state == k_walk_done) {
      store(slot, w.hit);
      active = false;
    }
  }
}

}  // namespace ge
