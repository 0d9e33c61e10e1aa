// rt_shade.h — the reference path tracer's shading on the device: the BSDF of brdf.slang
// (:39-264), the sampler of sampling.slang (:58-160), the lights of material.slang (:960-986) and
// brdf.slang (:370-383), the hit decode of material.slang (:244-302) and scene.slang
// (:120-183,566-610,789-796), and the environment of path_trace.slang (:166-176) with
// material.slang (:1084-1097). Single precision, in the shaders' operation order.
#pragma once
#pragma clang fp contract(off)

#include "rt_device.h"

namespace ge {

struct Surface {
  f3 albedo, normal, view;
  float roughness, metallic;
  float spec[4];
};

GE_DEV float brdf_alpha(float roughness) {
  const float r = clampr(roughness, 0.045f, 1.0f);
  return r * r;
}
GE_DEV f3 brdf_f0(f3 albedo, float metallic) { return lerp3(splat(0.04f), albedo, metallic); }
GE_DEV float d_ggx(float n_dot_h, float alpha) {
  const float a2 = alpha * alpha;
  const float d = n_dot_h * n_dot_h * (a2 - 1.0f) + 1.0f;
  return a2 / maxr(k_pi * d * d, 1e-9f);
}
GE_DEV float v_smith(float n_dot_v, float n_dot_l, float alpha) {
  const float a2 = alpha * alpha;
  const float v = n_dot_l * __builtin_sqrtf(n_dot_v * n_dot_v * (1.0f - a2) + a2);
  const float l = n_dot_v * __builtin_sqrtf(n_dot_l * n_dot_l * (1.0f - a2) + a2);
  return 0.5f / maxr(v + l, 1e-9f);
}
GE_DEV f3 schlick(f3 f0, float cos_theta) {
  const float m = clampr(1.0f - cos_theta, 0.0f, 1.0f);
  const float m2 = m * m;
  return f0 + (splat(1.0f) - f0) * (m2 * m2 * m);
}
GE_DEV f3 schlick_f90(f3 f0, float f90, float cos_theta) {
  const float m = clampr(1.0f - cos_theta, 0.0f, 1.0f);
  const float m2 = m * m;
  const float w = m2 * m2 * m;
  return f0 + (splat(f90) - f0) * w;
}
GE_DEV bool declares_specular(const Surface& s) {
  return s.spec[0] != 0.0f || s.spec[1] != 0.0f || s.spec[2] != 0.0f || s.spec[3] != 0.0f;
}
GE_DEV f3 dielectric_f0(const Surface& s) { return splat(0.04f) + f3{s.spec[0], s.spec[1], s.spec[2]}; }
GE_DEV float surface_f90(const Surface& s) { return 1.0f - s.spec[3] * (1.0f - s.metallic); }
GE_DEV f3 fresnel(const Surface& s, float cos_theta) {
  if (!declares_specular(s)) return schlick(brdf_f0(s.albedo, s.metallic), cos_theta);
  return schlick_f90(lerp3(dielectric_f0(s), s.albedo, s.metallic), surface_f90(s), cos_theta);
}
GE_DEV f3 diffuse_fresnel(const Surface& s, f3 f, float cos_theta) {
  if (!declares_specular(s)) return f;
  const f3 f0 = dielectric_f0(s);
  const float fmax = maxr(f0.x, maxr(f0.y, f0.z));
  return schlick_f90(lerp3(splat(fmax), s.albedo, s.metallic), surface_f90(s), cos_theta);
}
GE_DEV f3 brdf_eval(const Surface& s, f3 l) {
  const float n_dot_l = dot(s.normal, l);
  if (n_dot_l <= 0.0f) return splat(0.0f);
  const float n_dot_v = maxr(dot(s.normal, s.view), 1e-4f);
  const f3 half_sum = l + s.view;
  const f3 h = dot(half_sum, half_sum) > 1e-12f ? normalize(half_sum) : s.normal;
  const float n_dot_h = clampr(dot(s.normal, h), 0.0f, 1.0f);
  const float v_dot_h = clampr(dot(s.view, h), 0.0f, 1.0f);
  const float alpha = brdf_alpha(s.roughness);
  const f3 f = fresnel(s, v_dot_h);
  const f3 specular = f * (d_ggx(n_dot_h, alpha) * v_smith(n_dot_v, n_dot_l, alpha));
  const f3 diffuse = (splat(1.0f) - diffuse_fresnel(s, f, v_dot_h)) * ((1.0f - s.metallic) / k_pi) * s.albedo;
  return diffuse + specular;
}
GE_DEV f3 brdf_direct(const Surface& s, f3 l, f3 radiance) {
  return brdf_eval(s, l) * maxr(dot(s.normal, l), 0.0f) * radiance;
}
// brdf_eval's derivatives for a surface without its own specular record: per channel c,
// df_c/d albedo_c and df_c/d roughness (zero where brdf_alpha clamps).
GE_DEV void brdf_eval_grad(const Surface& s, f3 l, f3& df_da, f3& df_dr) {
  df_da = splat(0.0f);
  df_dr = splat(0.0f);
  const float n_dot_l = dot(s.normal, l);
  if (n_dot_l <= 0.0f) return;
  const float n_dot_v = maxr(dot(s.normal, s.view), 1e-4f);
  const f3 half_sum = l + s.view;
  const f3 h = dot(half_sum, half_sum) > 1e-12f ? normalize(half_sum) : s.normal;
  const float n_dot_h = clampr(dot(s.normal, h), 0.0f, 1.0f);
  const float v_dot_h = clampr(dot(s.view, h), 0.0f, 1.0f);
  const float rc = clampr(s.roughness, 0.045f, 1.0f);
  const float alpha = rc * rc;
  const float dalpha = (s.roughness > 0.045f && s.roughness < 1.0f) ? 2.0f * rc : 0.0f;
  const float m = s.metallic;
  const float mm = clampr(1.0f - v_dot_h, 0.0f, 1.0f);
  const float m2 = mm * mm;
  const float w5 = m2 * m2 * mm;
  const f3 f = fresnel(s, v_dot_h);
  const float d = d_ggx(n_dot_h, alpha);
  const float vis = v_smith(n_dot_v, n_dot_l, alpha);
  const float kd = (1.0f - m) / k_pi;
  const float df0 = m * (1.0f - w5);
  df_da = splat(df0 * (d * vis)) - splat(df0 * kd) * s.albedo + (splat(1.0f) - f) * kd;
  const float a2 = alpha * alpha;
  const float dd = n_dot_h * n_dot_h * (a2 - 1.0f) + 1.0f;
  const float dd_dalpha = k_pi * dd * dd > 1e-9f
                              ? 2.0f * alpha / (k_pi * dd * dd) - 4.0f * alpha * a2 * n_dot_h * n_dot_h / (k_pi * dd * dd * dd)
                              : 2.0f * alpha / 1e-9f;
  const float xv = n_dot_v * n_dot_v * (1.0f - a2) + a2;
  const float xl = n_dot_l * n_dot_l * (1.0f - a2) + a2;
  const float lambda = n_dot_l * __builtin_sqrtf(xv) + n_dot_v * __builtin_sqrtf(xl);
  const float dlambda = n_dot_l * alpha * (1.0f - n_dot_v * n_dot_v) / __builtin_sqrtf(xv) +
                        n_dot_v * alpha * (1.0f - n_dot_l * n_dot_l) / __builtin_sqrtf(xl);
  const float dv_dalpha = lambda > 1e-9f ? -0.5f * dlambda / (lambda * lambda) : 0.0f;
  df_dr = f * ((dd_dalpha * vis + d * dv_dalpha) * dalpha);
}

GE_DEV void onb(f3 n, f3& t, f3& b) {
  const float s = n.z >= 0.0f ? 1.0f : -1.0f;
  const float a = -1.0f / (s + n.z);
  const float bb = n.x * n.y * a;
  t = f3{1.0f + s * n.x * n.x * a, s * bb, -s * n.x};
  b = f3{bb, s + n.y * n.y * a, -n.y};
}
GE_DEV float luminance(f3 c) { return dot(c, f3{0.2126f, 0.7152f, 0.0722f}); }
GE_DEV f3 cosine_hemisphere(float ux, float uy) {
  const float r = __builtin_sqrtf(ux);
  const float phi = 2.0f * k_pi * uy;
  float sn, cs;
  sincos_f32(phi, &sn, &cs);
  return f3{r * cs, r * sn, __builtin_sqrtf(maxr(0.0f, 1.0f - ux))};
}
GE_DEV float ggx_g1(float n_dot_v, float alpha) {
  const float a2 = alpha * alpha;
  const float denom = n_dot_v + __builtin_sqrtf(a2 + (1.0f - a2) * n_dot_v * n_dot_v);
  return 2.0f * n_dot_v / maxr(denom, 1e-9f);
}
GE_DEV f3 ggx_vndf(f3 v, float alpha, float ux, float uy) {
  const f3 vh = normalize(f3{alpha * v.x, alpha * v.y, v.z});
  const float len_sq = vh.x * vh.x + vh.y * vh.y;
  const f3 t1 = len_sq > 0.0f ? f3{-vh.y, vh.x, 0.0f} * rsqrt_cr(len_sq) : f3{1.0f, 0.0f, 0.0f};
  const f3 t2 = cross(vh, t1);
  const float r = __builtin_sqrtf(ux);
  const float phi = 2.0f * k_pi * uy;
  float sn, cs;
  sincos_f32(phi, &sn, &cs);
  const float p1 = r * cs;
  float p2 = r * sn;
  const float s = 0.5f * (1.0f + vh.z);
  p2 = (1.0f - s) * __builtin_sqrtf(maxr(0.0f, 1.0f - p1 * p1)) + s * p2;
  const f3 nh = p1 * t1 + p2 * t2 + __builtin_sqrtf(maxr(0.0f, 1.0f - p1 * p1 - p2 * p2)) * vh;
  return normalize(f3{alpha * nh.x, alpha * nh.y, maxr(0.0f, nh.z)});
}
GE_DEV float ggx_vndf_pdf(f3 v, f3 l, float alpha) {
  if (v.z <= 0.0f || l.z <= 0.0f) return 0.0f;
  const f3 h = normalize(v + l);
  if (dot(v, h) <= 0.0f) return 0.0f;
  return ggx_g1(v.z, alpha) * d_ggx(h.z, alpha) / maxr(4.0f * v.z, 1e-9f);
}
GE_DEV float specular_probability(f3 albedo, float metallic, f3 f) {
  const float specular = luminance(f);
  const float diffuse = (1.0f - metallic) * luminance(albedo) * (1.0f - specular);
  const float total = specular + diffuse;
  if (total <= 1e-6f) return 0.5f;
  return clampr(specular / total, 0.05f, 0.95f);
}
GE_DEV float bsdf_pdf(f3 v, f3 l, float alpha, float p_specular) {
  return p_specular * ggx_vndf_pdf(v, l, alpha) + (1.0f - p_specular) * (maxr(l.z, 0.0f) / k_pi);
}
GE_DEV f3 reflect3(f3 i, f3 n) { return i - (n * 2.0f) * dot(i, n); }

GE_DEV float distance_attenuation(float distance_sq, float radius) {
  const float ratio = clampr(distance_sq / maxr(radius * radius, 1e-8f), 0.0f, 1.0f);
  const float window = 1.0f - ratio * ratio;
  return window * window / maxr(distance_sq, 1e-8f);
}
GE_DEV float spot_attenuation(f3 to_surface, f3 direction, float cos_inner, float cos_outer) {
  const float t = clampr((dot(to_surface, direction) - cos_outer) / maxr(cos_inner - cos_outer, 1e-4f), 0.0f, 1.0f);
  return t * t;
}
GE_DEV bool light_sample(const ge_resolve_light& light, f3 position, f3& direction, f3& radiance, float& distance,
                         float* attenuation_out = nullptr) {
  if (attenuation_out) *attenuation_out = 1.0f;
  if (light.direction_type[3] >= 1.5f) {
    direction = normalize(ld3(light.position_radius));
    radiance = ld3(light.color_intensity) * light.color_intensity[3];
    distance = 1e30f;
    return true;
  }
  const f3 to_light = ld3(light.position_radius) - position;
  const float distance_sq = dot(to_light, to_light);
  direction = f3{0.0f, 1.0f, 0.0f};
  radiance = splat(0.0f);
  distance = 0.0f;
  // At or beyond the radius the window is exactly zero (d^2 / r^2 rounds to at least 1, clamped to
  // 1), so the attenuation is zero: skip its two divisions. The shading loops visit every light
  // twice a vertex, and most are out of range.
  const float radius = light.position_radius[3];
  if (distance_sq >= maxr(radius * radius, 1e-8f)) return false;
  float attenuation = distance_attenuation(distance_sq, radius);
  if (attenuation <= 0.0f) return false;
  direction = normalize(to_light);
  if (light.direction_type[3] >= 0.5f)
    attenuation *= spot_attenuation(-direction, normalize(ld3(light.direction_type)), light.spot[0], light.spot[1]);
  radiance = ld3(light.color_intensity) * (light.color_intensity[3] * attenuation);
  if (attenuation_out) *attenuation_out = attenuation;
  distance = __builtin_sqrtf(distance_sq);
  return true;
}

GE_DEV f3 decode_normal_oct(uint32_t packed) {
  const float x = float(int32_t(packed << 16) >> 16) / 32767.0f;
  const float y = float(int32_t(packed) >> 16) / 32767.0f;
  f3 n = f3{x, y, 1.0f - __builtin_fabsf(x) - __builtin_fabsf(y)};
  if (n.z < 0.0f) {
    const float sx = x < 0.0f ? -1.0f : 1.0f, sy = y < 0.0f ? -1.0f : 1.0f;
    n.x = (1.0f - __builtin_fabsf(y)) * sx;
    n.y = (1.0f - __builtin_fabsf(x)) * sy;
  }
  return normalize(n);
}
GE_DEV f3 transform_normal(const ge_instance_desc& in, f3 n) {
  const f3 c0 = f3{in.rows[0][0], in.rows[1][0], in.rows[2][0]};
  const f3 c1 = f3{in.rows[0][1], in.rows[1][1], in.rows[2][1]};
  const f3 c2 = f3{in.rows[0][2], in.rows[1][2], in.rows[2][2]};
  const f3 x12 = cross(c1, c2);
  const f3 cof = x12 * n.x + cross(c2, c0) * n.y + cross(c0, c1) * n.z;
  return normalize(dot(c0, x12) < 0.0f ? -cof : cof);
}
// engine::relative_axis (core/math/world.h:181-195; scene.slang:120-126).
GE_DEV float relative_axis(int cell, float local, int eye_cell, float eye_local, float eye_residual) {
  const float cells = float(cell - eye_cell) * 64.0f;
  const float from_local = cells > 0.0f ? (cells - eye_local) + local : (cells + local) - eye_local;
  return from_local - eye_residual;
}

// A triangle fetched into frame space (fetch_triangle, material.slang:244-266): its corners off the
// mesh's 16-bit grid (load_quantized_position, scene.slang:126-170) through the instance's rows and
// eye-relative translation (instance_point, scene.slang:566-610), and its vertices' indices.
struct Corners {
  f3 p[3];
  uint32_t local[3];
};
template <class SceneV>
GE_DEV Corners fetch_corners(const SceneV& sv, const ge_frame_eye& eye, const ge_instance_desc& in,
                             uint32_t cluster_index, uint32_t triangle) {
  const ge_mesh_desc mesh = sv.meshes[in.mesh];
  const ge_cluster_desc cl = sv.clusters[cluster_index];
  const uint32_t packed = sv.triangles[cl.triangle_offset + triangle];
  const f3 T = f3{relative_axis(in.cell[0], in.rows[0][3], eye.cell[0], eye.local[0], eye.residual[0]),
                  relative_axis(in.cell[1], in.rows[1][3], eye.cell[1], eye.local[1], eye.residual[1]),
                  relative_axis(in.cell[2], in.rows[2][3], eye.cell[2], eye.local[2], eye.residual[2])};
  Corners c;
  const uint32_t* words = reinterpret_cast<const uint32_t*>(sv.quantized);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    c.local[k] = cl.vertex_offset + ((packed >> (8 * k)) & 0xffu);
    const uint32_t i = c.local[k] * 3;
    const uint32_t w0 = words[i >> 1], w1 = words[(i + 1) >> 1], w2 = words[(i + 2) >> 1];
    const uint32_t qx = (i & 1) ? (w0 >> 16) : (w0 & 0xffffu);
    const uint32_t qy = ((i + 1) & 1) ? (w1 >> 16) : (w1 & 0xffffu);
    const uint32_t qz = ((i + 2) & 1) ? (w2 >> 16) : (w2 & 0xffffu);
    const f3 m = f3{mesh.quant[0], mesh.quant[1], mesh.quant[2]} + f3{float(qx), float(qy), float(qz)} * mesh.quant[3];
    f3 d;
    d.x = in.rows[0][0] * m.x + in.rows[0][1] * m.y + in.rows[0][2] * m.z;
    d.y = in.rows[1][0] * m.x + in.rows[1][1] * m.y + in.rows[1][2] * m.z;
    d.z = in.rows[2][0] * m.x + in.rows[2][1] * m.y + in.rows[2][2] * m.z;
    c.p[k] = d + T;
  }
  return c;
}

struct Vertex {
  f3 position, normal, geometric;
};

// surface_at (material.slang:270-302): position, shading normal and geometric normal at
// barycentrics `b`, the geometric normal turned towards `towards`.
template <class SceneV>
GE_DEV Vertex surface_at(const SceneV& sv, const ge_instance_desc& in, const Corners& c, f3 b, f3 towards) {
  Vertex out;
  out.position = b.x * c.p[0] + b.y * c.p[1] + b.z * c.p[2];
  out.geometric = normalize(cross(c.p[1] - c.p[0], c.p[2] - c.p[0]));
  if (dot(out.geometric, towards - out.position) < 0.0f) out.geometric = -out.geometric;
  if (sv.attributes) {
    const f3 n0 = decode_normal_oct(sv.attributes[c.local[0]].normal_oct);
    const f3 n1 = decode_normal_oct(sv.attributes[c.local[1]].normal_oct);
    const f3 n2 = decode_normal_oct(sv.attributes[c.local[2]].normal_oct);
    const f3 n = b.x * n0 + b.y * n1 + b.z * n2;
    out.normal = transform_normal(in, n);
  } else {
    out.normal = out.geometric;
  }
  return out;
}

// ---- textured materials (material.slang:571-677, 746-767) -------------------------------------------
// A material after its level-0 textures.
struct MaterialSample {
  f3 albedo, emissive;
  float roughness, metallic;
};

// The three corners' UVs (decode_half2), zero without attributes.
template <class SceneV>
GE_DEV void corner_uvs(const SceneV& sv, const Corners& c, float uv[3][2]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const uint32_t h = sv.attributes ? sv.attributes[c.local[k]].uv_half2 : 0u;
    uv[k][0] = half_to_float(h & 0xffffu);
    uv[k][1] = half_to_float(h >> 16);
  }
}

// tangent_frame without a tangent stream: the frame the positions and UVs give, Gram-Schmidt against
// the normal, the bitangent negated; false where it degenerates.
GE_DEV bool tangent_frame(const Corners& c, const float uv[3][2], f3 normal, f3& tangent, f3& bitangent) {
  tangent = f3{1.0f, 0.0f, 0.0f};
  bitangent = f3{0.0f, 0.0f, 1.0f};
  const f3 edge1 = c.p[1] - c.p[0];
  const f3 edge2 = c.p[2] - c.p[0];
  const float du1 = uv[1][0] - uv[0][0], dv1 = uv[1][1] - uv[0][1];
  const float du2 = uv[2][0] - uv[0][0], dv2 = uv[2][1] - uv[0][1];
  const float area = du1 * dv2 - du2 * dv1;
  if (__builtin_fabsf(area) < 1e-12f) return false;
  const float r = 1.0f / area;
  f3 t = (edge1 * dv2 - edge2 * dv1) * r;
  f3 b = (edge2 * du1 - edge1 * du2) * r;
  t = t - normal * dot(normal, t);
  const float t_length = __builtin_sqrtf(dot(t, t));
  if (t_length < 1e-12f) return false;
  tangent = t / t_length;
  b = b - (normal * dot(normal, b) + tangent * dot(tangent, b));
  const float b_length = __builtin_sqrtf(dot(b, b));
  if (b_length < 1e-12f) return false;
  bitangent = b * (-1.0f / b_length);
  return true;
}

// sample_material_plain at (u, v) already through material_uv, and the normal map on `normal` where
// the material has one and the triangle a frame (path_trace.slang:326-349).
template <class SceneV>
GE_DEV MaterialSample shade_material(const SceneV& sv, const ge_resolve_material& m, const Corners& c, f3 b,
                                     f3& normal) {
  float uv[3][2];
  corner_uvs(sv, c, uv);
  const float su0 = (b.x * uv[0][0] + b.y * uv[1][0]) + b.z * uv[2][0];
  const float sv0 = (b.x * uv[0][1] + b.y * uv[1][1]) + b.z * uv[2][1];
  float su, sw;
  material_uv(m.flags, m.uv_transform, m.uv_offset, su0 * m.uv_scale, sv0 * m.uv_scale, su, sw);
  MaterialSample out;
  out.albedo = ld3(m.albedo);
  if (m.albedo_texture != GE_NO_TEXTURE) {
    const float4 t = sample_level0(sv.tex, m.albedo_texture, m.sampler, su, sw);
    out.albedo = out.albedo * f3{t.x, t.y, t.z};
  }
  out.roughness = m.albedo[3];
  out.metallic = m.emissive[3];
  if (m.metallic_roughness_texture != GE_NO_TEXTURE) {
    const float4 t = sample_level0(sv.tex, m.metallic_roughness_texture, slot_sampler(m, m.samplers_mr_normal, 0), su, sw);
    out.roughness = out.roughness * t.y;
    out.metallic = out.metallic * t.z;
  }
  out.emissive = ld3(m.emissive);
  if (m.emissive_texture != GE_NO_TEXTURE) {
    const float4 t =
        sample_level0(sv.tex, m.emissive_texture, slot_sampler(m, m.samplers_occlusion_emissive, 16), su, sw);
    out.emissive = out.emissive * f3{t.x, t.y, t.z};
  }
  if (m.normal_texture != GE_NO_TEXTURE && sv.attributes) {
    f3 t, bt;
    if (tangent_frame(c, uv, normal, t, bt)) {
      const float4 texel = sample_level0(sv.tex, m.normal_texture, slot_sampler(m, m.samplers_mr_normal, 16), su, sw);
      f3 tn;
      if (!(m.flags & GE_MATERIAL_NORMAL_RG)) {
        tn = f3{texel.x * 2.0f - 1.0f, texel.y * 2.0f - 1.0f, texel.z * 2.0f - 1.0f};
      } else {
        const float x = texel.x * 2.0f - 1.0f, y = texel.y * 2.0f - 1.0f;
        tn = f3{x, y, __builtin_sqrtf(clampr(1.0f - (x * x + y * y), 0.0f, 1.0f))};
      }
      tn.x = tn.x * m.normal_scale;
      tn.y = tn.y * m.normal_scale;
      const float length_sq = dot(tn, tn);
      if (!(length_sq <= 1e-12f)) {
        tn = tn * rsqrt_cr(length_sq);
        normal = normalize(t * tn.x + bt * tn.y + normal * tn.z);
      }
    }
  }
  return out;
}

// screen_barycentrics (material.slang:386-416), the barycentrics alone: where the ray through
// `pixel` (in pixels, y down) meets the triangle's plane, the edge functions on the plane's side
// of zero kept, the rest zeroed, over their sum.
GE_DEV f3 screen_barycentrics(const Corners& c, float px, float py, const float* vp, f3 eye, uint32_t width,
                              uint32_t height) {
  f3 cc[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const f3 v = c.p[i] - eye;
    cc[i].x = ((vp[0] * v.x + vp[4] * v.y) + vp[8] * v.z) + vp[12] * 0.0f;
    cc[i].y = ((vp[1] * v.x + vp[5] * v.y) + vp[9] * v.z) + vp[13] * 0.0f;
    cc[i].z = ((vp[3] * v.x + vp[7] * v.y) + vp[11] * v.z) + vp[15] * 0.0f;
  }
  const float ndx = 2.0f / float(width), ndy = -2.0f / float(height);
  const f3 p = f3{px * ndx + -1.0f, py * ndy + 1.0f, 1.0f};
  const f3 n0 = cross(cc[1], cc[2]);
  const f3 n1 = cross(cc[2], cc[0]);
  const f3 n2 = cross(cc[0], cc[1]);
  const f3 e = f3{dot(n0, p), dot(n1, p), dot(n2, p)};
  const float plane = e.x + e.y + e.z;
  const f3 s = plane < 0.0f ? -e : e;
  const f3 inside = f3{maxr(s.x, 0.0f), maxr(s.y, 0.0f), maxr(s.z, 0.0f)};
  const float total = inside.x + inside.y + inside.z;
  if (!(total > 0.0f)) return f3{1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 3.0f};
  return f3{inside.x / total, inside.y / total, inside.z / total};
}

}  // namespace ge
