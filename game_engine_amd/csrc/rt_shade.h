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
  sincos_cr(phi, &sn, &cs);
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
  sincos_cr(phi, &sn, &cs);
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
GE_DEV bool light_sample(const ge_resolve_light& light, f3 position, f3& direction, f3& radiance, float& distance) {
  if (light.direction_type[3] >= 1.5f) {
    direction = normalize(ld3(light.position_radius));
    radiance = ld3(light.color_intensity) * light.color_intensity[3];
    distance = 1e30f;
    return true;
  }
  const f3 to_light = ld3(light.position_radius) - position;
  const float distance_sq = dot(to_light, to_light);
  float attenuation = distance_attenuation(distance_sq, light.position_radius[3]);
  direction = f3{0.0f, 1.0f, 0.0f};
  radiance = splat(0.0f);
  distance = 0.0f;
  if (attenuation <= 0.0f) return false;
  direction = normalize(to_light);
  if (light.direction_type[3] >= 0.5f)
    attenuation *= spot_attenuation(-direction, normalize(ld3(light.direction_type)), light.spot[0], light.spot[1]);
  radiance = ld3(light.color_intensity) * (light.color_intensity[3] * attenuation);
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

}  // namespace ge
