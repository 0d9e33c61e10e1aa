// ref_brdf.cpp — TEST INFRASTRUCTURE ONLY. A C entry point for each function of the reference's
// own double-precision BSDF twin (domain/gfx/tests/brdf_reference.h, which mirrors brdf.slang),
// compiled from the reference's header where it lies (oracle/Makefile, target `ref`) into
// oracle/_ref/libbrdf_ref.so. The tests hold the oracle's BSDF to it.
//
// Array layouts match the oracle's probes: a surface is albedo[3] normal[3] view[3] roughness
// metallic specular[4] (the dielectric's F0 less 0.04 per channel, then one less its F90).
#include <brdf_reference.h>

namespace {
namespace ref = engine::brdf_ref;

ref::Dvec3 d3(const double* p) { return ref::Dvec3{p[0], p[1], p[2]}; }
void put(double* out, ref::Dvec3 v) {
  out[0] = v.x;
  out[1] = v.y;
  out[2] = v.z;
}
ref::Surface surface(const double* in) {
  ref::Surface s;
  s.albedo = d3(in);
  s.normal = d3(in + 3);
  s.view = d3(in + 6);
  s.roughness = in[9];
  s.metallic = in[10];
  s.specular = d3(in + 11);
  s.specular_w = in[14];
  return s;
}
}  // namespace

extern "C" {

double ge_ref_alpha(double roughness) { return ref::alpha_of(roughness); }
double ge_ref_d_ggx(double n_dot_h, double alpha) { return ref::d_ggx(n_dot_h, alpha); }
double ge_ref_v_smith(double n_dot_v, double n_dot_l, double alpha) { return ref::v_smith(n_dot_v, n_dot_l, alpha); }

// in: surface[15], light direction[3]
void ge_ref_eval(const double* in, double* out) { put(out, ref::eval(surface(in), d3(in + 15))); }

// in: surface[15], cos_theta
void ge_ref_fresnel(const double* in, double* out) { put(out, ref::fresnel_of(surface(in), in[15])); }

// in: surface[15], light direction[3], radiance[3]: the resolve's direct term (cosine included)
void ge_ref_direct(const double* in, double* out) {
  put(out, ref::direct(surface(in), d3(in + 15), d3(in + 18)));
}

double ge_ref_distance_attenuation(double distance_sq, double radius) {
  return ref::distance_attenuation(distance_sq, radius);
}

// in: sky[3]: the sky's radiance above the horizon (material.slang's sky_radiance)
void ge_ref_sky_radiance(const double* in, double* out) { put(out, ref::sky_radiance(d3(in))); }

// in: ground albedo[3], sun direction[3], sun intensity, sky[3]: the lit ground below the horizon
// (brdf.slang's brdf_ground_radiance through material.slang's ground_radiance)
void ge_ref_ground_radiance(const double* in, double* out) {
  put(out, ref::ground_radiance(d3(in), d3(in + 3), in[6], d3(in + 7)));
}

double ge_ref_spot_attenuation(const double* to_surface, const double* direction, double cos_inner, double cos_outer) {
  return ref::spot_attenuation(d3(to_surface), d3(direction), cos_inner, cos_outer);
}

}  // extern "C"
