"""ctypes mirror of include/ge_rt.h: the engine's GPU records and the backend's own.

Layouts are asserted against the engine's size tables on import (ClusterDesc 48, MeshDesc 64,
InstanceDesc 96, FrameEye 48, ResolveMaterial 128, ResolveLight 64, PathTraceParams 368,
RayVisibilityParams 256 — domain/gfx/tests/size_table.cpp:41,122,125,388,397,452 and
domain/geometry/tests/size_table.cpp:13 in the reference).
"""
import ctypes as C

import numpy as np

u32, i32, u64, f32, f64 = C.c_uint32, C.c_int32, C.c_uint64, C.c_float, C.c_double


class ClusterDesc(C.Structure):
    _fields_ = [("vertex_offset", u32), ("triangle_offset", u32), ("vertex_count", u32),
                ("triangle_count", u32), ("center", f32 * 3), ("radius", f32),
                ("cone_apex", f32 * 3), ("cone", u32)]


class VertexAttributes(C.Structure):
    _fields_ = [("normal_oct", u32), ("uv_half2", u32)]


class MeshDesc(C.Structure):
    _fields_ = [("quant", f32 * 4), ("quantized", u64), ("first_cluster", u32),
                ("cluster_count", u32), ("deform_pool", u64), ("deform_slots", u64),
                ("templates", u64), ("skin", u64)]


class InstanceDesc(C.Structure):
    _fields_ = [("rows", (f32 * 4) * 3), ("cell", i32 * 3), ("mesh", u32), ("material_base", u32),
                ("first_pair", u32), ("scale_max", f32), ("flags", u32), ("deform", u32),
                ("bounds_padding", f32), ("terrain", u32), ("band", u32)]


class FrameEye(C.Structure):
    _fields_ = [("cell", i32 * 3), ("masks", u32), ("local", f32 * 3), ("pad1", f32),
                ("residual", f32 * 3), ("pad2", f32)]


class ResolveMaterial(C.Structure):
    _fields_ = [("albedo", f32 * 4), ("emissive", f32 * 4), ("albedo_texture", u32),
                ("sampler", u32), ("uv_scale", f32), ("flags", u32),
                ("metallic_roughness_texture", u32), ("normal_texture", u32),
                ("normal_scale", f32), ("occlusion_texture", u32), ("emissive_texture", u32),
                ("occlusion_strength", f32), ("samplers_mr_normal", u32),
                ("samplers_occlusion_emissive", u32), ("uv_transform", f32 * 4),
                ("uv_offset", f32 * 4), ("specular", f32 * 4)]


class ResolveLight(C.Structure):
    _fields_ = [("position_radius", f32 * 4), ("color_intensity", f32 * 4),
                ("direction_type", f32 * 4), ("spot", f32 * 4)]


class PathTraceParams(C.Structure):
    _fields_ = [("clip_to_ray", f32 * 16), ("camera", f32 * 4), ("sky", f32 * 4), ("sun", f32 * 4),
                ("ground", f32 * 4), ("accum", u64), ("output", u64), ("clusters", u64),
                ("mesh", u64), ("triangles", u64), ("instances", u64), ("attributes", u64),
                ("materials", u64), ("cluster_materials", u64), ("visible", u64), ("lights", u64),
                ("width", u32), ("height", u32), ("light_count", u32), ("scene", u32),
                ("spp", u32), ("sample_base", u32), ("max_bounces", u32), ("seed", u32),
                ("flags", u32), ("ray_bias", f32), ("coverage", u64), ("background", u32),
                ("ray_bias_steps", f32), ("ground_detail", u64), ("sky_params", u64),
                ("eye", FrameEye), ("terrain_uv_x0_mm", i32), ("terrain_uv_z0_mm", i32),
                ("terrain_uv_per_mm", f32), ("extra_vertices", u32), ("hit_entries", u64),
                ("extra", u64)]


class RayVisibilityParams(C.Structure):
    _fields_ = [("view_proj", f32 * 16), ("clip_to_ray", f32 * 16), ("camera", f32 * 4),
                ("output", u64), ("instance_base", u64), ("width", u32), ("height", u32),
                ("scene", u32), ("caster_base", u32), ("visible", u64), ("instances", u64),
                ("meshes", u64), ("triangles", u64), ("clusters", u64), ("hit_entries", u64),
                ("masks", u64 * 4)]


class ShadowParams(C.Structure):
    _fields_ = [("sun", f32 * 4), ("offset", f32), ("offset_steps", f32), ("t_max", f32),
                ("count", u32), ("pad", u32 * 4)]


class ShadowResolveParams(C.Structure):
    _fields_ = [("view_proj", f32 * 16), ("camera", f32 * 4), ("sun", f32 * 4), ("width", u32), ("height", u32),
                ("shadow_bias", f32), ("shadow_bias_steps", f32), ("flags", u32), ("seed", u32), ("pad", u32 * 2)]


class Counters(C.Structure):
    _fields_ = [("closest_rays", u64), ("any_rays", u64), ("nodes", u64), ("triangles", u64),
                ("instances", u64), ("last_kernel_ms", f64), ("last_build_ms", f64), ("any_nodes", u64),
                ("any_triangles", u64), ("any_instances", u64), ("closest_ms", f64), ("any_ms", f64),
                ("closest_launches", u32), ("any_launches", u32), ("mask_tests", u64), ("any_mask_tests", u64)]


class SceneDesc(C.Structure):
    _fields_ = [("clusters", C.c_void_p), ("cluster_count", u32), ("vertex_count", u32),
                ("vertices", C.c_void_p), ("quantized", C.c_void_p), ("attributes", C.c_void_p),
                ("triangles", C.c_void_p), ("triangle_count", u32), ("mesh_count", u32),
                ("meshes", C.c_void_p), ("instances", C.c_void_p), ("instance_count", u32),
                ("material_count", u32), ("materials", C.c_void_p),
                ("cluster_materials", C.c_void_p), ("lights", C.c_void_p), ("light_count", u32),
                ("texture_count", u32), ("textures", C.c_void_p), ("texels", C.c_void_p),
                ("texel_bytes", u64), ("samplers", C.c_void_p), ("sampler_count", u32), ("pad", u32)]


class TextureDesc(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("format", u32), ("reserved", u32), ("offset", u64)]


class SamplerDesc(C.Structure):
    _fields_ = [("address_u", u32), ("address_v", u32), ("mag", u32), ("min", u32)]


# The radiance cache's records (no counterpart in the reference).
CACHE_INPUTS = 192
CACHE_WIDTH = 64
CACHE_OUTPUTS = 16
CACHE_LAYERS = 6
CACHE_WEIGHTS = 29696
CACHE_DIMS = (CACHE_INPUTS, CACHE_WIDTH, CACHE_WIDTH, CACHE_WIDTH, CACHE_WIDTH, CACHE_WIDTH, CACHE_OUTPUTS)
CACHE_QUERY_DTYPE = np.dtype([("position", "<f4", 3), ("direction", "<f4", 2), ("normal", "<f4", 2),
                              ("albedo", "<f4", 3), ("roughness", "<f4"), ("tag", "<u4")])
CACHE_SAMPLE_DTYPE = np.dtype([("query", CACHE_QUERY_DTYPE), ("radiance", "<f4", 3), ("path", "<u4")])
assert CACHE_QUERY_DTYPE.itemsize == 48 and CACHE_SAMPLE_DTYPE.itemsize == 64
assert sum(a * b for a, b in zip(CACHE_DIMS[:-1], CACHE_DIMS[1:])) == CACHE_WEIGHTS

class CacheTrainParams(C.Structure):
    _fields_ = [("samples", u32), ("steps", u32), ("learning_rate", f32), ("stride", u32)]


# numpy views of the backend's own records.
RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("instance_id", "<u4"),
                      ("geometry", "<u4"), ("primitive", "<u4"), ("key", "<u4"), ("pad", "<u4")])
NODE_DTYPE = np.dtype([("lo0", "<f4", 3), ("hi0", "<f4", 3), ("lo1", "<f4", 3), ("hi1", "<f4", 3),
                       ("child", "<u4", 2), ("parent", "<u4"), ("count", "<u4")])
WIDE_DTYPE = np.dtype([("origin", "<f4", 3), ("meta", "<u4"), ("child", "<u4", 4), ("lo", "<u4", 3), ("hi", "<u4", 3),
                       ("reserved", "<u4", 2)])
TRI_DTYPE = np.dtype([("v0", "<f4", 3), ("v1", "<f4", 3), ("v2", "<f4", 3), ("prim", "<u4"),
                      ("index", "<u4"), ("pad", "<u4")])
RECEIVER_DTYPE = np.dtype([("position", "<f4", 3), ("instance_id", "<u4"), ("geometric", "<f4", 3),
                           ("pad", "<u4")])

SIZES = {ClusterDesc: 48, VertexAttributes: 8, MeshDesc: 64, InstanceDesc: 96, FrameEye: 48,
         ResolveMaterial: 128, ResolveLight: 64, PathTraceParams: 368, RayVisibilityParams: 256,
         ShadowParams: 48, Counters: 120, ShadowResolveParams: 128, CacheTrainParams: 16,
         TextureDesc: 24, SamplerDesc: 16, SceneDesc: 152}
for _t, _n in SIZES.items():
    assert C.sizeof(_t) == _n, (_t.__name__, C.sizeof(_t), _n)
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 32
assert NODE_DTYPE.itemsize == 64 and WIDE_DTYPE.itemsize == 64 and TRI_DTYPE.itemsize == 48 and RECEIVER_DTYPE.itemsize == 32

OK = 0
DEVICE_PTR = 0x1
COUNT = 0x2
PT_MEGAKERNEL = 0x4
ADJOINT_REUSE_PRIMAL = 0x8
ALPHA_MASKS = 0x10
CULL_MASKED = 0x20
BVH_LEAF = 0x80000000
BVH_NONE = 0xFFFFFFFF
MISS = 0xFFFFFFFF
FRAME_STRUCTURE = 0x80000000  # bvh_info / download selector: the frame structure of instance (sel & ~bit)
PT_UNIFORM_SKY = 1
PT_PIXEL_CENTER = 2
PT_DITHER = 4
PT_ALPHA_MASKS = 8
BUILD_FULL, BUILD_SUBSET, BUILD_TOP, BUILD_REFIT = 0, 1, 2, 3
# Status codes (ge_rt_status).
OK, ERROR_INVALID_ARGUMENT, ERROR_INVALID_SCENE, ERROR_OUT_OF_MEMORY, ERROR_DEVICE, ERROR_UNSUPPORTED = 0, 1, 2, 3, 4, 5
SHADOW_HALF_RATE = 0x1
SHADOW_JITTER = 0x2
SHADOW_ALPHA_MASKS = 0x4
NO_TEXTURE = 0xFFFFFFFF
SAME_SAMPLER = 0xFFFF
MATERIAL_MIPPED = 1
MATERIAL_NORMAL_RG = 2
MATERIAL_UV_TRANSFORM = 4
MATERIAL_MASKED = 32
MATERIAL_STOCHASTIC = 131072
FORMAT_RGBA8_UNORM = 37
FORMAT_RGBA8_SRGB = 43
FILTER_NEAREST, FILTER_LINEAR = 0, 1
ADDRESS_REPEAT, ADDRESS_MIRRORED_REPEAT, ADDRESS_CLAMP_TO_EDGE, ADDRESS_CLAMP_TO_BORDER = 0, 1, 2, 3
SHADOW_NOT_TAKEN = 0xFFFFFFFF
SHADOW_STATIC_ENTRIES = 1 << 21  # the allocation-free shadow resolve's most entries


def shadow_entries(params):
    """Entries of a shadow resolve's output: pixels, or 2x2 blocks at half rate."""
    if params.flags & SHADOW_HALF_RATE:
        return ((params.width + 1) // 2) * ((params.height + 1) // 2)
    return params.width * params.height
INSTANCE_STATIC_RT = 128
