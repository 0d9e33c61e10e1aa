"""Python host over the C ABI of the HIP backend (game_engine_amd/lib/libge_rt.so).

`Context`, `Scene` and the trace calls map one to one onto include/ge_rt.h. `ReferenceRenderer`
mirrors the engine's `renderer::ReferenceRenderer` (systems/renderer/include/systems/renderer/
reference.h:51-170): `ReferenceSettings` in, a `ReferenceFrame` out (the accumulated linear
radiance mean with its sample count, the display bytes, the trace milliseconds), the batch loop of
`ceil(spp / batch)` dispatches with the same `(pixel, sample, seed)` streams, and an optional
progress callback that can stop the render early.

The backend has no CPU fallback: if the shared library is missing, or no GPU is present, every
call raises.
"""
import ctypes as C
import dataclasses
import math
import os
import time

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ERRORS = {0: "ok", 1: "invalid argument", 2: "invalid scene", 3: "out of memory", 4: "device error",
          5: "unsupported"}


class RtError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{ERRORS.get(status, status)}: {message}")
        self.status = status


def library_path():
    return os.path.join(_HERE, "lib", "libge_rt.so")


def lib():
    """Loads libge_rt.so; raises if it is not built (there is no fallback path)."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built: run __graft_entry__.build()")
        L = C.CDLL(path)
        vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int32
        sig = {
            "ge_rt_create": ([C.POINTER(vp), i32], i32),
            "ge_rt_destroy": ([vp], None),
            "ge_rt_last_error": ([vp, C.c_char_p, C.c_size_t], i32),
            "ge_rt_synchronize": ([vp], i32),
            "ge_rt_scene_upload": ([vp, C.POINTER(abi.SceneDesc), C.POINTER(vp)], i32),
            "ge_rt_scene_destroy": ([vp, vp], None),
            "ge_rt_bvh_build": ([vp, vp, u32, vp, vp, u32], i32),
            "ge_rt_scene_update_instances": ([vp, vp, u32, u32, vp], i32),
            "ge_rt_bvh_info": ([vp, vp, u32, C.POINTER(u32), C.POINTER(u32)], i32),
            "ge_rt_bvh_download": ([vp, vp, u32, vp, vp, vp, vp], i32),
            "ge_rt_bvh_download_wide": ([vp, vp, u32, vp], i32),
            "ge_rt_scene_set_visible": ([vp, vp, vp, u32], i32),
            "ge_rt_trace_closest": ([vp, vp, vp, u32, u32, vp, u32], i32),
            "ge_rt_trace_any": ([vp, vp, vp, u32, u32, vp, u32], i32),
            "ge_rt_ray_visibility": ([vp, vp, C.POINTER(abi.RayVisibilityParams), vp, u32], i32),
            "ge_rt_path_trace": ([vp, vp, C.POINTER(abi.PathTraceParams), vp, vp, u32], i32),
            "ge_rt_tonemap": ([vp, C.POINTER(abi.PathTraceParams), vp, vp, vp, u32], i32),
            "ge_rt_shadow_pass": ([vp, vp, C.POINTER(abi.ShadowParams), vp, vp, u32], i32),
            "ge_rt_shadow_resolve": ([vp, vp, C.POINTER(abi.ShadowResolveParams), vp, vp, u32], i32),
            "ge_rt_shadow_resolve_noalloc": ([vp, vp, C.POINTER(abi.ShadowResolveParams), vp, vp, u32], i32),
            "ge_rt_scene_update_materials": ([vp, vp, u32, u32, vp], i32),
            "ge_rt_scene_update_lights": ([vp, vp, u32, u32, vp], i32),
            "ge_rt_path_adjoint": ([vp, vp, C.POINTER(abi.PathTraceParams), vp, vp, vp, vp, vp, u32], i32),
            "ge_rt_cache_create": ([vp, vp, vp, u32, C.POINTER(vp)], i32),
            "ge_rt_cache_destroy": ([vp, vp], None),
            "ge_rt_cache_weights": ([vp, vp, vp, vp], i32),
            "ge_rt_cache_gradient": ([vp, vp, vp], i32),
            "ge_rt_cache_infer": ([vp, vp, vp, u32, vp, u32], i32),
            "ge_rt_cache_train": ([vp, vp, vp, u32, C.c_float, C.POINTER(C.c_double), u32], i32),
            "ge_rt_cache_collect": ([vp, vp, vp, C.POINTER(abi.PathTraceParams), vp, u32, C.POINTER(u32), u32], i32),
            "ge_rt_cache_samples": ([vp, vp, vp, u32, C.POINTER(u32), u32], i32),
            "ge_rt_cache_update": ([vp, vp, vp, C.POINTER(abi.PathTraceParams), u32, u32, C.c_float,
                                    C.POINTER(C.c_double), u32], i32),
            "ge_rt_path_trace_cached": ([vp, vp, vp, C.POINTER(abi.PathTraceParams), u32, vp, vp, vp,
                                         C.POINTER(C.c_double), u32], i32),
            "ge_rt_stats": ([vp, C.POINTER(abi.Counters)], i32),
            "ge_rt_reset_stats": ([vp], i32),
            "ge_clusters_read": ([vp, C.c_uint64, u32, C.POINTER(vp), C.c_char_p, C.c_size_t], i32),
            "ge_clusters_scene": ([vp, C.POINTER(abi.SceneDesc)], None),
            "ge_clusters_counts": ([vp, C.POINTER(u32), C.POINTER(u32)], None),
            "ge_clusters_destroy": ([vp], None),
            "ge_clusters_hash": ([vp, C.c_uint64], C.c_uint64),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _LIB = L
    return _LIB


EXPORTS = ("ge_rt_create", "ge_rt_destroy", "ge_rt_last_error", "ge_rt_synchronize", "ge_rt_scene_upload",
           "ge_rt_scene_destroy", "ge_rt_bvh_build", "ge_rt_scene_update_instances", "ge_rt_scene_set_visible",
           "ge_rt_bvh_info",
           "ge_rt_bvh_download", "ge_rt_bvh_download_wide", "ge_rt_trace_closest", "ge_rt_trace_any", "ge_rt_ray_visibility",
           "ge_rt_path_trace", "ge_rt_tonemap", "ge_rt_shadow_pass", "ge_rt_shadow_resolve",
           "ge_rt_shadow_resolve_noalloc", "ge_rt_stats",
           "ge_rt_reset_stats", "ge_rt_scene_update_materials", "ge_rt_scene_update_lights", "ge_rt_path_adjoint",
           "ge_rt_cache_create", "ge_rt_cache_destroy", "ge_rt_cache_weights", "ge_rt_cache_gradient", "ge_rt_cache_infer",
           "ge_rt_cache_train", "ge_rt_cache_collect", "ge_rt_cache_update", "ge_rt_path_trace_cached",
           "ge_rt_cache_samples", "ge_clusters_read", "ge_clusters_scene", "ge_clusters_counts",
           "ge_clusters_destroy", "ge_clusters_hash")


class ClustersFile:
    """A `.clusters` container read from memory (ge_clusters_read): its finest level as a scene
    descriptor for Context.upload. Raises ValueError with the reader's reason on a bad container."""

    def __init__(self, data, images=False):
        buf = (C.c_uint8 * len(data)).from_buffer_copy(bytes(data))
        err = C.create_string_buffer(512)
        h = C.c_void_p()
        st = lib().ge_clusters_read(buf, len(data), 1 if images else 0, C.byref(h), err, 512)
        if st != 0:
            raise ValueError(f"status {st}: {err.value.decode()}")
        self.h = h
        self._desc = abi.SceneDesc()
        lib().ge_clusters_scene(self.h, C.byref(self._desc))

    def desc(self):
        return self._desc

    def counts(self):
        a, b = C.c_uint32(), C.c_uint32()
        lib().ge_clusters_counts(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def close(self):
        if self.h:
            lib().ge_clusters_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def clusters_hash(data):
    """hash_bytes over `data` with the engine's seed: the container's content hash."""
    buf = (C.c_uint8 * max(len(data), 1)).from_buffer_copy(bytes(data) or b"\0")
    return lib().ge_clusters_hash(buf, len(data))


def _ptr(a):
    return None if a is None else (a if isinstance(a, int) else a.ctypes.data)


class Context:
    """One HIP device and its stream (ge_rt_create)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        st = lib().ge_rt_create(C.byref(h), device)
        if st != 0:
            raise RtError(st, f"ge_rt_create(device={device}) failed")
        self.h = h

    def check(self, status):
        if status != 0:
            buf = C.create_string_buffer(1024)
            lib().ge_rt_last_error(self.h, buf, 1024)
            raise RtError(status, buf.value.decode())

    def close(self):
        if self.h:
            lib().ge_rt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        self.check(lib().ge_rt_synchronize(self.h))

    def stats(self):
        c = abi.Counters()
        self.check(lib().ge_rt_stats(self.h, C.byref(c)))
        return {"closest": c.closest_rays, "any": c.any_rays, "nodes": c.nodes, "triangles": c.triangles,
                "instances": c.instances, "any_nodes": c.any_nodes, "any_triangles": c.any_triangles,
                "any_instances": c.any_instances, "kernel_ms": c.last_kernel_ms, "build_ms": c.last_build_ms,
                "closest_ms": c.closest_ms, "any_ms": c.any_ms, "closest_launches": c.closest_launches,
                "any_launches": c.any_launches, "mask_tests": c.mask_tests + c.any_mask_tests,
                "any_mask_tests": c.any_mask_tests}

    def reset_stats(self):
        self.check(lib().ge_rt_reset_stats(self.h))

    def upload(self, desc):
        return Scene(self, desc)


class Scene:
    """A scene on the device (ge_rt_scene_upload) and its structures (ge_rt_bvh_build)."""

    def __init__(self, ctx, desc):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(lib().ge_rt_scene_upload(ctx.h, C.byref(desc), C.byref(h)))
        self.h = h
        self.material_count = desc.material_count
        self.light_count = desc.light_count

    def close(self):
        if self.h:
            lib().ge_rt_scene_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build(self, flags=abi.BUILD_FULL, eye=None, meshes=None):
        m = None if meshes is None else np.ascontiguousarray(meshes, dtype=np.uint32)
        self.ctx.check(lib().ge_rt_bvh_build(self.ctx.h, self.h, flags, C.byref(eye) if eye else None, _ptr(m),
                                             0 if m is None else len(m)))
        return self

    def update_instances(self, first, records):
        self.ctx.check(lib().ge_rt_scene_update_instances(self.ctx.h, self.h, first, len(records), C.byref(records[0])))

    def set_visible(self, entries, check=True):
        """The frame's visible list (ge_rt_scene_set_visible): (n, 2) {instance, cluster} entries.
        Returns the status when `check` is false."""
        e = np.ascontiguousarray(entries, dtype=np.uint32).reshape(-1, 2)
        st = lib().ge_rt_scene_set_visible(self.ctx.h, self.h, _ptr(e) if len(e) else None, len(e))
        if check:
            self.ctx.check(st)
        return st

    def update_materials(self, first, records):
        self.ctx.check(lib().ge_rt_scene_update_materials(self.ctx.h, self.h, first, len(records), C.addressof(records)))

    def update_lights(self, first, records):
        self.ctx.check(lib().ge_rt_scene_update_lights(self.ctx.h, self.h, first, len(records), C.addressof(records)))

    def path_adjoint(self, params, d_image=None, sampling=None, image=True, grads=None, device=False,
                     megakernel=False, reuse_primal=False):
        """The path-replay adjoint (ge_rt_path_adjoint): (image, grad_materials, grad_lights).

        Host arrays by default: image (n, 4) float32 sample sums, gradients (materials, 4) and
        (1 + lights,) float64 when d_image (n, 4 float32, dLoss/dI) is given. With device=True,
        every argument is a device pointer (int) or None and nothing is returned. megakernel=True
        replays one path a lane with its vertices stored; the default is the wavefront.
        reuse_primal=True takes the primal pass from the context's last path_trace of the same
        parameters and scene when there is one (then no image: pass image=False or None)."""
        n = params.width * params.height
        flags = ((abi.DEVICE_PTR if device else 0) | (abi.PT_MEGAKERNEL if megakernel else 0) |
                 (abi.ADJOINT_REUSE_PRIMAL if reuse_primal else 0))
        if device:
            gm, gl = grads if grads is not None else (None, None)
            self.ctx.check(lib().ge_rt_path_adjoint(self.ctx.h, self.h, C.byref(params), d_image, sampling, image,
                                                    gm, gl, flags))
            return None
        img = np.zeros((n, 4), dtype=np.float32) if image else None
        gm = gl = None
        dptr = None
        if d_image is not None:
            d_image = np.ascontiguousarray(d_image, dtype=np.float32).reshape(n, 4)
            dptr = d_image.ctypes.data
            gm = np.zeros((self.material_count, 4), dtype=np.float64)
            gl = np.zeros(params.light_count + 1, dtype=np.float64)
        sptr = C.addressof(sampling) if sampling is not None else None
        self.ctx.check(lib().ge_rt_path_adjoint(self.ctx.h, self.h, C.byref(params), dptr, sptr, _ptr(img),
                                                _ptr(gm), _ptr(gl), flags))
        return img, gm, gl

    def bvh(self, mesh):
        n, p = C.c_uint32(), C.c_uint32()
        self.ctx.check(lib().ge_rt_bvh_info(self.ctx.h, self.h, mesh, C.byref(n), C.byref(p)))
        nodes = np.zeros(n.value, dtype=abi.NODE_DTYPE)
        morton = np.zeros(p.value, dtype=np.uint32)
        if mesh == 0xFFFFFFFF:
            prims = np.zeros(p.value, dtype=np.uint32)
            self.ctx.check(lib().ge_rt_bvh_download(self.ctx.h, self.h, mesh, _ptr(nodes), None, _ptr(prims), _ptr(morton)))
        else:
            prims = np.zeros(p.value, dtype=abi.TRI_DTYPE)
            self.ctx.check(lib().ge_rt_bvh_download(self.ctx.h, self.h, mesh, _ptr(nodes), _ptr(prims), None, _ptr(morton)))
        return nodes, prims, morton

    def bvh_wide(self, mesh):
        """The 4-wide nodes of mesh `mesh` (~0: the top level) at the binary nodes' indices."""
        n, p = C.c_uint32(), C.c_uint32()
        self.ctx.check(lib().ge_rt_bvh_info(self.ctx.h, self.h, mesh, C.byref(n), C.byref(p)))
        wide = np.zeros(n.value, dtype=abi.WIDE_DTYPE)
        self.ctx.check(lib().ge_rt_bvh_download_wide(self.ctx.h, self.h, mesh, _ptr(wide)))
        return wide

    def trace_closest(self, rays, mask=0xFF, count=False, hits=None, device=False, n=None, flags=0):
        """`flags` may add abi.ALPHA_MASKS or abi.CULL_MASKED (the masked-candidate rule)."""
        n = len(rays) if n is None else n
        hits = np.zeros(n, dtype=abi.HIT_DTYPE) if hits is None else hits
        flags |= (abi.COUNT if count else 0) | (abi.DEVICE_PTR if device else 0)
        self.ctx.check(lib().ge_rt_trace_closest(self.ctx.h, self.h, _ptr(rays), n, mask, _ptr(hits), flags))
        return hits

    def trace_any(self, rays, mask=0xFF, count=False, bits=None, device=False, n=None, flags=0):
        n = len(rays) if n is None else n
        bits = np.zeros((n + 31) // 32, dtype=np.uint32) if bits is None else bits
        flags |= (abi.COUNT if count else 0) | (abi.DEVICE_PTR if device else 0)
        self.ctx.check(lib().ge_rt_trace_any(self.ctx.h, self.h, _ptr(rays), n, mask, _ptr(bits), flags))
        return bits

    def ray_visibility(self, params, count=False, words=None, device=False):
        words = np.zeros(params.width * params.height, dtype=np.uint64) if words is None else words
        flags = (abi.COUNT if count else 0) | (abi.DEVICE_PTR if device else 0)
        self.ctx.check(lib().ge_rt_ray_visibility(self.ctx.h, self.h, C.byref(params), _ptr(words), flags))
        return words

    def path_trace(self, params, accum=None, coverage=None, count=False, device=False, megakernel=False):
        n = params.width * params.height
        if not device:
            accum = np.zeros((n, 4), dtype=np.float32) if accum is None else accum
            coverage = np.zeros(n, dtype=np.uint32) if coverage is None else coverage
        flags = ((abi.COUNT if count else 0) | (abi.DEVICE_PTR if device else 0) |
                 (abi.PT_MEGAKERNEL if megakernel else 0))
        self.ctx.check(lib().ge_rt_path_trace(self.ctx.h, self.h, C.byref(params), _ptr(accum), _ptr(coverage), flags))
        return accum, coverage

    def shadow_pass(self, params, receivers, count=False, bits=None, device=False, flags=0):
        bits = np.zeros((params.count + 31) // 32, dtype=np.uint32) if bits is None else bits
        flags |= (abi.COUNT if count else 0) | (abi.DEVICE_PTR if device else 0)
        self.ctx.check(lib().ge_rt_shadow_pass(self.ctx.h, self.h, C.byref(params), _ptr(receivers), _ptr(bits), flags))
        return bits

    def shadow_resolve(self, params, words, count=False, out=None, device=False):
        """The sun shadow pass over a visibility buffer: per pixel (or 2x2 block at half rate)
        pair << 1 | blocked, or abi.SHADOW_NOT_TAKEN."""
        out = np.zeros(abi.shadow_entries(params), dtype=np.uint32) if out is None else out
        flags = (abi.COUNT if count else 0) | (abi.DEVICE_PTR if device else 0)
        self.ctx.check(lib().ge_rt_shadow_resolve(self.ctx.h, self.h, C.byref(params), _ptr(words), _ptr(out), flags))
        return out


class RadianceCache:
    """The radiance cache (ge_rt_cache_*): an MLP on the GPU's matrix cores, trained online on the
    path tracer's own path suffixes, and the cached tracer that ends its paths into it."""

    def __init__(self, ctx, box_min, box_max, seed=0):
        self.ctx = ctx
        lo = (C.c_float * 3)(*[float(v) for v in box_min])
        hi = (C.c_float * 3)(*[float(v) for v in box_max])
        h = C.c_void_p()
        ctx.check(lib().ge_rt_cache_create(ctx.h, lo, hi, seed, C.byref(h)))
        self.h = h
        self.box_min = np.array(lo, dtype=np.float32)
        self.box_max = np.array(hi, dtype=np.float32)

    def close(self):
        if self.h:
            lib().ge_rt_cache_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def box_inv(self):
        """The normalisation the backend applies: 1 / (max - min) in float32."""
        return (np.float32(1.0) / (self.box_max - self.box_min)).astype(np.float32)

    def weights(self):
        w = np.zeros(abi.CACHE_WEIGHTS, dtype=np.float32)
        self.ctx.check(lib().ge_rt_cache_weights(self.ctx.h, self.h, w.ctypes.data, None))
        return w

    def gradient(self):
        """The last optimizer step's FP32 weight gradient (d mean loss / d weights)."""
        g = np.zeros(abi.CACHE_WEIGHTS, dtype=np.float32)
        self.ctx.check(lib().ge_rt_cache_gradient(self.ctx.h, self.h, g.ctypes.data))
        return g

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == abi.CACHE_WEIGHTS
        self.ctx.check(lib().ge_rt_cache_weights(self.ctx.h, self.h, None, w.ctypes.data))

    def infer(self, queries, out=None, device=False, n=None):
        n = len(queries) if n is None else n
        if not device:
            queries = np.ascontiguousarray(queries, dtype=abi.CACHE_QUERY_DTYPE)
            out = np.zeros((n, 3), dtype=np.float32) if out is None else out
        self.ctx.check(lib().ge_rt_cache_infer(self.ctx.h, self.h, _ptr(queries), n, _ptr(out),
                                               abi.DEVICE_PTR if device else 0))
        return out

    def train(self, samples, learning_rate, device=False, n=None):
        n = len(samples) if n is None else n
        if not device:
            samples = np.ascontiguousarray(samples, dtype=abi.CACHE_SAMPLE_DTYPE)
        loss = C.c_double()
        self.ctx.check(lib().ge_rt_cache_train(self.ctx.h, self.h, _ptr(samples), n, learning_rate, C.byref(loss),
                                               abi.DEVICE_PTR if device else 0))
        return loss.value

    def collect(self, scene, params, capacity):
        """Training samples of params' paths: (samples[:min(count, capacity)], count)."""
        out = np.zeros(capacity, dtype=abi.CACHE_SAMPLE_DTYPE)
        count = C.c_uint32()
        self.ctx.check(lib().ge_rt_cache_collect(self.ctx.h, scene.h, self.h, C.byref(params), out.ctypes.data,
                                                 capacity, C.byref(count), 0))
        return out[:min(count.value, capacity)], count.value

    def samples(self, capacity=1 << 20):
        """The samples the last gathering call kept, with their targets."""
        out = np.zeros(capacity, dtype=abi.CACHE_SAMPLE_DTYPE)
        count = C.c_uint32()
        self.ctx.check(lib().ge_rt_cache_samples(self.ctx.h, self.h, out.ctypes.data, capacity, C.byref(count), 0))
        return out[:count.value]

    def update(self, scene, params, samples=1 << 16, steps=4, learning_rate=1e-2):
        loss = C.c_double()
        self.ctx.check(lib().ge_rt_cache_update(self.ctx.h, scene.h, self.h, C.byref(params), samples, steps,
                                                learning_rate, C.byref(loss), 0))
        return loss.value

    def path_trace(self, scene, params, depth=1, accum=None, coverage=None, count=False, device=False, train=None,
                   want_loss=False):
        """The cached tracer (ge_rt_path_trace_cached): (accum, coverage), plus the training loss
        when `train` (an abi.CacheTrainParams) is given and want_loss is set."""
        n = params.width * params.height
        if not device:
            accum = np.zeros((n, 4), dtype=np.float32) if accum is None else accum
            coverage = np.zeros(n, dtype=np.uint32) if coverage is None else coverage
        flags = (abi.COUNT if count else 0) | (abi.DEVICE_PTR if device else 0)
        loss = C.c_double()
        self.ctx.check(lib().ge_rt_path_trace_cached(self.ctx.h, scene.h, self.h, C.byref(params), depth,
                                                     C.byref(train) if train is not None else None, _ptr(accum),
                                                     _ptr(coverage), C.byref(loss) if want_loss else None, flags))
        if want_loss:
            return accum, coverage, loss.value
        return accum, coverage


def tonemap(ctx, params, accum, coverage, device=False, output=None):
    n = params.width * params.height
    output = np.zeros(n, dtype=np.uint32) if output is None else output
    ctx.check(lib().ge_rt_tonemap(ctx.h, C.byref(params), _ptr(accum), _ptr(coverage), _ptr(output),
                                  abi.DEVICE_PTR if device else 0))
    return output


# ---- the engine's ReferenceRenderer, over the C ABI --------------------------------------------

@dataclasses.dataclass
class ReferenceSettings:
    """renderer::ReferenceSettings (reference.h:51-84): what a reference render is asked for."""
    spp: int = 64
    max_bounces: int = 3
    seed: int = 1
    batch: int = 8
    uniform_sky: bool = False
    pixel_center: bool = False


@dataclasses.dataclass
class ReferenceFrame:
    """renderer::ReferenceFrame (reference.h:86-98): the linear mean (rgb + count), RGBA8, timings."""
    width: int = 0
    height: int = 0
    samples: int = 0
    hdr: np.ndarray = None
    color: np.ndarray = None
    trace_ms: float = 0.0
    seconds: float = 0.0


class ReferenceRenderer:
    """renderer::ReferenceRenderer (reference.h:114-170), with the HIP backend as its device.

    `create` takes a context and an uploaded, built scene; `render` fills `PathTraceParams` from a
    frame template (camera, sky, sun, lights, ray bias — what the engine's `frame_lighting` and
    camera give it) and the settings, runs ceil(spp / batch) batches, then the tonemap and the
    readback. Errors are returned as (False, message), as the engine's bool + error string.
    """

    def __init__(self):
        self.ctx = None
        self.scene = None
        self.width = self.height = 0

    def create(self, ctx, scene, width, height):
        if ctx is None or scene is None:
            return False, "no device or scene"
        self.ctx, self.scene = ctx, scene
        return self.resize(width, height)

    def resize(self, width, height):
        if width <= 0 or height <= 0:
            return False, "empty extent"
        self.width, self.height = width, height
        return True, ""

    def destroy(self):
        self.ctx = self.scene = None

    def valid(self):
        return self.ctx is not None

    def render(self, frame_params, settings, progress=None):
        if not self.valid():
            return False, "not created", None
        p = abi.PathTraceParams.from_buffer_copy(frame_params)
        p.width, p.height = self.width, self.height
        p.max_bounces = settings.max_bounces
        p.seed = settings.seed
        p.flags = (abi.PT_UNIFORM_SKY if settings.uniform_sky else 0) | (abi.PT_PIXEL_CENTER if settings.pixel_center else 0)
        if settings.uniform_sky:  # the furnace: the sun and the analytic lights are off (reference.h:61-66)
            p.sun[3] = 0.0
            p.light_count = 0
        n = self.width * self.height
        accum = np.zeros((n, 4), dtype=np.float32)
        coverage = np.zeros(n, dtype=np.uint32)
        batch = max(1, settings.batch)
        total = max(0, settings.spp)
        start = time.perf_counter()
        trace_ms = 0.0
        done = 0
        try:
            while done < total:
                p.spp = min(batch, total - done)
                p.sample_base = done
                self.scene.path_trace(p, accum, coverage)
                trace_ms += self.ctx.stats()["kernel_ms"]
                done += p.spp
                if progress is not None and not progress(done, total):
                    break
            out = tonemap(self.ctx, p, accum, coverage)
        except RtError as e:
            return False, str(e), None
        hdr = accum.copy()
        nz = hdr[:, 3] > 0
        hdr[nz, :3] /= hdr[nz, 3:4]
        color = np.stack([(out >> s) & 255 for s in (0, 8, 16, 24)], axis=-1).astype(np.uint8)
        frame = ReferenceFrame(self.width, self.height, done, hdr, color, trace_ms, time.perf_counter() - start)
        return True, "", frame
