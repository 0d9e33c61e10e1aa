"""Shared test inputs: seeded rays and small fixture scenes."""
import numpy as np

from game_engine_amd import abi


def random_rays(n, lo, hi, seed, tmax=1e30):
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, dtype=abi.RAY_DTYPE)
    rays["origin"] = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays["dir"] = d.astype(np.float32)
    rays["tmin"] = 0.0
    rays["tmax"] = tmax
    return rays


def rays_towards(n, lo, hi, target_lo, target_hi, seed, tmax=1e30):
    """Rays from a box towards points of another box: most of them hit something."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    t = rng.uniform(target_lo, target_hi, (n, 3))
    d = t - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays = np.zeros(n, dtype=abi.RAY_DTYPE)
    rays["origin"] = o.astype(np.float32)
    rays["dir"] = d.astype(np.float32)
    rays["tmax"] = tmax
    return rays


def unpack_bits(bits, n):
    return np.unpackbits(bits.view(np.uint8), bitorder="little")[:n].astype(bool)


def box_triangles(center, half, n=4):
    """A box with each face split n x n, outward winding, as (2 n^2 * 6, 9) float32."""
    import itertools
    tris = []
    axes = [((1, 0, 0), (0, 1, 0), (0, 0, 1)), ((-1, 0, 0), (0, 0, 1), (0, 1, 0)), ((0, 1, 0), (0, 0, 1), (1, 0, 0)),
            ((0, -1, 0), (1, 0, 0), (0, 0, 1)), ((0, 0, 1), (1, 0, 0), (0, 1, 0)), ((0, 0, -1), (0, 1, 0), (1, 0, 0))]
    c, h = np.asarray(center, float), np.asarray(half, float)
    for nrm, u, w in axes:
        nrm, u, w = map(np.array, (nrm, u, w))
        g = lambda i, j: c + h * (nrm + u * (-1 + 2 * i / n) + w * (-1 + 2 * j / n))
        for i, j in itertools.product(range(n), range(n)):
            a, b, d, e = g(i, j), g(i + 1, j), g(i, j + 1), g(i + 1, j + 1)
            tris += [[a, b, e], [a, e, d]]
    return np.array(tris, dtype=np.float32).reshape(-1, 9)


def with_materials(synth_scene, albedo, roughness, metallic):
    """The synth scene's descriptor with every material replaced by one (a white furnace's)."""
    import ctypes as C
    from game_engine_amd import abi
    d = abi.SceneDesc.from_buffer_copy(synth_scene.desc())
    mats = (abi.ResolveMaterial * d.material_count)()
    for m in mats:
        m.albedo[:] = [*albedo, roughness]
        m.emissive[:] = [0, 0, 0, metallic]
        for f in ("albedo_texture", "metallic_roughness_texture", "normal_texture", "occlusion_texture",
                  "emissive_texture"):
            setattr(m, f, 0xFFFFFFFF)
        m.uv_scale = 1.0
    d.materials = C.addressof(mats)
    return d, mats


def far_from_origin(s, pt, rv, shift=(150000, 3, -200000), residual=(7e-4, -3e-4, 5e-4)):
    """The synth scene `s` and its frame (pt, rv), moved `shift` world cells (64 m each) from the
    world's origin and seen from a frame whose eye (FrameEye, cluster_cull.h:939-960) is the camera:
    every instance's cell moves by `shift`, the eye is the camera's world position as a cell, a
    float local and a small residual, and the frame's camera, lights and projection become
    eye-relative (the camera at the frame's origin). Changes the instances of `s` in place.

    Returns (desc, eye, pt, rv, keep): the scene descriptor with the eye-relative lights, the eye,
    the two parameter blocks, and what must stay alive while desc is used."""
    import ctypes as C

    inst = s.arrays()["instances"]
    inst["cell"] += np.asarray(shift, dtype=np.int32)
    cam = np.array(pt.camera[:3], dtype=np.float64)
    base = np.floor(cam / 64.0)
    eye = abi.FrameEye()
    eye.cell[:] = [int(base[a]) + int(shift[a]) for a in range(3)]
    eye.local[:] = [float(np.float32(cam[a] - 64.0 * base[a] - residual[a])) for a in range(3)]
    eye.residual[:] = [float(np.float32(r)) for r in residual]
    d = abi.SceneDesc.from_buffer_copy(s.desc())
    lights = (abi.ResolveLight * max(d.light_count, 1))()
    if d.light_count:
        C.memmove(lights, d.lights, C.sizeof(abi.ResolveLight) * d.light_count)
        for l in lights[: d.light_count]:
            if l.direction_type[3] != 2.0:  # a positioned light: from the eye
                l.position_radius[:3] = [float(np.float32(l.position_radius[a] - cam[a])) for a in range(3)]
        d.lights = C.addressof(lights)
    pt2 = abi.PathTraceParams.from_buffer_copy(pt)
    rv2 = abi.RayVisibilityParams.from_buffer_copy(rv)
    pt2.camera[:3] = [0.0, 0.0, 0.0]
    rv2.camera[:3] = [0.0, 0.0, 0.0]
    pt2.eye = eye
    # The projection of a camera at the frame's origin: the inverse of clip_to_ray (the rotation).
    m = np.array(pt.clip_to_ray, dtype=np.float64).reshape(4, 4).T
    rv2.view_proj[:] = [float(x) for x in np.linalg.inv(m).T.reshape(-1).astype(np.float32)]
    return d, eye, pt2, rv2, lights


def even_depth(nodes):
    """Indices of the binary nodes at even depth (the root's is 0): those that carry a wide node."""
    import numpy as np
    parent = nodes["parent"].astype(np.int64)
    depth = np.zeros(len(nodes), dtype=np.int64)
    for i in range(len(nodes)):
        p = parent[i]
        while p != 0xFFFFFFFF:
            depth[i] += 1
            p = parent[p]
    return np.nonzero((depth & 1) == 0)[0]
