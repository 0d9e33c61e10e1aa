"""The two headline workloads at their full sizes on the GPU, checked against the oracle.

- The 4M-triangle scene (4,096 instances), 1920 x 1080, jittered primaries, 2 spp, 8 bounces, the
  sun and 64 point lights: structures bit-exact, coverage bit-exact, radiance within 1e-4 relative
  (the FP32 oracle runs the shader's operation order), closest hits and occlusion bits of 200,000
  rays bit-exact, in the wavefront and megakernel forms.
- The 50.3M-triangle shadow scene (4096^2 fBm heightfield, 16,384 unique 1,024-triangle cluster
  groups), 3840 x 2160: a tenth of the groups moved and their structures and the instance
  structure rebuilt (GE_RT_BUILD_SUBSET), then the visibility words of every pixel and the jittered
  sun-shadow answers of every covered pixel, each bit-exact against the oracle's walk of the moved
  scene.

The oracle builds the 50M-triangle scene in about 35 s on 8 cores; each test prints its stages.
"""
import time

import numpy as np
import pytest

from game_engine_amd import abi, rt, synth
from oracle import binding as ob
from tests.helpers import even_depth, rays_towards

pytestmark = pytest.mark.gpu
TOP = 0xFFFFFFFF


def _log(msg, t0):
    print(f"[full size] {msg} ({time.perf_counter() - t0:.1f} s)", flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = rt.Context(0)
    yield c
    c.close()


def test_path_trace_4m_triangles_1080p(ctx):
    t0 = time.perf_counter()
    s = synth.Synth.config(2, 0, 4096)
    pt, _ = s.frame(1920, 1080)
    assert (pt.width, pt.height, pt.spp, pt.max_bounces, pt.light_count) == (1920, 1080, 2, 8, 64)
    g = ctx.upload(s.desc()).build()
    o = ob.Scene(s.desc())
    inst = len(s.arrays()["instances"])
    _log(f"{inst} instances of {s.triangle_count} unique triangles built on both sides", t0)
    gn, gp, gm = g.bvh(TOP)
    on, op, om = o.bvh(TOP)
    assert np.array_equal(gm, om) and gp.tobytes() == op.tobytes() and gn.tobytes() == on.tobytes()
    even = even_depth(on)
    assert g.bvh_wide(TOP)[even].tobytes() == o.bvh_wide(TOP)[even].tobytes(), "quantized wide nodes"
    lo, hi = s.bounds()
    rays = rays_towards(200000, lo, hi, lo, hi, 5)
    assert g.trace_closest(rays).tobytes() == o.trace_closest(rays).tobytes()
    assert np.array_equal(g.trace_any(rays, mask=0x01), o.trace_any(rays, mask=0x01))
    _log("instance structure, hits and occlusion bit-exact", t0)
    pt.sample_base = 6
    oa, oc = o.path_trace(pt)
    _log("oracle frame", t0)
    for megakernel in (False, True):
        ga, gc = g.path_trace(pt, megakernel=megakernel)
        assert np.array_equal(gc, oc)
        rel = np.abs(ga[:, :3] - oa[:, :3]) / np.maximum(np.abs(oa[:, :3]), 1e-6)
        exact = np.mean(np.all(ga == oa, axis=1))
        _log(f"megakernel={megakernel}: max rel err {rel.max():.3g}, bit-exact pixels {exact:.5f}", t0)
        assert rel.max() <= 1e-4
    g.close()
    o.close()


def test_shadow_pass_50m_triangles_after_a_move(ctx):
    t0 = time.perf_counter()
    s = synth.Synth.config(3, 0, 16384, 4096)
    assert s.triangle_count == 50315266
    g = ctx.upload(s.desc()).build()
    _log(f"{s.triangle_count} triangles built on the GPU in {ctx.stats()['build_ms']:.1f} ms", t0)
    o = ob.Scene(s.desc())
    _log("oracle built", t0)
    groups = len(s.arrays()["instances"]) - 1
    moving = groups // 10
    first = 1 + 3 * moving
    recs = s.move_instances(first, moving, 1003, 0.5, 0.2)
    g.update_instances(first, recs)
    g.build(abi.BUILD_SUBSET, meshes=[recs[i].mesh for i in range(moving)])
    o.update_instances(first, recs)
    gn, gp, gm = g.bvh(TOP)
    on, op, om = o.bvh(TOP)
    assert np.array_equal(gm, om) and gp.tobytes() == op.tobytes() and gn.tobytes() == on.tobytes()
    _log(f"{moving} groups moved, rebuilt; instance structures bit-exact", t0)
    pt, rv = s.frame(3840, 2160)
    words = g.ray_visibility(rv)
    want = o.ray_visibility(rv)
    assert np.count_nonzero(want) > 4_000_000
    assert np.array_equal(words, want)
    _log(f"visibility words of {len(words)} pixels bit-exact", t0)
    for flags, seed in ((abi.SHADOW_JITTER, 3), (0, 0)):
        p = synth.Synth.shadow_params(pt, rv, flags, seed)
        ref, _ = o.shadow_resolve(p, want)
        assert np.count_nonzero(ref != abi.SHADOW_NOT_TAKEN) > 4_000_000
        assert np.array_equal(g.shadow_resolve(p, words), ref)
        _log(f"shadow answers (flags {flags}) bit-exact", t0)
    g.close()
    o.close()
