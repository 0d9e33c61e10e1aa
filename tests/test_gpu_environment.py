"""The environment and the world-cell frame on the GPU against the oracle.

- The two closed-form plane scenes the oracle is pinned on (every escaped ray below the horizon, or
  above it with the sun off): radiance within 1e-4 relative, coverage bit-exact.
- The 4M-triangle layout at 512 instances moved 150,000 world cells from the origin and traced from
  an eye at its camera with a non-zero local and residual (FrameEye, cluster_cull.h:939-960;
  relative_axis, world.h:181-188, restated in the instance transforms the structures are built
  with and in the hit decode): the instance structure, hits, visibility words, radiance and the
  shadow resolve, each to the usual bar.
"""
import numpy as np
import pytest

from game_engine_amd import abi, rt, synth
from oracle import binding as ob
from tests.helpers import far_from_origin, rays_towards
from tests.test_oracle_environment import _plane_frame

pytestmark = pytest.mark.gpu
TOP = 0xFFFFFFFF


@pytest.fixture(scope="module")
def ctx():
    c = rt.Context(0)
    yield c
    c.close()


def _close(ga, oa):
    rel = np.abs(ga[:, :3] - oa[:, :3]) / np.maximum(np.abs(oa[:, :3]), 1e-6)
    return rel.max(), np.mean(np.all(ga == oa, axis=1))


@pytest.mark.parametrize("below", [True, False], ids=["ground", "sky"])
def test_plane_environment(ctx, below):
    sun = (0.3, 0.8, 0.5, 3.0) if below else (0.0, 1.0, 0.0, 0.0)
    s, desc, mats, pt = _plane_frame(below, (0.9, 0.5, 0.2), (0.25, 0.6, 0.35), sun, size=64, spp=4)
    g = ctx.upload(desc).build()
    o = ob.Scene(desc)
    for megakernel in (False, True):
        ga, gc = g.path_trace(pt, megakernel=megakernel)
        oa, oc = o.path_trace(pt)
        assert gc.all() and np.array_equal(gc, oc)
        err, exact = _close(ga, oa)
        print(f"plane ({'ground' if below else 'sky'}, megakernel={megakernel}): max rel err {err:.3g}, "
              f"bit-exact pixels {exact:.4f}")
        assert err <= 1e-4
    g.close()


def test_far_from_the_origin(ctx):
    s = synth.Synth.config(2, 0, 512)
    pt, rv = s.frame(128, 96)
    desc, eye, pt2, rv2, keep = far_from_origin(s, pt, rv)
    print(f"[scene] eye cell {list(eye.cell)}, local {list(eye.local)}, residual {list(eye.residual)}", flush=True)
    g = ctx.upload(desc).build(eye=eye)
    o = ob.Scene(desc, eye)
    gn, gp, gm = g.bvh(TOP)
    on, op, om = o.bvh(TOP)
    assert np.array_equal(gm, om) and gp.tobytes() == op.tobytes() and gn.tobytes() == on.tobytes()
    rays = rays_towards(20000, -60, 60, -40, 60, 31)  # from around the eye towards the scene, in the frame's space
    gh, oh = g.trace_closest(rays), o.trace_closest(rays)
    assert np.count_nonzero(oh["instance_id"] != abi.MISS) > 3000
    assert gh.tobytes() == oh.tobytes()
    assert np.array_equal(g.trace_any(rays, mask=0x01), o.trace_any(rays, mask=0x01))
    words = g.ray_visibility(rv2)
    assert np.count_nonzero(words) > 1000
    assert np.array_equal(words, o.ray_visibility(rv2))
    for megakernel in (False, True):
        ga, gc = g.path_trace(pt2, megakernel=megakernel)
        oa, oc = o.path_trace(pt2)
        assert np.array_equal(gc, oc)
        err, exact = _close(ga, oa)
        print(f"far from the origin (megakernel={megakernel}): max rel err {err:.3g}, bit-exact pixels {exact:.4f}")
        assert err <= 1e-4
    for flags in (0, abi.SHADOW_JITTER):
        p = synth.Synth.shadow_params(pt2, rv2, flags, 9)
        want, _ = o.shadow_resolve(p, words)
        assert np.count_nonzero(want != abi.SHADOW_NOT_TAKEN) > 1000
        assert np.array_equal(g.shadow_resolve(p, words), want)
    g.close()
