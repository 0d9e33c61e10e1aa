"""GPU parity of the frame's structures (instances that are not static, traced through the frame's
visible list: ge_rt_scene_set_visible), through the C ABI, against the oracle.

On the shadow-pass scene with its groups not static and a random visible list (some groups culled):
each frame structure (nodes, sorted triangles, codes, quantized wide nodes) and the instance
structure bit-exact; closest hits (geometry = the visible entry), occlusion bits and traversal
counters bit-exact; path-traced radiance within 1e-4, the visibility words and the jittered shadow
answers bit-exact. Then a second frame's list (GE_RT_BUILD_TOP) and moved groups refitted under the
same list (GE_RT_BUILD_REFIT), each against the oracle. A build with such instances and no list,
and a list that no longer fits the instances, are refused.
"""
import numpy as np
import pytest

from game_engine_amd import abi, rt, synth
from oracle import binding as ob
from tests.helpers import even_depth, rays_towards
from tests.test_oracle_visible import STATIC, _visible

pytestmark = pytest.mark.gpu
TOP = 0xFFFFFFFF
FRAME = abi.FRAME_STRUCTURE


@pytest.fixture(scope="module")
def ctx():
    c = rt.Context(0)
    yield c
    c.close()


def _structures_equal(g, o, sel):
    gn, gp, gm = g.bvh(sel)
    on, op, om = o.bvh(sel)
    assert np.array_equal(gm, om) and gp.tobytes() == op.tobytes() and gn.tobytes() == on.tobytes(), sel
    if len(on):
        even = even_depth(on)
        assert g.bvh_wide(sel)[even].tobytes() == o.bvh_wide(sel)[even].tobytes(), sel
    return len(gp)


def _frame_equal(ctx, s, g, o, k):
    n = len(s.arrays()["instances"])
    prims = sum(_structures_equal(g, o, FRAME | i) for i in range(1, n))
    _structures_equal(g, o, TOP)
    lo, hi = s.bounds()
    rays = rays_towards(30000, lo - 5, hi + 5, lo, hi, 10 + k)
    ctx.reset_stats()
    ob.reset_counters()
    gh = g.trace_closest(rays, count=True)
    oh = o.trace_closest(rays)
    assert gh.tobytes() == oh.tobytes()
    groups = np.count_nonzero((oh["instance_id"] != abi.MISS) & (oh["instance_id"] > 0))
    for mask in (0xFE, 0x01):
        assert np.array_equal(g.trace_any(rays, mask=mask, count=True), o.trace_any(rays, mask=mask))
    gs, os_ = ctx.stats(), ob.counters()
    for key in ("closest", "any", "nodes", "triangles", "instances", "any_nodes", "any_triangles", "any_instances"):
        assert gs[key] == os_[key], key
    pt, rv = s.frame(128, 72)
    pt.sample_base = k
    ga, gc = g.path_trace(pt)
    oa, oc = o.path_trace(pt)
    assert np.array_equal(gc, oc)
    rel = np.abs(ga - oa) / np.maximum(np.abs(oa), 1e-6)
    assert rel.max() <= 1e-4
    words = g.ray_visibility(rv)
    assert np.array_equal(words, o.ray_visibility(rv))
    p = synth.Synth.shadow_params(pt, rv, abi.SHADOW_JITTER, k)
    got = g.shadow_resolve(p, words)
    want, _ = o.shadow_resolve(p, words)
    assert np.array_equal(got, want)
    print(f"frame {k}: {prims} triangles in frame structures, {groups} group hits, radiance max rel err "
          f"{rel.max():.3g}", flush=True)


def test_frame_structures_against_the_oracle(ctx):
    s = synth.Synth.config(3, 0, 48, 129)
    inst = s.arrays()["instances"]
    inst["flags"][1:] &= ~np.uint32(STATIC)
    rng = np.random.default_rng(4)
    g = ctx.upload(s.desc())
    with pytest.raises(rt.RtError) as e:
        g.build()  # groups that are not static, and no visible list
    assert e.value.status == abi.ERROR_INVALID_SCENE
    entries = _visible(s, rng)
    g.set_visible(entries)
    g.build()
    o = ob.Scene(s.desc())
    assert o.set_visible(entries) == 0
    _frame_equal(ctx, s, g, o, 0)
    # The next frame: another list, the instance structure rebuilt over the new frame structures.
    entries = _visible(s, rng, keep=0.4)
    g.set_visible(entries)
    g.build(abi.BUILD_TOP)
    assert o.set_visible(entries) == 0
    _frame_equal(ctx, s, g, o, 1)
    # Moved groups under the same list: the instance structure refitted.
    recs = s.move_instances(1, 10, 77, 0.5, 0.2)
    g.update_instances(1, recs)
    g.build(abi.BUILD_REFIT)
    assert o.update_instances(1, recs, refit=True) == 0
    _frame_equal(ctx, s, g, o, 2)
    # Bad entries are refused and keep the list; a list that no longer fits the instances fails the build.
    assert g.set_visible(np.array([[len(inst), 0]], dtype=np.uint32), check=False) == abi.ERROR_INVALID_ARGUMENT
    assert any(int(i) == 1 for i, _ in entries)
    moved = (abi.InstanceDesc * 1)()
    moved[0] = recs[0]
    moved[0].mesh = int(inst[2]["mesh"])  # instance 1 now draws instance 2's mesh: its entries name another mesh's clusters
    g.update_instances(1, moved)
    with pytest.raises(rt.RtError) as e:
        g.build(abi.BUILD_TOP)
    assert e.value.status == abi.ERROR_INVALID_SCENE
