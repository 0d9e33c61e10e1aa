"""GPU parity: the HIP backend, called through the C ABI, against the CPU oracle on the same inputs.

Bars: bit-exact for the BVH (Morton codes, sorted triangles, node topology and boxes), hit records,
occlusion bits, visibility words, coverage and traversal counters; radiance within 1e-4 relative
per pixel (the FP32 oracle runs the same operation order, so in practice the sums are equal to the
bit, which the tests also report); display bytes within one code (pow is each side's own library's).
"""
import ctypes as C

import numpy as np
import pytest

from game_engine_amd import abi, rt, synth
from oracle import binding as ob
from tests.helpers import even_depth, random_rays, rays_towards, unpack_bits

pytestmark = pytest.mark.gpu
TOP = 0xFFFFFFFF
RAD_TOL = 1e-4


@pytest.fixture(scope="module")
def ctx():
    c = rt.Context(0)
    yield c
    c.close()


def _pair(ctx, s):
    print(f"[scene] {s.triangle_count} triangles", flush=True)
    return ctx.upload(s.desc()).build(), ob.Scene(s.desc())


@pytest.fixture(scope="module")
def c1(ctx):
    s = synth.Synth.config(1, 0)
    g, o = _pair(ctx, s)
    yield s, g, o
    g.close()


@pytest.fixture(scope="module")
def c2(ctx):
    s = synth.Synth.config(2, 0, 512)
    g, o = _pair(ctx, s)
    yield s, g, o
    g.close()


def _bvh_equal(g, o, mesh):
    gn, gp, gm = g.bvh(mesh)
    on, op, om = o.bvh(mesh)
    assert np.array_equal(gm, om), "Morton codes"
    assert gp.tobytes() == op.tobytes(), "sorted primitives"
    assert gn.tobytes() == on.tobytes(), "nodes"
    even = even_depth(on)
    assert g.bvh_wide(mesh)[even].tobytes() == o.bvh_wide(mesh)[even].tobytes(), "quantized wide nodes"
    return len(gn)


def test_bvh_c1_bit_exact(c1):
    s, g, o = c1
    assert _bvh_equal(g, o, 0) == 65535
    assert _bvh_equal(g, o, TOP) == 1


def test_bvh_c2_bit_exact(c2):
    s, g, o = c2
    for m in range(16):
        assert _bvh_equal(g, o, m) == 1023
    assert _bvh_equal(g, o, TOP) == 511


def test_closest_and_any_c1(c1):
    s, g, o = c1
    rays = np.concatenate([random_rays(20000, -20, 120, 1), rays_towards(20000, -30, 130, 0, 100, 2)])
    gh = g.trace_closest(rays)
    oh = o.trace_closest(rays)
    assert np.count_nonzero(oh["instance_id"] != abi.MISS) > 10000
    assert gh.tobytes() == oh.tobytes()
    ga = g.trace_any(rays)
    assert np.array_equal(ga, o.trace_any(rays))
    # Shorter rays: the any-hit interval matters.
    rays["tmax"] = np.float32(25.0)
    assert np.array_equal(g.trace_any(rays), o.trace_any(rays))


def test_closest_and_any_c2(c2):
    s, g, o = c2
    rays = rays_towards(30000, -20, 180, 20, 110, 3)
    gh = g.trace_closest(rays)
    oh = o.trace_closest(rays)
    assert np.count_nonzero(oh["instance_id"] != abi.MISS) > 10000
    assert gh.tobytes() == oh.tobytes()
    for mask in (0xFE, 0x01, 0x00):
        assert np.array_equal(g.trace_any(rays, mask=mask), o.trace_any(rays, mask=mask))


def test_counters_match_oracle_walk(ctx, c2):
    s, g, o = c2
    rays = rays_towards(5000, -20, 180, 20, 110, 4)
    ctx.reset_stats()
    ob.reset_counters()
    g.trace_closest(rays, count=True)
    o.trace_closest(rays)
    gs, os_ = ctx.stats(), ob.counters()
    for k in ("closest", "nodes", "triangles", "instances"):
        assert gs[k] == os_[k], k
    assert gs["any_nodes"] == os_["any_nodes"] == 0
    ctx.reset_stats()
    ob.reset_counters()
    g.trace_any(rays, count=True)
    o.trace_any(rays)
    gs, os_ = ctx.stats(), ob.counters()
    for k in ("any", "nodes", "any_nodes", "any_triangles", "any_instances"):
        assert gs[k] == os_[k], k
    assert gs["any_nodes"] == gs["nodes"] > 0


def test_ray_visibility_words(c1, c2):
    for s, g, o in (c1, c2):
        pt, rv = s.frame(256, 144)
        gw = g.ray_visibility(rv)
        ow = o.ray_visibility(rv)
        assert np.count_nonzero(ow) > 1000
        assert np.array_equal(gw, ow)


def _radiance_close(a, b):
    rel = np.abs(a[:, :3] - b[:, :3]) / np.maximum(np.abs(b[:, :3]), 1e-6)
    exact = np.mean(np.all(a == b, axis=1))
    return rel.max(), exact


@pytest.mark.parametrize("megakernel", [False, True], ids=["wavefront", "megakernel"])
def test_path_trace_c1(c1, megakernel):
    s, g, o = c1
    pt, _ = s.frame(128, 128)
    ga, gc = g.path_trace(pt, megakernel=megakernel)
    oa, oc = o.path_trace(pt)
    assert np.array_equal(gc, oc)
    err, exact = _radiance_close(ga, oa)
    print(f"C1 radiance: max rel err {err:.3g}, bit-exact pixels {exact:.4f}")
    assert err <= RAD_TOL


@pytest.mark.parametrize("megakernel", [False, True], ids=["wavefront", "megakernel"])
def test_path_trace_c2_lights(ctx, c2, megakernel):
    s, g, o = c2
    pt, _ = s.frame(96, 64)
    assert pt.light_count == 64
    ctx.reset_stats()
    ob.reset_counters()
    ga, gc = g.path_trace(pt, megakernel=megakernel, count=True)
    oa, oc = o.path_trace(pt)
    assert np.array_equal(gc, oc)
    err, exact = _radiance_close(ga, oa)
    print(f"C2 radiance: max rel err {err:.3g}, bit-exact pixels {exact:.4f}")
    assert err <= RAD_TOL
    gs, os_ = ctx.stats(), ob.counters()
    for k in ("closest", "any", "nodes", "triangles", "instances", "any_nodes", "any_triangles", "any_instances"):
        assert gs[k] == os_[k], k  # the same rays, walked the same way
    if not megakernel:
        # One walk launch a bounce (its closest-hit rays with the previous bounce's shadow rays) and
        # one for the last bounce's shadow rays, each timed; no separate any-hit launch.
        assert 1 <= gs["closest_launches"] <= pt.max_bounces + 2
        assert gs["any_launches"] == 0 and gs["any_ms"] == 0
        assert 0 < gs["closest_ms"] <= gs["kernel_ms"]


def test_path_trace_many_lights_regrows_shadow_queue(ctx):
    """Nine shadow rays a vertex overflow the wavefront's first shadow buffer (two a path): the
    shade pass is run again into a grown buffer and the picture is unchanged."""
    s = synth.Synth.config(1, 0)
    lights = (abi.ResolveLight * 8)()
    rng = np.random.default_rng(2)
    for l in lights:
        d = rng.normal(size=3)
        d[1] = abs(d[1]) + 0.5
        l.position_radius[:] = [*d, 0.0]
        l.color_intensity[:] = [1.0, 0.9, 0.8, 0.5]
        l.direction_type[:] = [0.0, 0.0, 0.0, 2.0]
    desc = abi.SceneDesc.from_buffer_copy(s.desc())
    desc.lights = C.addressof(lights)
    desc.light_count = 8
    g = ctx.upload(desc).build()
    o = ob.Scene(desc)
    pt, _ = s.frame(96, 96)
    pt.light_count = 8
    ga, gc = g.path_trace(pt)
    oa, oc = o.path_trace(pt)
    assert np.array_equal(gc, oc)
    assert np.array_equal(ga, oa)


def test_path_trace_progressive_and_deterministic(ctx, c1):
    s, g, o = c1
    pt, _ = s.frame(64, 64)
    pt.spp = 4
    one, cov1 = g.path_trace(pt)
    mk, _ = g.path_trace(pt, megakernel=True)
    assert one.tobytes() == mk.tobytes()  # the two integrator forms agree to the bit
    again, _ = g.path_trace(pt)
    assert one.tobytes() == again.tobytes()  # same seed, same bytes (reference_tests.cpp:424-497)
    acc = np.zeros_like(one)
    cov = np.zeros_like(cov1)
    for base in range(4):
        pt.spp, pt.sample_base = 1, base
        g.path_trace(pt, acc, cov)
    assert np.array_equal(cov, cov1)
    assert np.allclose(acc, one, rtol=1e-5, atol=1e-6)
    pt.seed = 7
    pt.spp, pt.sample_base = 4, 0
    other, _ = g.path_trace(pt)
    assert not np.array_equal(other, one)


def test_tonemap_bytes(ctx, c1):
    s, g, o = c1
    pt, _ = s.frame(64, 64)
    for flags in (0, abi.PT_DITHER, abi.PT_UNIFORM_SKY):
        pt.flags = flags
        acc, cov = o.path_trace(pt)
        gb = rt.tonemap(ctx, pt, acc, cov)
        ob_ = ob.tonemap(pt, acc, cov)
        diff = np.abs(np.stack([(gb >> k) & 255 for k in (0, 8, 16, 24)]).astype(int) -
                      np.stack([(ob_ >> k) & 255 for k in (0, 8, 16, 24)]).astype(int))
        assert diff.max() <= 1


def test_shadow_pass(c2):
    s, g, o = c2
    rays = rays_towards(20000, -20, 180, 20, 110, 5)
    hits = o.trace_closest(rays)
    m = hits["instance_id"] != abi.MISS
    rec = np.zeros(int(m.sum()), dtype=abi.RECEIVER_DTYPE)
    rec["position"] = rays["origin"][m] + rays["dir"][m] * hits["t"][m][:, None]
    rec["instance_id"] = hits["instance_id"][m]
    rec["geometric"] = -rays["dir"][m]
    pt, _ = s.frame(16, 16)
    p = abi.ShadowParams()
    p.sun[:] = list(pt.sun)
    p.offset, p.offset_steps, p.t_max, p.count = pt.ray_bias, 1.0, 1e30, len(rec)
    gb = g.shadow_pass(p, rec)
    obits = o.shadow_pass(p, rec)
    assert np.array_equal(gb, obits)
    frac = unpack_bits(obits, len(rec)).mean()
    assert 0.05 < frac < 0.95


def test_errors_are_reported(ctx, c1):
    s, g, o = c1
    pt, _ = s.frame(8, 8)
    pt.sky_params = 1  # the physical sky is not restated
    with pytest.raises(rt.RtError) as e:
        g.path_trace(pt)
    assert e.value.status == 5
    pt.sky_params = 0
    pt.light_count = 99
    with pytest.raises(rt.RtError):
        g.path_trace(pt)
    unbuilt = ctx.upload(s.desc())
    with pytest.raises(rt.RtError) as e:
        unbuilt.trace_closest(random_rays(4, 0, 1, 0))
    assert e.value.status == 2


@pytest.fixture(scope="module")
def c3(ctx):
    s = synth.Synth.config(3, 0, 48, 129)
    g, o = _pair(ctx, s)
    yield s, g, o
    g.close()


@pytest.mark.parametrize("flags", [0, abi.SHADOW_HALF_RATE, abi.SHADOW_JITTER],
                         ids=["full", "half", "jitter"])
def test_shadow_resolve(ctx, c1, c2, c3, flags):
    """The visibility-buffer sun shadow pass: every entry's pair and blocked bit, and the counters,
    equal to the oracle's."""
    for name, (s, g, o) in (("c1", c1), ("c2", c2), ("c3", c3)):
        pt, rv = s.frame(200, 120)
        words = o.ray_visibility(rv)
        p = synth.Synth.shadow_params(pt, rv, flags, 11)
        ctx.reset_stats()
        ob.reset_counters()
        got = g.shadow_resolve(p, words, count=True)
        want, _ = o.shadow_resolve(p, words)
        taken = want != abi.SHADOW_NOT_TAKEN
        print(f"{name}: {taken.sum()} rays, {(want[taken] & 1).sum()} blocked", flush=True)
        assert taken.sum() > 0
        assert np.array_equal(got, want), name
        gs, os_ = ctx.stats(), ob.counters()
        for k in ("any", "nodes", "triangles", "instances"):
            assert gs[k] == os_[k], (name, k)


def test_dynamic_frames_rebuild_moved_meshes(ctx):
    """Frames of the shadow-pass scene with a tenth of its instances moving: the moved instances'
    meshes rebuilt (GE_RT_BUILD_SUBSET) and the instance structure after them, then the visibility
    words and the jittered shadow pass, each equal to the oracle's after the same moves."""
    s = synth.Synth.config(3, 0, 400, 257)
    g, o = _pair(ctx, s)
    n = len(s.arrays()["instances"])
    moving = (n - 1) // 10
    for k in range(3):
        first = 1 + k * moving
        recs = s.move_instances(first, moving, 100 + k, 0.5, 0.2)
        g.update_instances(first, recs)
        o.update_instances(first, recs)
        meshes = [recs[i].mesh for i in range(moving)]
        g.build(abi.BUILD_SUBSET, meshes=meshes)
        _bvh_equal(g, o, TOP)
        for m in meshes[:4]:
            _bvh_equal(g, o, m)
        pt, rv = s.frame(256, 144)
        words = g.ray_visibility(rv)
        assert np.array_equal(words, o.ray_visibility(rv))
        p = synth.Synth.shadow_params(pt, rv, abi.SHADOW_JITTER, k)
        got = g.shadow_resolve(p, words)
        want, _ = o.shadow_resolve(p, words)
        print(f"frame {k}: {np.count_nonzero(want != abi.SHADOW_NOT_TAKEN)} shadow rays", flush=True)
        assert np.array_equal(got, want)
    g.close()
