"""The oracle's BSDF and light falloffs against the reference's own double-precision twin of
brdf.slang (domain/gfx/tests/brdf_reference.h), compiled from the reference's header into
oracle/_ref/ where the reference tree is present (skipped where it is not).

Both sides are double; they agree to rounding (1e-12 relative), so a dropped term, a wrong sign or
a swapped argument in the oracle fails here.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import binding as ob

REF = ob.ref_lib()
pytestmark = pytest.mark.skipif(REF is None, reason="oracle/_ref not built: no reference tree here")
TOL = 1e-12


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1)[:, None]


def _surfaces(rng, n):
    """Random surfaces with the view above the surface: albedo, normal, view, roughness, metallic,
    and a specular record (zero for half of them: the 4 % stand-in)."""
    normal = _unit(rng, n)
    view = _unit(rng, n)
    view = np.where((np.sum(view * normal, axis=1) < 0.05)[:, None], -view, view)
    keep = np.sum(view * normal, axis=1) > 0.05
    s = np.zeros((n, 15))
    s[:, 0:3] = rng.uniform(0, 1, (n, 3))
    s[:, 3:6] = normal
    s[:, 6:9] = view
    s[:, 9] = rng.uniform(0.0, 1.2, n)  # below the 0.045 clamp and above 1 too
    s[:, 10] = rng.choice([0.0, 1.0, 0.5], n)
    spec = rng.uniform(-0.03, 0.3, (n, 4))
    spec[: n // 2] = 0.0
    s[:, 11:15] = spec
    return s[keep]


def _ref3(fn, inp):
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    out = np.zeros(3)
    fn(inp.ctypes.data, out.ctypes.data)
    return out


def _close(a, b):
    return np.allclose(a, b, rtol=TOL, atol=TOL * max(1.0, float(np.max(np.abs(b)))))


def test_ggx_terms_and_alpha():
    rng = np.random.default_rng(1)
    for _ in range(400):
        nh, nv, nl = rng.uniform(0, 1, 3)
        r = rng.uniform(0, 1.2)
        alpha = REF.ge_ref_alpha(r)
        assert _close(ob.probe(9, [nh, alpha], ob.FP64)[0], REF.ge_ref_d_ggx(nh, alpha))
        assert _close(ob.probe(8, [nv, nl, alpha], ob.FP64)[0], REF.ge_ref_v_smith(nv, nl, alpha))


def test_fresnel_eval_and_direct():
    rng = np.random.default_rng(2)
    surfaces = _surfaces(rng, 600)
    lit = 0
    for s in surfaces:
        cos = rng.uniform(0, 1)
        assert _close(ob.probe(10, list(s) + [cos], ob.FP64)[:3], _ref3(REF.ge_ref_fresnel, list(s) + [cos]))
        light = _unit(rng, 1)[0]
        radiance = rng.uniform(0, 5, 3)
        if np.dot(light, s[3:6]) <= 0:
            continue
        lit += 1
        inp = list(s) + list(light)
        assert _close(ob.probe(3, inp, ob.FP64)[:3], _ref3(REF.ge_ref_eval, inp))
        inp = list(s) + list(light) + list(radiance)
        assert _close(ob.probe(16, inp, ob.FP64)[:3], _ref3(REF.ge_ref_direct, inp))
    assert lit > 200


def test_light_falloffs():
    rng = np.random.default_rng(3)
    for _ in range(300):
        # The light record is float32 (ResolveLight); round its fields as the oracle reads them.
        f = lambda v: np.float32(v).astype(np.float64)  # noqa: E731
        pos = f(rng.uniform(-10, 10, 3))
        radius = float(f(rng.uniform(1, 20)))
        color = f(rng.uniform(0, 1, 3))
        intensity = float(f(rng.uniform(0, 50)))
        point = rng.uniform(-10, 10, 3)
        spot = rng.random() < 0.5
        direction = f(_unit(rng, 1)[0])
        cos_outer = float(f(rng.uniform(0, 0.9)))
        cos_inner = float(f(rng.uniform(cos_outer, 1.0)))
        light = [*pos, radius, *color, intensity, *direction, 1.0 if spot else 0.0, cos_inner, cos_outer, 0, 0]
        out = ob.probe(12, light + list(point), ob.FP64)
        to_light = pos - point
        d2 = float(np.dot(to_light, to_light))
        att = REF.ge_ref_distance_attenuation(d2, radius)
        if att <= 0:
            assert out[0] == 0.0
            continue
        d = to_light / np.sqrt(d2)
        if spot:
            neg = np.ascontiguousarray(-d)
            dn = np.ascontiguousarray(direction / np.linalg.norm(direction))
            att *= REF.ge_ref_spot_attenuation(neg.ctypes.data, dn.ctypes.data, cos_inner, cos_outer)
        assert out[0] == 1.0
        assert _close(out[4:7], color * (intensity * att))
