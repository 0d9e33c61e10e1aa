"""The oracle's sampler and BSDF held to the reference's own sampling tests
(domain/gfx/tests/sampling_tests.cpp:140-369), re-derived: densities integrate to one, the VNDF
sampler draws from the density it reports, f cos / pdf = F G2 / G1 for a metal at every roughness,
the white furnace numbers, the generator's determinism and uniformity, and MIS sums to one.

These run the oracle in FP64 mode (its double-precision twin, like brdf_reference.h) and, where the
identity is exact, also in FP32 mode to a float tolerance.
"""
import math

import numpy as np
import pytest

from oracle import binding as ob

FP64, FP32 = ob.FP64, ob.FP32


def alpha_of(r):
    r = min(max(r, 0.045), 1.0)
    return r * r


def g1(ndv, a):
    return ob.probe(7, [ndv, a])[0]


def d_ggx(ndh, a):
    return ob.probe(9, [ndh, a])[0]


def vndf_density(v, h, a):
    """D_v(h) = G1(v) max(0, v.h) D(h) / v.z (Heitz 2018)."""
    vh = float(np.dot(v, h))
    if h[2] <= 0:
        return 0.0
    return g1(v[2], a) * max(0.0, vh) * d_ggx(h[2], a) / v[2]


def hemisphere_grid(nt, np_):
    """Midpoint quadrature over the hemisphere in (theta, phi): directions and solid-angle weights."""
    th = (np.arange(nt) + 0.5) * (math.pi / 2) / nt
    ph = (np.arange(np_) + 0.5) * (2 * math.pi) / np_
    T, P = np.meshgrid(th, ph, indexing="ij")
    dirs = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], axis=-1).reshape(-1, 3)
    w = (np.sin(T) * (math.pi / 2 / nt) * (2 * math.pi / np_)).reshape(-1)
    return dirs, w


def test_cosine_density_integrates_to_one():
    dirs, w = hemisphere_grid(256, 256)
    total = float(np.sum(np.maximum(dirs[:, 2], 0) / math.pi * w))
    assert total == pytest.approx(1.0, rel=1e-4)
    # The sampler's directions are unit and in the upper hemisphere.
    rng = np.random.default_rng(0)
    for u in rng.uniform(0, 1, (200, 2)):
        d = ob.probe(0, list(u))[:3]
        assert abs(np.linalg.norm(d) - 1) < 1e-12 and d[2] >= 0


@pytest.mark.parametrize("roughness", [0.25, 0.5, 1.0])
@pytest.mark.parametrize("ndv", [0.98, 0.6, 0.2])
def test_vndf_density_integrates_to_one(roughness, ndv):
    a = alpha_of(roughness)
    v = np.array([math.sqrt(1 - ndv * ndv), 0.0, ndv])
    dirs, w = hemisphere_grid(128, 128)
    vals = np.array([vndf_density(v, h, a) for h in dirs])
    assert float(np.sum(vals * w)) == pytest.approx(1.0, rel=5e-3)


def test_direction_pdf_is_the_half_vector_density_over_the_jacobian():
    v = np.array([0.3, 0.0, math.sqrt(1 - 0.09)])
    a = alpha_of(0.4)
    state = np.array([ob.lib().ge_oracle_rng_seed(3, 5, 1, 77)], dtype=np.uint32)
    import ctypes
    st = ctypes.c_uint32(int(state[0]))
    checked = 0
    for _ in range(64):
        u0 = ob.lib().ge_oracle_rng_next(ctypes.byref(st))
        u1 = ob.lib().ge_oracle_rng_next(ctypes.byref(st))
        h = ob.probe(1, [*v, a, u0, u1])[:3]
        l = h * (2 * np.dot(v, h)) - v
        if l[2] <= 0:
            continue
        pdf = ob.probe(2, [*v, *l, a])[0]
        assert pdf == pytest.approx(vndf_density(v, h, a) / (4 * np.dot(v, h)), rel=1e-9)
        checked += 1
    assert checked > 30


def test_vndf_histogram_matches_density():
    """Twelve elevation bands of sampled half vectors against the quadrature of the density."""
    a = alpha_of(0.5)
    ndv = 0.6
    v = np.array([math.sqrt(1 - ndv * ndv), 0.0, ndv])
    rng = np.random.default_rng(11)
    n = 20000
    hs = np.array([ob.probe(1, [*v, a, *u])[:3] for u in rng.uniform(0, 1, (n, 2))])
    bands = 12
    th = np.arccos(np.clip(hs[:, 2], -1, 1))
    hist = np.histogram(th, bins=bands, range=(0, math.pi / 2))[0] / n
    dirs, w = hemisphere_grid(96, 96)
    dens = np.array([vndf_density(v, h, a) for h in dirs]) * w
    thq = np.arccos(np.clip(dirs[:, 2], -1, 1))
    expect = np.histogram(thq, bins=bands, range=(0, math.pi / 2), weights=dens)[0]
    assert np.max(np.abs(hist - expect)) < 0.01


@pytest.mark.parametrize("mode", [FP64, FP32])
def test_specular_weight_is_f_g2_over_g1(mode):
    """For a metal the BSDF is F D V and the sample weight reduces to F G2 / G1 exactly."""
    rng = np.random.default_rng(21)
    tol = 1e-6 if mode == FP64 else 2e-4
    checked = 0
    for roughness in (0.05, 0.15, 0.4, 0.7, 1.0):
        a = alpha_of(roughness)
        for ndv in (0.99, 0.7, 0.3):
            v = np.array([math.sqrt(1 - ndv * ndv), 0.0, ndv])
            for u in rng.uniform(0, 1, (32, 2)):
                h = ob.probe(1, [*v, a, *u], mode)[:3]
                l = h * (2 * np.dot(v, h)) - v
                if l[2] <= 1e-6:
                    continue
                pdf = ob.probe(2, [*v, *l, a], mode)[0]
                if not pdf > 0:
                    continue
                surface = [1, 1, 1, 0, 0, 1, *v, roughness, 1.0, 0, 0, 0, 0]
                f = ob.probe(3, surface + list(l), mode)[0]
                weight = f * l[2] / pdf
                vs = ob.probe(8, [ndv, l[2], a])[0]
                g2 = 4 * ndv * l[2] * vs
                F = ob.probe(10, surface + [float(np.dot(v, h))])[0]
                expected = F * g2 / g1(ndv, a)
                assert weight == pytest.approx(expected, rel=tol)
                assert weight <= 1 + 1e-5
                checked += 1
    assert checked > 300


def directional_albedo(roughness, metallic, ndv, nt=192, np_=192):
    v = [math.sqrt(1 - ndv * ndv), 0.0, ndv]
    dirs, w = hemisphere_grid(nt, np_)
    surface = [1, 1, 1, 0, 0, 1, *v, roughness, metallic, 0, 0, 0, 0]
    total = 0.0
    for l, wi in zip(dirs, w):
        total += ob.probe(3, surface + list(l))[0] * l[2] * wi
    return total


def test_white_furnace():
    lam = directional_albedo(1.0, 0.0, 0.95)
    assert 0.94 < lam < 1.0  # the reference's 0.9726 (sampling_tests.cpp:273-319)
    assert lam == pytest.approx(0.9726, abs=2e-3)
    prev = 2.0
    for r in (0.3, 0.5, 0.75, 1.0):
        alb = directional_albedo(r, 1.0, 0.95)
        assert 0.25 < alb < 1.001
        assert alb < prev + 1e-6
        prev = alb
    assert 0.28 < prev < 0.36


def test_mis_sums_to_one():
    assert ob.probe(11, [1, 1])[0] == pytest.approx(0.5)
    assert ob.probe(11, [2, 0])[0] == pytest.approx(1.0)
    assert ob.probe(11, [0, 2])[0] == pytest.approx(0.0)
    assert ob.probe(11, [0, 0])[0] == pytest.approx(0.0)
    for a in (0.1, 1.0, 7.5):
        for b in (0.2, 3.0, 12.0):
            assert ob.probe(11, [a, b])[0] + ob.probe(11, [b, a])[0] == pytest.approx(1.0)


def test_onb_is_orthonormal():
    rng = np.random.default_rng(5)
    for n in rng.normal(size=(200, 3)):
        n = n / np.linalg.norm(n)
        out = ob.probe(6, list(n))
        t, b = out[:3], out[3:6]
        for x in (t, b):
            assert abs(np.linalg.norm(x) - 1) < 1e-12
        assert abs(np.dot(t, b)) < 1e-12 and abs(np.dot(t, n)) < 1e-12 and abs(np.dot(b, n)) < 1e-12


def test_light_falloff_window_and_spot():
    # A point light: windowed inverse square, zero at the radius with zero slope.
    light = [0, 10, 0, 20, 1, 1, 1, 100, 0, 0, 0, 0, 0, 0, 0, 0]
    out = ob.probe(12, light + [0, 0, 0])
    assert out[0] == 1.0 and np.allclose(out[1:4], [0, 1, 0])
    d2 = 100.0
    window = (1 - (d2 / 400) ** 2) ** 2
    assert out[4] == pytest.approx(100 * window / d2, rel=1e-12)
    assert out[7] == pytest.approx(10.0)
    assert ob.probe(12, light + [0, -15, 0])[0] == 0.0  # past the radius: no light
    # A directional light arrives everywhere with its colour and no falloff.
    sun = [1, 2, 2, 0, 1, 0.5, 0.25, 3, 0, 0, 0, 2, 0, 0, 0, 0]
    out = ob.probe(12, sun + [5, 5, 5])
    assert np.allclose(out[1:4], [1 / 3, 2 / 3, 2 / 3]) and np.allclose(out[4:7], [3, 1.5, 0.75])
    assert out[7] == pytest.approx(1e30)


def test_decode_normal_oct_inverts_the_packing():
    rng = np.random.default_rng(9)
    for n in rng.normal(size=(300, 3)):
        n = n / np.linalg.norm(n)
        l1 = np.abs(n).sum()
        p = n[:2] / l1
        if n[2] < 0:
            p = np.array([(1 - abs(p[1])) * (1 if p[0] >= 0 else -1), (1 - abs(p[0])) * (1 if p[1] >= 0 else -1)])
        x, y = (int(round(float(np.clip(c, -1, 1)) * 32767)) for c in p)
        packed = (x & 0xFFFF) | ((y & 0xFFFF) << 16)
        d = ob.probe(13, [packed])[:3]
        assert np.dot(d, n) > 1 - 1e-8


def test_transform_normal_is_the_cofactor_normal():
    rng = np.random.default_rng(3)
    for _ in range(50):
        m = rng.normal(size=(3, 3))
        n = rng.normal(size=3)
        out = ob.probe(15, list(m.reshape(-1)) + list(n))[:3]
        expect = np.linalg.inv(m).T @ n
        expect = expect / np.linalg.norm(expect)
        assert np.allclose(out, expect, atol=1e-9)  # inverse transpose, outward under a mirror too
