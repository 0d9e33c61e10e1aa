"""The oracle's closed-form pieces: the PCG stream and Morton codes against the committed golden
vectors (tests/golden/rng_morton.json, written from pure-Python closed forms by make_golden.py),
the stream's determinism and uniformity (domain/gfx/tests/sampling_tests.cpp:321-355), and the
own sin/cos against libm's double precision."""
import ctypes as C
import json
import math
import os
import struct

import numpy as np

from oracle import binding as ob

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "rng_morton.json")))


def f32_bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def test_pcg_hash_golden():
    for v, h in GOLDEN["pcg_hash"].items():
        assert ob.lib().ge_oracle_pcg_hash(int(v)) == h


def test_pixel_streams_golden():
    for s in GOLDEN["streams"]:
        st = ob.lib().ge_oracle_rng_seed(s["pixel"][0], s["pixel"][1], s["sample"], s["seed"])
        assert st == s["state0"]
        state = C.c_uint32(st)
        for bits in s["draw_f32_bits"]:
            assert f32_bits(ob.lib().ge_oracle_rng_next(C.byref(state))) == bits


def test_morton_golden():
    lo = (C.c_float * 3)(0.0, 0.0, 0.0)
    scale = (C.c_float * 3)(1.0, 1.0, 1.0)
    for e in GOLDEN["morton"]:
        q = e["q"]
        assert ob.lib().ge_oracle_morton3(q[0] + 0.25, q[1] + 0.25, q[2] + 0.25, lo, scale) == e["code"]


def test_stream_is_deterministic_and_uniform():
    seed = ob.lib().ge_oracle_rng_seed
    a, b = C.c_uint32(seed(13, 29, 7, 99)), C.c_uint32(seed(13, 29, 7, 99))
    for _ in range(8):
        assert ob.lib().ge_oracle_rng_next(C.byref(a)) == ob.lib().ge_oracle_rng_next(C.byref(b))
    assert seed(13, 29, 7, 99) != seed(13, 29, 8, 99)
    assert seed(13, 29, 7, 99) != seed(14, 29, 7, 99)
    assert seed(13, 29, 7, 99) != seed(13, 29, 7, 100)
    bins = np.zeros(8)
    total = 0.0
    for pixel in range(100):
        st = C.c_uint32(seed(pixel % 10, pixel // 10, 0, 5))
        for _ in range(1000):
            v = ob.lib().ge_oracle_rng_next(C.byref(st))
            assert 0.0 <= v < 1.0
            bins[int(v * 8)] += 1
            total += v
    assert abs(total / 100000 - 0.5) < 0.005
    assert np.all(bins > 100000 / 10) and np.all(bins < 100000 / 6)


def _ulps(got, ref):
    ulp = float(np.spacing(np.float32(abs(ref)))) if ref != 0 else 1e-45
    return abs(got - ref) / ulp


def test_sincos_is_within_one_ulp_of_libm():
    """The float sincos (Cody-Waite reduction, minimax polynomials) against libm's double sin and
    cos of the same float, the bound glibc states for its own sinf/cosf (1 ulp): a sample of
    [-64, 64], the sampler's [0, 2 pi), the floats nearest the multiples of pi/2 (the reduction's
    hardest cases) and the polynomials' ends. Every float of [-64, 64] (2.2e9 values) was checked
    when the coefficients were derived: worst 0.760 ulp (sin) and 0.749 ulp (cos), 99.8% of them
    correctly rounded; a dropped term, a wrong constant or a lost low part breaks the 0.8 bar below."""
    rng = np.random.default_rng(0)
    near = [np.float32(k * math.pi / 2) for k in range(-40, 41)]
    near += [np.nextafter(v, np.float32(np.inf)) for v in near] + [np.nextafter(v, np.float32(-np.inf)) for v in near]
    xs = np.concatenate([rng.uniform(-64, 64, 20000), np.float32(2 * math.pi) * rng.uniform(0, 1, 20000),
                         np.array(near), np.array([0.0, math.pi / 4, -math.pi / 4, 6.2831855, 1e-30, -1e-30])])
    xs = xs.astype(np.float32)
    s, c = C.c_float(), C.c_float()
    worst, rounded = 0.0, 0
    for x in xs:
        ob.lib().ge_oracle_sincos(float(x), C.byref(s), C.byref(c))
        es, ec = _ulps(s.value, math.sin(float(x))), _ulps(c.value, math.cos(float(x)))
        worst = max(worst, es, ec)
        rounded += (es <= 0.5) + (ec <= 0.5)
    assert worst <= 0.8, worst
    assert rounded >= 0.97 * 2 * len(xs)  # 98.5% on this sample


def test_det2_is_exact_in_sign_and_within_two_ulp():
    """The watertight test's fallback a*b - c*d (Kahan's algorithm) against the exact value in
    rationals: zero exactly when the value is, the sign always right, within 2 ulp (the bound of
    Jeannerod, Louvet and Muller 2013), on random and on cancelling inputs, where the plain float
    expression gets the sign wrong."""
    from fractions import Fraction
    rng = np.random.default_rng(3)
    f32 = lambda v: float(np.float32(v))  # noqa: E731
    cases = []
    for _ in range(3000):
        a, b, c = (f32(v) for v in rng.uniform(-10, 10, 3))
        d = f32(a * b / c) if c != 0 else 1.0  # a*b close to c*d: heavy cancellation
        cases.append((a, b, c, d))
        cases.append(tuple(f32(v) for v in rng.uniform(-10, 10, 4)))
    cases += [(1.0, 1.0, 1.0, 1.0), (3.0, 5.0, 5.0, 3.0), (0.1, 0.3, 0.3, 0.1)]
    wrong_plain = 0
    for a, b, c, d in cases:
        got = ob.lib().ge_oracle_det2(a, b, c, d)
        exact = Fraction(a) * Fraction(b) - Fraction(c) * Fraction(d)
        assert (got == 0) == (exact == 0), (a, b, c, d)
        assert (got > 0) == (exact > 0), (a, b, c, d)
        if exact != 0:
            assert _ulps(got, float(exact)) <= 2.0, (a, b, c, d)
        plain = np.float32(np.float32(a) * np.float32(b)) - np.float32(np.float32(c) * np.float32(d))
        wrong_plain += (plain > 0) != (exact > 0) or (plain == 0) != (exact == 0)
    assert wrong_plain > 100  # the inputs do reach the cases the fallback is for

