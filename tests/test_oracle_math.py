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


def test_sincos_is_correctly_rounded_within_one_ulp():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(0, 2 * math.pi, 20000), np.float32(2 * math.pi) * rng.uniform(0, 1, 2000),
                         np.array([0.0, math.pi / 4, math.pi / 2, math.pi, 3 * math.pi / 2, 6.2831855])]).astype(np.float32)
    s, c = C.c_float(), C.c_float()
    worst = 0.0
    for x in xs:
        ob.lib().ge_oracle_sincos(float(x), C.byref(s), C.byref(c))
        for got, ref in ((s.value, math.sin(float(x))), (c.value, math.cos(float(x)))):
            ulp = np.spacing(np.float32(abs(ref))) if ref != 0 else np.float32(1e-45)
            err = abs(got - ref) / float(ulp)
            worst = max(worst, err)
    assert worst <= 1.0, worst
