"""Writes the small golden fixtures under tests/golden/ from closed forms computed here in pure
Python (not by the oracle): the PCG hash and the pixel-stream draws of sampling.slang:27-51, and
30-bit Morton codes of quantized points. Run: python tests/golden/make_golden.py
"""
import json
import os
import struct

HERE = os.path.dirname(os.path.abspath(__file__))
M = 0xFFFFFFFF


def pcg_hash(v):
    state = (v * 747796405 + 2891336453) & M
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & M
    return (word >> 22) ^ word


def rng_seed(x, y, sample, seed):
    mixed = pcg_hash(((x * 73856093) & M) ^ ((y * 19349663) & M) ^ ((seed * 83492791) & M))
    return pcg_hash((mixed + sample * 2654435761) & M)


def rng_next(state):
    state = (state * 747796405 + 2891336453) & M
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & M
    word = (word >> 22) ^ word
    # float(word) rounded to f32, times 2^-32 (exact): the draw as its f32 bit pattern.
    f = struct.unpack("<f", struct.pack("<f", float(word)))[0] * 2.0 ** -32
    return state, struct.unpack("<I", struct.pack("<f", f))[0]


def morton(q):
    out = 0
    for b in range(10):
        out |= ((q[0] >> b) & 1) << (3 * b + 2) | ((q[1] >> b) & 1) << (3 * b + 1) | ((q[2] >> b) & 1) << (3 * b)
    return out


def main():
    hashes = {str(v): pcg_hash(v) for v in [0, 1, 2, 3, 255, 256, 65535, 1 << 31, M, 12345678, 2891336453]}
    streams = []
    for (x, y, s, seed) in [(0, 0, 0, 0), (13, 29, 7, 99), (1919, 1079, 1, 0), (5, 5, 1023, 77)]:
        st = rng_seed(x, y, s, seed)
        draws = []
        for _ in range(8):
            st, bits = rng_next(st)
            draws.append(bits)
        streams.append({"pixel": [x, y], "sample": s, "seed": seed, "state0": rng_seed(x, y, s, seed), "draw_f32_bits": draws})
    codes = [{"q": q, "code": morton(q)} for q in [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1023, 1023, 1023],
                                                  [512, 3, 1000], [77, 700, 5]]]
    with open(os.path.join(HERE, "rng_morton.json"), "w") as f:
        json.dump({"pcg_hash": hashes, "streams": streams, "morton": codes}, f, indent=1)


if __name__ == "__main__":
    main()
