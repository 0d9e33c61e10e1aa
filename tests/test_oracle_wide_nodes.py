"""The oracle's 64-byte wide nodes pinned by an independent decode (no GPU).

Each even-depth binary node becomes one 4-wide node whose slots are its children's children (a leaf
child taken as it is), and whose child boxes sit on a per-axis 8-bit grid. Decoded in float64 here
(every plane origin + q 2^e is exact in double), every occupied slot must:

- reference the binary structure's child (leaf slot or the grandchild's binary index), in order,
  with its occupied bit and, for a leaf, its leaf bit;
- hold its binary box with at least 2^-10 and less than 1 + 2^-10 grid steps to spare on every side
  (rounded outward, one step further only where the rounding left less than 2^-10 of a step);
- sit on the smallest grid exponent (>= -100) at which the node's planes, in steps, stay below 2^23
  and span at most 253 steps; empty slots carry no occupied bit, no reference and zero planes.

A quantizer with a floor and a ceiling exchanged, the margin step dropped, an origin off by one step
or a coarser grid than needed fails one of these. The walks over these nodes are held to brute force
on scenes chosen to stress the grid: triangles far from the origin, thin slivers beside a large
triangle, flat (zero-extent) boxes and rays aimed at shared edges.
"""
import numpy as np
import pytest

from game_engine_amd import abi, synth
from oracle import binding as ob
from tests.helpers import even_depth, random_rays, rays_towards

MARGIN = 2.0 ** -10


def expected_slots(nodes, i):
    """(references, boxes) of binary node i's wide node, from the binary structure alone."""
    refs, boxes = [], []
    nd = nodes[i]
    for c in range(2):
        ch = int(nd["child"][c])
        if ch == 0xFFFFFFFF:
            continue
        if ch & 0x80000000:
            refs.append(ch)
            boxes.append((nd[f"lo{c}"].astype(np.float64), nd[f"hi{c}"].astype(np.float64)))
            continue
        m = nodes[ch]
        for d in range(2):
            g = int(m["child"][d])
            if g == 0xFFFFFFFF:
                continue
            refs.append(g)
            boxes.append((m[f"lo{d}"].astype(np.float64), m[f"hi{d}"].astype(np.float64)))
    return refs, boxes


def grid_ok(L, H, e):
    """The grid condition at exponent e, exactly (float64 holds every value involved)."""
    s = 2.0 ** e
    k0, k1 = np.floor(L / s), np.ceil(H / s)
    return k1 - k0 <= 253 and abs(k0) < 2 ** 23 and abs(k1) < 2 ** 23


def check_wide(nodes, wide):
    """Every even-depth wide node against the binary structure."""
    assert len(wide) == len(nodes)
    for i in even_depth(nodes):
        w = wide[i]
        refs, boxes = expected_slots(nodes, i)
        n = len(refs)
        assert 1 <= n <= 4
        assert list(w["child"][:n]) == refs and all(c == 0xFFFFFFFF for c in w["child"][n:])
        meta = int(w["meta"])
        assert (meta >> 24) & 15 == (1 << n) - 1
        assert meta >> 28 == sum(1 << c for c in range(n) if refs[c] & 0x80000000)
        for a in range(3):
            e = ((meta >> (8 * a)) & 0xFF) - 127
            s = 2.0 ** e
            L = min(b[0][a] for b in boxes)
            H = max(b[1][a] for b in boxes)
            assert e >= -100 and grid_ok(L, H, e)
            assert e == -100 or not grid_ok(L, H, e - 1), "a finer grid was possible"
            origin = float(w["origin"][a])
            assert origin / s == np.floor(L / s) - 1  # one step below the node's lowest plane
            for c in range(4):
                qlo, qhi = (int(w["lo"][a]) >> (8 * c)) & 0xFF, (int(w["hi"][a]) >> (8 * c)) & 0xFF
                if c >= n:
                    assert qlo == qhi == 0
                    continue
                plo, phi = origin + qlo * s, origin + qhi * s
                lo, hi = boxes[c][0][a], boxes[c][1][a]
                assert MARGIN * s <= lo - plo < (1 + MARGIN) * s, (i, c, a, lo, plo, s)
                assert MARGIN * s <= phi - hi < (1 + MARGIN) * s, (i, c, a, hi, phi, s)


def scenes():
    rng = np.random.default_rng(11)
    out = {"config1": synth.Synth.config(1, 0), "config2": synth.Synth.config(2, 0, 64)}
    # Far from the origin: a cloud of small triangles around (3e5, -2e5, 7e4).
    base = rng.uniform(-50, 50, size=(400, 1, 3)) + np.array([3e5, -2e5, 7e4])
    far = (base + rng.normal(scale=0.3, size=(400, 3, 3))).astype(np.float32)
    out["far"] = synth.Synth.from_triangles(far, cluster_size=64)
    # Slivers beside a large triangle, and flat triangles in one plane (zero-extent boxes on z).
    big = np.array([[[-100, -100, 0], [100, -100, 0], [0, 100, 0]]], dtype=np.float64)
    sl = rng.uniform(-1, 1, size=(200, 1, 3)) + rng.normal(scale=1e-4, size=(200, 3, 3))
    flat = rng.uniform(-30, 30, size=(200, 3, 3))
    flat[:, :, 2] = 5.0
    out["slivers"] = synth.Synth.from_triangles(np.concatenate([big, sl, flat]).astype(np.float32), cluster_size=64)
    out["single"] = synth.Synth.from_triangles(np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]]], dtype=np.float32))
    return out


@pytest.mark.parametrize("name", ["config1", "config2", "far", "slivers", "single"])
def test_wide_nodes_hold_their_children_on_the_finest_grid(name):
    s = scenes()[name]
    o = ob.Scene(s.desc())
    for mesh in list(range(len(s.arrays()["meshes"]))) + [0xFFFFFFFF]:
        nodes, _, _ = o.bvh(mesh)
        check_wide(nodes, o.bvh_wide(mesh))


def test_quantized_walk_equals_brute_force_on_hard_geometry():
    """Far from the origin, slivers beside a large triangle, flat boxes: the wide walk over quantized
    boxes finds exactly brute force's hits and occlusion bits, rays aimed at the triangles' edges
    and vertices included."""
    rng = np.random.default_rng(5)
    for name in ("far", "slivers"):
        s = scenes()[name]
        o = ob.Scene(s.desc())
        lo, hi = s.bounds()
        rays = np.concatenate([random_rays(3000, lo, hi, 3), rays_towards(3000, lo - 5, hi + 5, lo, hi, 4)])
        tris = o.bvh(0)[1]
        pick = rng.integers(0, len(tris), 2000)
        w = rng.dirichlet([0.3, 0.3, 0.3], size=2000)
        w[::3, 0] = 0.0  # on an edge
        w[1::3] = np.eye(3)[rng.integers(0, 3, len(w[1::3]))]  # on a vertex
        w /= w.sum(axis=1, keepdims=True)
        target = (w[:, 0, None] * tris["v0"][pick] + w[:, 1, None] * tris["v1"][pick]
                  + w[:, 2, None] * tris["v2"][pick])
        origin = target + rng.normal(size=(2000, 3)) * (hi - lo) * 0.2
        aimed = np.zeros(2000, dtype=abi.RAY_DTYPE)
        aimed["origin"] = origin
        aimed["dir"] = (target - origin) / np.linalg.norm(target - origin, axis=1, keepdims=True)
        aimed["tmax"] = 1e30
        rays = np.concatenate([rays, aimed])
        hits = o.trace_closest(rays)
        assert hits.tobytes() == o.trace_closest(rays, brute=True).tobytes(), name
        assert np.array_equal(o.trace_any(rays), o.trace_any(rays, brute=True)), name
