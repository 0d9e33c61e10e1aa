"""The radiance cache on the GPU, through the C ABI.

Bars:
- the network (fused BF16 MFMA inference) against a plain PyTorch float32 reference of the same
  function (BF16-rounded weights and activations): within 3e-2 of the output scale, the BF16
  rounding of six layers' activations in another summation order;
- one training step against the reference's loss (1e-2 relative) and, Adam's first step being
  -lr g / |g|, the signs of the reference's gradient (97% of the components that matter);
- the cached tracer with a network that answers zero: bit-exact against the oracle's cached
  tracer (the same paths ended at the same vertices); with a real network, the oracle's radiance
  plus its throughput times the reference's answer, within the network's bar;
- the training samples: bit-exact against the oracle's (queries and targets), as a set;
- online training on the scene lowers the cached image's error against a plain render.
"""
import numpy as np
import pytest
import torch

from game_engine_amd import abi, rt, synth
from oracle import binding as ob
from tests import cache_reference as ref

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = rt.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def c2(ctx):
    s = synth.Synth.config(2, 0)
    print(f"[scene] {s.triangle_count} triangles, {s.desc().instance_count} instances", flush=True)
    g = ctx.upload(s.desc()).build()
    o = ob.Scene(s.desc())
    lo, hi = s.bounds()
    yield s, g, o, lo, hi
    g.close()


def _queries(n, seed):
    rng = np.random.default_rng(seed)
    q = np.zeros(n, dtype=abi.CACHE_QUERY_DTYPE)
    q["position"] = rng.uniform(0, 1, (n, 3))
    q["direction"] = rng.uniform(0, 1, (n, 2))
    q["normal"] = rng.uniform(0, 1, (n, 2))
    q["albedo"] = rng.uniform(0.05, 0.95, (n, 3))
    q["roughness"] = rng.uniform(0.05, 1.0, n)
    return q


def _frame(s, w=64, h=48, spp=2, bounces=4):
    pt, _ = s.frame(w, h)
    pt.spp = spp
    pt.max_bounces = bounces
    return pt


@pytest.mark.parametrize("n", [1, 33, 5000])
def test_inference_matches_the_torch_reference(ctx, c2, n):
    s, g, o, lo, hi = c2
    cache = rt.RadianceCache(ctx, lo, hi, seed=3)
    w = cache.weights()
    q = _queries(n, n)
    out = cache.infer(q)
    want = ref.forward(w, q).numpy()
    scale = max(np.abs(want).max(), 1e-3)
    print(f"[infer n={n}] max |diff| {np.abs(out - want).max():.3e} of scale {scale:.3e}", flush=True)
    assert np.abs(out - want).max() <= 3e-2 * scale
    cache.close()


def test_weights_round_trip(ctx, c2):
    s, g, o, lo, hi = c2
    cache = rt.RadianceCache(ctx, lo, hi, seed=1)
    w = np.random.default_rng(2).normal(0, 0.1, abi.CACHE_WEIGHTS).astype(np.float32)
    cache.set_weights(w)
    assert np.array_equal(cache.weights(), w)
    cache.close()


# Measured on MI355X: 3.8e-4 and 1.4e-4.
GRAD_REL_NORM, GRAD_MAX_COMPONENT = 2e-3, 1e-3


def test_training_step_matches_the_torch_reference(ctx, c2):
    s, g, o, lo, hi = c2
    cache = rt.RadianceCache(ctx, lo, hi, seed=5)
    w0 = cache.weights()
    n = 3000
    smp = np.zeros(n, dtype=abi.CACHE_SAMPLE_DTYPE)
    smp["query"] = _queries(n, 11)
    rng = np.random.default_rng(12)
    smp["radiance"] = rng.uniform(0.0, 2.0, (n, 3)).astype(np.float32)
    lr = 1e-3
    loss = cache.train(smp, lr)
    w1 = cache.weights()
    master = [t.clone().requires_grad_(True) for t in ref.layers(w0)]
    pred = ref.forward(None, smp["query"], master)
    tl = ref.relative_l2(pred, torch.as_tensor(smp["radiance"]))
    tl.backward()
    grad = torch.cat([t.grad.reshape(-1) for t in master]).numpy()
    print(f"[train] loss {loss:.6f} reference {tl.item():.6f}", flush=True)
    assert loss == pytest.approx(tl.item(), rel=1e-2)
    # The weight gradient itself against autograd's: both take BF16 weights and activations, the
    # backend also BF16 deltas, so they agree to BF16's precision, not to the bit.
    gw = cache.gradient()
    rel = np.linalg.norm(gw - grad) / np.linalg.norm(grad)
    worst = np.abs(gw - grad).max() / np.abs(grad).max()
    print(f"[train] weight gradient vs autograd: relative norm {rel:.3e}, max component {worst:.3e}", flush=True)
    assert rel <= GRAD_REL_NORM and worst <= GRAD_MAX_COMPONENT
    dw = w1 - w0
    big = np.abs(grad) > 1e-3 * np.abs(grad).max()
    agree = np.mean(np.sign(dw[big]) == -np.sign(grad[big]))
    print(f"[train] {big.sum()} components, sign agreement {agree:.4f}", flush=True)
    assert agree >= 0.97
    # Adam's first step moves each weight with a gradient by lr.
    assert np.allclose(np.abs(dw[big]), lr, rtol=1e-2)
    # The unused output rows never move.
    off = abi.CACHE_WEIGHTS - abi.CACHE_WIDTH * abi.CACHE_OUTPUTS
    rows = w1[off:].reshape(abi.CACHE_OUTPUTS, abi.CACHE_WIDTH)
    assert np.array_equal(rows[3:], w0[off:].reshape(abi.CACHE_OUTPUTS, abi.CACHE_WIDTH)[3:])
    cache.close()


def test_training_fits_a_smooth_field(ctx, c2):
    s, g, o, lo, hi = c2
    cache = rt.RadianceCache(ctx, lo, hi, seed=7)
    n = 4096
    losses = []
    for it in range(300):
        smp = np.zeros(n, dtype=abi.CACHE_SAMPLE_DTYPE)
        smp["query"] = _queries(n, 1000 + it)
        p = smp["query"]["position"].astype(np.float64)
        a = smp["query"]["albedo"].astype(np.float64)
        field = 0.6 + 0.4 * np.sin(2 * np.pi * p[:, :1]) * np.cos(np.pi * p[:, 1:2])
        smp["radiance"] = (field * a).astype(np.float32)
        losses.append(cache.train(smp, 5e-3))
    print(f"[fit] loss {losses[0]:.4f} -> {np.mean(losses[-10:]):.5f}", flush=True)
    assert np.mean(losses[-10:]) < 0.05 * losses[0]
    cache.close()


def _by_sample(values, n, spp):
    total = np.zeros((n, values.shape[1]), dtype=values.dtype)
    for si in range(spp):  # per pixel, in sample order, as the tracer sums them
        total = total + values[si * n:(si + 1) * n]
    return total


def test_cached_tracer_with_a_silent_cache_is_the_oracle(ctx, c2):
    s, g, o, lo, hi = c2
    cache = rt.RadianceCache(ctx, lo, hi, seed=0)
    cache.set_weights(np.zeros(abi.CACHE_WEIGHTS, dtype=np.float32))
    pt = _frame(s)
    acc, cov = cache.path_trace(g, pt, depth=1)
    rad, w, q = o.path_cached(pt, 1, cache.box_min, cache.box_inv())
    n = pt.width * pt.height
    assert w[:, 3].sum() > 0.1 * n * pt.spp
    assert np.array_equal(acc[:, :3], _by_sample(rad, n, pt.spp))
    assert np.array_equal(acc[:, 3], np.full(n, pt.spp, dtype=np.float32))
    _, plain_cov = g.path_trace(pt)
    assert np.array_equal(cov, plain_cov)
    # Resuming at sample_base > 0 adds into the accumulator as the plain tracer does.
    pt.sample_base = pt.spp
    acc2, _ = cache.path_trace(g, pt, depth=1, accum=acc.copy(), coverage=cov.copy())
    rad2, _, _ = o.path_cached(pt, 1, cache.box_min, cache.box_inv())
    assert np.array_equal(acc2[:, :3], acc[:, :3] + _by_sample(rad2, n, pt.spp))
    cache.close()


def test_cached_tracer_adds_the_network(ctx, c2):
    s, g, o, lo, hi = c2
    cache = rt.RadianceCache(ctx, lo, hi, seed=9)
    w = cache.weights()
    pt = _frame(s)
    acc, _ = cache.path_trace(g, pt, depth=1)
    rad, wt, q = o.path_cached(pt, 1, cache.box_min, cache.box_inv())
    pred = np.maximum(ref.forward(w, q).numpy().astype(np.float64), 0.0)
    asked = wt[:, 3:4] > 0
    want_path = rad.astype(np.float64) + np.where(asked, wt[:, :3].astype(np.float64) * pred, 0.0)
    n = pt.width * pt.height
    want = _by_sample(want_path, n, pt.spp)
    weight = _by_sample(np.where(asked, wt[:, :3].astype(np.float64), 0.0), n, pt.spp)
    scale = max(pred[asked[:, 0]].max(), 1e-3)
    err = np.abs(acc[:, :3] - want)
    print(f"[cached] max |diff| {err.max():.3e}, prediction scale {scale:.3e}", flush=True)
    assert (err <= 3e-2 * scale * weight + 1e-5 * (1 + np.abs(want))).all()
    cache.close()


@pytest.mark.parametrize("silent", [True, False])
def test_self_training_frame_matches_the_oracle(ctx, c2, silent):
    # One cached frame that trains: every third path goes on past the cache's vertex. The image and,
    # with a cache that answers zero, the samples and their targets are the oracle's to the bit;
    # with a real network the targets carry the reference's answer at the training path's end.
    s, g, o, lo, hi = c2
    cache = rt.RadianceCache(ctx, lo, hi, seed=6)
    if silent:
        cache.set_weights(np.zeros(abi.CACHE_WEIGHTS, dtype=np.float32))
    w = cache.weights()
    pt = _frame(s, spp=2, bounces=6)
    pt.sample_base = 5
    tp = abi.CacheTrainParams(samples=1 << 16, steps=1, learning_rate=0.0, stride=3)
    acc, _ = cache.path_trace(g, pt, depth=1, train=tp)
    assert np.array_equal(cache.weights(), w)  # a zero learning rate moves nothing
    v = o.path_cached_training(pt, 1, 3, pt.sample_base % 3, cache.box_min, cache.box_inv())
    n = pt.width * pt.height
    asked = v["weight"][:, 3:4] > 0
    pred = np.maximum(ref.forward(w, v["queries"]).numpy().astype(np.float64), 0.0)
    image = v["radiance"].astype(np.float64) + np.where(asked, v["weight"][:, :3] * pred, 0.0)
    sampled = v["sample_weight"][:, 3] > 0
    ended = v["train_weight"][:, 3:4] > 0
    pred2 = np.maximum(ref.forward(w, v["train_queries"]).numpy().astype(np.float64), 0.0)
    R = v["train_radiance"].astype(np.float64) + np.where(ended, v["train_weight"][:, :3] * pred2, 0.0)
    T1 = v["sample_weight"][:, :3]
    got = cache.samples()
    got = got[np.argsort(got["path"])]
    want = v["samples"][sampled]
    assert len(got) == sampled.sum() > 0
    assert np.array_equal(got["path"], want["path"])
    for f in ("position", "direction", "normal", "albedo", "roughness", "tag"):
        assert np.array_equal(got["query"][f], want["query"][f]), f
    if silent:
        assert np.array_equal(acc[:, :3], _by_sample(v["radiance"], n, pt.spp))
        Rf = v["train_radiance"][sampled]
        L1 = want["radiance"]
        tgt = np.where(T1[sampled] > 0, (Rf - L1) / np.where(T1[sampled] > 0, T1[sampled], 1), np.float32(0))
        assert np.array_equal(got["radiance"], tgt.astype(np.float32))
    else:
        scale = max(pred[asked[:, 0]].max(), pred2[ended[:, 0]].max(), 1e-3)
        err = np.abs(acc[:, :3] - _by_sample(image, n, pt.spp))
        assert (err <= 3e-2 * scale * _by_sample(np.where(asked, v["weight"][:, :3], 0.0), n, pt.spp)
                + 1e-5 * (1 + np.abs(image).max())).all()
        L1 = want["radiance"].astype(np.float64)
        t = T1[sampled].astype(np.float64)
        tgt = np.where(t > 0, (R[sampled] - L1) / np.where(t > 0, t, 1), 0.0)
        bound = 3e-2 * scale * np.where(t > 0, v["train_weight"][sampled, :3] / np.where(t > 0, t, 1), 0.0)
        assert (np.abs(got["radiance"] - tgt) <= bound + 1e-4 * (1 + np.abs(tgt))).all()
    cache.close()


def test_training_samples_match_the_oracle(ctx, c2):
    s, g, o, lo, hi = c2
    cache = rt.RadianceCache(ctx, lo, hi, seed=0)
    pt = _frame(s, spp=1, bounces=6)
    got, count = cache.collect(g, pt, 1 << 16)
    want, wcount = o.cache_samples(pt, cache.box_min, cache.box_inv())
    assert count == wcount and len(got) == count > 0
    key = lambda a: np.lexsort((a["query"]["tag"], a["path"]))  # noqa: E731
    got, want = got[key(got)], want[key(want)]
    assert np.array_equal(got["path"], want["path"])
    for f in ("position", "direction", "normal", "albedo", "roughness", "tag"):
        assert np.array_equal(got["query"][f], want["query"][f]), f
    assert np.array_equal(got["radiance"], want["radiance"])
    # Past the capacity the count still covers every vertex, and what is kept is a subset.
    half, c_all = cache.collect(g, pt, count // 2)
    assert c_all == count and len(half) == count // 2
    keys = set(zip(want["path"].tolist(), want["query"]["tag"].tolist()))
    assert set(zip(half["path"].tolist(), half["query"]["tag"].tolist())) <= keys
    cache.close()


@pytest.mark.parametrize("mode", ["separate", "in_frame"])
def test_online_training_lowers_the_cached_error(ctx, c2, mode):
    # Sun and sky only, pixel-centre rays: the camera vertex is the same in every sample, so the
    # error left is the light the cache stands in for. Training either on separate plain paths
    # (ge_rt_cache_update) or inside the cached frames themselves (self-training).
    s, g, o, lo, hi = c2
    cache = rt.RadianceCache(ctx, lo, hi, seed=4)
    view = _frame(s, w=96, h=64, spp=16, bounces=6)
    view.light_count = 0
    view.flags = abi.PT_PIXEL_CENTER
    plain = np.zeros((view.width * view.height, 4), dtype=np.float32)
    cov = np.zeros(view.width * view.height, dtype=np.uint32)
    for k in range(16):  # the plain tracer's 256-spp image of the view
        view.sample_base = k * view.spp
        g.path_trace(view, plain, cov)
    truth = plain[:, :3] / plain[:, 3:4]

    def cached_error():
        view.sample_base = 0
        view.spp = 16
        acc, _ = cache.path_trace(g, view, depth=1)
        return float(np.sqrt(np.mean((acc[:, :3] / acc[:, 3:4] - truth) ** 2)))

    before = cached_error()
    losses = []
    if mode == "separate":
        train = _frame(s, w=96, h=64, spp=1, bounces=6)
        train.light_count = 0
        for frame in range(150):
            train.sample_base = 1000 + frame
            losses.append(cache.update(g, train, samples=8192, steps=2, learning_rate=5e-3))
    else:
        tp = abi.CacheTrainParams(samples=8192, steps=2, learning_rate=5e-3, stride=0)
        view.spp = 4
        for frame in range(150):
            view.sample_base = 1000 + 4 * frame
            losses.append(cache.path_trace(g, view, depth=1, train=tp, want_loss=True)[2])
    after = cached_error()
    print(f"[online {mode}] loss {np.mean(losses[:5]):.4f} -> {np.mean(losses[-5:]):.4f}; "
          f"cached RMSE {before:.4f} -> {after:.4f}", flush=True)
    assert after < 0.5 * before
    cache.close()
