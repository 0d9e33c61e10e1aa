"""A plain PyTorch float32 reference of the radiance cache's network, for the GPU tests.

The same function as the HIP kernels: the query encoding in its documented input order
(include/ge_rt.h, game_engine_amd/csrc/rt_mlp.h), BF16 weights and activations (round to nearest
even), float32 products and sums, ReLU between layers, no biases; the relative L2 loss and its
gradient by autograd; Adam as the backend states it (beta1 0.9, beta2 0.99, epsilon 1e-8).
"""
import numpy as np
import torch

from game_engine_amd import abi


def _input_table():
    """For each of the 192 inputs: (kind, feature, band, cosine). kind 0 band, 1 feature, 2 one, 3 zero."""
    table = []
    for idx in range(abi.CACHE_INPUTS):
        kb, g, j = idx // 32, (idx % 32) // 8, idx % 8
        G = 6 * g + kb
        if G < 21:
            table.append((0, G // 3, 4 * (G % 3) + j // 2, j & 1))
        elif G == 21 and j < 4:
            table.append((1, 7 + j, 0, 0))
        elif G == 21 and j == 4:
            table.append((2, 0, 0, 0))
        else:
            table.append((3, 0, 0, 0))
    return table


TABLE = _input_table()


def features(queries):
    """(n, 11) float64 features of a CACHE_QUERY_DTYPE array."""
    q = queries
    return np.concatenate([q["position"], q["direction"], q["normal"], q["albedo"], q["roughness"][:, None]],
                          axis=1).astype(np.float64)


def encode(queries):
    """(n, 192) float32 encoded inputs (before the BF16 rounding)."""
    f = features(queries)
    out = np.zeros((len(f), abi.CACHE_INPUTS), dtype=np.float64)
    for idx, (kind, feat, band, cosine) in enumerate(TABLE):
        if kind == 0:
            a = (2.0 ** band) * np.pi * f[:, feat]
            out[:, idx] = np.cos(a) if cosine else np.sin(a)
        elif kind == 1:
            out[:, idx] = f[:, feat]
        elif kind == 2:
            out[:, idx] = 1.0
    return out.astype(np.float32)


def layers(w):
    """The flat weight vector as six [out][in] float32 tensors."""
    d = abi.CACHE_DIMS
    out, at = [], 0
    for i in range(abi.CACHE_LAYERS):
        n = d[i] * d[i + 1]
        out.append(torch.as_tensor(np.asarray(w[at:at + n], dtype=np.float32).reshape(d[i + 1], d[i])))
        at += n
    return out


def bf16(t):
    return t.to(torch.bfloat16).to(torch.float32)


def forward(w, queries, master=None):
    """The network's output rows 0..2, float32 (n, 3). `master` (a list of six float32 tensors that
    may require grad) stands in for the weights when given; they are rounded to BF16 on the way."""
    x = bf16(torch.as_tensor(encode(queries)))
    ws = master if master is not None else layers(w)
    h = x
    for i, wi in enumerate(ws):
        wq = wi + (bf16(wi) - wi).detach()  # BF16 values, gradients straight through
        z = h @ wq.T
        if i < len(ws) - 1:
            h = torch.relu(z)
            h = h + (bf16(h) - h).detach()
        else:
            h = z
    return h[:, :3]


def relative_l2(pred, target):
    """mean over samples of sum over rgb of (p - y)^2 / (p^2 + 0.01), p fixed in the denominator."""
    denom = pred.detach() ** 2 + 0.01
    return (((pred - target) ** 2) / denom).sum(dim=1).mean()
