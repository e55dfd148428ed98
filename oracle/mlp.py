"""CPU oracle of the learned proposal's network (test infrastructure).

fp64 forward / backward of the 2-layer MLP with the GPU's rounding points
restated: W1, w2 and the inputs x enter the first-layer product as bf16
(round-to-nearest-even), the W1 gradient multiplies bf16 dH by bf16 x; every
sum is exact fp64 here (fp32 accumulate on the GPU), so the two agree to
bf16-product precision.  Layout: W1 (256 x 16), b1 (256), w2 (256), b2.

Pinned by tests/test_learned.py: the backward equals central finite
differences of the fp64 loss (with the rounding switched off), and the GPU
kernels equal this oracle.
"""
import numpy as np

IN, H = 16, 256
P = H * IN + H + H + 1


def bf16(a):
    """Round fp32 values to bf16 (nearest even), returned as float64."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[:H * IN].reshape(H, IN)
    b1 = p[H * IN:H * IN + H]
    w2 = p[H * IN + H:H * IN + 2 * H]
    return W1, b1, w2, p[-1]


def forward(p, X, rounding=True):
    W1, b1, w2, b2 = unpack(p)
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    Hp = r(X) @ r(W1).T + b1
    y = np.maximum(Hp, 0.0) @ r(w2) + b2
    return Hp, y


def loss_and_grad(p, X, t, rounding=True):
    """Mean binary cross-entropy of the logits and its gradient."""
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, w2, b2 = unpack(p)
    Hp, y = forward(p, X, rounding)
    B = len(t)
    loss = float(np.mean(np.log1p(np.exp(-np.abs(y))) + np.maximum(y, 0) - t * y))
    dy = (1.0 / (1.0 + np.exp(-y)) - t) / B
    act = np.maximum(Hp, 0.0)
    dH = (dy[:, None] * r(w2)[None, :]) * (Hp > 0)
    g = np.zeros(P)
    g[:H * IN] = (r(dH).T @ r(X)).ravel()
    g[H * IN:H * IN + H] = dH.sum(axis=0)
    g[H * IN + H:H * IN + 2 * H] = dy @ act
    g[-1] = dy.sum()
    return loss, g
