"""numpy restatement of the capacity MLP (test infrastructure only).

Same layout as the HIP kernels: W1[64][8] b1 W2[64][64] b2 W3[64][64] b3
W4[3][64] b4, bfloat16-rounded weights, float32 activations. Used to check
the HIP forward pass and the HIP gradients of the masked Poisson negative
log-likelihood (rate 2^y, y clamped at 30).
"""

from __future__ import annotations

import numpy as np

SHAPES = [(64, 8), (64,), (64, 64), (64,), (64, 64), (64,), (3, 64), (3,)]


def unpack(p):
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(p[o:o + n], np.float64).reshape(s))
        o += n
    assert o == p.size
    return out


def bf16(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(np.shape(x))


def _weights(p, rounding):
    return [(bf16(a) if rounding else a).astype(np.float64) for a in unpack(p)]


def forward(p, x, rounding=True):
    W1, b1, W2, b2, W3, b3, W4, b4 = _weights(p, rounding)
    x = x.astype(np.float64)
    a1 = np.maximum(x @ W1.T + b1, 0)
    a2 = np.maximum(a1 @ W2.T + b2, 0)
    a3 = np.maximum(a2 @ W3.T + b3, 0)
    return a3 @ W4.T + b4, (x, a1, a2, a3)


LN2 = float(np.log(2.0))
MAX_LOG_RATE = 30.0


def loss_and_grad(p, x, t, m, rounding=True):
    """Masked Poisson NLL (mean over the batch) of rates 2^y against counts t,
    and its gradient in the parameter layout, computed in float64
    (bfloat16-rounded weights unless rounding=False)."""
    W1, b1, W2, b2, W3, b3, W4, b4 = _weights(p, rounding)
    y, (x0, a1, a2, a3) = forward(p, x, rounding)
    B = x.shape[0]
    t = np.asarray(t, np.float64)
    m = np.asarray(m, np.float64)
    yc = np.minimum(y, MAX_LOG_RATE)
    lam = np.exp2(yc)
    loss = float((m * (lam - t * yc * LN2)).sum() / B)
    dy = np.where(y > MAX_LOG_RATE, 0.0, m * (lam - t) * LN2 / B)
    gW4 = dy.T @ a3
    gb4 = dy.sum(0)
    d3 = (dy @ W4) * (a3 > 0)
    gW3 = d3.T @ a2
    gb3 = d3.sum(0)
    d2 = (d3 @ W3) * (a2 > 0)
    gW2 = d2.T @ a1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (a1 > 0)
    gW1 = d1.T @ x0
    gb1 = d1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4)])
    return loss, g
