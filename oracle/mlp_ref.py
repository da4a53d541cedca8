"""CPU restatement of the learned load balancer's MLP (fp64 numpy).

TEST INFRASTRUCTURE ONLY: the checker of libspz's MFMA forward/backward
kernels (no reference counterpart: the reference has no learned component).
`forward`/`grad` follow the device arithmetic step by step -- bf16 weights,
bf16 activations between layers, fp32 biases -- but accumulate in fp64, so a
device result differs only by fp32 accumulation order.  `grad_exact` is the
same network without any bf16 rounding, checked against finite differences.
"""
from __future__ import annotations

import numpy as np

IN, H = 640, 256
OFF_W1 = 0
OFF_B1 = OFF_W1 + IN * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
N_PARAMS = OFF_B3 + 1


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(IN, H), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def features(boards, thr, h):
    """x[t*N2 + c] = 1 if tile t at cell c; x[625] = h/32; x[626] = (thr-h)/32."""
    boards = np.asarray(boards, dtype=np.int64)
    n, n2 = boards.shape
    x = np.zeros((n, IN))
    rows = np.repeat(np.arange(n), n2)
    cells = np.tile(np.arange(n2), n)
    x[rows, boards.reshape(-1) * n2 + cells] = 1.0
    x[:, 625] = np.asarray(h) / 32.0
    x[:, 626] = (np.asarray(thr) - np.asarray(h)) / 32.0
    return x


def forward(p, x, rounding=True):
    r = bf16 if rounding else (lambda v: np.asarray(v, dtype=np.float64))
    W1, b1, W2, b2, w3, b3 = unpack(p)
    b1 = np.float32(b1).astype(np.float64) if rounding else b1
    z1 = r(x) @ r(W1) + b1
    h1 = r(np.maximum(z1, 0))
    z2 = h1 @ r(W2) + b2
    h2 = r(np.maximum(z2, 0))
    y = h2 @ r(w3) + b3
    return y, (x, z1, h1, z2, h2)


def grad(p, x, target, rounding=True):
    """loss = mean (y - t)^2 and d loss / d params (flat)."""
    r = bf16 if rounding else (lambda v: np.asarray(v, dtype=np.float64))
    W1, b1, W2, b2, w3, b3 = unpack(p)
    y, (x, z1, h1, z2, h2) = forward(p, x, rounding)
    n = len(y)
    d = y - np.asarray(target, dtype=np.float64)
    loss = float(np.mean(d * d))
    dy = 2 * d / n
    g = np.zeros(N_PARAMS)
    g[OFF_W3:OFF_B3] = h2.T @ dy
    g[OFF_B3] = dy.sum()
    dz2 = np.outer(dy, r(w3)) * (z2 > 0)
    dz2b = r(dz2)
    g[OFF_W2:OFF_B2] = (h1.T @ dz2b).reshape(-1)
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dh1 = dz2b @ r(W2).T
    dz1 = dh1 * (z1 > 0)
    dz1b = r(dz1)
    g[OFF_W1:OFF_B1] = (r(x).T @ dz1b).reshape(-1)
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    return loss, g
