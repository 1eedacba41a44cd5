"""FP64 restatement of the learned-initialisation MLP (test infrastructure).

No reference counterpart: the learned warm start is new.  This file fixes
what the GPU kernels must compute:
  patches  16 nearest neighbours of a point among the indexed points
           (exact (d², index) order, the point itself dropped), offsets
           (p_j − p_i)/r with r the 16th-neighbour distance, 48 values
  network  y = W3 g(W2 g(W1 x + b1) + b2) + b3, g = tanh-form GELU
           (Hendrycks & Gimpel 2016), ŷ = y/|y|
  loss     mean_i (1 − ŷ_i · t_i)
Pinned by tests/test_oracle_mlp.py: the analytic gradient against central
finite differences of the loss, GELU against its erf form, patch
invariants (scale and translation invariance, sortedness).
"""
from __future__ import annotations

import numpy as np

from . import knn_brute

IN, HID, OUT, K = 48, 128, 3, 16
SIZES = [(HID, IN), (HID,), (HID, HID), (HID,), (OUT, HID), (OUT,)]
N_PARAMS = sum(int(np.prod(s)) for s in SIZES)


def unpack(theta):
    out, o = [], 0
    for s in SIZES:
        n = int(np.prod(s))
        out.append(np.asarray(theta[o:o + n], np.float64).reshape(s))
        o += n
    return out


def gelu(x):
    c = np.sqrt(2.0 / np.pi)
    return 0.5 * x * (1.0 + np.tanh(c * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    c = np.sqrt(2.0 / np.pi)
    t = np.tanh(c * (x + 0.044715 * x ** 3))
    return 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * c * (1 + 3 * 0.044715 * x * x)


def forward(theta, X):
    W1, b1, W2, b2, W3, b3 = unpack(theta)
    z1 = X @ W1.T + b1
    h1 = gelu(z1)
    z2 = h1 @ W2.T + b2
    h2 = gelu(z2)
    y = h2 @ W3.T + b3
    return y, (X, z1, h1, z2, h2)


def loss_grad(theta, X, T, line=False):
    """mean(1 − ŷ·t), or with line=True mean(1 − (ŷ·t)²), and its gradient
    with respect to theta (FP64)."""
    y, (X, z1, h1, z2, h2) = forward(theta, X)
    W1, b1, W2, b2, W3, b3 = unpack(theta)
    B = len(X)
    ny = np.linalg.norm(y, axis=1, keepdims=True)
    yh = y / ny
    c = (yh * T).sum(1, keepdims=True)
    loss = float((1 - c * c).mean()) if line else float((1 - c).mean())
    dy = -(2 * c if line else 1.0) * (T - c * yh) / ny / B
    gW3 = dy.T @ h2
    gb3 = dy.sum(0)
    dz2 = (dy @ W3) * gelu_grad(z2)
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2) * gelu_grad(z1)
    gW1 = dz1.T @ X
    gb1 = dz1.sum(0)
    g = np.concatenate([a.ravel() for a in [gW1, gb1, gW2, gb2, gW3, gb3]])
    return loss, g


def loss_only(theta, X, T, line=False):
    y, _ = forward(theta, X)
    yh = y / np.linalg.norm(y, axis=1, keepdims=True)
    c = (yh * T).sum(1)
    return float((1 - c * c).mean()) if line else float((1 - c).mean())


def patches(P):
    """16-neighbour patches of every point of P (float32 arithmetic, as the
    kernel computes them, before BF16 rounding)."""
    P = np.ascontiguousarray(P, np.float32)
    idx, d2 = knn_brute(P, P, K + 1, return_d2=True)
    n = len(P)
    out = np.zeros((n, IN), np.float32)
    for i in range(n):
        sel, last = [], np.float32(0)
        skipped = False
        for j in range(K + 1):
            q = idx[i, j]
            if q < 0:
                break
            if not skipped and q == i:
                skipped = True
                continue
            if len(sel) < K:
                sel.append(q)
                last = d2[i, j]
        r = np.sqrt(np.float32(last)) if last > 0 else np.float32(1)
        inv = np.float32(1) / np.float32(r)
        off = (P[sel] - P[i]) * inv
        out[i, : 3 * len(sel)] = off.ravel()
    return out


def patches_from(S, Q):
    """16-neighbour patches of the points Q over the points S (float32
    arithmetic): 17 nearest by (d², index), the first at distance 0 dropped
    (else the 17th), offsets over the distance to the 16th kept neighbour."""
    S = np.ascontiguousarray(S, np.float32)
    Q = np.ascontiguousarray(Q, np.float32)
    idx, d2 = knn_brute(S, Q, K + 1, return_d2=True)
    out = np.zeros((len(Q), IN), np.float32)
    for i in range(len(Q)):
        keep = list(range(K + 1))
        if d2[i, 0] == 0:
            keep.pop(0)
        keep = [j for j in keep[:K] if idx[i, j] >= 0]
        last = d2[i, keep[-1]] if keep else np.float32(0)
        r = np.sqrt(np.float32(last)) if last > 0 else np.float32(1)
        inv = np.float32(1) / np.float32(r)
        off = (S[idx[i, keep]] - Q[i]) * inv
        out[i, : 3 * len(keep)] = off.ravel()
    return out


def to_bf16(x):
    """Round-to-nearest-even BF16 (returned as float32 values)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def bf16_bits(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
