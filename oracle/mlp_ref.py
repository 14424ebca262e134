"""Reference network of the learned curvature closure -- TEST INFRASTRUCTURE.

FP64 numpy forward and backward of the 27 -> 64 -> 64 -> 1 GELU MLP.  With
bf16=True the operands of every matrix product are rounded to bfloat16
(round to nearest even) at the same points as the device kernels (weights,
first-layer activations, second- and first-layer deltas; the FP32 inputs
enter as the sum of two BF16 numbers, hi = bf16(x) and lo = bf16(x - hi)),
so the device gradients can be checked to accumulation-order precision; with
bf16=False it is the exact FP64 network, for finite-difference checks.
"""
from __future__ import annotations

import numpy as np

C = 0.7978845608028654


def bf16(x):
    a = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    a = (a + 0x7FFF + ((a >> 16) & 1)) & 0xFFFF0000
    return a.astype(np.uint32).view(np.float32).astype(np.float64)


def split_bf16(x):
    """FP32 x as hi + lo with both parts BF16 (about 16 significant bits)."""
    x32 = np.asarray(x, dtype=np.float32)
    hi = bf16(x32)
    lo = bf16((x32 - hi.astype(np.float32)).astype(np.float32))
    return hi + lo


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(C * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(C * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * C * (1.0 + 3 * 0.044715 * x * x)


def split(p):
    p = np.asarray(p, dtype=np.float64)
    o = np.cumsum([0, 64 * 27, 64, 64 * 64, 64, 64, 1])
    return (p[o[0]:o[1]].reshape(64, 27), p[o[1]:o[2]], p[o[2]:o[3]].reshape(64, 64), p[o[3]:o[4]],
            p[o[4]:o[5]], p[o[5]:o[6]])


def forward(p, x, bf=True):
    W1, b1, W2, b2, w3, b3 = split(p)
    r = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    X = split_bf16(x) if bf else np.asarray(x, dtype=np.float64)
    Z1 = X @ r(W1).T + b1
    H1 = r(gelu(Z1)) if bf else gelu(Z1)
    Z2 = H1 @ r(W2).T + b2
    H2 = gelu(Z2)
    y = H2 @ r(w3) + b3[0]
    return y, (X, Z1, H1, Z2, H2)


def loss_and_grad(p, x, t, bf=True):
    """Mean squared error over the samples and its gradient (FP64)."""
    W1, b1, W2, b2, w3, b3 = split(p)
    r = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    n = x.shape[0]
    y, (X, Z1, H1, Z2, H2) = forward(p, x, bf)
    e = y - t
    dy = 2.0 * e / n
    w3r = r(w3)
    dZ2 = dy[:, None] * w3r[None, :] * gelu_grad(Z2)
    D2 = r(dZ2)
    gw3 = H2.T @ dy
    gb2 = dZ2.sum(0)
    gb3 = dy.sum()
    gW2 = D2.T @ H1
    dH1 = D2 @ r(W2)
    dZ1 = dH1 * gelu_grad(Z1)
    D1 = r(dZ1)
    gb1 = dZ1.sum(0)
    gW1 = D1.T @ X
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return float((e * e).mean()), g
