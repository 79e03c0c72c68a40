"""FP64 numpy oracle of the LM-seeding MLP (test infrastructure only).

The network has no reference counterpart; it is pinned by mathematics: the
backward pass here is checked against central finite differences of the
forward pass (tests/test_mlp_oracle.py), and the device kernels are compared
with this restatement.  `bf16=True` rounds exactly the values the device
stores in BF16 (the input, the weights, the hidden activations and the
back-propagated pre-activation gradients) to nearest-even BF16 while keeping
every sum in FP64, so the comparison isolates the device's accumulation
order and GEMM layout from the BF16 format itself.

Layout: flat parameters W0 (1536x256), b0, W1 (256x256), b1, W2, b2,
W3 (256x68), b3; y = x W + b, GELU(z) = z Phi(z) after the three hidden layers.
Loss = 0.5/M sum_i sum_j wt_j (y_ij - t_ij)^2.
"""
from __future__ import annotations

import math

import numpy as np

IN, HIDDEN, OUT = 1536, 256, 68
LAYERS = ((IN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, OUT))
NPARAMS = sum(i * o + o for i, o in LAYERS)

_erf = np.vectorize(math.erf, otypes=[np.float64])


def bf16_round(x) -> np.ndarray:
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    out, o = [], 0
    for i, n in LAYERS:
        W = p[o:o + i * n].reshape(i, n)
        o += i * n
        out.append((W, p[o:o + n]))
        o += n
    return out


def gelu(z):
    return 0.5 * z * (1 + _erf(z / math.sqrt(2)))


def gelu_grad(z):
    return 0.5 * (1 + _erf(z / math.sqrt(2))) + z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)


def forward(p, X, bf16=False):
    """Returns (Y, cache).  Y = last pre-activation (FP64)."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    p64 = np.asarray(p, dtype=np.float64)
    layers = unpack(p64)
    h = r(X)
    cache = []
    for li, (W, b) in enumerate(layers):
        z = h @ r(W) + b
        cache.append((h, z))
        h = r(gelu(z)) if li < len(layers) - 1 else z
    return h, cache


def loss_grad(p, X, Tg, wt, bf16=False):
    """(loss, flat gradient) of the weighted MSE, FP64 back-propagation."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    M = X.shape[0]
    Y, cache = forward(p, X, bf16)
    E = Y - np.asarray(Tg, dtype=np.float64)
    wt = np.asarray(wt, dtype=np.float64)
    loss = 0.5 * np.sum(wt * E * E) / M
    # the device rounds the output gradient to FP32 before the backward pass
    dH = (wt * E / M).astype(np.float32).astype(np.float64) if bf16 else wt * E / M
    layers = unpack(np.asarray(p, dtype=np.float64))
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        W, b = layers[li]
        h, z = cache[li]
        dZ = dH * gelu_grad(z) if li < len(layers) - 1 else dH
        db = dZ.sum(axis=0)
        dZr = r(dZ)
        dW = h.T @ dZr
        grads.append((dW, db))
        dH = dZr @ r(W).T
    g = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in reversed(grads)])
    return loss, g


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (bias-corrected), FP64; returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
