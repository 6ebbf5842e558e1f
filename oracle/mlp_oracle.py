"""FP64 CPU oracle of the learned smoothing-length network (TEST INFRASTRUCTURE ONLY: imported by tests/,
the smoke check and nothing in the product).

The network is new code (the reference keeps h fixed, ie/crates/engine/src/hydrostatic.rs:58-61), so there is no
reference algorithm to restate: this file fixes the definition the device kernels must follow.
  forward:  a1 = gelu(W1 x + b1), a2 = gelu(W2 a1 + b2), y = W3 . a2 + b3
  gelu(z) = z/2 (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)))   (tanh form)
  loss    = mean over samples of (y - t)^2 / 2
Weights are BF16 on the device (round to nearest even of the FP32 master); `bf16_round` reproduces that, so the
oracle evaluates the same function the device evaluates, in FP64.

Pinned by (tests/test_oracle_mlp.py): central finite differences of the FP64 loss; closed forms for one-path
networks; gelu against its erf-free definition at known points; bf16 rounding against bit patterns.
"""
from __future__ import annotations

import numpy as np

N_IN, N_H = 6, 32
NPARAM = N_H * N_IN + N_H + N_H * N_H + N_H + N_H + 1
_K = np.sqrt(2.0 / np.pi)
_C = 0.044715


def bf16_round(w) -> np.ndarray:
    """FP32 -> BF16 (round to nearest, ties to even) -> FP32, elementwise."""
    u = np.ascontiguousarray(w, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(_K * (z + _C * z ** 3)))


def gelu_grad(z):
    t = np.tanh(_K * (z + _C * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * _K * (1.0 + 3.0 * _C * z * z)


def unpack(w):
    w = np.asarray(w, np.float64)
    o = 0
    W1 = w[o:o + N_H * N_IN].reshape(N_H, N_IN); o += N_H * N_IN
    b1 = w[o:o + N_H]; o += N_H
    W2 = w[o:o + N_H * N_H].reshape(N_H, N_H); o += N_H * N_H
    b2 = w[o:o + N_H]; o += N_H
    W3 = w[o:o + N_H]; o += N_H
    b3 = w[o]
    return W1, b1, W2, b2, W3, b3


def forward(w, x):
    """y for rows x (n, 6); w: flat parameters (used as given, in FP64)."""
    W1, b1, W2, b2, W3, b3 = unpack(w)
    x = np.asarray(x, np.float64)
    a1 = gelu(x @ W1.T + b1)
    a2 = gelu(a1 @ W2.T + b2)
    return a2 @ W3 + b3


def loss_and_grad(w, x, t):
    """Mean (y - t)^2 / 2 and its gradient with respect to the flat parameters, in FP64."""
    W1, b1, W2, b2, W3, b3 = unpack(w)
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    n = x.shape[0]
    z1 = x @ W1.T + b1
    a1 = gelu(z1)
    z2 = a1 @ W2.T + b2
    a2 = gelu(z2)
    y = a2 @ W3 + b3
    dy = (y - t) / n
    gW3 = a2.T @ dy
    gb3 = dy.sum()
    dz2 = (dy[:, None] * W3[None, :]) * gelu_grad(z2)
    gW2 = dz2.T @ a1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2) * gelu_grad(z1)
    gW1 = dz1.T @ x
    gb1 = dz1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return 0.5 * np.mean((y - t) ** 2), g
