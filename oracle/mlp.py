"""TEST INFRASTRUCTURE — the potential surrogate's network in numpy, for the gradient checks of the device kernels
(csrc/kernels_mlp.hpp). Arithmetic in float64; the operands of the hidden layers' matrix products are rounded to
bf16 (round to nearest even) as the device's matrix cores take them, so the device and this restatement differ by
the FP32 accumulation alone. Layout of the parameter vector: W1 (6 x 128), b1, W2 (128 x 128), b2, W3, b3, W4
(128), b4."""

from __future__ import annotations

import numpy as np

IN, WIDTH = 6, 128
SHAPES = [(IN, WIDTH), (WIDTH,), (WIDTH, WIDTH), (WIDTH,), (WIDTH, WIDTH), (WIDTH,), (WIDTH,), (1,)]
COUNT = sum(int(np.prod(s)) for s in SHAPES)


def split(p: np.ndarray) -> list[np.ndarray]:
    out, k = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(p[k:k + n], dtype=np.float64).reshape(s))
        k += n
    return out


def join(parts: list[np.ndarray]) -> np.ndarray:
    return np.concatenate([np.ravel(a) for a in parts])


def bf16(a: np.ndarray) -> np.ndarray:
    """float32 values rounded to bf16 (nearest even), as float64."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_slope(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p: np.ndarray, x: np.ndarray):
    w1, b1, w2, b2, w3, b3, w4, b4 = split(p)
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ w1 + b1
    a1 = silu(z1)
    z2 = bf16(a1) @ bf16(w2) + b2
    a2 = silu(z2)
    z3 = bf16(a2) @ bf16(w3) + b3
    a3 = silu(z3)
    y = a3 @ w4 + b4[0]
    return y, (x, z1, a1, z2, a2, z3, a3)


def loss_and_gradient(p: np.ndarray, x: np.ndarray, t: np.ndarray):
    """Mean squared error over the rows and its gradient, with the device's bf16 operands."""
    w1, b1, w2, b2, w3, b3, w4, b4 = split(p)
    y, (x, z1, a1, z2, a2, z3, a3) = forward(p, x)
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(d * d))
    dy = 2.0 * d / len(d)
    g_w4 = a3.T @ dy
    g_b4 = np.array([dy.sum()])
    dz3 = np.outer(dy, w4) * silu_slope(z3)
    g_w3 = bf16(a2).T @ bf16(dz3)
    g_b3 = dz3.sum(axis=0)
    dz2 = (bf16(dz3) @ bf16(w3).T) * silu_slope(z2)
    g_w2 = bf16(a1).T @ bf16(dz2)
    g_b2 = dz2.sum(axis=0)
    dz1 = (bf16(dz2) @ bf16(w2).T) * silu_slope(z1)
    g_w1 = x.T @ dz1
    g_b1 = dz1.sum(axis=0)
    return loss, join([g_w1, g_b1, g_w2, g_b2, g_w3, g_b3, g_w4, g_b4])


def initial_parameters(seed: int) -> np.ndarray:
    """He-scaled normal weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for s in SHAPES:
        if len(s) == 2:
            parts.append(rng.normal(0.0, np.sqrt(2.0 / s[0]), s))
        elif s == (WIDTH,) and len(parts) == 6:  # W4
            parts.append(rng.normal(0.0, np.sqrt(1.0 / WIDTH), s))
        else:
            parts.append(np.zeros(s))
    return join(parts).astype(np.float32)
