"""CPU reference of the matrix-core perceptron (felis_amd/csrc/felis_mlp.hip).

TEST INFRASTRUCTURE ONLY: tests compare the GPU network's outputs and
gradients with this numpy restatement. It rounds to BF16 at the same points
the kernels do (GEMM operands: inputs, weights, activations, output and
hidden gradients) and accumulates in FP32/FP64, so the two agree to the
accumulation order.
"""
from __future__ import annotations

import numpy as np


def bf16(a):
    """Round to BF16 (nearest even) and return as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7F800000) == 0x7F800000
    nan &= (u & 0x7FFFFF) != 0
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, ((u >> 16) | 0x40) << 16, r)
    return r.astype(np.uint32).view(np.float32)


def gelu(z):
    c = 0.7978845608
    return 0.5 * z * (1 + np.tanh(c * (z + 0.044715 * z ** 3)))


def gelu_d(z):
    c = 0.7978845608
    u = c * (z + 0.044715 * z ** 3)
    t = np.tanh(u)
    return 0.5 * (1 + t) + 0.5 * z * (1 - t * t) * c * (1 + 3 * 0.044715 * z * z)


def _id(a):
    return np.asarray(a, np.float64)


def forward(Ws, bs, X, rounding=True):
    """Returns the output and the per-layer (H_in BF16, Z) for backward.
    rounding=False keeps every operand in FP64 (the exact function whose
    derivative the gradients restate)."""
    r = bf16 if rounding else _id
    H = r(X).astype(np.float64)
    cache = []
    for l, (W, b) in enumerate(zip(Ws, bs)):
        Z = H @ r(W).astype(np.float64) + b
        cache.append((H, Z))
        if l < len(Ws) - 1:
            H = r(gelu(Z)).astype(np.float64)
        else:
            H = Z
    return H, cache


def gradients(Ws, bs, X, T, wrow=None, rounding=True):
    """Mean squared error, dW and db as the kernels form them."""
    r = bf16 if rounding else _id
    Y, cache = forward(Ws, bs, X, rounding)
    M, N = Y.shape
    w = np.ones(M) if wrow is None else np.asarray(wrow, np.float64)
    loss = float(np.sum(w[:, None] * (Y - T) ** 2) / (M * N))
    dZf = 2 * w[:, None] * (Y - T) / (M * N)
    dZ = r(dZf).astype(np.float64)
    gWs, gbs = [None] * len(Ws), [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        H, Z = cache[l]
        if l < len(Ws) - 1:
            dZf = dH * gelu_d(Z)
            dZ = r(dZf).astype(np.float64)
        gWs[l] = H.T @ dZ
        gbs[l] = dZf.sum(0)
        if l > 0:
            dH = dZ @ r(Ws[l]).astype(np.float64).T
    return loss, gWs, gbs
