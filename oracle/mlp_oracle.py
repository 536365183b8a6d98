"""CPU oracle of the warm-start MLP (TEST INFRASTRUCTURE ONLY: the checker the
HIP kernels of libpdal_ws.so are compared against; nothing in the product
imports it).

Same architecture and parameter layout as include/pdal_warmstart.h: dense
layers, ReLU between hidden layers, none after the last, W_l stored out x in,
loss = mean over batch and outputs of (prediction - target)^2.  With
`bf16=True` every GEMM input (activations, weights, back-propagated
gradients) is rounded to BF16 (round to nearest even) before an FP64
product, which is what the device does ahead of its FP32 MFMA accumulation;
with `bf16=False` the whole computation is FP64 and is pinned by finite
differences in the tests.
"""
from __future__ import annotations

import numpy as np


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest BF16 (ties to even); returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def split_params(p: np.ndarray, widths):
    out, off = [], 0
    for l in range(len(widths) - 1):
        n_in, n_out = widths[l], widths[l + 1]
        W = p[off:off + n_out * n_in].reshape(n_out, n_in)
        off += n_out * n_in
        b = p[off:off + n_out]
        off += n_out
        out.append((W, b))
    return out


def forward(p, widths, X, bf16=True):
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    H = [np.asarray(X, dtype=np.float64)]
    layers = split_params(np.asarray(p, dtype=np.float64), widths)
    for l, (W, b) in enumerate(layers):
        Z = r(H[-1].astype(np.float32) if bf16 else H[-1]) @ r(W.astype(np.float32) if bf16 else W).T + b
        if bf16:
            Z = Z.astype(np.float32).astype(np.float64)  # the device stores activations in FP32
        H.append(np.maximum(Z, 0.0) if l + 1 < len(layers) else Z)
    return H


def loss_grad(p, widths, X, Y, bf16=True):
    """(loss, gradient in the parameter layout)."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    f32 = (lambda a: a.astype(np.float32)) if bf16 else (lambda a: a)
    layers = split_params(np.asarray(p, dtype=np.float64), widths)
    H = forward(p, widths, X, bf16)
    P = H[-1]
    D = P - np.asarray(Y, dtype=np.float64)
    n = D.size
    loss = float(np.sum(D * D) / n)
    G = 2.0 * D / n
    if bf16:
        G = G.astype(np.float32).astype(np.float64)
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        dW = r(f32(G)).T @ r(f32(H[l]))
        db = G.sum(axis=0)
        grads.append((dW, db))
        if l > 0:
            G = (r(f32(G)) @ r(f32(W))) * (H[l] > 0)
            if bf16:
                G = G.astype(np.float32).astype(np.float64)
    grads.reverse()
    flat = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
    return loss, flat


def init_params(widths, rng: np.random.Generator) -> np.ndarray:
    parts = []
    for l in range(len(widths) - 1):
        a = np.sqrt(6.0 / widths[l])
        parts.append(rng.uniform(-a, a, widths[l + 1] * widths[l]))
        parts.append(np.zeros(widths[l + 1]))
    return np.concatenate(parts)
