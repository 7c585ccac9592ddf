"""FP64 numpy restatement of the selector MLP — TEST INFRASTRUCTURE ONLY
(the checker of the HIP training/inference kernels in
portblas_amd/csrc/mlp.hip; never called by the library).

Network: layer l maps d_l -> d_{l+1}, Z = X W^T + b, ReLU on hidden layers,
linear output; loss = sum(mask * (y - t)^2) / max(1, sum(mask)).
Pinned in tests/test_mlp.py by central finite differences.
"""
from __future__ import annotations

import numpy as np


def forward(ws, bs, x):
    """Returns the list of activations [x, a_1, ..., y] (float64)."""
    acts = [np.asarray(x, dtype=np.float64)]
    for l, (w, b) in enumerate(zip(ws, bs)):
        z = acts[-1] @ np.asarray(w, np.float64).T + np.asarray(b, np.float64)
        if l + 1 < len(ws):
            z = np.maximum(z, 0.0)
        acts.append(z)
    return acts


def loss(ws, bs, x, t, mask=None):
    y = forward(ws, bs, x)[-1]
    m = np.ones_like(y) if mask is None else np.asarray(mask, np.float64)
    return float(np.sum(m * (y - t) ** 2) / max(1.0, float(np.sum(m))))


def backward(ws, bs, x, t, mask=None):
    """Gradients (dW list, db list) and the loss."""
    acts = forward(ws, bs, x)
    y = acts[-1]
    m = np.ones_like(y) if mask is None else np.asarray(mask, np.float64)
    denom = max(1.0, float(np.sum(m)))
    g = 2.0 * m * (y - t) / denom
    dws, dbs = [None] * len(ws), [None] * len(ws)
    for l in range(len(ws) - 1, -1, -1):
        dws[l] = g.T @ acts[l]
        dbs[l] = g.sum(axis=0)
        if l > 0:
            g = (g @ np.asarray(ws[l], np.float64)) * (acts[l] > 0)
    return dws, dbs, float(np.sum(m * (y - t) ** 2) / denom)


def bf16_round(a):
    """Round-to-nearest-even to BF16, returned as float32."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out)


def forward_bf16(ws, bs, x):
    """BF16 inference: weights and each layer's input rounded to BF16, FP32
    accumulation (the arithmetic of pb_mlp_forward_bf16)."""
    h = np.asarray(x, np.float32)
    for l, (w, b) in enumerate(zip(ws, bs)):
        z = bf16_round(h).astype(np.float64) @ bf16_round(w).astype(np.float64).T + np.asarray(b, np.float64)
        h = z.astype(np.float32)
        if l + 1 < len(ws):
            h = np.maximum(h, 0.0)
    return h
