"""The learned band proposal's network and features restated in numpy: the checker for the device's bf16 MFMA
forward and backward kernels and its k-mer featurizer (test infrastructure only, as the rest of oracle/).

The network has no counterpart in the reference. Its arithmetic is stated here from the mathematics: bf16 is
float32 with its low 16 bits rounded off to nearest even; each layer multiplies bf16 inputs by bf16 weights and
accumulates in float32; ReLU between; the loss is the mean squared error of log(1 + cost). The backward pass is the
chain rule, written out; `numeric_gradient` checks it by central differences on a float64 copy of the same network.
"""

from __future__ import annotations

import numpy as np

STRIDE = 288
HIDDEN = 256
FEATURES = 258


def bf16(x: np.ndarray) -> np.ndarray:
    """float32 rounded to bfloat16 (nearest even), returned as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def split(weights: np.ndarray):
    """The flat weight vector as (w1, b1, w2, b2, w3, b3)."""
    h = HIDDEN
    at = 0
    out = []
    for size, shape in ((STRIDE * h, (STRIDE, h)), (h, (h,)), (h * h, (h, h)), (h, (h,)), (h, (h,)), (1, (1,))):
        out.append(weights[at:at + size].reshape(shape))
        at += size
    return out


def forward(weights: np.ndarray, x: np.ndarray):
    """The device's forward pass: (y, z1, a1, z2, a2), bf16 where the device keeps bf16."""
    w1, b1, w2, b2, w3, b3 = split(weights.astype(np.float32))
    xb = bf16(x)
    z1 = (xb @ bf16(w1)).astype(np.float32) + b1
    a1 = bf16(np.maximum(z1, 0))
    z2 = (a1 @ bf16(w2)).astype(np.float32) + b2
    a2 = bf16(np.maximum(z2, 0))
    y = (a2 @ bf16(w3)).astype(np.float32) + b3[0]
    return y, z1, a1, z2, a2


def gradients(weights: np.ndarray, x: np.ndarray, t: np.ndarray):
    """The device's backward pass: the flat gradient of the mean squared error, and the loss."""
    w1, b1, w2, b2, w3, b3 = split(weights.astype(np.float32))
    y, z1, a1, z2, a2 = forward(weights, x)
    count = len(t)
    dy = (2.0 * (y - t) / count).astype(np.float32)
    w3b = bf16(w3)
    dz2 = bf16(np.where(z2 > 0, dy[:, None] * w3b[None, :], 0))
    dw3 = a2.T @ dy
    db3 = np.array([dy.sum()], dtype=np.float32)
    dw2 = a1.T @ dz2
    db2 = dz2.sum(axis=0)
    dz1 = bf16(np.where(z1 > 0, dz2 @ bf16(w2).T, 0))
    dw1 = bf16(x).T @ dz1
    db1 = dz1.sum(axis=0)
    flat = np.concatenate([dw1.ravel(), db1, dw2.ravel(), db2, dw3, db3]).astype(np.float32)
    return flat, float(np.mean((y - t) ** 2))


def exact_loss(weights: np.ndarray, x: np.ndarray, t: np.ndarray) -> float:
    """The same network in float64 with no rounding: what the gradients differentiate, up to the bf16 steps."""
    w1, b1, w2, b2, w3, b3 = [w.astype(np.float64) for w in split(weights)]
    a1 = np.maximum(x @ w1 + b1, 0)
    a2 = np.maximum(a1 @ w2 + b2, 0)
    y = a2 @ w3 + b3[0]
    return float(np.mean((y - t) ** 2))


def exact_gradients(weights: np.ndarray, x: np.ndarray, t: np.ndarray) -> np.ndarray:
    """The float64, unrounded chain rule (the same formulas as `gradients`)."""
    w1, b1, w2, b2, w3, b3 = [w.astype(np.float64) for w in split(weights)]
    z1 = x @ w1 + b1
    a1 = np.maximum(z1, 0)
    z2 = a1 @ w2 + b2
    a2 = np.maximum(z2, 0)
    y = a2 @ w3 + b3[0]
    dy = 2.0 * (y - t) / len(t)
    dz2 = np.where(z2 > 0, dy[:, None] * w3[None, :], 0)
    dz1 = np.where(z1 > 0, dz2 @ w2.T, 0)
    return np.concatenate([(x.T @ dz1).ravel(), dz1.sum(0), (a1.T @ dz2).ravel(), dz2.sum(0), a2.T @ dy,
                           [dy.sum()]])


def numeric_gradient(weights: np.ndarray, x: np.ndarray, t: np.ndarray, index: int, step: float = 1e-4) -> float:
    """Central difference of `exact_loss` in parameter `index`."""
    w = weights.astype(np.float64).copy()
    w[index] += step
    up = exact_loss(w, x, t)
    w[index] -= 2 * step
    down = exact_loss(w, x, t)
    return (up - down) / (2 * step)


CODES = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3}


def features(ref: bytes, qry: bytes) -> np.ndarray:
    """A pair's features as the device computes them: 4-mer count differences (reference less query) over the
    2-bit codes A, C, G, T = 0..3, scaled by 8 / sqrt(1 + n + m), then (log(1 + n) - 7) / 2 and the same of m."""
    out = np.zeros(STRIDE, dtype=np.float32)
    hist = np.zeros(256, dtype=np.int64)
    for text, delta in ((ref, 1), (qry, -1)):
        codes = np.array([CODES[b] for b in text], dtype=np.int64)
        if len(codes) >= 4:
            kmers = codes[:-3] * 64 + codes[1:-2] * 16 + codes[2:-1] * 4 + codes[3:]
            hist += delta * np.bincount(kmers, minlength=256)
    n, m = len(ref), len(qry)
    out[:256] = hist.astype(np.float32) * np.float32(8.0 / np.sqrt(1.0 + n + m))
    out[256] = (np.log(1.0 + n) - 7.0) * 0.5
    out[257] = (np.log(1.0 + m) - 7.0) * 0.5
    return out
