"""ORACLE -- test infrastructure only.

FP64 numpy restatement of the schedule network (16 -> 64 -> 64 -> 4, ReLU,
softmax cross-entropy) used to check the library's GPU forward and training
kernels.  The network has no reference counterpart (GraphHET schedules BFS
with the fixed rule of GPU_BFS_pre_csr.cuh:91), so the oracle is pinned by
mathematics instead: the analytic backward agrees with FP64 central
differences, and the forward reduces to hand-computable cases (zero weights
give the biases; a one-hot path through a single unit gives a product of
weights).

Parameter layout (the library's): W1[64][16], b1[64], W2[64][64], b2[64],
W3[4][64], b3[4], output-major.  Every function evaluates the network at the
FP16 rounding of the parameters, as the library does.

Only tests/ import this module; the product (graphhet_amd) never does.
"""
from __future__ import annotations

import numpy as np

IN, H, OUT = 16, 64, 4
SHAPES = [(H, IN), (H,), (H, H), (H,), (OUT, H), (OUT,)]
NPARAMS = sum(int(np.prod(s)) for s in SHAPES)


def unpack(p: np.ndarray) -> list[np.ndarray]:
    out, o = [], 0
    for s in SHAPES:
        k = int(np.prod(s))
        out.append(np.asarray(p[o:o + k], np.float64).reshape(s))
        o += k
    return out


def fp16(p: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even FP16 values of FP32 parameters, as FP64."""
    return np.asarray(p, np.float32).astype(np.float16).astype(np.float64)


def forward(p16: np.ndarray, x: np.ndarray):
    """Logits and the activations the backward needs (FP64)."""
    W1, b1, W2, b2, W3, b3 = unpack(p16)
    x = np.asarray(x, np.float64).reshape(-1, IN)
    z1 = x @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    z3 = h2 @ W3.T + b3
    return z3, (x, z1, h1, z2, h2)


def loss(p16: np.ndarray, x: np.ndarray, labels: np.ndarray) -> float:
    z3, _ = forward(p16, x)
    m = z3.max(axis=1, keepdims=True)
    lse = np.log(np.exp(z3 - m).sum(axis=1)) + m[:, 0]
    return float(np.mean(lse - z3[np.arange(len(labels)), labels]))


def grad(p16: np.ndarray, x: np.ndarray, labels: np.ndarray) -> np.ndarray:
    """Analytic gradient of the mean cross-entropy w.r.t. the parameters."""
    W1, b1, W2, b2, W3, b3 = unpack(p16)
    z3, (x, z1, h1, z2, h2) = forward(p16, x)
    B = x.shape[0]
    e = np.exp(z3 - z3.max(axis=1, keepdims=True))
    dz3 = e / e.sum(axis=1, keepdims=True)
    dz3[np.arange(B), labels] -= 1.0
    dz3 /= B
    gW3, gb3 = dz3.T @ h2, dz3.sum(axis=0)
    dz2 = (dz3 @ W3) * (z2 > 0)
    gW2, gb2 = dz2.T @ h1, dz2.sum(axis=0)
    dz1 = (dz2 @ W2) * (z1 > 0)
    gW1, gb1 = dz1.T @ x, dz1.sum(axis=0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])


def grad_fd(p16: np.ndarray, x: np.ndarray, labels: np.ndarray, idx, h: float = 1e-6) -> np.ndarray:
    """FP64 central differences of the loss at parameter indices `idx`."""
    out = []
    for i in idx:
        a, b = p16.copy(), p16.copy()
        a[i] += h
        b[i] -= h
        out.append((loss(a, x, labels) - loss(b, x, labels)) / (2 * h))
    return np.asarray(out)


def adam_step(master: np.ndarray, m1: np.ndarray, m2: np.ndarray, g: np.ndarray, step: int, lr: float,
              beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
    """One Adam update (bias-corrected), FP64."""
    m1 = beta1 * m1 + (1 - beta1) * g
    m2 = beta2 * m2 + (1 - beta2) * g * g
    w = master - lr * (m1 / (1 - beta1 ** step)) / (np.sqrt(m2 / (1 - beta2 ** step)) + eps)
    return w, m1, m2
