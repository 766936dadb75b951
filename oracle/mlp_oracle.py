"""CPU oracle of the warm-start network (n_in -> h1 -> h2 -> n_out MLP, ReLU). TEST INFRASTRUCTURE ONLY.

Imported by tests/ as the oracle of the GPU network (taral-lp_amd/csrc/mlp.hip); never part of the product.
The learned warm start has no counterpart in the reference, so there is no reference file to follow: the
algorithm is textbook back-propagation of the mean squared error and Adam (Kingma & Ba, 2015, Algorithm 1).

Two modes share one code path:
  * rounding=False: FP64 everywhere. Its gradient is pinned by central finite differences of the loss
    (tests/test_mlp_oracle.py), which fails for a transposed operand, a dropped bias or a wrong ReLU mask.
  * rounding=True: the GPU's number format. bf16 (round to nearest even) is applied at exactly the places the
    GPU kernels round — the inputs, the weight copies, the hidden activations and the back-propagated errors
    dY, dZ2, dZ1 before each matrix product — and every product is taken in FP64 from those bf16 values (the
    bf16 x bf16 products are exact in FP32 too, so only the FP32 accumulation order of the GPU differs) and
    stored as FP32. Biases, pre-activations, outputs, losses, bias gradients and the Adam state are FP32.

Parameter vector layout (FP32): W1 [n_in x h1], b1 [h1], W2 [h1 x h2], b2 [h2], W3 [h2 x n_out], b3 [n_out].
"""
from __future__ import annotations

import numpy as np


def bf16(a) -> np.ndarray:
    """Round FP32 values to bf16 (nearest, ties to even; finite inputs), returned as FP32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def param_count(shape) -> int:
    n_in, h1, h2, n_out = shape
    return n_in * h1 + h1 + h1 * h2 + h2 + h2 * n_out + n_out


def unpack(p, shape):
    """(W1, b1, W2, b2, W3, b3) views of the flat parameter vector."""
    n_in, h1, h2, n_out = shape
    p = np.asarray(p)
    if p.size != param_count(shape):
        raise ValueError("parameter vector does not match the shape")
    out, q = [], 0
    for dims in ((n_in, h1), (h1,), (h1, h2), (h2,), (h2, n_out), (n_out,)):
        k = int(np.prod(dims))
        out.append(p[q:q + k].reshape(dims))
        q += k
    return out


def _mode(rounding):
    if rounding:
        return bf16, (lambda a: np.asarray(a, dtype=np.float32))
    return (lambda a: np.asarray(a, dtype=np.float64)), (lambda a: np.asarray(a, dtype=np.float64))


def _mm(a, b):
    return np.asarray(a, dtype=np.float64) @ np.asarray(b, dtype=np.float64)


def forward(p, shape, X, rounding=True) -> dict:
    """Intermediates of one forward pass: Xb, Z1 (= H1, post-ReLU), H1b, Z2, H2b, Y."""
    W1, b1, W2, b2, W3, b3 = unpack(p, shape)
    rnd, st = _mode(rounding)
    Xb = rnd(X)
    Z1 = st(np.maximum(_mm(Xb, rnd(W1)) + b1, 0.0))
    H1b = rnd(Z1)
    Z2 = st(np.maximum(_mm(H1b, rnd(W2)) + b2, 0.0))
    H2b = rnd(Z2)
    Y = st(_mm(H2b, rnd(W3)) + b3)
    return {"Xb": Xb, "Z1": Z1, "H1b": H1b, "Z2": Z2, "H2b": H2b, "Y": Y}


def loss_grad(p, shape, X, T, rounding=True):
    """Mean squared error over batch x n_out and its gradient (flat vector, FP32 or FP64 by mode)."""
    _, _, W2, _, W3, _ = unpack(p, shape)
    rnd, st = _mode(rounding)
    f = forward(p, shape, X, rounding)
    B, n_out = np.shape(T)
    Y = f["Y"]
    if rounding:
        d = (Y - np.asarray(T, dtype=np.float32)).astype(np.float32)
        dY = (np.float32(2.0) / np.float32(B * n_out)) * d
    else:
        d = Y - np.asarray(T, dtype=np.float64)
        dY = 2.0 * d / (B * n_out)
    loss = float(np.sum(np.asarray(d, dtype=np.float64) ** 2) / (B * n_out))
    dYb = rnd(dY)
    gW3 = st(_mm(f["H2b"].T, dYb))
    gb3 = st(np.sum(np.asarray(dY, dtype=np.float64), axis=0))
    dZ2 = st(_mm(dYb, rnd(W3).T) * (f["Z2"] > 0))
    dZ2b = rnd(dZ2)
    gW2 = st(_mm(f["H1b"].T, dZ2b))
    gb2 = st(np.sum(np.asarray(dZ2, dtype=np.float64), axis=0))
    dZ1 = st(_mm(dZ2b, rnd(W2).T) * (f["Z1"] > 0))
    dZ1b = rnd(dZ1)
    gW1 = st(_mm(f["Xb"].T, dZ1b))
    gb1 = st(np.sum(np.asarray(dZ1, dtype=np.float64), axis=0))
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, g


def loss(p, shape, X, T, rounding=False) -> float:
    Y = forward(p, shape, X, rounding)["Y"]
    return float(np.mean((np.asarray(Y, dtype=np.float64) - np.asarray(T, dtype=np.float64)) ** 2))


def adam_step(p, g, m1, m2, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update in FP32 (bias-corrected moments; step counts from 1). Returns (p, m1, m2)."""
    f = np.float32
    p, g, m1, m2 = (np.asarray(a, dtype=np.float32) for a in (p, g, m1, m2))
    c1, c2 = f(1.0 - f(b1) ** f(step)), f(1.0 - f(b2) ** f(step))
    m1 = f(b1) * m1 + f(1.0 - b1) * g
    m2 = f(b2) * m2 + f(1.0 - b2) * g * g
    p = p - f(lr) * (m1 / c1) / (np.sqrt(m2 / c2) + f(eps))
    return p.astype(np.float32), m1.astype(np.float32), m2.astype(np.float32)
