"""Pins of the CPU oracle of the warm-start network (oracle/mlp_oracle.py); CPU only.

The network has no counterpart in the reference, so the oracle is pinned by what the mathematics fixes: its
bf16 rounding against torch's own float32 -> bfloat16 conversion and hand-worked ties, its FP64 gradient against
central finite differences of the loss (a transposed operand, a dropped bias or a wrong ReLU mask fails), its
bf16 mode against the FP64 mode at bf16 accuracy, and Adam's first steps against their closed form.
"""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo


def test_bf16_matches_torch_conversion():
    rng = np.random.default_rng(0)
    a = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 20000),
                        np.float32([0.0, -0.0, 1.0, -3.5, 65504.0, 1e-38])]).astype(np.float32)
    ours = mo.bf16(a)
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))


def test_bf16_ties_to_even():
    one = np.float32(1.0)
    assert mo.bf16(np.float32([1 + 2.0 ** -8]))[0] == one            # tie between 1 and 1+2^-7: even is 1
    assert mo.bf16(np.float32([1 + 3 * 2.0 ** -8]))[0] == np.float32(1 + 2.0 ** -6)  # tie: even is 1+2^-6
    assert mo.bf16(np.float32([1 + 2.0 ** -8 + 2.0 ** -20]))[0] == np.float32(1 + 2.0 ** -7)  # above the tie


SHAPE = (16, 32, 16, 5)


def _case(seed=1, B=7):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(mo.param_count(SHAPE)) * 0.3
    X = rng.standard_normal((B, SHAPE[0]))
    T = rng.standard_normal((B, SHAPE[3]))
    return p, X, T


def test_fp64_gradient_matches_finite_differences():
    p, X, T = _case()
    loss, g = mo.loss_grad(p, SHAPE, X, T, rounding=False)
    assert loss == pytest.approx(mo.loss(p, SHAPE, X, T), rel=1e-14)
    rng = np.random.default_rng(2)
    # several coordinates of every block: W1, b1, W2, b2, W3, b3
    sizes = [16 * 32, 32, 32 * 16, 16, 16 * 5, 5]
    starts = np.cumsum([0] + sizes[:-1])
    h = 1e-6
    for s0, k in zip(starts, sizes):
        for j in rng.choice(k, size=min(k, 6), replace=False):
            e = np.zeros_like(p)
            e[s0 + j] = h
            fd = (mo.loss(p + e, SHAPE, X, T) - mo.loss(p - e, SHAPE, X, T)) / (2 * h)
            assert g[s0 + j] == pytest.approx(fd, rel=1e-5, abs=1e-9), (s0, j)


def test_bf16_mode_is_fp64_mode_to_bf16_accuracy():
    p, X, T = _case(3, 9)
    l64, g64 = mo.loss_grad(p, SHAPE, X, T, rounding=False)
    l16, g16 = mo.loss_grad(p, SHAPE, X, T, rounding=True)
    assert g16.dtype == np.float32
    assert abs(l16 - l64) <= 0.02 * l64
    assert np.linalg.norm(g16 - g64) <= 0.03 * np.linalg.norm(g64)
    # and the rounding is really applied: the two modes differ
    assert np.linalg.norm(g16 - g64) > 1e-6 * np.linalg.norm(g64)


def test_adam_first_steps_closed_form():
    # With bias correction, step 1 moves every coordinate by lr * g / (|g| + eps) -> lr * sign(g).
    g = np.float32([0.5, -2.0, 1e-3])
    p, m1, m2 = mo.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01)
    assert np.allclose(p, -0.01 * np.sign(g), rtol=1e-4)
    # step 2 with the same gradient: m-hat = g, v-hat = g^2 -> same move again
    p2, _, _ = mo.adam_step(p, g, m1, m2, 2, 0.01)
    assert np.allclose(p2, -0.02 * np.sign(g), rtol=1e-4)


def test_unpack_layout():
    p = np.arange(mo.param_count(SHAPE), dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = mo.unpack(p, SHAPE)
    assert W1.shape == (16, 32) and W1[0, 1] == 1 and W1[1, 0] == 32
    assert b1[0] == 16 * 32 and W2[0, 0] == 16 * 32 + 32 and b3[-1] == p[-1]
    with pytest.raises(ValueError):
        mo.unpack(p[:-1], SHAPE)
