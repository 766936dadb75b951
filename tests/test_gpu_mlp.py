"""The warm-start network on the GPU (bf16 MFMA forward/backward kernels, Adam) against the CPU oracle.

The oracle (oracle/mlp_oracle.py) rounds to bf16 at the same places as the kernels and takes every product in
FP64 from those values, so the only difference left is the FP32 accumulation order inside the MFMA tiles (plus
the rare bf16 rounding flip it causes in a hidden activation). Tolerances: outputs within 2e-3 of their scale,
gradients within 5e-3 in norm per parameter block; the GPU must sit much closer to the bf16-mode oracle than to
the FP64-mode oracle, which shows that the rounding happens where the oracle says. Non-square, padded shapes
(batch 37, n_out 20) pin the MFMA lane layout: a transposed operand or a wrong row/column mapping fails.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as mo
from taral_pdhg import ffi

pytestmark = pytest.mark.gpu

SHAPE = (32, 48, 32, 20)


def _net(shape=SHAPE, seed=0):
    net = ffi.Mlp(*shape, seed=seed)
    rng = np.random.default_rng(seed + 1)
    p = (rng.standard_normal(mo.param_count(shape)) * 0.25).astype(np.float32)
    net.set_params(p)
    return net, p


def test_params_round_trip_and_init():
    net = ffi.Mlp(*SHAPE, seed=5)
    p0 = net.get_params()
    assert p0.size == mo.param_count(SHAPE) == net.param_count()
    W1, b1, W2, b2, W3, b3 = mo.unpack(p0, SHAPE)
    lim = np.sqrt(6 / 32)
    assert np.all(np.abs(W1) <= lim) and np.std(W1) > 0.3 * lim and not b1.any() and not b3.any()
    q = np.arange(p0.size, dtype=np.float32) / p0.size
    net.set_params(q)
    assert np.array_equal(net.get_params(), q)


def test_forward_matches_oracle():
    net, p = _net()
    X = np.random.default_rng(7).standard_normal((37, SHAPE[0])).astype(np.float32)
    y = net.forward(X)
    y16 = mo.forward(p, SHAPE, X, rounding=True)["Y"]
    y64 = mo.forward(p.astype(np.float64), SHAPE, X.astype(np.float64), rounding=False)["Y"]
    scale = np.max(np.abs(y64))
    e16, e64 = np.max(np.abs(y - y16)), np.max(np.abs(y - y64))
    print("forward err vs bf16 oracle", e16 / scale, "vs fp64 oracle", e64 / scale)
    assert e16 <= 2e-3 * scale
    assert e16 <= 0.25 * e64
    # a batch of one and a batch that is a multiple of 16
    for B in (1, 64):
        Xb = X[:B] if B <= 37 else np.tile(X, (2, 1))[:B]
        assert np.max(np.abs(net.forward(Xb) - mo.forward(p, SHAPE, Xb)["Y"])) <= 2e-3 * scale


def test_gradients_match_oracle():
    net, p = _net(seed=3)
    rng = np.random.default_rng(11)
    X = rng.standard_normal((37, SHAPE[0])).astype(np.float32)
    T = rng.standard_normal((37, SHAPE[3])).astype(np.float32)
    loss, g = net.loss_grad(X, T)
    l16, g16 = mo.loss_grad(p, SHAPE, X, T, rounding=True)
    l64, g64 = mo.loss_grad(p.astype(np.float64), SHAPE, X.astype(np.float64), T.astype(np.float64),
                            rounding=False)
    assert loss == pytest.approx(l16, rel=1e-4)
    blocks = [SHAPE[0] * SHAPE[1], SHAPE[1], SHAPE[1] * SHAPE[2], SHAPE[2], SHAPE[2] * SHAPE[3], SHAPE[3]]
    q = 0
    for name, k in zip(("W1", "b1", "W2", "b2", "W3", "b3"), blocks):
        a, b, c = g[q:q + k], g16[q:q + k], g64[q:q + k]
        e16, e64 = np.linalg.norm(a - b), np.linalg.norm(a - c)
        print(name, "grad err vs bf16 oracle", e16 / np.linalg.norm(b), "vs fp64", e64 / np.linalg.norm(b))
        assert e16 <= 5e-3 * np.linalg.norm(b), name
        q += k
    assert np.linalg.norm(g - g16) <= 0.25 * np.linalg.norm(g - g64)


def test_one_adam_step_matches_oracle():
    net, p = _net(seed=4)
    rng = np.random.default_rng(12)
    X = rng.standard_normal((64, SHAPE[0])).astype(np.float32)
    T = rng.standard_normal((64, SHAPE[3])).astype(np.float32)
    lr = 1e-3
    net.train(X, T, batch=64, epochs=1, lr=lr, seed=0)   # one full batch: one Adam step
    _, g = mo.loss_grad(p, SHAPE, X, T, rounding=True)
    z = np.zeros_like(p)
    p1, _, _ = mo.adam_step(p, g, z, z, 1, lr)
    d = np.abs(net.get_params() - p1)
    # step 1 moves each coordinate by ~lr * sign(g): gradients within rounding noise of zero may take the other
    # sign, everything else agrees closely
    assert np.max(d) <= 2.01 * lr
    assert np.mean(d <= 1e-2 * lr) >= 0.99


def test_training_fits_a_teacher_network():
    shape = (32, 512, 512, 257)   # the warm-start network's shape: 256 basis coefficients + log primal weight
    rng = np.random.default_rng(21)
    teacher = (rng.standard_normal(mo.param_count(shape)) * 0.05).astype(np.float32)
    X = rng.standard_normal((2048, 32)).astype(np.float32)
    T = mo.forward(teacher, shape, X, rounding=False)["Y"].astype(np.float32)
    net = ffi.Mlp(*shape, seed=1)
    l0, _ = net.loss_grad(X[:256], T[:256])
    last = net.train(X, T, batch=256, epochs=40, lr=1e-3, seed=2)
    l1, _ = net.loss_grad(X[:256], T[:256])
    print("teacher fit: loss", l0, "->", last, l1)
    assert l1 < 0.1 * l0


def test_bad_arguments():
    net = ffi.Mlp(*SHAPE)
    with pytest.raises(ValueError):
        net.forward(np.zeros((3, 31), np.float32))
    with pytest.raises(ffi.TpdhgError) as e:
        net.train(np.zeros((0, 32), np.float32), np.zeros((0, 20), np.float32))
    assert e.value.code == ffi.ERR_ARG
