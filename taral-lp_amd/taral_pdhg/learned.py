"""Learned warm start for scenario families of one LP (no reference counterpart).

A family shares the matrix of a base LP; a scenario theta (one value per factor) rescales fixed subsets of b and
c (gen.ScenarioFamily). The warm start maps theta to a start point (x0, y0) and an initial primal weight:

  * the solutions (x, y) of the training scenarios, solved by cold-start PDHG on the GPU, are compressed by PCA
    (mean + the leading ``n_bases`` right singular vectors of the centred solution matrix, on the host);
  * a bf16 MFMA MLP (ffi.Mlp: n_factors -> 512 -> 512 -> n_bases + 1, trained on the GPU by the library's own
    forward and backward kernels) predicts the standardised PCA coefficients and the standardised log primal
    weight from theta / sigma;
  * the prediction is decoded to x0 (clipped to the bounds), y0 and (optionally) omega0, and handed to the batched
    solver with the cold-start fallback on: an instance whose predicted point has a worse relative KKT than the
    cold start starts cold, so plain PDHG both checks and backs up the network.
"""
from __future__ import annotations

import time

import numpy as np

from . import ffi, gen

# the scenarios of a family share A, so the batch keeps ||A_s|| from its first solve (reuse_norm)
FAST = dict(tol=1e-6, precision_bits=32, halpern=1, refine=1, reuse_norm=1)


class ScenarioSolver:
    """A batch of ``size`` instances of one scenario family kept on the GPU (the shared matrix stored once);
    ``solve`` loads scenarios into the instances with set_vectors and solves them together."""

    def __init__(self, fam: gen.ScenarioFamily, size: int, **opts):
        o = dict(FAST)
        o.update(opts)
        self.fam = fam
        self.size = int(size)
        self.models = [ffi.Model.from_lp(fam.base) for _ in range(self.size)]
        self.batch = ffi.Batch(self.models, **o)

    def solve(self, thetas, x0=None, y0=None, omega0=None) -> list:
        """Solves the scenarios ``thetas`` (at most ``size`` rows; unused instances keep the base vectors and are
        solved too, their results dropped). Warm starts: arrays with one row per scenario, or None."""
        thetas = np.atleast_2d(thetas)
        k = thetas.shape[0]
        if k > self.size:
            raise ValueError("more scenarios than instances")
        for i in range(k):
            self.batch.set_vectors(i, *self.fam.vectors(thetas[i]))
        xs = ys = pw = None
        if x0 is not None:
            xs = [x0[i] if i < k else None for i in range(self.size)]
            ys = [y0[i] if i < k else None for i in range(self.size)]
            pw = np.zeros(self.size)
            if omega0 is not None:
                pw[:k] = omega0
        res = self.batch.solve(x0=xs, y0=ys, primal_weight=pw)
        return res[:k]


def solve_family(fam, thetas, batch_size=256, warm=None, log=None, **opts):
    """Solves every scenario (chunks of ``batch_size`` in one batch); warm = (X0, Y0, omega0) or None.
    Returns (results, seconds spent in the solves)."""
    thetas = np.atleast_2d(thetas)
    size = min(batch_size, thetas.shape[0])
    solver = ScenarioSolver(fam, size, **opts)
    out, spent = [], 0.0
    for a in range(0, thetas.shape[0], size):
        b = min(a + size, thetas.shape[0])
        t = time.perf_counter()
        if warm is None:
            out += solver.solve(thetas[a:b])
        else:
            out += solver.solve(thetas[a:b], warm[0][a:b], warm[1][a:b], warm[2][a:b])
        spent += time.perf_counter() - t
        if log:
            log(f"solved scenarios {a}..{b - 1} in {time.perf_counter() - t:.1f}s")
    return out, spent


class WarmStartModel:
    """PCA bases of the training solutions + the GPU MLP from scenario parameters to basis coefficients and the
    log primal weight."""

    def __init__(self, fam: gen.ScenarioFamily, n_bases: int = 256, hidden=(512, 512), sigma: float = 0.1,
                 seed: int = 0):
        self.fam = fam
        self.n, self.m = fam.base.n, fam.base.m
        self.n_bases = int(n_bases)
        self.hidden = tuple(int(h) for h in hidden)
        self.sigma = float(sigma)
        self.seed = int(seed)
        self.net = None

    def fit(self, thetas, X, Y, omega, epochs: int = 300, batch: int = 256, lr: float = 1e-3) -> dict:
        thetas = np.asarray(thetas, dtype=np.float64)
        S = np.hstack([np.asarray(X, dtype=np.float64), np.asarray(Y, dtype=np.float64)])
        N = S.shape[0]
        self.mean = S.mean(axis=0)
        C = S - self.mean
        # economy SVD through the N x N Gram matrix (N << n + m)
        G = C @ C.T
        evals, evecs = np.linalg.eigh(G)
        order = np.argsort(evals)[::-1]
        k = min(self.n_bases, N - 1, int(np.sum(evals > 1e-12 * max(evals.max(), 1e-300))))
        k = max(k, 1)
        ev, U = evals[order[:k]], evecs[:, order[:k]]
        sv = np.sqrt(np.maximum(ev, 1e-300))
        self.bases = (U.T @ C) / sv[:, None]          # k x (n + m), orthonormal rows
        coef = C @ self.bases.T                        # N x k
        self.k = k
        self.explained = float(ev.sum() / max(np.trace(G), 1e-300))
        self.cm, self.cs = coef.mean(axis=0), coef.std(axis=0) + 1e-12
        lw = np.log(np.asarray(omega, dtype=np.float64))
        self.lm, self.ls = float(lw.mean()), float(lw.std() + 1e-12)
        T = np.hstack([(coef - self.cm) / self.cs, ((lw - self.lm) / self.ls)[:, None]]).astype(np.float32)
        Xin = (thetas / self.sigma).astype(np.float32)
        n_in = Xin.shape[1]
        if n_in % 16:
            raise ValueError("the number of scenario factors must be a multiple of 16")
        self.net = ffi.Mlp(n_in, self.hidden[0], self.hidden[1], k + 1, seed=self.seed)
        t = time.perf_counter()
        loss0, _ = self.net.loss_grad(Xin[:min(N, 1024)], T[:min(N, 1024)])
        last = self.net.train(Xin, T, batch=min(batch, N), epochs=epochs, lr=lr, seed=self.seed)
        self.train_s = time.perf_counter() - t
        return dict(samples=N, bases=k, explained_variance=self.explained, loss_initial=loss0, loss_final=last,
                    train_s=self.train_s)

    def predict(self, thetas, with_omega: bool = False):
        """(X0 [N x n], Y0 [N x m], omega0 [N]) for scenarios ``thetas``. omega0 is the predicted primal weight
        with ``with_omega``, else 0 (the solver's default initial weight): on the planted families the final primal
        weights of the training solves spread over four decades and predicting them lengthened the slowest solves,
        so the default leaves the weight to the solver."""
        out = self.net.forward((np.atleast_2d(thetas) / self.sigma).astype(np.float32)).astype(np.float64)
        coef = out[:, :self.k] * self.cs + self.cm
        S = self.mean + coef @ self.bases
        X0 = np.clip(S[:, :self.n], self.fam.base.col_lo, self.fam.base.col_up)
        Y0 = S[:, self.n:]
        omega = np.exp(out[:, self.k] * self.ls + self.lm) if with_omega else np.zeros(X0.shape[0])
        return X0, Y0, omega

    def loss(self, thetas, X, Y, omega) -> float:
        """Mean squared error of the network's standardised outputs on scenarios with known solutions."""
        S = np.hstack([X, Y]) - self.mean
        coef = S @ self.bases.T
        T = np.hstack([(coef - self.cm) / self.cs,
                       ((np.log(omega) - self.lm) / self.ls)[:, None]]).astype(np.float32)
        loss, _ = self.net.loss_grad((np.atleast_2d(thetas) / self.sigma).astype(np.float32), T)
        return loss
