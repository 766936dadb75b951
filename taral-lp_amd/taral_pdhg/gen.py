"""Seeded synthetic LP instances for the restarted PDHG solver.

Two families:

* ``planted_lp`` — random sparse equality LPs with a known optimum (the benchmark family). A has
  ``nnz_per_col`` distinct uniform-random rows per column with N(0,1) values; x* has half its entries 0 and
  the rest U(0,1); b = A x*; y* ~ N(0,1); c = A^T y* + s with s_j ~ U(0,1) where x*_j = 0 and 0 elsewhere;
  bounds 0 <= x <= 10. x* is primal feasible, (y*, s) is dual feasible with s'x* = 0, so c'x* = b'y* is the
  optimal objective.
* ``packing_mps`` / ``transport_mps`` — the reference's seeded crossover-study instances, produced by the
  same numpy call sequence as gpu/gen_lp.py:46-101 and written as the same free-format MPS text
  (gpu/gen_lp.py:20-43, ``%.12g`` numbers), so that the reference's published PDHG ledgers
  (results/gpu_pdhg/synth_*.csv) apply to them unchanged.

Every generator returns an :class:`LPModel` holding the column-wise model exactly as the MPS reader would
produce it (the ``*_mps`` ones also give the MPS text).
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class LPModel:
    """Column-wise LP  min c'x + obj_const  s.t.  row_lo <= A x <= row_up,  col_lo <= x <= col_up.

    Mirrors the fields of the reference ``Model`` (src/taral.hpp:24-35) that the PDHG path reads.
    ``colptr`` is int64 [n+1], ``rowidx`` int32 [nnz] (ascending within each column), ``val`` float64.
    """

    m: int
    n: int
    colptr: np.ndarray
    rowidx: np.ndarray
    val: np.ndarray
    cost: np.ndarray
    col_lo: np.ndarray
    col_up: np.ndarray
    row_lo: np.ndarray
    row_up: np.ndarray
    obj_const: float = 0.0
    maximize: bool = False
    name: str = "LP"
    planted_objective: float | None = None
    x_star: np.ndarray | None = None
    y_star: np.ndarray | None = None

    @property
    def nnz(self) -> int:
        return int(self.colptr[-1])

    def csr(self):
        """(ptr int32 [m+1], idx int32, val float64) of A by rows, columns ascending within a row."""
        cols = np.repeat(np.arange(self.n, dtype=np.int32), np.diff(self.colptr))
        order = np.lexsort((cols, self.rowidx))
        ptr = np.zeros(self.m + 1, dtype=np.int64)
        np.cumsum(np.bincount(self.rowidx, minlength=self.m), out=ptr[1:])
        return ptr.astype(np.int32), cols[order].astype(np.int32), self.val[order]

    def matvec(self, x):
        cols = np.repeat(np.arange(self.n), np.diff(self.colptr))
        return np.bincount(self.rowidx, weights=self.val * x[cols], minlength=self.m)

    def rmatvec(self, y):
        cols = np.repeat(np.arange(self.n), np.diff(self.colptr))
        return np.bincount(cols, weights=self.val * y[self.rowidx], minlength=self.n)


def _distinct_rows(rng, m, n, k):
    """k distinct uniform row indices per column (sorted), by redrawing the columns that hit a repeat."""
    rows = np.sort(rng.integers(0, m, (n, k), dtype=np.int64), axis=1)
    bad = np.nonzero((rows[:, 1:] == rows[:, :-1]).any(axis=1))[0]
    while bad.size:
        redraw = np.sort(rng.integers(0, m, (bad.size, k), dtype=np.int64), axis=1)
        rows[bad] = redraw
        still = (redraw[:, 1:] == redraw[:, :-1]).any(axis=1)
        bad = bad[still]
    return rows


def planted_lp(m: int, n: int, nnz_per_col, seed: int, lo: float = 0.0, up: float = 10.0,
               name: str | None = None) -> LPModel:
    """The planted equality-LP family. ``nnz_per_col`` is an int (fixed) or an array of per-column counts."""
    rng = np.random.default_rng(seed)
    if np.isscalar(nnz_per_col):
        k = int(nnz_per_col)
        rows = _distinct_rows(rng, m, n, k)
        counts = np.full(n, k, dtype=np.int64)
        rowidx = rows.reshape(-1)
    else:
        counts = np.asarray(nnz_per_col, dtype=np.int64)
        parts = []
        for k in np.unique(counts):
            sel = np.nonzero(counts == k)[0]
            parts.append((sel, _distinct_rows(rng, m, sel.size, int(k))))
        rowidx = np.empty(int(counts.sum()), dtype=np.int64)
        starts = np.concatenate([[0], np.cumsum(counts)])
        for sel, rws in parts:
            k = rws.shape[1]
            pos = (starts[sel][:, None] + np.arange(k)[None, :]).reshape(-1)
            rowidx[pos] = rws.reshape(-1)
    colptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=colptr[1:])
    nnz = int(colptr[-1])
    val = rng.standard_normal(nnz)
    xstar = np.zeros(n)
    nz = rng.permutation(n)[n // 2:]
    xstar[nz] = rng.random(nz.size)
    cols = np.repeat(np.arange(n), counts)
    b = np.bincount(rowidx, weights=val * xstar[cols], minlength=m)
    ystar = rng.standard_normal(m)
    s = np.where(xstar == 0, rng.random(n), 0.0)
    c = np.bincount(cols, weights=val * ystar[rowidx], minlength=n) + s
    p_star = float(np.dot(c, xstar))
    return LPModel(m=m, n=n, colptr=colptr, rowidx=rowidx.astype(np.int32), val=val, cost=c,
                   col_lo=np.full(n, lo), col_up=np.full(n, up), row_lo=b.copy(), row_up=b.copy(),
                   name=name or f"PLANTED_{m}x{n}", planted_objective=p_star, x_star=xstar, y_star=ystar)


# ------------------------------------------------------------------ gpu/gen_lp.py-equivalent instances
def _g12(v):
    """%.12g text of each value, as gpu/gen_lp.py:20-21 writes numbers."""
    return np.char.mod('%.12g', v)


def _parsed(v):
    """What the MPS reader makes of a %.12g-formatted value (correctly rounded decimal to binary)."""
    return np.array([float('%.12g' % t) for t in np.asarray(v, dtype=np.float64).tolist()], dtype=np.float64)


def packing_mps(N: int, K: int = 5, seed: int = 1, text: bool = True):
    """(LPModel, MPS text) of ``gen_lp.py packing N --k K --seed seed`` (gpu/gen_lp.py:71-101); the text is None
    with text=False (the model is the same: its values are those the text reads back as)."""
    rng = np.random.default_rng(seed)
    M = max(1, N // 2)
    rows = np.sort(rng.integers(0, M, (N, K)), axis=1)
    dup = np.zeros_like(rows, dtype=bool)
    dup[:, 1:] = rows[:, 1:] == rows[:, :-1]
    vals = rng.uniform(0.1, 1.0, (N, K))
    vals[dup] = 0
    x0 = rng.random(N)
    ax0 = np.zeros(M)
    np.add.at(ax0, rows.ravel(), (vals * x0[:, None]).ravel())
    b = ax0 * rng.uniform(1.0, 1.5, M) + 0.1
    p = rng.uniform(0.5, 1.5, N)
    keep = ~dup
    counts = keep.sum(axis=1)
    colptr = np.zeros(N + 1, dtype=np.int64)
    np.cumsum(counts, out=colptr[1:])
    rowidx = rows[keep].astype(np.int32)
    val = _parsed(vals[keep])
    model = LPModel(m=M, n=N, colptr=colptr, rowidx=rowidx, val=val, cost=_parsed(-p),
                    col_lo=np.zeros(N), col_up=np.ones(N), row_lo=np.full(M, -np.inf), row_up=_parsed(b),
                    name=f"PACKING_{N}")
    if not text:
        return model, None
    # MPS text, line for line as gpu/gen_lp.py:24-43 writes it
    rn = np.array([f'R{i}' for i in range(M)])
    out = [f'NAME PACKING_{N}\nROWS\n N OBJ\n', ''.join(f' L {r}\n' for r in rn), 'COLUMNS\n']
    cn = np.char.add('X', np.arange(N).astype(str))
    cn_all = np.concatenate([cn[:, None], np.repeat(cn, K).reshape(-1, K)], axis=1)
    rn_all = np.concatenate([np.full((N, 1), 'OBJ'), rn[rows]], axis=1)
    v_all = np.concatenate([-p[:, None], vals], axis=1)
    mask = np.concatenate([np.ones((N, 1), bool), keep], axis=1)
    lines = np.char.add(np.char.add(np.char.add(np.char.add('    ', cn_all[mask]), ' '),
                                    np.char.add(rn_all[mask], ' ')), _g12(v_all[mask]))
    out.append('\n'.join(lines.tolist()) + '\n')
    out.append('RHS\n' + ''.join(f'    RHS {r} {v:.12g}\n' for r, v in zip(rn, b)))
    ub = np.char.add(np.char.add(np.char.add(' UP BND ', cn), ' '), _g12(np.ones(N)))
    out.append('BOUNDS\n' + '\n'.join(ub.tolist()) + '\n' + 'ENDATA\n')
    return model, ''.join(out)


def transport_mps(S: int, D: int, seed: int = 1):
    """(LPModel, MPS text) of ``gen_lp.py transport S D --seed seed`` (gpu/gen_lp.py:46-68)."""
    rng = np.random.default_rng(seed)
    ps, pd = rng.random((S, 2)), rng.random((D, 2))
    demand = rng.uniform(1, 10, D)
    supply = rng.uniform(1, 10, S)
    supply *= 1.25 * demand.sum() / supply.sum()
    ii = np.repeat(np.arange(S), D)
    jj = np.tile(np.arange(D), S)
    cost = np.sqrt(((ps[ii] - pd[jj]) ** 2).sum(axis=1))
    n = S * D
    colptr = np.arange(0, 2 * n + 1, 2, dtype=np.int64)
    rowidx = np.stack([ii, S + jj], axis=1).reshape(-1).astype(np.int32)
    model = LPModel(m=S + D, n=n, colptr=colptr, rowidx=rowidx, val=np.ones(2 * n), cost=_parsed(cost),
                    col_lo=np.zeros(n), col_up=np.full(n, np.inf),
                    row_lo=np.concatenate([np.full(S, -np.inf), _parsed(demand)]),
                    row_up=np.concatenate([_parsed(supply), np.full(D, np.inf)]),
                    name=f"TRANSPORT_{S}x{D}")
    sn = [f'S{i}' for i in range(S)]
    dn = [f'D{j}' for j in range(D)]
    out = [f'NAME TRANSPORT_{S}x{D}\nROWS\n N OBJ\n',
           ''.join(f' {t} {r}\n' for t, r in zip(['L'] * S + ['G'] * D, sn + dn)), 'COLUMNS\n']
    cn = np.char.add('X', (ii * D + jj).astype(str))
    cn3 = np.repeat(cn, 3)
    rn3 = np.stack([np.full(n, 'OBJ'), np.array(sn)[ii], np.array(dn)[jj]], axis=1).ravel()
    v3 = np.stack([cost, np.ones(n), np.ones(n)], axis=1).ravel()
    lines = np.char.add(np.char.add(np.char.add(np.char.add('    ', cn3), ' '), np.char.add(rn3, ' ')), _g12(v3))
    out.append('\n'.join(lines.tolist()) + '\n')
    out.append('RHS\n' + ''.join(f'    RHS {r} {v:.12g}\n' for r, v in zip(sn + dn, list(supply) + list(demand))))
    out.append('ENDATA\n')
    return model, ''.join(out)


def _r(v) -> str:
    """Shortest round-trip decimal of a float (exact on reading back)."""
    return repr(float(v))


def write_mps(model: LPModel, path: str) -> None:
    """Free-format MPS of a model (equality rows as E, boxed rows as G with RANGES, bounds as LO/UP/FR/MI)."""
    m, n = model.m, model.n
    with open(path, 'w') as f:
        f.write(f'NAME {model.name}\n')
        if model.maximize:
            f.write('OBJSENSE\n MAX\n')
        f.write('ROWS\n N OBJ\n')
        kinds, rhs, rng_ = [], [], {}
        for i in range(m):
            lo, up = model.row_lo[i], model.row_up[i]
            if lo == up:
                kinds.append('E'), rhs.append(lo)
            elif np.isfinite(lo) and np.isfinite(up):
                kinds.append('G'), rhs.append(lo)
                rng_[i] = up - lo
            elif np.isfinite(lo):
                kinds.append('G'), rhs.append(lo)
            else:
                kinds.append('L'), rhs.append(up)
        f.write(''.join(f' {k} R{i}\n' for i, k in enumerate(kinds)))
        f.write('COLUMNS\n')
        for j in range(n):
            f.write(f'    X{j} OBJ {_r(model.cost[j])}\n')
            for k in range(model.colptr[j], model.colptr[j + 1]):
                f.write(f'    X{j} R{model.rowidx[k]} {_r(model.val[k])}\n')
        f.write('RHS\n')
        for i in range(m):
            if rhs[i] != 0 and np.isfinite(rhs[i]):
                f.write(f'    RHS R{i} {_r(rhs[i])}\n')
        if model.obj_const:
            f.write(f'    RHS OBJ {_r(-model.obj_const)}\n')
        if rng_:
            f.write('RANGES\n' + ''.join(f'    RNG R{i} {_r(v)}\n' for i, v in rng_.items()))
        f.write('BOUNDS\n')
        for j in range(n):
            lo, up = model.col_lo[j], model.col_up[j]
            if lo == -np.inf and up == np.inf:
                f.write(f' FR BND X{j}\n')
                continue
            if lo == -np.inf:
                f.write(f' MI BND X{j}\n')
            elif lo != 0:
                f.write(f' LO BND X{j} {_r(lo)}\n')
            if up != np.inf:
                f.write(f' UP BND X{j} {_r(up)}\n')
        f.write('ENDATA\n')


class ScenarioFamily:
    """Scenario LPs that share the base LP's matrix: each of ``n_factors`` factors owns a fixed random subset of
    ``frac`` of the rows and of the columns, and a scenario theta (one value per factor) multiplies the row bounds
    of its rows and the costs of its columns by exp(theta_k) (overlapping subsets multiply). The base rows of the
    planted family are equalities, so a scenario moves b = row_lo = row_up; the planted optimum no longer holds.
    """

    def __init__(self, base: LPModel, n_factors: int = 32, frac: float = 0.0015, seed: int = 0):
        rng = np.random.default_rng(seed)
        self.base = base
        self.n_factors = int(n_factors)
        kr, kc = max(1, int(round(frac * base.m))), max(1, int(round(frac * base.n)))
        self.row_sets = [rng.choice(base.m, kr, replace=False) for _ in range(self.n_factors)]
        self.col_sets = [rng.choice(base.n, kc, replace=False) for _ in range(self.n_factors)]

    def thetas(self, count: int, seed: int, sigma: float = 0.1) -> np.ndarray:
        """``count`` scenarios, theta_k ~ N(0, sigma^2) i.i.d."""
        return np.random.default_rng(seed).normal(0.0, sigma, size=(int(count), self.n_factors))

    def vectors(self, theta):
        """(cost, col_lo, col_up, row_lo, row_up) of scenario ``theta``."""
        theta = np.asarray(theta, dtype=np.float64)
        if theta.shape != (self.n_factors,):
            raise ValueError(f"theta must have {self.n_factors} entries")
        rs, cs = np.ones(self.base.m), np.ones(self.base.n)
        for k in range(self.n_factors):
            e = np.exp(theta[k])
            rs[self.row_sets[k]] *= e
            cs[self.col_sets[k]] *= e
        b = self.base
        return b.cost * cs, b.col_lo.copy(), b.col_up.copy(), b.row_lo * rs, b.row_up * rs

    def model(self, theta, name: str = "SCENARIO") -> LPModel:
        """The scenario as a stand-alone LPModel (same matrix arrays)."""
        c, l, u, rl, ru = self.vectors(theta)
        b = self.base
        return LPModel(m=b.m, n=b.n, colptr=b.colptr, rowidx=b.rowidx, val=b.val, cost=c, col_lo=l, col_up=u,
                       row_lo=rl, row_up=ru, name=name)


def _planted_from_coo(m, n, rows, cols, vals, rng, lo=0.0, up=10.0, name="LP") -> LPModel:
    """The planted construction (x*: half zeros, half U(0,1); b = A x*; y* ~ N(0,1); s_j ~ U(0,1) where x*_j = 0;
    c = A^T y* + s; equality rows; lo <= x <= up) on a matrix given as COO entries."""
    order = np.lexsort((rows, cols))
    rows, cols, vals = rows[order], cols[order], vals[order]
    colptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(cols, minlength=n), out=colptr[1:])
    xstar = np.zeros(n)
    nz = rng.permutation(n)[n // 2:]
    xstar[nz] = rng.random(nz.size)
    b = np.bincount(rows, weights=vals * xstar[cols], minlength=m)
    ystar = rng.standard_normal(m)
    s = np.where(xstar == 0, rng.random(n), 0.0)
    c = np.bincount(cols, weights=vals * ystar[rows], minlength=n) + s
    return LPModel(m=m, n=n, colptr=colptr, rowidx=rows.astype(np.int32), val=vals, cost=c,
                   col_lo=np.full(n, lo), col_up=np.full(n, up), row_lo=b.copy(), row_up=b.copy(), name=name,
                   planted_objective=float(np.dot(c, xstar)), x_star=xstar, y_star=ystar)


def block_angular_lp(blocks: int = 64, rows_per_block: int = 5000, cols_per_block: int = 8000,
                     nnz_per_col: int = 8, linking_rows: int = 200, link_frac: float = 0.02,
                     seed: int = 2) -> LPModel:
    """Block-angular planning LP: ``blocks`` diagonal blocks (rows_per_block x cols_per_block, ``nnz_per_col``
    N(0,1) entries per column at distinct rows of its block) plus ``linking_rows`` rows that each touch a random
    ``link_frac`` of the columns of every block with coefficients +-1; right-hand side and costs planted as in
    planted_lp. Rows: blocks * rows_per_block block rows, then the linking rows."""
    rng = np.random.default_rng(seed)
    n = blocks * cols_per_block
    m = blocks * rows_per_block + linking_rows
    local = _distinct_rows(rng, rows_per_block, n, nnz_per_col)              # per column, rows within its block
    block_of = np.arange(n) // cols_per_block
    r_blk = (local + (block_of * rows_per_block)[:, None]).reshape(-1)
    c_blk = np.repeat(np.arange(n), nnz_per_col)
    v_blk = rng.standard_normal(r_blk.size)
    per_block = max(1, int(round(link_frac * cols_per_block)))
    r_l, c_l = [], []
    for k in range(linking_rows):
        for p in range(blocks):
            sel = rng.choice(cols_per_block, per_block, replace=False) + p * cols_per_block
            c_l.append(sel)
            r_l.append(np.full(per_block, blocks * rows_per_block + k))
    r_l, c_l = np.concatenate(r_l), np.concatenate(c_l)
    v_l = np.where(rng.random(r_l.size) < 0.5, -1.0, 1.0)
    rows = np.concatenate([r_blk, r_l]).astype(np.int64)
    cols = np.concatenate([c_blk, c_l]).astype(np.int64)
    vals = np.concatenate([v_blk, v_l])
    return _planted_from_coo(m, n, rows, cols, vals, rng, name=f"BLOCKANGULAR_{blocks}x{rows_per_block}x{cols_per_block}")


def uneven_batch(count: int = 512, m_lo: int = 500, m_hi: int = 20000, ratio=(1.5, 3.0), nnz=(6, 12),
                 seed: int = 3) -> list:
    """``count`` planted LPs of uneven size: m ~ U[m_lo, m_hi], n = round(m * U[ratio]), one nonzeros-per-column
    count per LP drawn from U{nnz[0]..nnz[1]}; LP k uses seed (seed * 100003 + k)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        m = int(rng.integers(m_lo, m_hi + 1))
        n = int(round(m * rng.uniform(*ratio)))
        kk = int(rng.integers(nnz[0], nnz[1] + 1))
        out.append(planted_lp(m, n, kk, seed=seed * 100003 + k, name=f"BATCH{k}_{m}x{n}x{kk}"))
    return out
