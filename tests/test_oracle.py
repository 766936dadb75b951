"""Pins the CPU oracle (oracle/pdhg_oracle.c) before anything is compared with it.

* Driver trajectory: the reference's published FP64 ledgers (tests/golden/ledger_pdhg.json) on the
  gpu/gen_lp.py instances, regenerated bit for bit: iteration and restart counts exactly, objective and the
  relative residuals to the digits the ledgers print.
* End-to-end optimum: the planted closed form c'x* = b'y* of the 2,000 x 5,000 planted LP, and the
  reference's FP64 dual simplex optimum for it (tests/golden/cfg0_dual_simplex.json).
* Pieces: scaling is the exact identity A_s = diag(Dr) A diag(Dc) with the Ruiz/Pock-Chambolle invariants;
  SpMV against dense products; KKT scalars against brute-force sums written from their definitions; the
  power iteration against numpy's SVD.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle
from taral_pdhg import gen

LEDGER = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ledger_pdhg.json")))


def _instance(case):
    if case["family"] == "packing":
        return gen.packing_mps(*case["args"])[0]
    return gen.transport_mps(*case["args"])[0]


FAST = [c for c in LEDGER["cases"] if c["case"] in ("packing_10000", "transport_100x100")]


@pytest.mark.parametrize("case", FAST, ids=lambda c: f"{c['case']}-{c['tol']:g}")
def test_oracle_reproduces_reference_ledger(case):
    md = _instance(case)
    assert (md.m, md.n, md.nnz) == (case["rows"], case["cols"], case["nnz"])
    r = oracle.OracleLP(md).solve(tol=case["tol"])
    assert r["status"] == "near_optimal"
    assert (r["iterations"], r["restarts"]) == (case["iterations"], case["restarts"])
    assert float("%.12g" % r["objective"]) == case["objective"]
    assert "%.6e" % r["rel_primal"] == "%.6e" % case["rel_primal"]
    assert "%.6e" % r["rel_gap"] == "%.6e" % case["rel_gap"]
    assert float("%.6e" % r["max_row_violation"]) == pytest.approx(case["max_row_violation"], rel=1e-6)


def test_oracle_packing_100000_ledger():
    case = next(c for c in LEDGER["cases"] if c["case"] == "packing_100000" and c["tol"] == 1e-4)
    md = _instance(case)
    assert (md.m, md.n, md.nnz) == (case["rows"], case["cols"], case["nnz"])
    r = oracle.OracleLP(md).solve(tol=case["tol"])
    assert (r["iterations"], r["restarts"]) == (case["iterations"], case["restarts"])
    assert float("%.12g" % r["objective"]) == case["objective"]


def test_oracle_transport_316x316_ledger():
    """synth_transport_tol1e-4.csv:7: 31,616 iterations, 21 restarts (about 20 s on one core)."""
    case = next(c for c in LEDGER["cases"] if c["case"] == "transport_316x316")
    md = _instance(case)
    assert (md.m, md.n, md.nnz) == (case["rows"], case["cols"], case["nnz"])
    r = oracle.OracleLP(md).solve(tol=case["tol"])
    assert (r["iterations"], r["restarts"]) == (case["iterations"], case["restarts"])
    assert float("%.12g" % r["objective"]) == case["objective"]


def _row_rel_norm(md, x):
    """||row violation_i / (1 + |b_i|)||_2 on the original model, b_i the larger finite |row bound|
    (gpu/pdhg.cu:137-139, 748, 793)."""
    ax = md.matvec(x)
    viol = np.maximum(md.row_lo - ax, 0) + np.maximum(ax - md.row_up, 0)
    lo, up = md.row_lo, md.row_up
    b = np.where(np.isfinite(lo) & np.isfinite(up), np.maximum(abs(lo), abs(up)),
                 np.where(np.isfinite(lo), abs(lo), np.where(np.isfinite(up), abs(up), 0.0)))
    return float(np.linalg.norm(viol / (1 + b)))


@pytest.mark.parametrize("case", ["packing_10000", "transport_100x100"])
def test_oracle_row_rel_stop_test(case):
    """--row-rel K adds the per-row test ||viol_i / (1 + |b_i|)|| <= K tol: off, the returned point breaks it at
    K = 1 (so the test binds on these LPs); on, the returned point meets it."""
    md = gen.packing_mps(10000)[0] if case.startswith("packing") else gen.transport_mps(100, 100)[0]
    O = oracle.OracleLP(md)
    off, on = O.solve(tol=1e-4), O.solve(tol=1e-4, row_rel=1.0)
    assert _row_rel_norm(md, off["x"] * O.vec("Dc")) > 1e-4
    assert _row_rel_norm(md, on["x"] * O.vec("Dc")) <= 1e-4
    assert on["status"] == "near_optimal" and on["iterations"] > off["iterations"]


def test_planted_optimum_and_dual_simplex(cfg0, golden_dir):
    g = json.load(open(os.path.join(golden_dir, "cfg0_dual_simplex.json")))
    p_star = cfg0.planted_objective
    # the closed form itself: c'x* = b'y* (x* feasible, s'x* = 0)
    assert np.dot(cfg0.row_lo, cfg0.y_star) == pytest.approx(p_star, rel=1e-12)
    assert g["objective"] == pytest.approx(p_star, rel=1e-9)
    r = oracle.OracleLP(cfg0).solve(tol=1e-6)
    assert r["status"] == "near_optimal"
    assert abs(r["objective"] - p_star) / (1 + abs(p_star)) < 2e-6


def _tiny(seed=3, m=7, n=11, density=0.4):
    rng = np.random.default_rng(seed)
    A = np.where(rng.random((m, n)) < density, rng.standard_normal((m, n)) * 10 ** rng.uniform(-2, 2, (m, n)), 0.0)
    A[0, 0] = 0.5 if A[0, 0] == 0 else A[0, 0]
    cols = [np.nonzero(A[:, j])[0] for j in range(n)]
    colptr = np.concatenate([[0], np.cumsum([len(c) for c in cols])]).astype(np.int64)
    rowidx = np.concatenate(cols).astype(np.int32)
    val = np.concatenate([A[c, j] for j, c in enumerate(cols)])
    lo = rng.uniform(-2, 0, m)
    up = lo + rng.uniform(0, 3, m)
    up[1], lo[2] = np.inf, -np.inf
    md = gen.LPModel(m=m, n=n, colptr=colptr, rowidx=rowidx, val=val, cost=rng.standard_normal(n),
                     col_lo=np.where(rng.random(n) < 0.2, -np.inf, -1.0), col_up=np.where(rng.random(n) < 0.2, np.inf, 2.0),
                     row_lo=lo, row_up=up)
    return md, A


def test_scaling_identity_and_invariants():
    md, A = _tiny()
    o = oracle.OracleLP(md)
    ptr, idx, val = o.csr("A")
    As = np.zeros_like(A)
    for i in range(md.m):
        As[i, idx[ptr[i]:ptr[i + 1]]] = val[ptr[i]:ptr[i + 1]]
    Dr, Dc = o.vec("Dr"), o.vec("Dc")
    np.testing.assert_allclose(As, Dr[:, None] * A * Dc[None, :], rtol=1e-13, atol=0)
    # after Ruiz + Pock-Chambolle every nonempty row and column has l1 norm close to 1 in the scaled matrix
    rn, cn = np.abs(As).sum(1), np.abs(As).sum(0)
    assert np.all((rn[rn > 0] > 0.2) & (rn[rn > 0] < 5)) and np.all((cn[cn > 0] > 0.2) & (cn[cn > 0] < 5))
    tp, ti, tv = o.csr("AT")
    AsT = np.zeros((md.n, md.m))
    for j in range(md.n):
        AsT[j, ti[tp[j]:tp[j + 1]]] = tv[tp[j]:tp[j + 1]]
    assert np.array_equal(AsT, As.T)
    np.testing.assert_array_equal(o.vec("c"), md.cost * Dc)
    np.testing.assert_array_equal(o.vec("rl"), md.row_lo * Dr)
    assert o.S.bnorm == pytest.approx(np.sqrt(sum(max(abs(a) if np.isfinite(a) else 0, abs(b) if np.isfinite(b) else 0) ** 2
                                                  for a, b in zip(md.row_lo, md.row_up))))


def test_spmv_against_dense():
    md, A = _tiny(seed=5, m=40, n=60, density=0.2)
    o = oracle.OracleLP(md)
    ptr, idx, val = o.csr("A")
    v = np.random.default_rng(1).standard_normal(md.n)
    y = np.zeros(md.m)
    import ctypes as C
    oracle.lib().or_spmv(md.m, ptr.ctypes.data_as(C.POINTER(C.c_int32)), idx.ctypes.data_as(C.POINTER(C.c_int32)),
                         val.ctypes.data_as(C.POINTER(C.c_double)), v.ctypes.data_as(C.POINTER(C.c_double)),
                         y.ctypes.data_as(C.POINTER(C.c_double)))
    Dr, Dc = o.vec("Dr"), o.vec("Dc")
    np.testing.assert_allclose(y, (Dr[:, None] * A * Dc[None, :]) @ v, rtol=1e-12, atol=1e-14)


def _kkt_by_definition(o, Ax, z, x, ATz, xlast, zlast, f):
    """The ten KKT scalars written out from their definitions (gpu/pdhg.cu:236-271)."""
    rl, ru, Dr, c, l, u, Dc = (o.vec(k) for k in ("rl", "ru", "Dr", "c", "l", "u", "Dc"))
    ax, zz = Ax * f, z * f
    viol = np.maximum(rl - ax, 0) + np.maximum(ax - ru, 0)
    b = np.array([max(abs(a) if np.isfinite(a) else 0, abs(q) if np.isfinite(q) else 0) for a, q in zip(rl, ru)])
    rel = (viol / Dr) / (1 + b / Dr)
    drow = -np.sum(np.where((zz > 0) & np.isfinite(ru), zz * np.where(np.isfinite(ru), ru, 0), 0)) \
        - np.sum(np.where((zz < 0) & np.isfinite(rl), zz * np.where(np.isfinite(rl), rl, 0), 0))
    xx = x * f
    d = c + ATz * f
    lo_ok, up_ok = np.isfinite(l), np.isfinite(u)
    dcol = np.sum(np.where((d > 0) & lo_ok, d * np.where(lo_ok, l, 0), 0)) + \
        np.sum(np.where((d < 0) & up_ok, d * np.where(up_ok, u, 0), 0))
    res = np.where((d > 0) & ~lo_ok, d, 0) + np.where((d < 0) & ~up_ok, -d, 0)
    return np.array([np.sum(viol ** 2), np.sum((viol / Dr) ** 2), np.sum(rel ** 2), drow, np.sum((zz - zlast) ** 2),
                     np.sum(res ** 2), np.sum((res / Dc) ** 2), np.dot(c, xx), dcol, np.sum((xx - xlast) ** 2)])


def test_eval_against_definitions():
    md, A = _tiny(seed=7, m=30, n=45, density=0.25)
    o = oracle.OracleLP(md)
    rng = np.random.default_rng(2)
    vecs = [rng.standard_normal(k) * 3 for k in (md.m, md.m, md.n, md.n, md.n, md.m)]
    for f in (1.0, 0.25):
        got = o.eval(*vecs, f=f)
        want = _kkt_by_definition(o, *vecs, f)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_power_norm_against_svd():
    md, A = _tiny(seed=9, m=25, n=35, density=0.3)
    o = oracle.OracleLP(md)
    As = o.vec("Dr")[:, None] * A * o.vec("Dc")[None, :]
    smax = np.linalg.svd(As, compute_uv=False)[0]
    assert o.power_norm(iters=2000) == pytest.approx(smax, rel=1e-5)


@pytest.mark.parametrize("fp32", [False, True])
def test_threaded_block_equals_single_thread(fp32):
    """The CPU baseline's threaded iteration block (the reference Pool's partition, a barrier per phase) against
    the single-thread block: no phase reduces across rows, so every state vector must agree bit for bit; a wrong
    partition (rows skipped or done twice) or a missing barrier (a phase reading the other's half-written
    vector) changes them. Thread counts include uneven splits and more threads than rows."""
    md = gen.planted_lp(300, 700, 6, seed=5)
    o = oracle.OracleLP(md)
    want = o.block(0.4, 0.6, 40, fp32=fp32)
    for threads in (2, 3, 8):
        got = o.block(0.4, 0.6, 40, fp32=fp32, threads=threads)
        for k in want:
            np.testing.assert_array_equal(got[k], want[k], err_msg=f"{k} with {threads} threads")
    tiny, _ = _tiny(seed=3, m=5, n=9, density=0.5)
    ot = oracle.OracleLP(tiny)
    x0 = np.linspace(0.0, 1.0, tiny.n)
    want = ot.block(0.3, 0.5, 12, x0=x0, fp32=fp32)
    got = ot.block(0.3, 0.5, 12, x0=x0, fp32=fp32, threads=16)
    for k in want:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)


@pytest.mark.parametrize("shape", [(2_000, 5_000, 10), (65_000, 130_000, 10)])
def test_library_scaling_equals_oracle(shape):
    """The library's host preparation (tpdhg_scaled_build: transposes, Ruiz x10, Pock-Chambolle, vectors) against
    the oracle's restatement of gpu/pdhg.cu:49-133, bit for bit. The larger LP has over 2^20 nonzeros, where the
    library's counting sorts and row/column measures run in chunks on several threads."""
    from taral_pdhg import ffi
    m, n, k = shape
    md = gen.planted_lp(m, n, k, seed=3)
    S = ffi.Scaled(ffi.Model.from_lp(md))
    o = oracle.OracleLP(md)
    for which in ("A", "AT"):
        for got, want in zip(S.csr(which), o.csr(which)):
            np.testing.assert_array_equal(np.asarray(got), np.asarray(want), err_msg=which)
    for name in ("Dr", "Dc", "c", "l", "u", "rl", "ru"):
        np.testing.assert_array_equal(S.vec(name), o.vec(name), err_msg=name)
