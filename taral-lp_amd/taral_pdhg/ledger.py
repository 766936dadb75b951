"""Benchmark ledger in the reference's CSV schema (gpu/bench_pdhg.py:23-27), so new results line up with the
reference's results/gpu_pdhg/*.csv rows.

    python -m taral_pdhg.ledger CASE [CASE ...] --out ledger.csv [--tol 1e-4] [--repeats 3] [--fast]

A CASE is an MPS file, or `packing:N` / `transport:SxD` for the reference's generated cases (gpu/gen_lp.py, restated
in gen.py), built in memory under the names its ledgers use (results/gpu_pdhg/synth_*.csv: packing_N,
transport_SxD); such a case has no MPS parse in its total_s. Each case is read once, solved ``--repeats`` times on the GPU (a fresh tpdhg_solve each time, like the reference
harness's fresh process per run), rechecked on the original model by the library's C++ recheck (the chk_* columns,
gpu/vector_check.py's quantities) and compared with HiGHS. The reference harness calls highspy, which this image
does not have; the HiGHS that scipy bundles (scipy.optimize.linprog, method "highs") gives highs_objective and
highs_s instead. A run whose status is not near_optimal is not repeated (gpu/bench_pdhg.py:141-143).
"""
from __future__ import annotations

import argparse
import csv
import os
import statistics
import time

import numpy as np

from . import ffi, gen

FIELDS = ['case', 'rows', 'cols', 'nnz', 'device', 'threads', 'precision', 'tol', 'status', 'objective',
          'highs_objective', 'objective_rel_error_vs_highs', 'rel_primal', 'rel_dual', 'rel_gap', 'max_row_violation',
          'iterations', 'restarts', 'setup_s', 'solve_s', 'total_s', 'total_s_min', 'total_s_max', 'repeats',
          'device_mem_bytes', 'highs_s', 'load_avg_1m', 'chk_objective', 'chk_row_viol_rel', 'chk_bound_viol',
          'chk_dual_bound', 'chk_dual_obj_loose', 'chk_dual_infeas_rel', 'chk_dual_gap_rel', 'chk_dual_sign',
          'chk_row_viol_rel_max_repeats', 'objective_identical_repeats', 'chk_error']


def rel_err(obj, ref):
    """|obj - ref| / max(1, |ref|) (gpu/bench_pdhg.py:112-115)."""
    if obj is None or ref is None:
        return None
    return abs(obj - ref) / max(1.0, abs(ref))


def highs_objective(model: ffi.Model):
    """(objective, seconds) of the model by scipy's HiGHS, or (None, seconds) when it does not solve."""
    from scipy.optimize import linprog
    from scipy.sparse import csc_matrix
    info = model.info()
    a = model.arrays()
    m, n = info.m, info.n
    A = csc_matrix((a["val"], a["rowidx"], a["colptr"]), shape=(m, n)).tocsr()
    sign = -1.0 if info.maximize else 1.0
    eq = a["row_lo"] == a["row_up"]
    up = ~eq & np.isfinite(a["row_up"])
    lo = ~eq & np.isfinite(a["row_lo"])
    A_ub = None
    b_ub = None
    if up.any() or lo.any():
        from scipy.sparse import vstack
        A_ub = vstack([A[up], -A[lo]]).tocsr()
        b_ub = np.concatenate([a["row_up"][up], -a["row_lo"][lo]])
    bounds = [(None if not np.isfinite(l_) else l_, None if not np.isfinite(u_) else u_)
              for l_, u_ in zip(a["col_lo"], a["col_up"])]
    t = time.perf_counter()
    res = linprog(sign * a["cost"], A_ub=A_ub, b_ub=b_ub, A_eq=A[eq] if eq.any() else None,
                  b_eq=a["row_lo"][eq] if eq.any() else None, bounds=bounds, method="highs")
    wall = time.perf_counter() - t
    if res.status != 0:
        return None, wall
    return sign * res.fun + info.obj_const, wall


def load_case(spec: str):
    """(name, Model) of a CASE argument: an MPS file or a generated packing / transport case."""
    kind, _, arg = spec.partition(":")
    if kind == "packing" and arg:
        return f"packing_{int(arg)}", ffi.Model.from_lp(gen.packing_mps(int(arg), text=False)[0])
    if kind == "transport" and arg:
        s, d = (int(v) for v in arg.split("x"))
        return f"transport_{s}x{d}", ffi.Model.from_lp(gen.transport_mps(s, d)[0])
    return os.path.splitext(os.path.basename(spec))[0], ffi.Model.from_mps(spec)


def case_row(path: str, tol: float, repeats: int, fast: bool, time_limit: float, highs: bool = True) -> dict:
    name, model = load_case(path)
    info = model.info()
    opts = dict(tol=tol, time_limit=time_limit)
    if fast:
        opts.update(precision_bits=32, halpern=1, refine=1)
    runs = [ffi.solve(model, **opts)]
    if runs[0]["status"] == "near_optimal":
        runs += [ffi.solve(model, **opts) for _ in range(repeats - 1)]
    chks = [ffi.recheck(model, r["x"], r["y"]) for r in runs]
    r0, c0 = runs[0], chks[0]
    ref, highs_s = highs_objective(model) if highs else (None, None)
    med = lambda k: statistics.median(r[k] for r in runs)
    return dict(
        case=name, rows=info.m, cols=info.n, nnz=r0["nnz"], device="gpu",
        threads=0, precision="fp32" if fast else "fp64", tol=tol, status=r0["status"], objective=r0["objective"],
        highs_objective=ref, objective_rel_error_vs_highs=rel_err(r0["objective"], ref),
        rel_primal=r0["rel_primal"], rel_dual=r0["rel_dual"], rel_gap=r0["rel_gap"],
        max_row_violation=r0["max_row_violation"], iterations=r0["iterations"], restarts=r0["restarts"],
        setup_s=med("setup_s"), solve_s=med("solve_s"), total_s=med("total_s"),
        total_s_min=min(r["total_s"] for r in runs), total_s_max=max(r["total_s"] for r in runs), repeats=len(runs),
        device_mem_bytes=r0["device_mem_bytes"], highs_s=highs_s, load_avg_1m=os.getloadavg()[0],
        chk_objective=c0["objective"], chk_row_viol_rel=c0["row_viol_rel"], chk_bound_viol=c0["bound_viol"],
        chk_dual_bound=c0["dual_bound"], chk_dual_obj_loose=c0["dual_obj_loose"],
        chk_dual_infeas_rel=c0["dual_infeas_rel"], chk_dual_gap_rel=c0["dual_gap_rel"],
        chk_dual_sign=c0["dual_sign"], chk_row_viol_rel_max_repeats=max(c["row_viol_rel"] for c in chks),
        objective_identical_repeats=len({r["objective"] for r in runs}) <= 1, chk_error=None)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("cases", nargs="+")
    ap.add_argument("--out", required=True)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--time-limit", type=float, default=3600.0)
    ap.add_argument("--fast", action="store_true", help="FP32 Halpern + FP64 refinement instead of the reference scheme")
    ap.add_argument("--no-highs", action="store_true")
    a = ap.parse_args(argv)
    rows = []
    for path in a.cases:
        row = case_row(path, a.tol, a.repeats, a.fast, a.time_limit, highs=not a.no_highs)
        rows.append(row)
        e = row["objective_rel_error_vs_highs"]
        print(f"{row['case']:16s} {row['status']:14s} it {row['iterations']} relerr "
              f"{'n/a' if e is None else f'{e:.2e}'} total {row['total_s']:.3f} solve {row['solve_s']:.3f}", flush=True)
    with open(a.out, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for row in rows:
            w.writerow({k: ("" if row.get(k) is None else row[k]) for k in FIELDS})


if __name__ == "__main__":
    main()
