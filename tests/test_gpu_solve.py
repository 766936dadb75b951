"""Full solves on the GPU against the reference's ledgers, the oracle and the planted optimum.

FP64 (the reference's default): iteration and restart counts equal to the reference's published ledgers
(identical on its GPU and CPU backends), objective to 1e-9 relative of the ledger's 12 printed digits.
The planted 2,000 x 5,000 LP: same counts as the oracle, objective within the requested tolerance of the
planted optimum c'x* and of the reference dual simplex's optimum. FP32 (the reference's --fp32): reaches
1e-4 with an objective within 1e-4 of the planted optimum.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle
from taral_pdhg import ffi, gen

pytestmark = pytest.mark.gpu
LEDGER = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ledger_pdhg.json")))
@pytest.mark.parametrize("case", LEDGER["cases"], ids=lambda c: f"{c['case']}-{c['tol']:g}")
def test_ledger_trajectory_fp64(case):
    md = gen.packing_mps(*case["args"])[0] if case["family"] == "packing" else gen.transport_mps(*case["args"])[0]
    print(f"solve {case['case']} tol={case['tol']}", flush=True)
    r = ffi.solve(ffi.Model.from_lp(md), tol=case["tol"], precision_bits=64)
    assert r["status"] == "near_optimal"
    assert (r["iterations"], r["restarts"]) == (case["iterations"], case["restarts"])
    assert r["objective"] == pytest.approx(case["objective"], rel=1e-9)
    assert r["rel_primal"] == pytest.approx(case["rel_primal"], rel=1e-4)
    assert (r["m"], r["n"], r["nnz"]) == (case["rows"], case["cols"], case["nnz"])


@pytest.mark.parametrize("case", ["packing_10000", "transport_100x100"])
def test_row_rel_trajectory_matches_oracle(case):
    """--row-rel 1 (gpu/pdhg.cu:748, 793): the per-row relative primal test binds on these LPs (the oracle test
    shows it), so the GPU's FP64 trajectory must follow the oracle's longer one exactly."""
    md = gen.packing_mps(10000)[0] if case.startswith("packing") else gen.transport_mps(100, 100)[0]
    print(f"row_rel {case}", flush=True)
    r = ffi.solve(ffi.Model.from_lp(md), tol=1e-4, row_rel=1.0, precision_bits=64)
    o = oracle.OracleLP(md).solve(tol=1e-4, row_rel=1.0)
    assert r["status"] == "near_optimal" == o["status"]
    assert (r["iterations"], r["restarts"]) == (o["iterations"], o["restarts"])
    assert r["objective"] == pytest.approx(o["objective"], rel=1e-9)


@pytest.mark.parametrize("tol", [1e-4, 1e-6])
def test_planted_fp64_matches_oracle(cfg0, golden_dir, tol):
    print(f"cfg0 fp64 tol={tol}", flush=True)
    r = ffi.solve(ffi.Model.from_lp(cfg0), tol=tol, precision_bits=64)
    o = oracle.OracleLP(cfg0).solve(tol=tol)
    assert r["status"] == "near_optimal"
    assert (r["iterations"], r["restarts"]) == (o["iterations"], o["restarts"])
    assert r["objective"] == pytest.approx(o["objective"], rel=1e-9)
    p_star = cfg0.planted_objective
    ds = json.load(open(os.path.join(golden_dir, "cfg0_dual_simplex.json")))["objective"]
    for ref in (p_star, ds):
        assert abs(r["objective"] - ref) / (1 + abs(ref)) <= 3 * tol
    # the returned point is feasible to the reported residual on the original model
    ax = cfg0.matvec(r["x"])
    assert np.linalg.norm(ax - cfg0.row_lo) / (1 + np.linalg.norm(cfg0.row_lo)) <= 1.01 * tol
    assert np.all(r["x"] >= -1e-12) and np.all(r["x"] <= 10 + 1e-12)


def test_planted_fp32(cfg0):
    """The reference's --fp32 (FP32 iterates and products, FP64 reductions) against the oracle's FP32 mode: the
    summation orders differ (row blocks here, CSR order there), so the iterates differ in their last FP32 bits, but
    both are deterministic and the restart decisions are not near ties on this LP: the same iterations and
    restarts, objectives within the tolerance of each other."""
    print("cfg0 fp32 tol=1e-4", flush=True)
    r = ffi.solve(ffi.Model.from_lp(cfg0), tol=1e-4, precision_bits=32)
    o = oracle.OracleLP(cfg0).solve(tol=1e-4, fp32=True)
    print(f"  gpu it={r['iterations']} restarts={r['restarts']} obj={r['objective']:.10g}; oracle it={o['iterations']} "
          f"restarts={o['restarts']} obj={o['objective']:.10g}", flush=True)
    assert r["status"] == "near_optimal" == o["status"]
    assert (r["iterations"], r["restarts"]) == (o["iterations"], o["restarts"])
    assert abs(r["objective"] - o["objective"]) / (1 + abs(o["objective"])) <= 1e-4
    assert abs(r["objective"] - cfg0.planted_objective) / (1 + abs(cfg0.planted_objective)) <= 3e-4


def test_warm_start_at_optimum_stops_at_first_evaluation(cfg0):
    r = ffi.solve(ffi.Model.from_lp(cfg0), x0=cfg0.x_star, y0=-cfg0.y_star, tol=1e-6, precision_bits=64)
    assert r["status"] == "near_optimal" and r["iterations"] == 64


def test_iteration_limit_and_json(cfg0):
    r = ffi.solve(ffi.Model.from_lp(cfg0), tol=1e-9, max_iter=100, precision_bits=64)
    assert r["status"] == "iteration_limit" and r["iterations"] == 100
    j = json.loads(r["json"])
    for k in ("solver", "status", "objective", "rel_primal", "rel_dual", "rel_gap", "tol", "iterations", "restarts",
              "point", "max_row_violation", "max_bound_violation", "setup_s", "parse_s", "scale_s", "device_init_s",
              "transfer_s", "solve_s", "power_iter_s", "total_s", "device", "backend", "threads", "precision", "rows",
              "cols", "nnz", "norm_a", "primal_weight", "eval_freq", "graph", "device_mem_bytes"):
        assert k in j, k
    assert j["status"] == "iteration_limit" and j["rows"] == 2000 and j["backend"] == "gpu"


@pytest.mark.parametrize("fast", [False, True])
@pytest.mark.parametrize("name,optimum", [("ref_steel_production.mps", 192000.0), ("ref_transportation.mps", 153.675)])
def test_literature_lps(golden_dir, name, optimum, fast):
    """examples/literature_lp/*.mps to the optima of results/literature_lp_20261003/results.json, on the reference
    scheme (FP64) and on the fast path (FP32 iterates + FP64 refinement)."""
    md = ffi.Model.from_mps(os.path.join(golden_dir, "mps", name))
    opts = dict(precision_bits=32, halpern=1, refine=1) if fast else dict(precision_bits=64)
    r = ffi.solve(md, tol=1e-8, **opts)
    assert r["status"] == "near_optimal"
    assert r["objective"] == pytest.approx(optimum, rel=1e-6)
