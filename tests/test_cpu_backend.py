"""The CPU backend (the reference's `--device cpu`, gpu/pdhg.cu:517-741): the reference scheme on host threads.

Pinned the way the GPU backend is: the reference's published FP64 ledger trajectories (iterations, restarts and
the 12-digit objective; the reference's GPU and CPU backends agree on them), the oracle's trajectory and the
planted optimum on the 2,000 x 5,000 LP, results that do not depend on the thread count (rows are summed serially
by their owning thread, reductions over fixed chunks), and the CLI's --device cpu output. Runs without a GPU.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle
from taral_pdhg import ffi, gen

LEDGER = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ledger_pdhg.json")))
CASES = [c for c in LEDGER["cases"] if c["case"] in ("packing_10000", "transport_100x100")
         or (c["case"] == "packing_100000" and c["tol"] == 1e-4)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['case']}-{c['tol']:g}")
def test_cpu_backend_ledger_trajectory(case):
    md = gen.packing_mps(*case["args"], text=False)[0] if case["family"] == "packing" \
        else gen.transport_mps(*case["args"])[0]
    r = ffi.solve(ffi.Model.from_lp(md), tol=case["tol"], precision_bits=64, cpu_threads=4)
    assert r["status"] == "near_optimal"
    assert (r["iterations"], r["restarts"]) == (case["iterations"], case["restarts"])
    assert r["objective"] == pytest.approx(case["objective"], rel=1e-9)
    j = json.loads(r["json"])
    assert j["backend"] == "cpu" and j["threads"] == 4 and j["device"] == "cpu x4 threads"


def test_cpu_backend_matches_oracle_and_planted(cfg0, golden_dir):
    r = ffi.solve(ffi.Model.from_lp(cfg0), tol=1e-6, precision_bits=64, cpu_threads=2)
    o = oracle.OracleLP(cfg0).solve(tol=1e-6)
    assert r["status"] == "near_optimal" == o["status"]
    assert (r["iterations"], r["restarts"]) == (o["iterations"], o["restarts"])
    assert r["objective"] == pytest.approx(o["objective"], rel=1e-9)
    ds = json.load(open(os.path.join(golden_dir, "cfg0_dual_simplex.json")))["objective"]
    for ref in (cfg0.planted_objective, ds):
        assert abs(r["objective"] - ref) / (1 + abs(ref)) <= 3e-6


@pytest.mark.parametrize("bits", [64, 32])
def test_cpu_backend_does_not_depend_on_thread_count(cfg0, bits):
    runs = [ffi.solve(ffi.Model.from_lp(cfg0), tol=1e-4, precision_bits=bits, cpu_threads=t) for t in (1, 3, 8)]
    for r in runs[1:]:
        assert (r["iterations"], r["restarts"]) == (runs[0]["iterations"], runs[0]["restarts"])
        assert np.array_equal(r["x"], runs[0]["x"]) and np.array_equal(r["y"], runs[0]["y"])
    if bits == 32:  # the reference's --fp32: FP32 iterates reach 1e-4 near the planted optimum
        p = cfg0.planted_objective
        assert runs[0]["status"] == "near_optimal" and abs(runs[0]["objective"] - p) / (1 + abs(p)) <= 3e-4


def test_cli_device_cpu(transport_100x100, tmp_path):
    mps = tmp_path / "t.mps"
    mps.write_text(transport_100x100[1])
    js, sol = tmp_path / "t.json", tmp_path / "t.sol"
    r = subprocess.run([ffi.CLI_PATH, str(mps), "--device", "cpu", "--threads", "2", "--tol", "1e-4", "--json",
                        str(js), "--sol", str(sol)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = json.load(open(js))
    assert out["backend"] == "cpu" and out["threads"] == 2 and out["status"] == "near_optimal"
    assert (out["iterations"], out["restarts"]) == (5440, 12)  # synth_transport_tol1e-4.csv:2
    assert "device cpu x2 threads" in r.stdout and len(sol.read_text().splitlines()) == 10000
    # automatic thread count: one thread below 200,000 nonzeros (gpu/pdhg.cu:986-989)
    r = subprocess.run([ffi.CLI_PATH, str(mps), "--device", "cpu", "--tol", "1e-4", "--json", str(js)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and json.load(open(js))["threads"] == 1


LITERATURE = [("ref_steel_production.mps", 192000.0), ("ref_transportation.mps", 153.675)]


@pytest.mark.parametrize("name,optimum", LITERATURE)
def test_cpu_backend_literature_lps(golden_dir, name, optimum):
    """examples/literature_lp/*.mps, optima from results/literature_lp_20261003/results.json (the reference's exact
    dual simplex, HiGHS agreeing)."""
    md = ffi.Model.from_mps(os.path.join(golden_dir, "mps", name))
    r = ffi.solve(md, tol=1e-8, precision_bits=64, cpu_threads=1)
    assert r["status"] == "near_optimal"
    assert r["objective"] == pytest.approx(optimum, rel=1e-6)
