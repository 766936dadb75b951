"""The `.sol` / `.dual` wire format into the reference's crossover (src/main.cpp:661-683, src/crossover.cpp:15-60).

tests/golden/make_crossover_fixture.py (container only) fed the library's answer files for two LPs to the reference
engine `taral --method dual --warm-sol --warm-dual` and recorded the outcome in tests/golden/crossover/. Here (CPU
only): the build still writes those answer files bit for bit (CPU backend, deterministic), the reference read them
and finished optimal from the basis they identify in far fewer dual simplex iterations than from a cold start, and
its optimum agrees with scipy's HiGHS and with the PDHG objective to the PDHG tolerance.
"""
import json
import os
import subprocess

import numpy as np
import pytest
from scipy.optimize import linprog

from taral_pdhg import ffi, gen

HERE = os.path.join(os.path.dirname(__file__), "golden", "crossover")
OUTCOME = json.load(open(os.path.join(HERE, "outcome.json")))
MODELS = {"transport_20x20": lambda: gen.transport_mps(20, 20), "packing_2000": lambda: gen.packing_mps(2000)}


@pytest.mark.parametrize("case", OUTCOME["cases"], ids=lambda c: c["case"])
def test_answer_files_and_reference_crossover(case, tmp_path):
    md, text = MODELS[case["case"]]()
    mps = tmp_path / "m.mps"
    mps.write_text(text)
    sol, dual = tmp_path / "m.sol", tmp_path / "m.dual"
    r = subprocess.run([ffi.CLI_PATH, str(mps), "--device", "cpu", "--tol", "1e-6", "--sol", str(sol), "--dual",
                        str(dual)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    name = case["case"]
    assert sol.read_text() == open(os.path.join(HERE, f"{name}.sol")).read()
    assert dual.read_text() == open(os.path.join(HERE, f"{name}.dual")).read()
    cold, warm = case["reference_cold"], case["reference_warm"]
    assert cold["status"] == warm["status"] == "optimal" and cold["exit"] == warm["exit"] == 0
    assert warm["iterations"] <= 1 < cold["iterations"] and warm["crossover_interior"] is not None
    assert warm["objective"] == pytest.approx(cold["objective"], rel=1e-10)
    # the optimum the reference reached, against scipy's HiGHS on the same model
    A = np.zeros((md.m, md.n))
    for j in range(md.n):
        A[md.rowidx[md.colptr[j]:md.colptr[j + 1]], j] = md.val[md.colptr[j]:md.colptr[j + 1]]
    lo, up = md.row_lo, md.row_up
    A_ub = np.vstack([A[np.isfinite(up)], -A[np.isfinite(lo)]])
    b_ub = np.concatenate([up[np.isfinite(up)], -lo[np.isfinite(lo)]])
    h = linprog(md.cost, A_ub=A_ub, b_ub=b_ub, bounds=list(zip(md.col_lo, md.col_up)), method="highs")
    assert h.status == 0 and warm["objective"] == pytest.approx(h.fun, rel=1e-8)
    assert case["pdhg"]["objective"] == pytest.approx(h.fun, rel=3e-6)
