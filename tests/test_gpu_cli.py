"""The CLI on the GPU: --recheck keys, .sol/.dual written and read back as a warm start, --batch lists."""
import json
import os
import subprocess

import pytest

from taral_pdhg import ffi

pytestmark = pytest.mark.gpu


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_recheck_and_warm_files_round_trip(transport_100x100, tmp_path):
    mps = _write(tmp_path, "t.mps", transport_100x100[1])
    js, sol, dual = (str(tmp_path / f) for f in ("t.json", "t.sol", "t.dual"))
    r = subprocess.run([ffi.CLI_PATH, mps, "--device", "gpu", "--tol", "1e-6", "--fast", "--recheck", "--json", js,
                        "--sol", sol, "--dual", dual], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.load(open(js))
    assert out["status"] == "near_optimal" and "recheck objective" in r.stdout
    assert out["chk_objective"] == pytest.approx(out["objective"], rel=1e-12)
    assert out["chk_row_viol_rel"] < 1e-5 and out["chk_bound_viol"] < 1e-9
    # transport is an equality/inequality LP with x >= 0: the dual bound is within the solve's gap of the optimum
    assert out["chk_dual_obj_loose"] == pytest.approx(41.5037970614, rel=1e-4)
    r2 = subprocess.run([ffi.CLI_PATH, mps, "--device", "gpu", "--tol", "1e-6", "--fast", "--warm-sol", sol,
                         "--warm-dual", dual, "--json", js], capture_output=True, text=True, timeout=120)
    assert r2.returncode == 0, r2.stderr
    w = json.load(open(js))
    assert w["status"] == "near_optimal" and w["iterations"] == 0 and w["warm_start_used"] is True


def test_batch_list(packing_10000, transport_100x100, tmp_path):
    a = _write(tmp_path, "p.mps", packing_10000[1])
    b = _write(tmp_path, "t.mps", transport_100x100[1])
    lst = _write(tmp_path, "list.txt", f"{a}\n# skipped\n{b}\n")
    js = str(tmp_path / "out.jsonl")
    r = subprocess.run([ffi.CLI_PATH, "--batch", lst, "--tol", "1e-6", "--fp32", "--refine", "--recheck", "--json", js],
                       capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    rows = [json.loads(line) for line in open(js)]
    assert [x["case"] for x in rows] == [a, b]
    for x in rows:
        assert x["status"] == "near_optimal" and max(x["rel_primal"], x["rel_dual"], x["rel_gap"]) <= 1e-6
        assert "chk_dual_gap_rel" in x
    # the same optimum as a single fast solve (the CLI's batch picks its own row-block stage, so the trajectories
    # may differ in summation order; both stop at 1e-6 relative KKT)
    single = ffi.solve(ffi.Model.from_mps(b), tol=1e-6, precision_bits=32, halpern=1, refine=1)
    assert rows[1]["objective"] == pytest.approx(single["objective"], rel=1e-5)


def test_row_rel_flag(packing_10000, tmp_path):
    """`--row-rel 1` on the CLI takes the per-row stop test's longer trajectory (the oracle's 6,080 iterations and 9
    restarts at 1e-4 on packing_10000, tests/test_oracle.py) instead of the default 640 / 4."""
    mps = _write(tmp_path, "p.mps", packing_10000[1])
    js = str(tmp_path / "p.json")
    r = subprocess.run([ffi.CLI_PATH, mps, "--device", "gpu", "--tol", "1e-4", "--row-rel", "1", "--json", js],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.load(open(js))
    assert out["status"] == "near_optimal" and (out["iterations"], out["restarts"]) == (6080, 9)
