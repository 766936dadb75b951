"""Ledger counterpart (taral_pdhg.ledger): schema, the HiGHS reference objective, and (GPU) a full ledger row.

The HiGHS objective (scipy's bundled HiGHS) is pinned by the reference ledger's exact optimum of transport_100x100
(41.5037970614, results/gpu_pdhg/synth_transport_tol1e-4.csv:2) and by the planted optimum of a seeded LP.
"""
import csv
import json
import os

import pytest

from taral_pdhg import ffi, gen, ledger


def test_schema_is_the_reference_ledgers():
    assert ledger.FIELDS[:3] == ["case", "rows", "cols"] and ledger.FIELDS[-1] == "chk_error"
    assert len(ledger.FIELDS) == 38 and len(set(ledger.FIELDS)) == 38
    assert ledger.rel_err(101.0, 100.0) == pytest.approx(0.01) and ledger.rel_err(0.5, 0.0) == 0.5
    assert ledger.rel_err(None, 1.0) is None


def test_highs_reference_objective(transport_100x100, tmp_path):
    md, text = transport_100x100
    p = tmp_path / "transport_100x100.mps"
    p.write_text(text)
    obj, secs = ledger.highs_objective(ffi.Model.from_mps(str(p)))
    assert obj == pytest.approx(41.5037970614, rel=1e-10) and secs > 0
    pl = gen.planted_lp(200, 450, 5, seed=3)
    obj, _ = ledger.highs_objective(ffi.Model.from_lp(pl))
    assert obj == pytest.approx(pl.planted_objective, rel=1e-7, abs=1e-7)


@pytest.mark.gpu
def test_ledger_row_on_the_gpu(transport_100x100, tmp_path):
    md, text = transport_100x100
    p = tmp_path / "transport_100x100.mps"
    p.write_text(text)
    out = tmp_path / "ledger.csv"
    ledger.main([str(p), "--out", str(out), "--tol", "1e-4", "--repeats", "2"])
    rows = list(csv.DictReader(open(out)))
    assert len(rows) == 1 and list(rows[0].keys()) == ledger.FIELDS
    r = rows[0]
    # the reference ledger's row for this case at 1e-4: 5440 iterations, 12 restarts, objective 41.5017277764
    assert r["status"] == "near_optimal" and r["iterations"] == "5440" and r["restarts"] == "12"
    assert float(r["objective"]) == pytest.approx(41.5017277764, rel=1e-9)
    assert float(r["objective_rel_error_vs_highs"]) < 1e-4
    assert r["repeats"] == "2" and r["objective_identical_repeats"] == "True"
    assert float(r["chk_row_viol_rel"]) < 1e-3 and float(r["chk_bound_viol"]) < 1e-6
    assert r["chk_dual_obj_loose"] != ""


def test_generated_cases_are_named_as_the_reference_ledgers(transport_100x100):
    """`packing:N` / `transport:SxD` build the reference's generated cases in memory, named as its ledgers name them
    (results/gpu_pdhg/synth_*.csv), with the dimensions the ledgers record."""
    ledger_file = os.path.join(os.path.dirname(__file__), "golden", "ledger_pdhg.json")
    dims = {c["case"]: (c["rows"], c["cols"], c["nnz"]) for c in json.load(open(ledger_file))["cases"]}
    name, model = ledger.load_case("transport:100x100")
    i = model.info()
    assert name == "transport_100x100" and (i.m, i.n, i.nnz) == dims[name]
    md, _ = transport_100x100
    assert (md.m, md.n, md.nnz) == dims[name]
    name, model = ledger.load_case("packing:10000")
    i = model.info()
    assert name == "packing_10000" and (i.m, i.n, i.nnz) == dims[name]


@pytest.mark.gpu
def test_ledger_generated_case_on_the_gpu(tmp_path):
    """A generated case through the ledger tool reproduces the reference ledger's row (packing_10000 at 1e-4:
    results/gpu_pdhg/synth_packing_tol1e-4.csv:2)."""
    ledger_file = os.path.join(os.path.dirname(__file__), "golden", "ledger_pdhg.json")
    want = next(c for c in json.load(open(ledger_file))["cases"] if c["case"] == "packing_10000" and c["tol"] == 1e-4)
    out = tmp_path / "generated.csv"
    ledger.main(["packing:10000", "--out", str(out), "--tol", "1e-4", "--repeats", "1", "--no-highs"])
    r = list(csv.DictReader(open(out)))[0]
    assert r["case"] == "packing_10000" and r["status"] == "near_optimal"
    assert (int(r["iterations"]), int(r["restarts"])) == (want["iterations"], want["restarts"])
    assert float(r["objective"]) == pytest.approx(want["objective"], rel=1e-9)
