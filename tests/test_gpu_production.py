"""The production kernel paths and the full-size configurations on the GPU.

Small LPs default to the row-block kernels, so the layouts the large matrices use (column-sorted blocks, sorted
column slabs with their combine epilogues) are exercised here in two ways:

* the reference's published ledger trajectories (FP64, eval_freq 64; results/gpu_pdhg/synth_*.csv) with the
  layout forced: TPDHG_LAYOUT=sorted (k_sorted with the reference scheme's EpiRow/EpiCol) and TPDHG_LAYOUT=slab
  (k_sorted_slab + k_combine_epi with EpiRow/EpiCol on every matrix of 16+ columns). Iterations and restarts
  must equal the ledger's, the objective its 12 printed digits (1e-9 relative);
* the ledger's 1M- and 2M-column packing rows (synth_packing_tol1e-4.csv:12,17; synth_packing_tol1e-6.csv:12,17)
  on the default layout selection, which lays their matrices out in sorted slabs (A: 1M / 2M gathered columns)
  and sorted blocks, with the halved FP64 blocks;
* the ledger's transport_1000x1000 row (synth_transport_tol1e-4.csv:12), whose A (2,000 rows of 1,000 nonzeros
  over 1M columns; each supply row fills one column slab and is empty in the others) goes through the timed choice
  between slabs and one piece (choose_layout), on that choice and on forced slabs;
* the fast path (FP32 iterates + FP64 refinement) at the full sizes of the configurations: the 1M x 2M planted LP,
  the block-angular LP, the 512-LP batch and scenarios of the 50k x 100k scenario base. Each result is rechecked
  by an independent original-space KKT computation (tests/kkt_check.py) at the requested 1e-6 and, where the
  instance is planted, against the planted optimum c'x* = b'y*.
"""
import functools
import json
import os

import numpy as np
import pytest

from kkt_check import rel_kkt
from taral_pdhg import ffi, gen

pytestmark = pytest.mark.gpu
LEDGER = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ledger_pdhg.json")))
FAST = dict(tol=1e-6, precision_bits=32, halpern=1, refine=1)

LARGE_PACKING = [  # results/gpu_pdhg/synth_packing_tol1e-4.csv:12,17 and synth_packing_tol1e-6.csv:12,17
    dict(N=1000000, tol=1e-4, iterations=640, restarts=4, objective=-687667.841846),
    dict(N=1000000, tol=1e-6, iterations=6848, restarts=9, objective=-687666.110124),
    dict(N=2000000, tol=1e-4, iterations=640, restarts=4, objective=-1374066.47575),
    dict(N=2000000, tol=1e-6, iterations=6848, restarts=9, objective=-1374063.42034),
]


LARGE_TRANSPORT = dict(S=1000, D=1000, tol=1e-4, iterations=28032, restarts=22, objective=133.882893894)


def _ledger_model(case):
    if case["family"] == "packing":
        return gen.packing_mps(*case["args"], text=False)[0]
    return gen.transport_mps(*case["args"])[0]


@pytest.mark.parametrize("layout", ["sorted", "slab"])
@pytest.mark.parametrize("case", LEDGER["cases"], ids=lambda c: f"{c['case']}-{c['tol']:g}")
def test_ledger_trajectory_forced_layout(case, layout, monkeypatch):
    monkeypatch.setenv("TPDHG_LAYOUT", layout)
    md = _ledger_model(case)
    print(f"ledger {case['case']} tol={case['tol']} layout={layout}", flush=True)
    r = ffi.solve(ffi.Model.from_lp(md), tol=case["tol"], precision_bits=64)
    assert r["status"] == "near_optimal"
    assert (r["iterations"], r["restarts"]) == (case["iterations"], case["restarts"])
    assert r["objective"] == pytest.approx(case["objective"], rel=1e-9)
    assert r["rel_primal"] == pytest.approx(case["rel_primal"], rel=1e-4)


@functools.lru_cache(maxsize=1)
def _packing(N):
    return gen.packing_mps(N, text=False)[0]


@pytest.mark.timeout(600)
@pytest.mark.parametrize("case", LARGE_PACKING, ids=lambda c: f"packing_{c['N']}-{c['tol']:g}")
def test_ledger_trajectory_large_packing(case, monkeypatch):
    monkeypatch.delenv("TPDHG_LAYOUT", raising=False)
    monkeypatch.delenv("TPDHG_SLABS", raising=False)
    md = _packing(case["N"])
    dims = next(d for d in LEDGER["generator_dims"] if d["args"] == [case["N"]])
    assert (md.m, md.n, md.nnz) == (dims["rows"], dims["cols"], dims["nnz"])
    print(f"ledger packing_{case['N']} tol={case['tol']} (default layout: sorted slabs)", flush=True)
    r = ffi.solve(ffi.Model.from_lp(md), tol=case["tol"], precision_bits=64)
    print(f"  {r['status']} it={r['iterations']} restarts={r['restarts']} obj={r['objective']:.12g} "
          f"solve={r['solve_s']:.2f}s", flush=True)
    assert r["status"] == "near_optimal"
    assert (r["iterations"], r["restarts"]) == (case["iterations"], case["restarts"])
    assert r["objective"] == pytest.approx(case["objective"], rel=1e-9)


def _check_fast(md, r, tol=1e-6, obj_rel=1e-5):
    assert r["status"] == "near_optimal", r["status"]
    p, d, g = rel_kkt(md, r["x"], r["y"])
    assert max(p, d, g) <= tol * 1.0001, (p, d, g)
    assert r["rel_primal"] == pytest.approx(p, rel=1e-5, abs=1e-12)
    assert r["rel_gap"] == pytest.approx(g, rel=1e-5, abs=1e-12)
    assert np.all(r["x"] >= md.col_lo - 1e-9) and np.all(r["x"] <= md.col_up + 1e-9)
    if md.planted_objective is not None:
        ps = md.planted_objective
        assert abs(r["objective"] - ps) / (1 + abs(ps)) <= obj_rel


@pytest.mark.timeout(900)
def test_fast_full_size_cfg1(monkeypatch):
    """The 1M x 2M planted LP (20 nonzeros per column, 40M nonzeros) on its production layout."""
    monkeypatch.delenv("TPDHG_LAYOUT", raising=False)
    md = gen.planted_lp(1_000_000, 2_000_000, 20, seed=1)
    print("fast 1M x 2M planted LP, tol 1e-6", flush=True)
    r = ffi.solve(ffi.Model.from_lp(md), progress_interval=30.0, **FAST)
    print(f"  {r['status']} it={r['iterations']} refinements={r['refinements']} solve={r['solve_s']:.1f}s", flush=True)
    _check_fast(md, r)


@pytest.mark.timeout(300)
def test_fast_full_size_block_angular(monkeypatch):
    monkeypatch.delenv("TPDHG_LAYOUT", raising=False)
    md = gen.block_angular_lp()
    assert (md.m, md.n) == (320_200, 512_000)
    print("fast block-angular 320,200 x 512,000", flush=True)
    r = ffi.solve(ffi.Model.from_lp(md), **FAST)
    print(f"  {r['status']} it={r['iterations']} solve={r['solve_s']:.2f}s", flush=True)
    _check_fast(md, r)


@pytest.mark.timeout(600)
def test_fast_full_size_batch_512(monkeypatch):
    """All 512 LPs of the uneven batch in one batched solve; every instance rechecked."""
    monkeypatch.delenv("TPDHG_LAYOUT", raising=False)
    lps = gen.uneven_batch(count=512)
    print("batch of 512 uneven planted LPs", flush=True)
    b = ffi.Batch([ffi.Model.from_lp(md) for md in lps], progress_interval=30.0, **FAST)
    res = b.solve()
    b.close()
    print(f"  max iterations {max(r['iterations'] for r in res)}", flush=True)
    for md, r in zip(lps, res):
        _check_fast(md, r)


@pytest.mark.timeout(600)
def test_fast_full_size_scenarios(monkeypatch):
    """Cold solves of four scenarios of the 50,000 x 100,000 scenario base (6 nonzeros per column) in one batch
    sharing the matrix; the planted optimum does not hold for rescaled scenarios, so each is rechecked by KKT."""
    monkeypatch.delenv("TPDHG_LAYOUT", raising=False)
    fam = gen.ScenarioFamily(gen.planted_lp(50_000, 100_000, 6, seed=4), seed=4)
    th = fam.thetas(4, seed=2)
    models = [ffi.Model.from_lp(fam.model(t)) for t in th]
    print("four scenarios of the 50k x 100k base", flush=True)
    b = ffi.Batch(models, progress_interval=30.0, **FAST)
    assert b.info()["unique_matrices"] == 1
    res = b.solve()
    b.close()
    for t, r in zip(th, res):
        print(f"  {r['status']} it={r['iterations']}", flush=True)
        _check_fast(fam.model(t), r)


@pytest.fixture(scope="module")
def transport_1000x1000():
    c = LARGE_TRANSPORT
    return ffi.Model.from_lp(gen.transport_mps(c["S"], c["D"])[0])


@pytest.mark.timeout(600)
@pytest.mark.parametrize("layout", ["default", "slab"])
def test_ledger_trajectory_large_transport(layout, transport_1000x1000, monkeypatch):
    """results/gpu_pdhg/synth_transport_tol1e-4.csv:12 on the timed layout choice and on forced slabs."""
    monkeypatch.delenv("TPDHG_SLABS", raising=False)
    if layout == "slab":
        monkeypatch.setenv("TPDHG_LAYOUT", "slab")
    else:
        monkeypatch.delenv("TPDHG_LAYOUT", raising=False)
    c = LARGE_TRANSPORT
    print(f"ledger transport_1000x1000 tol={c['tol']} layout={layout}", flush=True)
    r = ffi.solve(transport_1000x1000, tol=c["tol"], precision_bits=64)
    print(f"  {r['status']} it={r['iterations']} restarts={r['restarts']} obj={r['objective']:.12g} "
          f"solve={r['solve_s']:.2f}s", flush=True)
    assert r["status"] == "near_optimal"
    assert (r["iterations"], r["restarts"]) == (c["iterations"], c["restarts"])
    assert r["objective"] == pytest.approx(c["objective"], rel=1e-9)


@pytest.mark.timeout(300)
def test_fast_transport_weight_rule_4():
    """omega_rule 4 (the weight guard's two decades counted from the first restart's weight) on the transport LP
    whose weight the default guard holds on its lower bound: a 1e-6 solution, rechecked by the independent KKT
    computation, in far fewer iterations than under rule 3 (measured: 28,032 against 112,448)."""
    c = LARGE_TRANSPORT
    md = gen.transport_mps(c["S"], c["D"])[0]
    print("fast transport_1000x1000, omega_rule 4", flush=True)
    r = ffi.solve(ffi.Model.from_lp(md), omega_rule=4, **FAST)
    print(f"  {r['status']} it={r['iterations']} solve={r['solve_s']:.2f}s obj={r['objective']:.10g}", flush=True)
    _check_fast(md, r)
    assert r["iterations"] < 60000
    assert r["objective"] == pytest.approx(133.89572585, rel=1e-5)  # HiGHS, synth_transport_tol1e-6.csv:12
