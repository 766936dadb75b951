"""Batched fast path on the GPU: many LPs packed block-diagonally and solved together.

Each instance must take exactly the decisions a single fast-path solve takes. Row blocks, epilogues and the
reduction partition are the single-instance ones (stage of 2048 nonzeros), so the batch is compared with
single solves on the same layout (column-sorted blocks by default, and the unsorted row blocks) bit for bit:
status, iterations, restarts, recentrings and the returned point. Beyond that: the
planted optima, uneven sizes with inequality rows, a shared-matrix scenario family (stored once), warm starts
with the cold-start fallback, and the autotuned stage size.
"""
import os

import numpy as np
import pytest

from kkt_check import rel_kkt
from taral_pdhg import ffi, gen

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["sorted", "unsorted"], autouse=True)
def layout(request, monkeypatch):
    """Both row-block layouts (TPDHG_LAYOUT): the batch and the single solves it is compared with use the same one."""
    monkeypatch.setenv("TPDHG_LAYOUT", request.param)
    return request.param

# a solve that fails to converge ends at the time limit (and reports progress) instead of running for an hour
FAST = dict(tol=1e-6, precision_bits=32, halpern=1, refine=1, time_limit=120, progress_interval=10)


def _mixed():
    lps = [gen.planted_lp(150, 330, 4, seed=11), gen.planted_lp(400, 900, 6, seed=12),
           gen.planted_lp(900, 2000, 8, seed=13)]
    packing, _ = gen.packing_mps(2000)
    return lps + [packing]


@pytest.mark.parametrize("metric", [1, 0])
def test_batch_matches_single_solves_bit_for_bit(metric):
    """Both restart measures (restart_metric 1: the fixed-point residual, the default; 0: the weighted KKT error)."""
    lps = _mixed()
    models = [ffi.Model.from_lp(md) for md in lps]
    batch = ffi.Batch(models, cap=2048, restart_metric=metric, **FAST)
    assert batch.info()["cap"] == 2048 and batch.info()["unique_matrices"] == len(lps)
    rb = batch.solve()
    for md, model, b in zip(lps, models, rb):
        s = ffi.solve(model, restart_metric=metric, **FAST)
        print(md.name, b["status"], b["iterations"], s["iterations"], b["objective"], s["objective"])
        assert b["status"] == s["status"] == "near_optimal"
        assert (b["iterations"], b["restarts"], b["refinements"]) == (s["iterations"], s["restarts"], s["refinements"])
        assert np.array_equal(b["x"], s["x"]) and np.array_equal(b["y"], s["y"])
        assert b["objective"] == s["objective"]
        assert max(rel_kkt(md, b["x"], b["y"])) <= 1.0001e-6
        if md.planted_objective is not None:
            assert abs(b["objective"] - md.planted_objective) <= 1e-4 * (1 + abs(md.planted_objective))


def test_batch_fp64_iterates_match_single_solves():
    lps = [gen.planted_lp(200, 450, 5, seed=21), gen.planted_lp(350, 700, 5, seed=22)]
    models = [ffi.Model.from_lp(md) for md in lps]
    rb = ffi.Batch(models, cap=2048, tol=1e-6, precision_bits=64, halpern=1).solve()
    for model, b in zip(models, rb):
        s = ffi.solve(model, tol=1e-6, precision_bits=64, halpern=1)
        assert (b["status"], b["iterations"], b["restarts"]) == (s["status"], s["iterations"], s["restarts"])
        assert np.array_equal(b["x"], s["x"])


def test_scenario_family_shares_the_matrix():
    base = gen.planted_lp(500, 1000, 6, seed=31)
    fam = gen.ScenarioFamily(base, n_factors=8, frac=0.02, seed=31)
    th = fam.thetas(4, seed=5)
    lps = [fam.model(t) for t in th]
    models = [ffi.Model.from_lp(md) for md in lps]
    batch = ffi.Batch(models, cap=2048, **FAST)
    info = batch.info()
    assert info["unique_matrices"] == 1 and info["nnz_stored"] * 4 == info["nnz_total"]
    rb = batch.solve()
    for md, model, b in zip(lps, models, rb):
        s = ffi.solve(model, **FAST)
        assert (b["status"], b["iterations"]) == (s["status"], s["iterations"]) == ("near_optimal", s["iterations"])
        assert np.array_equal(b["x"], s["x"])
        assert max(rel_kkt(md, b["x"], b["y"])) <= 1.0001e-6
    # set_vectors moves instance 0 to scenario 1: it must then solve exactly like instance 1
    c, l, u, rl, ru = fam.vectors(th[1])
    batch.set_vectors(0, c, l, u, rl, ru)
    r2 = batch.solve()
    assert r2[0]["iterations"] == rb[1]["iterations"] and np.array_equal(r2[0]["x"], rb[1]["x"])


def test_batch_warm_starts_and_fallback():
    lps = [gen.planted_lp(300, 700, 5, seed=41), gen.planted_lp(300, 700, 5, seed=42)]
    models = [ffi.Model.from_lp(md) for md in lps]
    cold = ffi.Batch(models, cap=2048, **FAST).solve()
    batch = ffi.Batch(models, cap=2048, warm_fallback=1, **FAST)
    bad_x = np.full(lps[1].n, 10.0)     # every column at its upper bound: far worse than the cold start
    bad_y = np.full(lps[1].m, 1e3)
    r = batch.solve(x0=[cold[0]["x"], bad_x], y0=[cold[0]["y"], bad_y],
                    primal_weight=[cold[0]["primal_weight"], 1.0])
    assert r[0]["warm_start_used"] == 1 and r[0]["iterations"] == 0 and r[0]["status"] == "near_optimal"
    assert r[1]["warm_start_used"] == 0 and r[1]["warm_kkt"] > r[1]["cold_kkt"]
    assert r[1]["iterations"] == cold[1]["iterations"] and np.array_equal(r[1]["x"], cold[1]["x"])


def test_autotuned_stage_and_uneven_sizes():
    rng = np.random.default_rng(7)
    lps = []
    for k in range(12):
        m = int(rng.integers(100, 1500))
        n = int(round(m * rng.uniform(1.5, 3.0)))
        lps.append(gen.planted_lp(m, n, int(rng.integers(6, 13)), seed=100 + k))
    models = [ffi.Model.from_lp(md) for md in lps]
    batch = ffi.Batch(models, **FAST)
    info = batch.info()
    print("autotune", info)
    assert info["cap"] in (1024, 2048, 4096, 8192) and min(info["tune_ms"]) > 0
    for md, b in zip(lps, batch.solve()):
        assert b["status"] == "near_optimal"
        assert max(rel_kkt(md, b["x"], b["y"])) <= 1.0001e-6
        assert abs(b["objective"] - md.planted_objective) <= 1e-4 * (1 + abs(md.planted_objective))


def test_batch_iteration_limit_stops_every_instance():
    lps = [gen.planted_lp(300, 700, 5, seed=51), gen.planted_lp(200, 400, 5, seed=52)]
    rb = ffi.Batch([ffi.Model.from_lp(md) for md in lps], cap=2048, max_iter=128, **FAST).solve()
    for b in rb:
        assert b["status"] == "iteration_limit" and b["iterations"] == 128


def _batch_run(lps, monkeypatch, cap=2048, **env):
    for k, v in env.items():
        monkeypatch.setenv(k, str(v))
    b = ffi.Batch([ffi.Model.from_lp(md) for md in lps], cap=cap, **FAST)
    res, info = b.solve(), b.info()
    b.close()
    for k in env:
        monkeypatch.delenv(k)
    return res, info


def _same(ra, rb):
    for a, b in zip(ra, rb):
        assert (a["status"], a["iterations"], a["restarts"], a["refinements"]) == \
               (b["status"], b["iterations"], b["restarts"], b["refinements"])
        assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["y"], b["y"])


def test_persistent_blocks_match_lockstep(monkeypatch):
    """Small live sets run each 64-iteration block as one cooperative launch (k_bpersist: row phase, grid barrier,
    column phase, grid barrier, per step) with the lockstep launches' block code: bit-identical results."""
    lps = _mixed()
    lock, il = _batch_run(lps, monkeypatch, TPDHG_PERSIST_MAX=0)
    pers, ip = _batch_run(lps, monkeypatch, TPDHG_PERSIST_MAX=256)
    print(f"lockstep blocks {il['lockstep_blocks']}; persistent {ip['persistent_blocks']} / lockstep "
          f"{ip['lockstep_blocks']}", flush=True)
    assert il["persistent_blocks"] == 0 and il["persist_max"] == 0
    assert ip["persistent_blocks"] > 0 and ip["persist_max"] > 0
    _same(lock, pers)


def test_persistent_switches_in_when_the_live_set_shrinks(monkeypatch):
    """Uneven LPs at a fixed stage (the autotuner times stages, so two batches may pick different ones and sum in
    different orders): lockstep while many blocks are live, persistent once few are."""
    lps = gen.uneven_batch(count=24, m_hi=6000)
    lock, il = _batch_run(lps, monkeypatch, cap=1024, TPDHG_PERSIST_MAX=0)
    pers, ip = _batch_run(lps, monkeypatch, cap=1024, TPDHG_PERSIST_MAX=160)
    assert il["cap"] == ip["cap"] == 1024
    print(f"persistent {ip['persistent_blocks']} / lockstep {ip['lockstep_blocks']} blocks", flush=True)
    assert ip["persistent_blocks"] > 0 and ip["lockstep_blocks"] > 0
    _same(lock, pers)
    for md, r in zip(lps, pers):
        assert max(rel_kkt(md, r["x"], r["y"])) <= 1.0001e-6


def test_persistent_barrier_time_out_restores_and_replays(monkeypatch):
    """Failure path forced from the host: every grid barrier waits for one arrival too many and gives up after 64
    polls, so each workgroup sets the error flag and exits; the host restores the block's snapshot, replays it with
    the lockstep launches and turns the persistent path off. Results equal the lockstep solve's bit for bit."""
    lps = _mixed()
    lock, _ = _batch_run(lps, monkeypatch, TPDHG_PERSIST_MAX=0)
    fb, info = _batch_run(lps, monkeypatch, TPDHG_PERSIST_MAX=256, TPDHG_PERSIST_TEST_UNREACHABLE=1,
                          TPDHG_PERSIST_SPINS=64)
    assert info["persist_max"] == 0 and info["persistent_blocks"] == 0 and info["lockstep_blocks"] > 0
    _same(lock, fb)


def test_stage_follows_the_live_set(monkeypatch):
    """With an autotuned stage the batch moves to smaller row-block stages as its live set shrinks (resident
    per-stage pools): the stage changes, every instance still meets 1e-6 by the independent KKT recheck, and a
    second solve of the same batch (starting all live at the tuned stage again) repeats the first bit for bit.
    The tuner's pick is given as 8,192 (TPDHG_TUNED_STAGE): timed, it may pick 1,024, below which nothing moves."""
    monkeypatch.setenv("TPDHG_TUNED_STAGE", "8192")
    lps = gen.uneven_batch(count=16, m_hi=8000)
    b = ffi.Batch([ffi.Model.from_lp(md) for md in lps], **FAST)
    r1 = b.solve()
    info = b.info()
    r2 = b.solve()
    b.close()
    print(f"stage {info['cap']} switches {info['stage_switches']}", flush=True)
    assert info["stage_switches"] > 0 and info["cap"] < 8192
    _same(r1, r2)
    for md, r in zip(lps, r1):
        assert r["status"] == "near_optimal" and max(rel_kkt(md, r["x"], r["y"])) <= 1.0001e-6


def test_probe_leaves_later_solves_unchanged():
    """tpdhg_batch_probe (the benchmark's roofline timing) iterates all instances on the batch's own device vectors;
    every solve re-initialises them, so a solve after a probe repeats the solve before it bit for bit. The probe
    reports one row and one column launch per iteration, each timed."""
    b = ffi.Batch([ffi.Model.from_lp(md) for md in _mixed()], **FAST)
    r1 = b.solve()
    p = b.probe(8)
    r2 = b.solve()
    b.close()
    assert p["row_launches"] == 8 and p["col_launches"] == 8 and p["row_ms"] > 0 and p["col_ms"] > 0
    _same(r1, r2)
