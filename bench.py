#!/usr/bin/env python3
"""Benchmark: LP solves/sec of the restarted PDHG solver on one MI355X, with the iteration-kernel roofline.

Workloads (--workload):
  cfg1 (default)  the single large planted LP: m = 1,000,000 equality rows, n = 2,000,000 columns, 20 nonzeros per
                  column at uniform-random rows, N(0,1) values, 0 <= x <= 10, seed 1.
  cfg2            the block-angular planning LP (64 blocks of 5,000 x 8,000 plus 200 linking rows), planted.
  cfg3            a batch of 512 independent planted LPs of uneven size, solved together by the batched fast path.
  cfg4            a scenario family of one planted LP (base 5,000 x 10,000 by default, --scenario-m): the learned
                  warm start is trained on cold solves of 2,048 scenarios, then one step predicts and solves the
                  512 held-out scenarios (batches of 512 sharing the matrix), with the cold-start fallback.
  cfg0, mid       the 2,000 x 5,000 and 100,000 x 200,000 planted LPs.
Every solve runs to 1e-6 relative KKT (original-space primal residual, dual residual and gap, the reference's
stopping test). A "step" is one full solve from a cold start of problems already resident in HBM (power iteration,
PDHG iterations, FP64 evaluations, download and the original-space check); for cfg3 one step solves the whole
batch and the value counts every LP. The default mode is the fast path (FP32 reflected Halpern PDHG iterates inside
FP64 iterative refinement); one cfg1 solve takes minutes, so the defaults time one step without warm-up.

Printed JSON line: the contract's fields, plus
  roofline: single LPs: the dominant iteration kernel (A^T z fused with the primal step, or A x fused with the dual
            step), algorithmic bytes per launch / mean launch time from HIP events on the solver stream in a
            separate capped solve; batches: one batched iteration with every instance live (both launches),
            timed by the batch's own stage autotuner; against the 8 TB/s HBM peak;
  cpu_baseline: the CPU oracle (FP64, its iterations on up to 16 host threads partitioned as the reference's CPU
            backend) timed on a bounded sample (about two seconds of PDHG iterations), scaled to solves/sec with
            the GPU solve's iteration counts;
  fp64_reference: the plain FP64 cold start of the reference algorithm (single LPs: the same LP, capped at
            --fp64-time-limit; batches: the first --fp64-sample LPs one at a time).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "taral-lp_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
METRIC = "LP solves/sec to 1e-6 rel KKT (learned+FP32 vs plain FP64 PDHG); SpMV HBM GB/s"

WORKLOADS = {  # name: (kind, description, generator)
    "cfg1": ("single", "planted equality LP m=1,000,000 n=2,000,000 nnz/col=20 seed=1, 0<=x<=10",
             lambda gen, a: gen.planted_lp(1_000_000, 2_000_000, 20, seed=1)),
    "cfg0": ("single", "planted equality LP m=2,000 n=5,000 nnz/col=10 seed=0, 0<=x<=10",
             lambda gen, a: gen.planted_lp(2_000, 5_000, 10, seed=0)),
    "mid": ("single", "planted equality LP m=100,000 n=200,000 nnz/col=20 seed=1, 0<=x<=10",
            lambda gen, a: gen.planted_lp(100_000, 200_000, 20, seed=1)),
    "cfg2": ("single", "block-angular planning LP: 64 blocks of 5,000x8,000 (8 N(0,1)/col) + 200 linking rows "
                       "(2% of each block's columns, +-1), planted, m=320,200 n=512,000 seed=2",
             lambda gen, a: gen.block_angular_lp()),
    "cfg3": ("batch", "independent planted LPs of uneven size, m~U[500,20000], n=round(m*U[1.5,3]), "
                      "nnz/col~U{6..12} per LP, seed=3",
             lambda gen, a: gen.uneven_batch(count=a.batch_count)),
    "cfg4": ("scenarios", "scenario family of one planted LP (6 nnz/col, seed 4): 32 factors theta~N(0,0.1^2), "
                          "each rescaling a fixed 0.15% of b and of c by exp(theta)",
             lambda gen, a: gen.ScenarioFamily(gen.planted_lp(a.scenario_m, 2 * a.scenario_m, 6, seed=4), seed=4)),
}


def kernel_bytes(m, n, nnz, s):
    """Algorithmic HBM bytes of one launch of each fused iteration kernel (s = bytes per value):
    row kernel: val+idx (s+4)/nnz, row pointers 4(m+1), x read once s*n, and per row z, Ax, rl, ru, Axs, zs
    read and Ax, z, Axs, zs written (10 s per row); column kernel likewise over A^T with 10 s per column."""
    row = (s + 4) * nnz + 4 * (m + 1) + s * n + 10 * s * m
    col = (s + 4) * nnz + 4 * (n + 1) + s * m + 10 * s * n
    return row, col


MODES = {  # name: solver options
    "fast": dict(precision_bits=32, halpern=1, refine=1),   # FP32 reflected Halpern + FP64 refinement
    "fp64": dict(precision_bits=64),                        # the reference algorithm, plain FP64 cold start
    "fp32": dict(precision_bits=32),                        # the reference algorithm with FP32 iterates
    "halpern-fp64": dict(precision_bits=64, halpern=1),
}


def halpern_kernel_bytes(m, n, nnz, s):
    """Same for the fast path's kernels: row kernel reads w once (s*n) and per row z, zh, rl, ru, z0 and writes z
    (6 s per row); column kernel reads z once (s*m) and per column x, x+, x0, c, l, u and writes x, x+, w
    (9 s per column; the A^T z store of a block's last step is not counted)."""
    row = (s + 4) * nnz + 4 * (m + 1) + s * n + 6 * s * m
    col = (s + 4) * nnz + 4 * (n + 1) + s * m + 9 * s * n
    return row, col


def mode_options(mode):
    if mode not in MODES:
        raise SystemExit(f"unknown mode {mode}")
    return dict(MODES[mode])


CPU_THREADS = 16  # host threads for the CPU baseline: 16 cores per GPU


def cpu_threads():
    return max(1, min(CPU_THREADS, os.cpu_count() or 1))


def cpu_baseline(md, iterations, sample_iters):
    """Oracle FP64 on the host's cores (the reference CPU backend's row partition and barriers, one thread per
    core, at most CPU_THREADS): time PDHG iterations of the same problem (at least `sample_iters`, doubled until the
    sample takes two seconds of wall time), scale to one solve of `iterations` iterations."""
    from oracle import oracle
    nt = cpu_threads()
    o = oracle.OracleLP(md)
    t0 = time.perf_counter()
    o.block(0.5, 0.5, 0, threads=nt)  # start: two products
    t_start = time.perf_counter() - t0
    k = max(1, sample_iters)
    while True:
        t0 = time.perf_counter()
        o.block(0.5, 0.5, k, threads=nt)
        t = time.perf_counter() - t0 - t_start
        if t >= 2.0 or k >= 1 << 16:
            break
        k *= 2
    per_iter = max(t, 1e-9) / k
    return dict(value=1.0 / (per_iter * iterations), unit="solves/s", cores=nt, kind="port",
                sample=f"CPU oracle (oracle/pdhg_oracle.c, FP64, {nt} threads partitioned as the reference's "
                       f"CPU backend): {k} PDHG iterations of the same LP took {t:.3f}s ({per_iter * 1e3:.3f} "
                       f"ms/iter); scaled to the GPU solve's {iterations} iterations", ms_per_iter=per_iter * 1e3)


def roofline_entry(b_row, b_col, row_avg, col_avg, label=None):
    if col_avg >= row_avg:
        dom, dom_bytes, dom_avg = "col (A^T z + primal step)", b_col, col_avg
    else:
        dom, dom_bytes, dom_avg = "row (A x + dual step)", b_row, row_avg
    if label:
        dom = label
    achieved = dom_bytes / (dom_avg * 1e-3) / 1e9 if dom_avg > 0 else 0.0
    return dict(bound="hbm", kernel=dom, achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None, bytes_per_launch=int(dom_bytes),
                avg_launch_ms=round(dom_avg, 5), row_kernel_avg_ms=round(row_avg, 5), col_kernel_avg_ms=round(col_avg, 5),
                row_gbs=round(b_row / (row_avg * 1e-3) / 1e9, 1) if row_avg else 0,
                col_gbs=round(b_col / (col_avg * 1e-3) / 1e9, 1) if col_avg else 0)


def measured_traffic(workload, mode, kernel, side=None):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of
    this workload (copy-kernel calibrated), or None when there is none. `side` names the summary's line directly
    (batches: "iter<stage>", one batched iteration at that stage); otherwise it is the kernel's step, row or col."""
    path = os.path.join(ROOT, "profiles", f"pdhg_{workload}_{mode}_traffic.md")
    if side is None:
        side = "col" if kernel.startswith("col") else "row"
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == side:
                return int(parts[2])
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def solver_opts(a, mode_name, **extra):
    o = dict(tol=a.tol, time_limit=a.time_limit, max_iter=a.max_iter, use_graph=0 if a.no_graph else 1,
             verbose=int(a.verbose), progress_interval=a.progress, **mode_options(mode_name))
    if a.omega_rule is not None:
        o["omega_rule"] = a.omega_rule
    if a.reflection is not None and o.get("halpern"):
        o["reflection"] = a.reflection
    o.update(extra)
    return o


def rel_kkt_of(r):
    return max(r["rel_primal"], r["rel_dual"], r["rel_gap"])


def run_single(a, ffi, gen, desc, md):
    model = ffi.Model.from_lp(md)
    mode = a.mode

    def timed(mode_name, steps, warmup, **extra):
        t0 = time.perf_counter()
        sess = ffi.Session(model, **solver_opts(a, mode_name, **extra))
        setup = time.perf_counter() - t0
        log(f"{mode_name}: session setup {setup:.1f}s")
        for w in range(warmup):
            r = sess.solve()
            log(f"{mode_name} warmup {w}: {r['status']} it={r['iterations']} solve={r['solve_s']:.3f}s")
        results = []
        t0 = time.perf_counter()
        for s_ in range(steps):
            r = sess.solve()
            results.append(r)
            log(f"{mode_name} step {s_}: {r['status']} it={r['iterations']} solve={r['solve_s']:.3f}s")
        wall = time.perf_counter() - t0
        sess.close()
        return wall, results, setup

    wall, res, setup = timed(mode, a.steps, a.warmup)
    last = res[-1]
    value = a.steps / wall
    # kernel roofline: a separate capped solve with HIP-event timing of every iteration launch (no graphs)
    _, prof, _ = timed(mode, 1, 0, profile_kernels=1, max_iter=min(a.max_iter, a.profile_iters))
    q = prof[-1]
    s_val = 8 if MODES[mode]["precision_bits"] == 64 else 4
    b_row, b_col = (halpern_kernel_bytes if MODES[mode].get("halpern") else kernel_bytes)(md.m, md.n, q["nnz"], s_val)
    roofline = roofline_entry(b_row, b_col, q["row_kernel_ms"] / max(q["row_launches"], 1),
                              q["col_kernel_ms"] / max(q["col_launches"], 1))
    roofline["profiled_iterations"] = q["iterations"]
    roofline["traffic"] = measured_traffic(a.workload, mode, roofline["kernel"])
    out = {
        "metric": METRIC, "value": round(value, 6), "unit": "solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(1e3 * wall / a.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": DTYPES[mode], "data": "synthetic",
        "config": {"workload": desc, "m": md.m, "n": md.n, "nnz": last["nnz"], "tol": a.tol, "mode": mode,
                   "parallelism": "single GPU"},
        "iterations": last["iterations"], "restarts": last["restarts"], "refinements": last["refinements"],
        "status": last["status"], "rel_kkt": rel_kkt_of(last),
        "iters_per_s": round(last["iterations"] / last["solve_s"], 1), "session_setup_s": round(setup, 3),
        "device": last["device"], "roofline": roofline,
    }
    if md.planted_objective is not None:
        p_star = md.planted_objective
        out["objective_rel_error_vs_planted"] = abs(last["objective"] - p_star) / (1 + abs(p_star))
    if mode != "fp64" and not a.no_fp64_compare:
        # the plain FP64 cold start of the reference algorithm, capped at --fp64-time-limit
        saved = a.time_limit
        a.time_limit = a.fp64_time_limit
        w64, r64, _ = timed("fp64", 1, 0)
        a.time_limit = saved
        q = r64[-1]
        out["fp64_reference"] = dict(status=q["status"], iterations=q["iterations"], wall_s=round(w64, 3),
                                     rel_kkt=rel_kkt_of(q), time_cap_s=a.fp64_time_limit)
        if q["status"] == "near_optimal":
            out["fp64_solves_per_s"] = round(1 / w64, 6)
            out["speedup_vs_fp64"] = round(value * w64, 4)
    if not a.no_cpu_baseline:
        log("cpu baseline sample")
        out["cpu_baseline"] = cpu_baseline(md, last["iterations"], a.cpu_sample_iters)
    return out


def run_batch(a, ffi, gen, desc, lps):
    import numpy as np
    if a.mode != "fast":
        raise SystemExit("batch workloads run the fast path (--mode fast)")
    models = [ffi.Model.from_lp(md) for md in lps]
    t0 = time.perf_counter()
    batch = ffi.Batch(models, cap=a.cap, **solver_opts(a, "fast"))
    setup = time.perf_counter() - t0
    info = batch.info()
    log(f"batch of {len(lps)}: setup {setup:.1f}s, stage {info['cap']}, autotune ms {info['tune_ms']}")
    for w in range(a.warmup):
        batch.solve(points=False)
    t0 = time.perf_counter()
    for s_ in range(a.steps):
        res = batch.solve(points=True)
        log(f"batch step {s_}: {sum(r['status'] == 'near_optimal' for r in res)} of {len(res)} near_optimal")
    wall = time.perf_counter() - t0
    value = len(lps) * a.steps / wall
    its = np.array([r["iterations"] for r in res])
    # roofline: one batched iteration with every instance live (row launch + column launch), from the autotune
    # timing of the chosen stage size (HIP events around 24 iterations on the batch stream)
    b_it = 0
    for md in lps:
        r_, c_ = halpern_kernel_bytes(md.m, md.n, md.nnz, 4)
        b_it += r_ + c_
    ti = {1024: 0, 2048: 1, 4096: 2, 8192: 3}.get(info["cap"])
    it_ms = info["tune_ms"][ti] / 24 if ti is not None and info["tune_ms"][ti] > 0 else 0.0
    achieved = b_it / (it_ms * 1e-3) / 1e9 if it_ms else 0.0
    roofline = dict(bound="hbm", kernel="one batched iteration (row + column launch, all instances live)",
                    achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4),
                    traffic=measured_traffic(a.workload, "fast", None, side=f"iter{info['cap']}"),
                    bytes_per_launch=int(b_it), avg_launch_ms=round(it_ms, 5))
    errs = [abs(r["objective"] - md.planted_objective) / (1 + abs(md.planted_objective)) for r, md in zip(res, lps)]
    out = {
        "metric": METRIC, "value": round(value, 6), "unit": "solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(1e3 * wall / a.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": DTYPES["fast"], "data": "synthetic",
        "config": {"workload": f"{len(lps)} {desc}", "count": len(lps), "nnz_total": info["nnz_total"],
                   "tol": a.tol, "mode": "fast (batched)", "stage": info["cap"], "parallelism": "single GPU"},
        "status_counts": {s: sum(r["status"] == s for r in res) for s in sorted({r["status"] for r in res})},
        "iterations": {"min": int(its.min()), "median": int(np.median(its)), "max": int(its.max())},
        "max_rel_kkt": max(rel_kkt_of(r) for r in res), "max_objective_rel_error_vs_planted": max(errs),
        "session_setup_s": round(setup, 3), "device": res[0]["device"], "roofline": roofline,
    }
    if not a.no_fp64_compare and a.fp64_sample > 0:
        # the plain FP64 cold start of the reference algorithm, one LP at a time, on the first LPs of the batch
        k = min(a.fp64_sample, len(lps))
        t0 = time.perf_counter()
        ok = 0
        for i in range(k):
            r = ffi.solve(models[i], **solver_opts(a, "fp64", time_limit=a.fp64_time_limit))
            ok += r["status"] == "near_optimal"
            log(f"fp64 {lps[i].name}: {r['status']} it={r['iterations']} {r['solve_s']:.2f}s")
        t64 = time.perf_counter() - t0
        out["fp64_reference"] = dict(sample=f"first {k} LPs of the batch, one at a time", converged=ok, wall_s=round(t64, 3),
                                     solves_per_s=round(k / t64, 4), time_cap_s=a.fp64_time_limit)
        out["speedup_vs_fp64"] = round(value / (k / t64), 3)
    if not a.no_cpu_baseline:
        log("cpu baseline sample")
        i_med = int(np.argsort([md.nnz for md in lps])[len(lps) // 2])
        cb = cpu_baseline(lps[i_med], max(int(its[i_med]), 1), a.cpu_sample_iters)
        # CPU time of the whole batch: every LP's iterations at the sample's per-nonzero iteration cost
        per_nnz = cb["ms_per_iter"] * 1e-3 / lps[i_med].nnz
        total = sum(int(r["iterations"]) * md.nnz * per_nnz for r, md in zip(res, lps))
        cb["value"] = len(lps) / total
        cb["sample"] += f"; whole batch = sum of each LP's GPU iteration count x its nonzeros at that per-nonzero cost"
        out["cpu_baseline"] = cb
    return out


def run_scenarios(a, ffi, gen, desc, fam):
    import numpy as np
    from taral_pdhg import learned
    if a.mode != "fast":
        raise SystemExit("scenario workloads run the fast path (--mode fast)")
    th_tr, th_te = fam.thetas(a.train_scenarios, seed=1), fam.thetas(a.test_scenarios, seed=2)
    bs = min(a.scenario_batch, a.test_scenarios)
    extra = {} if a.omega_rule is None else {"omega_rule": a.omega_rule}
    # training data: cold-start PDHG (the batched fast path) on the training scenarios
    rtr, s_tr = learned.solve_family(fam, th_tr, batch_size=a.scenario_batch,
                                     log=lambda s_: log(f"training data: {s_}"), progress_interval=a.progress,
                                     **extra)
    X = np.array([r["x"] for r in rtr]); Y = np.array([r["y"] for r in rtr])
    W = np.array([r["primal_weight"] for r in rtr])
    wm = learned.WarmStartModel(fam, seed=1)
    fit = wm.fit(th_tr, X, Y, W, epochs=a.epochs)
    log(f"network trained on {len(th_tr)} scenarios: {fit}")
    solver = learned.ScenarioSolver(fam, bs, progress_interval=a.progress, warm_fallback=1, **extra)
    info = solver.batch.info()

    def held_out(warm):
        res, t0 = [], time.perf_counter()
        for i0 in range(0, len(th_te), bs):
            th = th_te[i0:i0 + bs]
            if warm:
                p = wm.predict(th)
                res += solver.solve(th, p[0], p[1], p[2])
            else:
                res += solver.solve(th)
        return res, time.perf_counter() - t0

    cold, s_cold = held_out(False)
    log(f"held-out cold: {len(th_te) / s_cold:.2f} solves/s")
    for w in range(a.warmup):
        held_out(True)
    t0 = time.perf_counter()
    for s_ in range(a.steps):
        res, _ = held_out(True)
        log(f"held-out learned step {s_}: {sum(r['status'] == 'near_optimal' for r in res)} near_optimal")
    wall = time.perf_counter() - t0
    value = len(th_te) * a.steps / wall
    it_w = np.array([r["iterations"] for r in res]); it_c = np.array([r["iterations"] for r in cold])
    # algorithmic bytes of one batched iteration: the shared matrix (A and A^T: values, indices, row pointers) once,
    # every scenario's vectors
    m0, n0, z0 = fam.base.m, fam.base.n, fam.base.nnz
    r_, c_ = halpern_kernel_bytes(m0, n0, z0, 4)
    b_mat = 2 * 8 * z0 + 4 * (m0 + 1) + 4 * (n0 + 1)
    b_it = b_mat + bs * (r_ + c_ - b_mat)
    ti = {1024: 0, 2048: 1, 4096: 2, 8192: 3}.get(info["cap"])
    it_ms = info["tune_ms"][ti] / 24 if ti is not None and info["tune_ms"][ti] > 0 else 0.0
    achieved = b_it / (it_ms * 1e-3) / 1e9 if it_ms else 0.0
    out = {
        "metric": METRIC, "value": round(value, 6), "unit": "solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(1e3 * wall / a.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": DTYPES["fast"] + "; network bf16 (MFMA) / f32",
        "data": "synthetic",
        "config": {"workload": f"{desc}; base m={fam.base.m} n={fam.base.n}; {len(th_tr)} training and "
                               f"{len(th_te)} held-out scenarios", "tol": a.tol,
                   "mode": "learned warm start + fast path, batches of %d" % bs, "parallelism": "single GPU"},
        "learned": {"training_solves_s": round(s_tr, 3), "train_s": round(fit["train_s"], 3),
                    "bases": fit["bases"], "explained_variance": round(fit["explained_variance"], 4),
                    "warm_used": int(sum(r["warm_start_used"] for r in res)),
                    "median_warm_kkt": float(np.median([r["warm_kkt"] for r in res])),
                    "median_cold_kkt": float(np.median([r["cold_kkt"] for r in res]))},
        "cold_batched_solves_per_s": round(len(th_te) / s_cold, 4),
        "learned_gain_vs_cold_batched": round(value * s_cold / len(th_te), 3),
        "iterations": {"learned_mean": float(it_w.mean()), "cold_mean": float(it_c.mean()),
                       "learned_max": int(it_w.max()), "cold_max": int(it_c.max())},
        "status_counts": {s2: sum(r["status"] == s2 for r in res) for s2 in sorted({r["status"] for r in res})},
        "max_rel_kkt": max(rel_kkt_of(r) for r in res),
        "roofline": dict(bound="hbm", kernel=f"one batched iteration of {bs} scenarios (row + column launch, all "
                                             "live): the shared matrix counted once, every scenario's vectors",
                         achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                         frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None, bytes_per_launch=int(b_it),
                         avg_launch_ms=round(it_ms, 5)),
    }
    if not a.no_fp64_compare and a.fp64_sample > 0:
        k = min(a.fp64_sample, len(th_te))
        t0 = time.perf_counter()
        ok = 0
        for i in range(k):
            r = ffi.solve(ffi.Model.from_lp(fam.model(th_te[i])), **solver_opts(a, "fp64", time_limit=a.fp64_time_limit))
            ok += r["status"] == "near_optimal"
            log(f"fp64 scenario {i}: {r['status']} it={r['iterations']} {r['solve_s']:.2f}s")
        t64 = time.perf_counter() - t0
        out["fp64_reference"] = dict(sample=f"first {k} held-out scenarios, one at a time, cold start", converged=ok,
                                     wall_s=round(t64, 3), solves_per_s=round(k / t64, 4),
                                     time_cap_s=a.fp64_time_limit)
        out["speedup_vs_fp64"] = round(value / (k / t64), 3)
    if not a.no_cpu_baseline:
        log("cpu baseline sample")
        out["cpu_baseline"] = cpu_baseline(fam.model(th_te[0]), max(int(it_c.mean()), 1), a.cpu_sample_iters)
        out["cpu_baseline"]["sample"] += " (the held-out scenarios' mean cold iteration count)"
    return out


DTYPES = {"fp64": "f64", "fp32": "f32 (f64 reductions)", "fast": "f32 iterates + f64 refinement",
          "halpern-fp64": "f64"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--workload", default="cfg1", choices=sorted(WORKLOADS))
    ap.add_argument("--mode", default="fast", choices=sorted(MODES))
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--time-limit", type=float, default=600.0)
    ap.add_argument("--max-iter", type=int, default=2_000_000_000,
                    help="iteration cap per solve (profiling runs; the reported value then is not a full solve)")
    ap.add_argument("--profile-iters", type=int, default=4096,
                    help="iterations of the separate HIP-event-timed solve behind the roofline figure")
    ap.add_argument("--no-fp64-compare", action="store_true")
    ap.add_argument("--fp64-time-limit", type=float, default=120.0,
                    help="time cap of each plain FP64 reference-algorithm solve reported beside the timed mode")
    ap.add_argument("--fp64-sample", type=int, default=16, help="batch workloads: LPs solved one by one in FP64")
    ap.add_argument("--batch-count", type=int, default=512, help="batch workloads: number of LPs")
    ap.add_argument("--cap", type=int, default=0, help="batch row-block stage (0 = autotune)")
    ap.add_argument("--reflection", type=float, default=None,
                    help="fast path: Halpern reflection coefficient in [1, 2) (default: the library's)")
    ap.add_argument("--omega-rule", type=int, default=None,
                    help="fast path: primal-weight guard (0 reference update, 1 bounded, 2 balance, 3 both; "
                         "default: the library's, 3)")
    ap.add_argument("--scenario-m", type=int, default=5000,
                    help="scenario workloads: rows of the base LP (n = 2m); the configuration's 50,000 needs ~1.5M "
                         "iterations per scenario, so the default is a tenth of it")
    ap.add_argument("--train-scenarios", type=int, default=2048)
    ap.add_argument("--test-scenarios", type=int, default=512)
    ap.add_argument("--scenario-batch", type=int, default=512)
    ap.add_argument("--epochs", type=int, default=100)
    ap.add_argument("--verbose", action="store_true", help="solver progress on stderr every 50 evaluations")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch kernels directly instead of replaying HIP graphs (used under rocprofv3)")
    ap.add_argument("--progress", type=float, default=30.0,
                    help="seconds between the solver's progress lines on stderr (0 = none)")
    ap.add_argument("--cpu-sample-iters", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="only time a plain device-to-device copy of this many GiB (3 times) and exit: the "
                         "calibration kernel for the FETCH_SIZE / WRITE_SIZE counters")
    a = ap.parse_args()
    if a.copy_probe_gib > 0:
        import torch
        nbytes = int(a.copy_probe_gib * (1 << 30))
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / 3
        print(json.dumps({"copy_bytes": nbytes, "ms": dt * 1e3, "GBps_read_plus_write": 2 * nbytes / dt / 1e9}))
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this solver runs on one GPU; sharding the rows of A and A^T over several GPUs (with "
              "all-gathers of the iterate vectors) is not built", file=sys.stderr)
        sys.exit(2)
    from taral_pdhg import ffi, gen
    kind, desc, make = WORKLOADS[a.workload]
    t = time.perf_counter()
    data = make(gen, a)
    log(f"generated {a.workload} in {time.perf_counter() - t:.1f}s")
    run = {"single": run_single, "batch": run_batch, "scenarios": run_scenarios}[kind]
    out = run(a, ffi, gen, desc, data)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
