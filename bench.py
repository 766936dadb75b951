#!/usr/bin/env python3
"""Benchmark: LP solves/sec of the restarted PDHG solver on one MI355X, with the iteration-kernel roofline.

Workloads (--workload):
  cfg3 (default)  a batch of 512 independent planted LPs of uneven size (m ~ U[500, 20000], n = round(m U[1.5, 3]),
                  6-12 nonzeros per column per LP, seed 3), solved together by the batched fast path: the largest
                  configuration of the set that one GPU holds and the batched path the >= 5x target names.
  cfg1            the single large planted LP: m = 1,000,000 equality rows, n = 2,000,000 columns, 20 nonzeros per
                  column at uniform-random rows, N(0,1) values, 0 <= x <= 10, seed 1 (one fast solve takes about
                  two minutes).
  cfg2            the block-angular planning LP (64 blocks of 5,000 x 8,000 plus 200 linking rows), planted.
  cfg4            a scenario family of one planted LP (--scenario-m rows, n = 2m): the learned warm start is trained
                  on cold solves of --train-scenarios scenarios, then one step predicts and solves the held-out
                  scenarios (batches sharing the matrix) with the cold-start fallback.
  cfg0, mid       the 2,000 x 5,000 and 100,000 x 200,000 planted LPs.
Every solve runs to 1e-6 relative KKT (original-space primal residual, dual residual and gap: the reference's
stopping test). A "step" is one full solve from a cold start (cfg4: the learned start) of problems already resident
in HBM: power iteration, PDHG iterations, FP64 evaluations, download and the original-space check; for cfg3 one step
solves the whole batch. `value` counts only solves that ended near_optimal.

Wall-clock budget (--budget-s, from process start): the FP64 comparison and the CPU baseline run first, each capped;
warm-up then stops early once it has used a fifth of what is left, and timed steps stop when the next one would
not fit. `steps` and `warmup` report what ran (`steps_requested`, `warmup_requested` what was asked).

Printed JSON line: the contract's fields, plus
  roofline: the dominant iteration kernel's algorithmic bytes per launch / its mean launch time from HIP events on
            the solver stream. Single LPs: a separate capped solve with an event pair around every launch. Batches:
            `tpdhg_batch_probe`, all instances live, each row and column launch timed (a separately named kernel
            instantiation, so a kernel trace of the same command reports exactly these launches);
  cpu_baseline: the CPU oracle (FP64, the reference algorithm, its iteration block on the host threads partitioned
            as the reference's CPU backend) timed on a bounded sample of PDHG iterations and extrapolated to whole
            solves with the plain FP64 GPU solves' iteration counts (stated in `sample`);
  fp64_reference: the plain FP64 cold start of the reference algorithm on the GPU (single LPs: the same LP; batches:
            a stratified sample of the LPs, one at a time), each capped; both legs run the power iteration per solve.
"""
import argparse
import json
import os
import sys
import time

T_START = time.perf_counter()
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
DEFAULT_STEPS = {"cfg3": 3, "cfg1": 1, "cfg2": 3, "cfg4": 1, "cfg0": 5, "mid": 3}


def kernel_bytes(m, n, nnz, s):
    """Algorithmic HBM bytes of one launch of each fused iteration kernel of the reference scheme (s = bytes per
    value): row kernel: val+idx (s+4)/nnz, row pointers 4(m+1), x read once s*n, and per row z, Ax, rl, ru, Axs, zs
    read and Ax, z, Axs, zs written (10 s per row); column kernel likewise over A^T with 10 s per column."""
    row = (s + 4) * nnz + 4 * (m + 1) + s * n + 10 * s * m
    col = (s + 4) * nnz + 4 * (n + 1) + s * m + 10 * s * n
    return row, col


def halpern_kernel_bytes(m, n, nnz, s):
    """Same for the fast path's kernels: row kernel reads w once (s*n) and per row z, zh, rl, ru, z0 and writes z
    (6 s per row); column kernel reads z once (s*m) and per column x, x+, x0, c, l, u and writes x, x+, w
    (9 s per column; the A^T z store of a block's last step is not counted)."""
    row = (s + 4) * nnz + 4 * (m + 1) + s * n + 6 * s * m
    col = (s + 4) * nnz + 4 * (n + 1) + s * m + 9 * s * n
    return row, col


MODES = {  # name: solver options
    "fast": dict(precision_bits=32, halpern=1, refine=1),   # FP32 reflected Halpern + FP64 refinement
    "fp64": dict(precision_bits=64),                        # the reference algorithm, plain FP64 cold start
    "fp32": dict(precision_bits=32),                        # the reference algorithm with FP32 iterates
    "halpern-fp64": dict(precision_bits=64, halpern=1),
}
DTYPES = {"fp64": "f64", "fp32": "f32 (f64 reductions)", "fast": "f32 iterates + f64 refinement",
          "halpern-fp64": "f64"}
CPU_THREADS = 16  # host threads for the CPU baseline: the CPU share of one GPU on the GPU hosts


def log(msg):
    print(f"[bench {time.perf_counter() - T_START:7.1f}s] {msg}", file=sys.stderr, flush=True)


class Budget:
    def __init__(self, total_s):
        self.total = float(total_s)

    def left(self):
        return self.total - (time.perf_counter() - T_START)


def cpu_threads():
    return max(1, min(CPU_THREADS, os.cpu_count() or 1))


def cpu_iter_cost(md, sample_s, threads):
    """Seconds per PDHG iteration of the CPU oracle (FP64, the reference's iteration block on `threads` host threads
    partitioned as its CPU backend) on `md`: iterations doubled until one sample takes `sample_s` seconds."""
    from oracle import oracle
    o = oracle.OracleLP(md)
    t0 = time.perf_counter()
    o.block(0.5, 0.5, 0, threads=threads)  # start: two products
    t_start = time.perf_counter() - t0
    k = 2
    while True:
        t0 = time.perf_counter()
        o.block(0.5, 0.5, k, threads=threads)
        t = time.perf_counter() - t0 - t_start
        if t >= sample_s or k >= 1 << 16:
            return max(t, 1e-9) / k, k, t
        k *= 2


def rel_kkt_of(r):
    return max(r["rel_primal"], r["rel_dual"], r["rel_gap"])


def solver_opts(a, mode_name, **extra):
    o = dict(tol=a.tol, time_limit=a.time_limit, max_iter=a.max_iter, use_graph=0 if a.no_graph else 1,
             verbose=int(a.verbose), progress_interval=a.progress, **MODES[mode_name])
    if a.omega_rule is not None:
        o["omega_rule"] = a.omega_rule
    if a.restart_metric is not None:
        o["restart_metric"] = a.restart_metric
    if a.reflection is not None and o.get("halpern"):
        o["reflection"] = a.reflection
    o.update(extra)
    return o


def timed_steps(step, steps, warmup, budget, count_ok):
    """Warm-up (at most `warmup` steps, stopping once it has used a fifth of the budget left when it began), then up
    to `steps` timed steps while the next one still fits the budget (at least one). `step()` returns its results;
    `count_ok(results)` the near_optimal solves among them. Returns (wall, timed results, warm-ups run, solves ok)."""
    w_done, longest = 0, 0.0
    w_cap = max(0.0, budget.left()) / 5
    t_w0 = time.perf_counter()
    for w in range(warmup):
        if w > 0 and time.perf_counter() - t_w0 + longest > w_cap:
            break
        t = time.perf_counter()
        step()
        longest = max(longest, time.perf_counter() - t)
        w_done += 1
        log(f"warm-up {w} done ({time.perf_counter() - t:.2f}s)")
    results, ok = [], 0
    t0 = time.perf_counter()
    while len(results) < steps:
        if results and budget.left() < 1.25 * longest:
            log(f"budget: stopping after {len(results)} of {steps} timed steps")
            break
        t = time.perf_counter()
        r = step()
        longest = max(longest, time.perf_counter() - t)
        results.append(r)
        ok += count_ok(r)
        log(f"timed step {len(results) - 1} done ({time.perf_counter() - t:.2f}s)")
    return time.perf_counter() - t0, results, w_done, ok


def contract_line(a, value, wall, n_steps, w_done, dtype, config):
    return {"metric": METRIC, "value": round(value, 6), "unit": "solves/s", "n_gpus": 1, "steps": n_steps,
            "warmup": w_done, "steps_requested": a.steps, "warmup_requested": a.warmup,
            "ms_per_step": round(1e3 * wall / max(n_steps, 1), 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}


def roofline_from(kind, row_bytes, col_bytes, t, traffic_of):
    """Roofline entry of the dominant iteration kernel from HIP-event totals t (row/col ms and launch counts)."""
    row_avg = t["row_ms"] / max(t["row_launches"], 1)
    col_avg = t["col_ms"] / max(t["col_launches"], 1)
    if col_avg >= row_avg:
        side, dom, b, avg = "col", "column step (A^T z fused with the primal step)", col_bytes, col_avg
    else:
        side, dom, b, avg = "row", "row step (A x fused with the dual step)", row_bytes, row_avg
    achieved = b / (avg * 1e-3) / 1e9 if avg > 0 else 0.0
    return dict(bound="hbm", kernel=f"{kind}: {dom}", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic_of(side), bytes_per_launch=int(b),
                avg_launch_ms=round(avg, 5), launches=int(t[f"{side}_launches"]),
                row_avg_ms=round(row_avg, 5), col_avg_ms=round(col_avg, 5),
                row_gbs=round(row_bytes / (row_avg * 1e-3) / 1e9, 1) if row_avg else 0,
                col_gbs=round(col_bytes / (col_avg * 1e-3) / 1e9, 1) if col_avg else 0)


def measured_traffic(workload, side):
    """HBM bytes per launch of the `side` ("row" or "col") kernel from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (copy-kernel calibrated), or None when there is none."""
    path = os.path.join(ROOT, "profiles", f"pdhg_{workload}_fast_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == side:
                return int(parts[2])
    return None


def fp64_sessions(a, ffi, models, label):
    """Plain FP64 reference-algorithm solves, one model at a time, each on its own session (setup untimed), each
    capped at --fp64-time-limit, the whole leg at --fp64-budget-s. Returns per-model (status, iterations, seconds)."""
    out, t_leg = [], time.perf_counter()
    for i, md in enumerate(models):
        left = a.fp64_budget_s - (time.perf_counter() - t_leg)
        if left <= 1:
            log(f"fp64: leg budget spent after {i} of {len(models)}")
            break
        s = ffi.Session(md, **solver_opts(a, "fp64", time_limit=min(a.fp64_time_limit, left)))
        t = time.perf_counter()
        r = s.solve()
        dt = time.perf_counter() - t
        s.close()
        out.append((r["status"], int(r["iterations"]), dt))
        log(f"fp64 {label(i)}: {r['status']} it={r['iterations']} {dt:.2f}s")
    return out


def run_single(a, ffi, gen, desc, md, budget):
    model = ffi.Model.from_lp(md)
    mode = a.mode
    out_extra = {}
    # plain FP64 reference first (bounded), then the CPU sample, then the timed mode with what is left
    fp64 = None
    if mode != "fp64" and not a.no_fp64_compare:
        fp64 = fp64_sessions(a, ffi, [model], lambda i: md.name)
        if fp64:
            st, it, dt = fp64[0]
            out_extra["fp64_reference"] = dict(status=st, iterations=it, wall_s=round(dt, 3),
                                               time_cap_s=a.fp64_time_limit, power_iteration="included")
            if st == "near_optimal":
                out_extra["fp64_solves_per_s"] = round(1 / dt, 6)
    t = time.perf_counter()
    sess = ffi.Session(model, **solver_opts(a, mode))
    setup = time.perf_counter() - t
    log(f"{mode}: session setup {setup:.1f}s")

    def step():
        r = sess.solve()
        log(f"{mode}: {r['status']} it={r['iterations']} solve={r['solve_s']:.3f}s")
        return r

    wall, res, w_done, ok = timed_steps(step, a.steps, a.warmup, budget,
                                        lambda r: int(r["status"] == "near_optimal"))
    last = res[-1]
    value = ok / wall
    sess.close()
    # kernel roofline: a separate capped solve with HIP-event timing of every iteration launch (no graphs)
    sp = ffi.Session(model, **solver_opts(a, mode, profile_kernels=1, max_iter=min(a.max_iter, a.profile_iters)))
    q = sp.solve()
    sp.close()
    s_val = 8 if MODES[mode]["precision_bits"] == 64 else 4
    b_row, b_col = (halpern_kernel_bytes if MODES[mode].get("halpern") else kernel_bytes)(md.m, md.n, q["nnz"], s_val)
    t = dict(row_ms=q["row_kernel_ms"], col_ms=q["col_kernel_ms"], row_launches=q["row_launches"],
             col_launches=q["col_launches"])
    roofline = roofline_from("one LP", b_row, b_col, t, lambda side: measured_traffic(a.workload, side))
    config = {"workload": desc, "m": md.m, "n": md.n, "nnz": last["nnz"], "tol": a.tol, "mode": mode,
              "parallelism": "single GPU"}
    out = contract_line(a, value, wall, len(res), w_done, DTYPES[mode], config)
    out.update({"solves_near_optimal": ok, "iterations": last["iterations"], "restarts": last["restarts"],
                "refinements": last["refinements"], "status": last["status"], "rel_kkt": rel_kkt_of(last),
                "iters_per_s": round(last["iterations"] / last["solve_s"], 1), "session_setup_s": round(setup, 3),
                "power_iteration": "included in every timed solve", "device": last["device"],
                "roofline": roofline})
    out.update(out_extra)
    if "fp64_solves_per_s" in out:
        out["speedup_vs_fp64"] = round(value / out["fp64_solves_per_s"], 4)
    if md.planted_objective is not None:
        p_star = md.planted_objective
        out["objective_rel_error_vs_planted"] = abs(last["objective"] - p_star) / (1 + abs(p_star))
    if not a.no_cpu_baseline:
        nt = cpu_threads()
        per, k, ts = cpu_iter_cost(md, a.cpu_sample_s, nt)
        use64 = fp64 and fp64[0][0] == "near_optimal"
        its = fp64[0][1] if use64 else last["iterations"]
        src = "the plain FP64 GPU solve's" if use64 else "the timed GPU solve's (the FP64 solve did not finish)"
        out["cpu_baseline"] = dict(
            value=1.0 / (per * its), unit="solves/s", cores=nt, kind="port", host_nproc=os.cpu_count(),
            extrapolated=True, ms_per_iter=round(per * 1e3, 4),
            sample=f"CPU oracle (oracle/pdhg_oracle.c, FP64, the reference algorithm's iteration block on {nt} host "
                   f"threads partitioned as its CPU backend): {k} iterations of the same LP took {ts:.2f}s; "
                   f"extrapolated to {its} iterations ({src} count)")
    return out


def stratified(nnzs, k, seed=7):
    """k LP indices, one from each of k equal slices of the LPs sorted by nonzeros (seeded pick inside a slice)."""
    import numpy as np
    order = np.argsort(nnzs, kind="stable")
    rng = np.random.default_rng(seed)
    cuts = np.linspace(0, len(order), k + 1).astype(int)
    return [int(order[rng.integers(cuts[i], cuts[i + 1])]) for i in range(k) if cuts[i + 1] > cuts[i]]


def run_batch(a, ffi, gen, desc, lps, budget):
    import numpy as np
    if a.mode != "fast":
        raise SystemExit("batch workloads run the fast path (--mode fast)")
    models = [ffi.Model.from_lp(md) for md in lps]
    nnzs = [md.nnz for md in lps]
    k = len(lps) if a.fp64_sample <= 0 else min(a.fp64_sample, len(lps))
    sample = list(range(len(lps))) if k == len(lps) else stratified(nnzs, k)
    fp64 = None
    if not a.no_fp64_compare:
        fp64 = fp64_sessions(a, ffi, [models[i] for i in sample], lambda j: lps[sample[j]].name)
    t0 = time.perf_counter()
    batch = ffi.Batch(models, cap=a.cap, **solver_opts(a, "fast"))
    setup = time.perf_counter() - t0
    info = batch.info()
    log(f"batch of {len(lps)}: setup {setup:.1f}s, stage {info['cap']}, autotune ms {info['tune_ms']}")
    cpu = None
    if not a.no_cpu_baseline:  # per-iteration CPU cost on the smallest, median and largest LP of the FP64 sample
        nt = cpu_threads()
        ref = sample if fp64 is None else sample[:len(fp64)]
        by = sorted(ref, key=lambda i: nnzs[i])
        picks = sorted({by[0], by[len(by) // 2], by[-1]})
        costs = {}
        for i in picks:
            per, kk, ts = cpu_iter_cost(lps[i], a.cpu_sample_s / len(picks), nt)
            costs[i] = per
            log(f"cpu {lps[i].name}: {per * 1e3:.3f} ms/iter ({kk} iterations in {ts:.2f}s)")
        cpu = (nt, costs)

    def step():
        res = batch.solve(points=True)
        log(f"batch: {sum(r['status'] == 'near_optimal' for r in res)} of {len(res)} near_optimal")
        return res

    wall, runs, w_done, ok = timed_steps(step, a.steps, a.warmup, budget,
                                        lambda res: sum(r["status"] == "near_optimal" for r in res))
    res = runs[-1]
    value = ok / wall
    its = np.array([r["iterations"] for r in res])
    info_after = batch.info()
    # roofline: one batched iteration with every instance live, both launches timed apart
    t = batch.probe(a.probe_iters)
    b_row = b_col = 0
    for md, r in zip(lps, res):
        r_, c_ = halpern_kernel_bytes(md.m, md.n, r["nnz"], 4)
        b_row += r_
        b_col += c_
    roofline = roofline_from(f"{len(lps)} LPs all live, stage {info['cap']}", b_row, b_col, t,
                             lambda side: None if a.batch_only else measured_traffic(a.workload, side))
    errs = [abs(r["objective"] - md.planted_objective) / (1 + abs(md.planted_objective)) for r, md in zip(res, lps)]
    config = {"workload": f"{len(lps)} {desc}", "count": len(lps), "nnz_total": info["nnz_total"], "tol": a.tol,
              "mode": "fast (batched)", "stage": info["cap"], "parallelism": "single GPU"}
    out = contract_line(a, value, wall, len(runs), w_done, DTYPES["fast"], config)
    out.update({
        "solves_near_optimal": ok, "solves_attempted": len(lps) * len(runs),
        "status_counts": {s: sum(r["status"] == s for r in res) for s in sorted({r["status"] for r in res})},
        "iterations": {"min": int(its.min()), "median": int(np.median(its)), "max": int(its.max()),
                       "slowest": lps[int(its.argmax())].name},
        "max_rel_kkt": max(rel_kkt_of(r) for r in res), "max_objective_rel_error_vs_planted": max(errs),
        "session_setup_s": round(setup, 3), "power_iteration": "included in every timed solve",
        "stage_switches": info_after["stage_switches"], "device": res[0]["device"], "roofline": roofline})
    if fp64:
        done = [sample[j] for j in range(len(fp64))]
        tot = sum(dt for _, _, dt in fp64)
        conv = sum(st == "near_optimal" for st, _, _ in fp64)
        rate = len(fp64) / tot
        out["fp64_reference"] = dict(
            sample=(f"all {len(lps)} LPs" if len(sample) == len(lps) else
                    f"{len(sample)} of {len(lps)} LPs, one from each nonzero-count slice (stratified, seed 7)") +
                   (f" ({len(fp64)} solved before the leg budget ran out)" if len(fp64) < len(sample) else "") +
                   ", solved one at a time (each on its own resident session; power iteration included)",
            lps=[lps[i].name for i in done], converged=conv, wall_s=round(tot, 3), solves_per_s=round(rate, 4),
            time_cap_s=a.fp64_time_limit,
            note="solves that hit the cap count at the cap, so the FP64 rate is an upper bound when converged < sample")
        out["speedup_vs_fp64"] = round(value / rate, 3)
    if cpu:
        nt, costs = cpu
        xs = np.array([nnzs[i] for i in costs], dtype=float)
        ys = np.array([costs[i] for i in costs])
        slope, icpt = np.polyfit(xs, ys, 1) if len(xs) > 1 else (ys[0] / xs[0], 0.0)
        cost = lambda i: max(icpt + slope * nnzs[i], 1e-9)  # noqa: E731
        if fp64:
            est = sum(it * cost(sample[j]) for j, (_, it, _) in enumerate(fp64)) / len(fp64)
            src = f"the plain FP64 GPU solves' iteration counts on the {len(fp64)} sampled LPs"
        else:
            est = float(np.mean([int(r["iterations"]) * cost(i) for i, r in enumerate(res)]))
            src = "the fast-path GPU solves' iteration counts"
        out["cpu_baseline"] = dict(
            value=1.0 / est, unit="solves/s", cores=nt, kind="port", host_nproc=os.cpu_count(), extrapolated=True,
            sample=f"CPU oracle (oracle/pdhg_oracle.c, FP64, the reference algorithm's iteration block on {nt} host "
                   f"threads partitioned as its CPU backend): seconds per iteration timed on "
                   f"{', '.join(lps[i].name for i in costs)} and fitted linearly in nonzeros; per-LP solve time = "
                   f"that cost x {src}; value = 1 / mean per-LP time")
    return out


def run_scenarios(a, ffi, gen, desc, fam, budget):
    import numpy as np
    from taral_pdhg import learned
    if a.mode != "fast":
        raise SystemExit("scenario workloads run the fast path (--mode fast)")
    th_tr, th_te = fam.thetas(a.train_scenarios, seed=1), fam.thetas(a.test_scenarios, seed=2)
    bs = min(a.scenario_batch, a.test_scenarios)
    extra = {} if a.omega_rule is None else {"omega_rule": a.omega_rule}
    if a.restart_metric is not None:
        extra["restart_metric"] = a.restart_metric
    fp64 = None
    if not a.no_fp64_compare and a.fp64_sample > 0:
        # plain FP64 cold starts of held-out scenarios on one session of the base model (its norm reused, as the
        # batched scenario solver reuses it)
        s = ffi.Session(ffi.Model.from_lp(fam.base), **solver_opts(a, "fp64", reuse_norm=1))
        fp64, t_leg = [], time.perf_counter()
        for i in range(min(a.fp64_sample, len(th_te))):
            left = a.fp64_budget_s - (time.perf_counter() - t_leg)
            if left <= 1:
                break
            s.set_vectors(*fam.vectors(th_te[i]))
            t = time.perf_counter()
            r = s.solve()
            fp64.append((r["status"], int(r["iterations"]), time.perf_counter() - t))
            log(f"fp64 scenario {i}: {r['status']} it={r['iterations']} {fp64[-1][2]:.2f}s")
        s.close()
    rtr, s_tr = learned.solve_family(fam, th_tr, batch_size=a.scenario_batch,
                                     log=lambda s_: log(f"training data: {s_}"), progress_interval=a.progress,
                                     **extra)
    X = np.array([r["x"] for r in rtr])
    Y = np.array([r["y"] for r in rtr])
    W = np.array([r["primal_weight"] for r in rtr])
    wm = learned.WarmStartModel(fam, seed=1)
    fit = wm.fit(th_tr, X, Y, W, epochs=a.epochs)
    log(f"network trained on {len(th_tr)} scenarios: {fit}")
    solver = learned.ScenarioSolver(fam, bs, progress_interval=a.progress, warm_fallback=1, **extra)

    def held_out(warm):
        res = []
        for i0 in range(0, len(th_te), bs):
            th = th_te[i0:i0 + bs]
            if warm:
                p = wm.predict(th)
                res += solver.solve(th, p[0], p[1], p[2])
            else:
                res += solver.solve(th)
        return res

    t = time.perf_counter()
    cold = held_out(False)
    s_cold = time.perf_counter() - t
    log(f"held-out cold: {len(th_te) / s_cold:.2f} solves/s")
    wall, runs, w_done, ok = timed_steps(lambda: held_out(True), a.steps, a.warmup, budget,
                                        lambda res: sum(r["status"] == "near_optimal" for r in res))
    res = runs[-1]
    value = ok / wall
    it_w = np.array([r["iterations"] for r in res])
    it_c = np.array([r["iterations"] for r in cold])
    t_probe = solver.batch.probe(a.probe_iters)
    m0, n0, z0 = fam.base.m, fam.base.n, fam.base.nnz
    r_, c_ = halpern_kernel_bytes(m0, n0, z0, 4)
    # the shared matrix (values, indices, row pointers) is read once per launch; every scenario's vectors
    b_row = (8 * z0 + 4 * (m0 + 1)) + bs * (r_ - 8 * z0 - 4 * (m0 + 1))
    b_col = (8 * z0 + 4 * (n0 + 1)) + bs * (c_ - 8 * z0 - 4 * (n0 + 1))
    roofline = roofline_from(f"{bs} scenarios all live (the shared matrix counted once)", b_row, b_col, t_probe,
                             lambda side: None)
    config = {"workload": f"{desc}; base m={fam.base.m} n={fam.base.n}; {len(th_tr)} training and {len(th_te)} "
                          f"held-out scenarios", "tol": a.tol,
              "mode": "learned warm start + fast path, batches of %d" % bs, "parallelism": "single GPU"}
    out = contract_line(a, value, wall, len(runs), w_done, DTYPES["fast"] + "; network bf16 (MFMA) / f32", config)
    out.update({
        "solves_near_optimal": ok,
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
        "max_rel_kkt": max(rel_kkt_of(r) for r in res), "roofline": roofline})
    if fp64:
        tot = sum(dt for _, _, dt in fp64)
        out["fp64_reference"] = dict(sample=f"first {len(fp64)} held-out scenarios, one at a time, cold start",
                                     converged=sum(st == "near_optimal" for st, _, _ in fp64), wall_s=round(tot, 3),
                                     solves_per_s=round(len(fp64) / tot, 4), time_cap_s=a.fp64_time_limit)
        out["speedup_vs_fp64"] = round(value / (len(fp64) / tot), 3)
    if not a.no_cpu_baseline:
        nt = cpu_threads()
        per, k, ts = cpu_iter_cost(fam.model(th_te[0]), a.cpu_sample_s, nt)
        its = int(np.mean([it for _, it, _ in fp64])) if fp64 else int(it_c.mean())
        out["cpu_baseline"] = dict(
            value=1.0 / (per * its), unit="solves/s", cores=nt, kind="port", host_nproc=os.cpu_count(),
            extrapolated=True,
            sample=f"CPU oracle (FP64, {nt} threads): {k} iterations of a held-out scenario took {ts:.2f}s; "
                   f"extrapolated to {its} iterations (mean of the {'FP64 GPU solves' if fp64 else 'cold solves'})")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cfg3", choices=sorted(WORKLOADS))
    ap.add_argument("--mode", default="fast", choices=sorted(MODES))
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--budget-s", type=float, default=480.0,
                    help="wall-clock budget of the whole run from process start (warm-up and timed steps stop in time)")
    ap.add_argument("--time-limit", type=float, default=600.0)
    ap.add_argument("--max-iter", type=int, default=2_000_000_000,
                    help="iteration cap per solve (profiling runs; the reported value then is not a full solve)")
    ap.add_argument("--profile-iters", type=int, default=4096,
                    help="single LPs: iterations of the separate HIP-event-timed solve behind the roofline figure")
    ap.add_argument("--probe-iters", type=int, default=24, help="batches: all-live iterations timed for the roofline")
    ap.add_argument("--no-fp64-compare", action="store_true")
    ap.add_argument("--fp64-time-limit", type=float, default=30.0, help="cap of each plain FP64 reference solve")
    ap.add_argument("--fp64-budget-s", type=float, default=60.0, help="cap of the whole FP64 comparison leg")
    ap.add_argument("--fp64-sample", type=int, default=16,
                    help="batches: LPs solved one at a time in FP64 (stratified by nonzeros; 0 = all)")
    ap.add_argument("--batch-count", type=int, default=512, help="batch workloads: number of LPs")
    ap.add_argument("--cap", type=int, default=0, help="batch row-block stage (0 = autotune)")
    ap.add_argument("--batch-only", default="", help="batch workloads: comma-separated LP indices to keep")
    ap.add_argument("--reflection", type=float, default=None,
                    help="fast path: Halpern reflection coefficient in [1, 2) (default: the library's)")
    ap.add_argument("--omega-rule", type=int, default=None,
                    help="fast path: primal-weight guard (0 reference update, 1 bounded, 2 balance, 3 both, 4 both "
                         "with the bound anchored at the first restart's weight; default: the library's, 3)")
    ap.add_argument("--restart-metric", type=int, default=None,
                    help="fast path: what the restart tests measure (1 the fixed-point residual, 0 the weighted KKT "
                         "error; default: the library's, 1)")
    ap.add_argument("--scenario-m", type=int, default=5000,
                    help="scenario workloads: rows of the base LP (n = 2m); the configuration's 50,000 needs ~1.5M "
                         "iterations per scenario")
    ap.add_argument("--train-scenarios", type=int, default=2048)
    ap.add_argument("--test-scenarios", type=int, default=512)
    ap.add_argument("--scenario-batch", type=int, default=512)
    ap.add_argument("--epochs", type=int, default=100)
    ap.add_argument("--verbose", action="store_true", help="solver progress on stderr every 50 evaluations")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch kernels directly instead of replaying HIP graphs")
    ap.add_argument("--progress", type=float, default=30.0,
                    help="seconds between the solver's progress lines on stderr (0 = none)")
    ap.add_argument("--cpu-sample-s", type=float, default=12.0, help="seconds of CPU oracle iterations timed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ceiling", action="store_true", help="skip the streaming read/copy ceiling measurement")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this solver runs on one GPU; sharding the rows of A and A^T over several GPUs (with "
              "all-gathers of the iterate vectors) is not built", file=sys.stderr)
        sys.exit(2)
    if a.steps is None:
        a.steps = DEFAULT_STEPS[a.workload]
    if a.steps < 1 or a.warmup < 0:
        raise SystemExit("--steps must be >= 1 and --warmup >= 0")
    budget = Budget(a.budget_s)
    from taral_pdhg import ffi, gen
    kind, desc, make = WORKLOADS[a.workload]
    t = time.perf_counter()
    data = make(gen, a)
    if a.batch_only:  # a subset of a batch workload's LPs (by index), e.g. its slowest member alone
        if kind != "batch":
            raise SystemExit("--batch-only applies to batch workloads")
        pick = [int(v) for v in a.batch_only.split(",")]
        data = [data[i] for i in pick]
        desc += f"; only LPs {a.batch_only} of {a.batch_count}"
    log(f"generated {a.workload} in {time.perf_counter() - t:.1f}s")
    run = {"single": run_single, "batch": run_batch, "scenarios": run_scenarios}[kind]
    out = run(a, ffi, gen, desc, data, budget)
    if not a.no_ceiling:  # the box's streaming ceiling with the library's own read and copy kernels (1 GiB)
        out["hbm_ceiling"] = {k: round(v, 1) for k, v in ffi.stream_bench(1 << 30, 10).items()}
    out["budget_s"] = a.budget_s
    out["bench_wall_s"] = round(time.perf_counter() - T_START, 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
