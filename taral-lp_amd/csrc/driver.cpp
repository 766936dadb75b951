// Host driver of the restarted averaged PDHG (gpu/pdhg.cu:743-849) over the HIP device backend, and the
// solve pipeline with its original-space report (gpu/pdhg.cu:851-911).
#include "driver.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>

namespace tpdhg {

namespace {
using Clock = std::chrono::steady_clock;
double since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

struct Rel {
    double p, d, g;
    double worst() const { return std::max(p, std::max(d, g)); }
};

// True once every o.progress_s seconds (never when 0): the solve loop then prints a progress line on stderr.
bool progress_due(const Options& o, Clock::time_point& last) {
    if (!(o.progress_s > 0) || since(last) < o.progress_s) return false;
    last = Clock::now();
    return true;
}

// Original-space relative residuals of one evaluation (gpu/pdhg.cu:791-796).
Rel rel_of(const double* k, const Scaled& P, double row_rel) {
    const double po = k[POBJ], du = k[DROW] + k[DCOL];
    const double prow = row_rel > 0 ? std::sqrt(k[PR2R]) / row_rel : 0;
    return Rel{std::max(std::sqrt(k[PR2O]) / (1 + P.bnorm), prow), std::sqrt(k[DR2O]) / (1 + P.cnorm),
               std::fabs(po - du) / (1 + std::fabs(po) + std::fabs(du))};
}
}  // namespace

LP flatten(const Model& md) {
    LP lp;
    lp.name = md.name;
    lp.m = int(md.row_lo.size()), lp.n = int(md.cols.size());
    lp.colptr.assign(size_t(lp.n) + 1, 0);
    for (int j = 0; j < lp.n; ++j) lp.colptr[size_t(j) + 1] = lp.colptr[size_t(j)] + int64_t(md.cols[size_t(j)].size());
    lp.rowidx.resize(size_t(lp.colptr.back())), lp.val.resize(size_t(lp.colptr.back()));
    for (int j = 0; j < lp.n; ++j) {
        size_t k = size_t(lp.colptr[size_t(j)]);
        for (const Entry& e : md.cols[size_t(j)]) lp.rowidx[k] = e.index, lp.val[k] = e.value, ++k;
    }
    lp.cost = md.cost, lp.col_lo = md.col_lo, lp.col_up = md.col_up, lp.row_lo = md.row_lo, lp.row_up = md.row_up;
    lp.row_names = md.row_names, lp.col_names = md.col_names, lp.is_int = md.is_int;
    lp.qobj_terms = md.qobj.size(), lp.sentinel_bounds = md.sentinel_bounds;
    lp.obj_const = md.obj_const, lp.maximize = md.maximize;
    return lp;
}

const char* input_error(const LP& lp) {
    Model probe;  // the gate reads only these two properties
    probe.is_int = lp.is_int;
    if (lp.qobj_terms) probe.qobj.push_back({0, 0, 1.0});
    return pdhg_input_error(probe);
}

Outcome run_pdhg(Device& b, const Scaled& P, const Options& o, double time_budget, const double* x0, const double* z0,
                 double omega0, double norm_a) {
    Outcome r;
    auto t0 = Clock::now(), t_progress = t0;
    b.set_profile(o.profile_kernels);
    r.norm_a = norm_a > 0 ? norm_a : b.power_norm(200);
    r.power_s = since(t0);
    const double eta = r.norm_a > 0 ? 0.95 / r.norm_a : 1.0;
    double cs = 0, bs = 0;  // initial primal weight from the scaled cost and row bounds (gpu/pdhg.cu:768-779)
    for (double v : P.c) cs += v * v;
    for (int i = 0; i < P.m; ++i) {
        const double v = row_bound_measure(P.rl[size_t(i)], P.ru[size_t(i)]);
        bs += v * v;
    }
    cs = std::sqrt(cs), bs = std::sqrt(bs);
    double w = cs > 1e-10 && bs > 1e-10 ? cs / bs : 1.0;
    if (omega0 > 0) w = omega0;
    b.set_steps(eta / w, eta * w);
    b.start(x0, z0);

    auto wkkt = [&](const double* k) {  // KKT error with primal weight, scaled space (gpu/pdhg.cu:783-786)
        const double g = k[POBJ] - (k[DROW] + k[DCOL]);
        return std::sqrt(w * k[PR2S] + k[DR2S] / w + g * g);
    };
    auto rel = [&](const double* k) { return rel_of(k, P, o.row_rel); };
    double kc[NKKT], ka[NKKT];
    b.eval2(1.0, kc, nullptr);
    double k_last = wkkt(kc), k_prev = kInf;
    int64_t it = 0, n_since = 0;
    bool avg_final = false;
    Rel best{kInf, kInf, kInf};
    for (;;) {
        const int steps = int(std::min<int64_t>(o.eval_freq, o.max_iter - it));
        if (steps <= 0) break;
        b.iterate(steps);
        it += steps, n_since += steps;
        const double f = 1.0 / double(n_since);
        b.eval2(f, kc, ka);
        const Rel rc = rel(kc), ra = rel(ka);
        const bool use_avg = ra.worst() < rc.worst();
        best = use_avg ? ra : rc, avg_final = use_avg;
        if ((o.verbose && (it / o.eval_freq) % 50 == 0) || progress_due(o, t_progress))
            std::fprintf(stderr, "it %ld  w %.3e  cur %.2e %.2e %.2e  avg %.2e %.2e %.2e  pobj %.10g\n", long(it), w,
                         rc.p, rc.d, rc.g, ra.p, ra.d, ra.g, kc[POBJ]);
        if (!std::isfinite(rc.worst()) && !std::isfinite(ra.worst())) {
            r.status = "numerical_failure";
            break;
        }
        if (best.worst() <= o.tol) {
            r.status = "near_optimal";
            break;
        }
        if (since(t0) > time_budget) {
            r.status = "time_limit";
            break;
        }
        if (it >= o.max_iter) break;
        // Adaptive restart to the better of current and average (gpu/pdhg.cu:827-842).
        const double wc = wkkt(kc), wa = wkkt(ka);
        const bool cand_avg = wa < wc;
        const double* kk = cand_avg ? ka : kc;
        const double kcand = std::min(wa, wc);
        const bool restart = double(n_since) >= 0.36 * double(it) || kcand <= 0.2 * k_last ||
                             (kcand <= 0.8 * k_last && kcand > k_prev);
        k_prev = kcand;
        if (restart) {
            const double dx = std::sqrt(kk[DX2]), dz = std::sqrt(kk[DZ2]);
            b.restart(cand_avg, f);
            if (dx > 1e-10 && dz > 1e-10) {
                w = std::exp(0.5 * std::log(dz / dx) + 0.5 * std::log(w));
                b.set_steps(eta / w, eta * w);
            }
            k_last = wkkt(kk), k_prev = kInf, n_since = 0, r.restarts++;
        }
    }
    r.iterations = it;
    r.rel_primal = best.p, r.rel_dual = best.d, r.rel_gap = best.g, r.omega = w, r.from_average = avg_final;
    r.x.resize(size_t(P.n)), r.z.resize(size_t(P.m));
    b.download(avg_final, 1.0 / double(std::max<int64_t>(n_since, 1)), r.x.data(), r.z.data());
    r.iterate_s = since(t0) - r.power_s;
    r.kernels = b.kernel_times();
    return r;
}

Outcome run_fast(FastDevice& b, const Scaled& P, const Options& o, double time_budget, const double* x0, const double* z0,
                 double omega0, double norm_a) {
    Outcome r;
    auto t0 = Clock::now(), t_progress = t0;
    b.set_profile(o.profile_kernels);
    r.norm_a = norm_a > 0 ? norm_a : b.power_norm(200);
    r.power_s = since(t0);
    const double eta = r.norm_a > 0 ? 0.95 / r.norm_a : 1.0;
    double cs = 0, bs = 0;
    for (double v : P.c) cs += v * v;
    for (int i = 0; i < P.m; ++i) {
        const double v = row_bound_measure(P.rl[size_t(i)], P.ru[size_t(i)]);
        bs += v * v;
    }
    cs = std::sqrt(cs), bs = std::sqrt(bs);
    double w = cs > 1e-10 && bs > 1e-10 ? cs / bs : 1.0;
    if (omega0 > 0) w = omega0;
    b.set_base(x0, z0);
    auto wkkt = [&](const double* k) {
        const double g = k[POBJ] - (k[DROW] + k[DCOL]);
        return std::sqrt(w * k[PR2S] + k[DR2S] / w + g * g);
    };
    auto rel = [&](const double* k) { return rel_of(k, P, o.row_rel); };
    double kk[NKKT];
    int64_t it = 0, n_since = 0;
    b.begin_round();
    b.set_steps(eta / w, eta * w);
    b.restart();
    b.eval(kk);
    Rel best = rel(kk);
    double k_last = wkkt(kk), k_prev = kInf;
    // Each recentring folds the inner point into the FP64 base and rebuilds the inner problem with exact FP64
    // products; the next inner solve then only spans one restart cycle, so its FP32 rounding stays relative to
    // that cycle's (shrinking) distance instead of to the distance from the start.
    auto recentre = [&]() {
        b.end_round();
        b.begin_round();
        b.set_steps(eta / w, eta * w);
        b.restart();
    };
    if (best.worst() <= o.tol) r.status = "near_optimal";
    while (r.status != "near_optimal") {
        const int steps = int(std::min<int64_t>(o.eval_freq, o.max_iter - it));
        if (steps <= 0) break;
        b.iterate(steps);
        it += steps, n_since += steps;
        b.eval(kk);
        const Rel rc = rel(kk);
        best = rc;
        if ((o.verbose && (it / o.eval_freq) % 50 == 0) || progress_due(o, t_progress))
            std::fprintf(stderr, "it %ld  recentred %d  w %.3e  p %.2e d %.2e g %.2e  pobj %.10g\n", long(it),
                         r.refinements, w, rc.p, rc.d, rc.g, kk[POBJ]);
        if (!std::isfinite(rc.worst())) {
            r.status = "numerical_failure";
            break;
        }
        if (rc.worst() <= o.tol) {
            // Confirm with exact FP64 products: recentre at this point and evaluate there.
            recentre();
            b.eval(kk);
            best = rel(kk);
            if (best.worst() <= o.tol) {
                r.status = "near_optimal";
                break;
            }
            r.refinements++;  // not yet: carry on from the recentred point
            k_last = wkkt(kk), k_prev = kInf, n_since = 0;
            continue;
        }
        if (since(t0) > time_budget) {
            r.status = "time_limit";
            break;
        }
        if (it >= o.max_iter) break;
        // Restart tests of gpu/pdhg.cu:827-842 with the current point as the candidate.
        const double kc = wkkt(kk);
        const bool restart =
            double(n_since) >= 0.36 * double(it) || kc <= 0.2 * k_last || (kc <= 0.8 * k_last && kc > k_prev);
        k_prev = kc;
        if (restart) {
            const double dx = std::sqrt(kk[DX2]), dz = std::sqrt(kk[DZ2]);
            if (dx > 1e-10 && dz > 1e-10) w = std::exp(0.5 * std::log(dz / dx) + 0.5 * std::log(w));
            if (o.refine && r.refinements < o.max_refinements) {
                recentre();
                r.refinements++;
            } else {
                b.set_steps(eta / w, eta * w);
                b.restart();
            }
            k_last = wkkt(kk), k_prev = kInf, n_since = 0, r.restarts++;
        }
    }
    r.iterations = it;
    r.rel_primal = best.p, r.rel_dual = best.d, r.rel_gap = best.g, r.omega = w, r.from_average = false;
    r.x.resize(size_t(P.n)), r.z.resize(size_t(P.m));
    b.download(r.x.data(), r.z.data());
    r.iterate_s = since(t0) - r.power_s;
    r.kernels = b.kernel_times();
    return r;
}

Session::Session(const LP& lp, const Options& o, Timings tm) : lp_(lp), o_(o), tm_(tm) {
    created_ = Clock::now();
    auto t = Clock::now();
    P_ = build_scaled_csc(lp.m, lp.n, lp.colptr.data(), lp.rowidx.data(), lp.val.data(), lp.cost.data(),
                          lp.col_lo.data(), lp.col_up.data(), lp.row_lo.data(), lp.row_up.data(),
                          lp.maximize ? -1.0 : 1.0);
    tm_.scale = since(t);
    t = Clock::now();
    if (device_count() == 0)
        throw std::runtime_error("no HIP device visible: the PDHG solver runs on an AMD GPU only (there is no CPU path)");
    tm_.init = since(t);
    t = Clock::now();
    if (o.halpern) {
        FastOptions fo;
        fo.precision_bits = o.precision_bits, fo.eval_freq = o.eval_freq, fo.use_graph = o.graph, fo.rho = o.rho;
        fast_ = make_fast_device(P_, fo);
    } else {
        DeviceOptions dopt;
        dopt.precision_bits = o.precision_bits;
        dopt.eval_freq = o.eval_freq;
        dopt.use_graph = o.graph;
        dev_ = make_device(P_, dopt);
    }
    tm_.transfer = since(t);
    setup_s_ = tm_.parse + tm_.scale + tm_.init + tm_.transfer;
}

Report Session::solve(const double* x0_orig, const double* y0_orig) {
    const LP& lp = lp_;
    const Scaled& P = P_;
    Report rep;
    auto t_call = Clock::now();
    rep.tm = tm_;
    rep.setup_s = setup_s_;
    rep.device_name = fast_ ? fast_->name() : dev_->name();
    rep.device_mem_bytes = fast_ ? fast_->mem_used() : dev_->mem_used();
    rep.nnz = P.A.nnz();
    std::vector<double> xs0, zs0;
    if (x0_orig) {
        xs0.resize(size_t(lp.n));
        for (int j = 0; j < lp.n; ++j) xs0[size_t(j)] = x0_orig[j] / P.Dc[size_t(j)];
    }
    if (y0_orig) {
        zs0.resize(size_t(lp.m));
        for (int i = 0; i < lp.m; ++i) zs0[size_t(i)] = y0_orig[i] / P.Dr[size_t(i)];
    }
    const double* xw = x0_orig ? xs0.data() : nullptr;
    const double* zw = y0_orig ? zs0.data() : nullptr;
    bool warm = xw || zw;
    double warm_kkt = 0, cold_kkt = 0;
    if (warm && o_.warm_fallback) {  // keep the warm start only when its relative KKT beats the cold start's
        warm_kkt = probe_kkt(xw, zw);
        cold_kkt = probe_kkt(nullptr, nullptr);
        if (!(warm_kkt < cold_kkt)) warm = false, xw = zw = nullptr;
    }
    if (fast_)
        rep.out = run_fast(*fast_, P, o_, o_.time_limit - setup_s_, xw, zw, warm ? o_.omega0 : 0, norm_a_);
    else
        rep.out = run_pdhg(*dev_, P, o_, o_.time_limit - setup_s_, xw, zw, warm ? o_.omega0 : 0, norm_a_);
    norm_a_ = rep.out.norm_a;  // A_s is fixed for the session's lifetime: later solves skip the power iteration
    rep.out.warm_used = warm, rep.out.warm_kkt = warm_kkt, rep.out.cold_kkt = cold_kkt;
    rep.solve_s = since(t_call);
    // Unscale and check the point on the ORIGINAL model in FP64 (gpu/pdhg.cu:866-876).
    rep.x.resize(size_t(lp.n)), rep.y.resize(size_t(lp.m));
    for (int j = 0; j < lp.n; ++j) rep.x[size_t(j)] = rep.out.x[size_t(j)] * P.Dc[size_t(j)];
    for (int i = 0; i < lp.m; ++i) rep.y[size_t(i)] = P.Dr[size_t(i)] * rep.out.z[size_t(i)];
    long double obj = lp.obj_const;
    for (int j = 0; j < lp.n; ++j) obj += (long double)lp.cost[size_t(j)] * rep.x[size_t(j)];
    std::vector<double> ax(static_cast<size_t>(lp.m), 0.0);
    double bviol = 0, rviol = 0;
    for (int j = 0; j < lp.n; ++j) {
        const double xj = rep.x[size_t(j)];
        for (int64_t k = lp.colptr[size_t(j)]; k < lp.colptr[size_t(j) + 1]; ++k)
            ax[size_t(lp.rowidx[size_t(k)])] += lp.val[size_t(k)] * xj;
        bviol = std::max(bviol, std::max(lp.col_lo[size_t(j)] - xj, xj - lp.col_up[size_t(j)]));
    }
    for (int i = 0; i < lp.m; ++i)
        rviol = std::max(rviol, std::max(lp.row_lo[size_t(i)] - ax[size_t(i)], ax[size_t(i)] - lp.row_up[size_t(i)]));
    rep.objective = double(obj), rep.max_row_violation = rviol, rep.max_bound_violation = bviol;
    rep.total_s = setup_s_ + since(t_call);
    return rep;
}

double Session::probe_kkt(const double* xs, const double* zs) {
    double kk[NKKT];
    if (fast_) {
        fast_->set_base(xs, zs);
        fast_->begin_round();
        fast_->eval(kk);
    } else {
        dev_->start(xs, zs);
        dev_->eval2(1.0, kk, nullptr);
    }
    return rel_of(kk, P_, o_.row_rel).worst();
}

void Session::set_vectors(LP& lp, const double* cost, const double* col_lo, const double* col_up, const double* row_lo,
                          const double* row_up) {
    if (&lp != &lp_) throw std::runtime_error("set_vectors: not this session's model");
    lp.cost.assign(cost, cost + lp.n), lp.col_lo.assign(col_lo, col_lo + lp.n), lp.col_up.assign(col_up, col_up + lp.n);
    lp.row_lo.assign(row_lo, row_lo + lp.m), lp.row_up.assign(row_up, row_up + lp.m);
    scale_vectors(P_, lp.cost.data(), lp.col_lo.data(), lp.col_up.data(), lp.row_lo.data(), lp.row_up.data(),
                  lp.maximize ? -1.0 : 1.0);
    if (fast_) fast_->set_vectors(P_.c.data(), P_.l.data(), P_.u.data(), P_.rl.data(), P_.ru.data());
    else dev_->set_vectors(P_.c.data(), P_.l.data(), P_.u.data(), P_.rl.data(), P_.ru.data());
}

Report solve_lp(const LP& lp, const Options& o, Timings tm, Clock::time_point t_start, const double* x0_orig,
                const double* y0_orig) {
    Session s(lp, o, tm);
    Report r = s.solve(x0_orig, y0_orig);
    r.total_s = since(t_start);
    return r;
}

std::string report_json(const Report& r, const LP& lp, const Options& o) {
    // Keys and formats of gpu/pdhg.cu:883-895, then this build's additions.
    char buf[4096];
    const Outcome& q = r.out;
    const double iters_per_s = r.solve_s > 0 ? double(q.iterations) / r.solve_s : 0;
    std::snprintf(
        buf, sizeof buf,
        "{\"solver\": \"taral-pdhg-mi355x (near-optimal)\", \"status\": \"%s\", \"objective\": %.17g, "
        "\"rel_primal\": %.6e, \"rel_dual\": %.6e, \"rel_gap\": %.6e, \"tol\": %.3g, \"iterations\": %ld, "
        "\"restarts\": %d, \"point\": \"%s\", \"max_row_violation\": %.6e, \"max_bound_violation\": %.6e, "
        "\"setup_s\": %.6f, \"parse_s\": %.6f, \"scale_s\": %.6f, \"device_init_s\": %.6f, \"transfer_s\": %.6f, "
        "\"solve_s\": %.6f, \"power_iter_s\": %.6f, \"total_s\": %.6f, \"device\": \"%s\", \"backend\": \"%s\", "
        "\"threads\": %d, \"precision\": \"%s\", \"rows\": %d, \"cols\": %d, \"nnz\": %ld, \"norm_a\": %.6e, "
        "\"primal_weight\": %.6e, \"eval_freq\": %d, \"graph\": %s, \"device_mem_bytes\": %zu, "
        "\"precision_mode\": \"%s\", \"refinements\": %d, \"iters_per_s\": %.6e, \"solves_per_s\": %.6e, "
        "\"warm_start_used\": %s}\n",
        q.status.c_str(), r.objective, q.rel_primal, q.rel_dual, q.rel_gap, o.tol, long(q.iterations), q.restarts,
        q.from_average ? "average" : "current", r.max_row_violation, r.max_bound_violation, r.setup_s, r.tm.parse,
        r.tm.scale, r.tm.init, r.tm.transfer, r.solve_s, q.power_s, r.total_s, r.device_name.c_str(), "gpu", 0,
        o.precision_bits == 32 ? "fp32" : "fp64", lp.m, lp.n, long(r.nnz), q.norm_a, q.omega, o.eval_freq,
        o.graph ? "true" : "false", r.device_mem_bytes,
        (std::string(o.halpern ? "halpern-" : "") +
         (o.precision_bits == 32 ? (o.refine ? "fp32+fp64-refinement" : "fp32") : (o.refine ? "fp64+refinement" : "fp64")))
            .c_str(),
        q.refinements, iters_per_s,
        r.solve_s > 0 ? 1.0 / r.solve_s : 0.0, q.warm_used ? "true" : "false");
    return buf;
}

}  // namespace tpdhg
