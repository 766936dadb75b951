// Host driver of the restarted averaged PDHG (gpu/pdhg.cu:743-849) over the HIP device backend, and the
// solve pipeline with its original-space report (gpu/pdhg.cu:851-911).
#include "driver.hpp"
#include "parallel.hpp"

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

// Primal weight after a restart. Rule 0 is the reference's smoothed update exp(0.5 ln(dz/dx) + 0.5 ln w)
// (gpu/pdhg.cu:839-840), skipped when either distance is below 1e-10. The fast path's optional guards:
//   1  keep the weight within two decades of the solve's starting weight w_start;
//   2  residual balancing: when the relative primal residual exceeds three times the larger of the dual residual
//      and the gap, the weight may not fall (a small weight makes the primal step large and the dual step small,
//      which leaves the primal residual behind); in the opposite case it may not rise;
//   3  both;
//   4  both, with the bound's anchor w_start moved to the weight the first restart's update gives (the callers
//      move it): ||c||/||b|| can sit decades from the weight an LP wants (the transport LP 1000 x 1000: 42.4
//      against the 0.07 the reference update reaches), and the bound around it then holds the weight on its edge
//      (that LP: 112,448 iterations under rule 3, 28,032 under rule 4).
double next_omega(double w, double dx, double dz, double p, double dg, double w_start, int rule) {
    double nw = w;
    if (dx > 1e-10 && dz > 1e-10) nw = std::exp(0.5 * std::log(dz / dx) + 0.5 * std::log(w));
    if (rule == 2 || rule == 3 || rule == 4) {
        if (p > 3 * dg) nw = std::max(nw, w);
        else if (dg > 3 * p) nw = std::min(nw, w);
    }
    if (rule == 1 || rule == 3 || rule == 4) nw = std::min(std::max(nw, w_start * 1e-2), w_start * 1e2);
    return nw;
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
    double w_start = w;  // the weight guard's anchor (rule 4 moves it to the first restart's weight)
    bool anchored = false;
    b.set_base(x0, z0);
    // The restart tests' measure of the current point (restart_metric): the fixed-point residual of the last
    // evaluation, ||u - T u||_P = sqrt(||dx||^2 / tau + ||dz||^2 / sigma + 2 dz' A dx) with tau = eta / w and
    // sigma = eta w (the norm in which a PDHG step is nonexpansive; Halpern iterations drive it down at a known
    // rate between restarts), or the reference's weighted KKT error.
    auto measure = [&](const double* k) {  // what the restart tests compare
        if (o.restart_metric == 1) {
            double f[3];
            b.fixed_point(f);
            return std::sqrt(std::max(0.0, f[0] * w / eta + f[1] / (eta * w) + 2 * f[2]));
        }
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
    double k_last = measure(kk), k_prev = kInf;
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
            k_last = measure(kk), k_prev = kInf, n_since = 0;
            continue;
        }
        if (since(t0) > time_budget) {
            r.status = "time_limit";
            break;
        }
        if (it >= o.max_iter) break;
        // The restart tests of gpu/pdhg.cu:827-842 (decay to 0.2; to 0.8 without progress; 36% of the iterations)
        // on the restart measure of the current point, the candidate.
        const double kc = measure(kk);
        const bool restart =
            double(n_since) >= 0.36 * double(it) || kc <= 0.2 * k_last || (kc <= 0.8 * k_last && kc > k_prev);
        k_prev = kc;
        if (restart) {
            const double dx = std::sqrt(kk[DX2]), dz = std::sqrt(kk[DZ2]);
            w = next_omega(w, dx, dz, rc.p, std::max(rc.d, rc.g), w_start, o.omega_rule);
            if (o.omega_rule == 4 && !anchored) w_start = w, anchored = true;
            if (o.refine && r.refinements < o.max_refinements) {
                recentre();
                r.refinements++;
            } else {
                b.set_steps(eta / w, eta * w);
                b.restart();
            }
            k_last = measure(kk), k_prev = kInf, n_since = 0, r.restarts++;
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
    if (o.cpu_threads >= 0) {  // the reference's CPU backend: reference scheme only
        if (o.halpern || o.refine)
            throw std::invalid_argument("the fast path (Halpern iterations, refinement) runs on the GPU only");
        DeviceOptions dopt;
        dopt.precision_bits = o.precision_bits, dopt.eval_freq = o.eval_freq, dopt.use_graph = false;
        threads_ = cpu_threads_for(P_.A.nnz(), o.cpu_threads);
        dev_ = make_cpu_device(P_, dopt, threads_);
        tm_.transfer = since(t);
        setup_s_ = tm_.parse + tm_.scale + tm_.init + tm_.transfer;
        return;
    }
    if (device_count() == 0)
        throw std::runtime_error("no HIP device visible: the PDHG solver runs on an AMD GPU only (--device cpu runs "
                                 "the reference scheme on host threads)");
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

// Unscale and check the point on the ORIGINAL model in FP64 (gpu/pdhg.cu:866-876).
void finish_report(const LP& lp, const Scaled& P, Report& rep) {
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
}

Report Session::solve(const double* x0_orig, const double* y0_orig) {
    const LP& lp = lp_;
    const Scaled& P = P_;
    Report rep;
    auto t_call = Clock::now();
    rep.tm = tm_;
    rep.setup_s = setup_s_;
    rep.device_name = fast_ ? fast_->name() : dev_->name();
    rep.backend = threads_ > 0 ? "cpu" : "gpu", rep.threads = threads_;
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
        rep.out = run_fast(*fast_, P, o_, o_.time_limit - setup_s_, xw, zw, warm ? o_.omega0 : 0,
                           o_.reuse_norm ? norm_a_ : 0);
    else
        rep.out = run_pdhg(*dev_, P, o_, o_.time_limit - setup_s_, xw, zw, warm ? o_.omega0 : 0,
                           o_.reuse_norm ? norm_a_ : 0);
    norm_a_ = rep.out.norm_a;  // A_s is fixed for the session's lifetime (reused only with reuse_norm)
    rep.out.warm_used = warm, rep.out.warm_kkt = warm_kkt, rep.out.cold_kkt = cold_kkt;
    rep.solve_s = since(t_call);
    finish_report(lp, P, rep);
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

std::vector<Outcome> run_batch(BatchDevice& b, const std::vector<const Scaled*>& P, const Options& o,
                               double time_budget, const std::vector<const double*>& x0,
                               const std::vector<const double*>& z0, const std::vector<double>& omega0,
                               std::vector<double>* norms) {
    const int B = b.count();
    std::vector<Outcome> r(static_cast<size_t>(B));
    auto t0 = Clock::now(), t_progress = t0;
    b.set_profile(o.profile_kernels);
    std::vector<double> na = norms && int(norms->size()) == B ? *norms : b.power_norm(200);
    if (norms) *norms = na;
    const double power_s = since(t0);
    std::vector<double> eta(static_cast<size_t>(B)), w(static_cast<size_t>(B)), tau(static_cast<size_t>(B)),
        sig(static_cast<size_t>(B));
    for (int i = 0; i < B; ++i) {
        const Scaled& Q = *P[size_t(i)];
        r[size_t(i)].norm_a = na[size_t(i)], r[size_t(i)].power_s = power_s;
        eta[size_t(i)] = na[size_t(i)] > 0 ? 0.95 / na[size_t(i)] : 1.0;
        double cs = 0, bs = 0;  // as run_fast
        for (double v : Q.c) cs += v * v;
        for (int k = 0; k < Q.m; ++k) {
            const double v = row_bound_measure(Q.rl[size_t(k)], Q.ru[size_t(k)]);
            bs += v * v;
        }
        cs = std::sqrt(cs), bs = std::sqrt(bs);
        w[size_t(i)] = cs > 1e-10 && bs > 1e-10 ? cs / bs : 1.0;
        if (size_t(i) < omega0.size() && omega0[size_t(i)] > 0) w[size_t(i)] = omega0[size_t(i)];
    }
    std::vector<double> w_start = w;  // per instance; rule 4 moves it to the first restart's weight as run_fast
    auto push_steps = [&]() {
        for (int i = 0; i < B; ++i) tau[size_t(i)] = eta[size_t(i)] / w[size_t(i)], sig[size_t(i)] = eta[size_t(i)] * w[size_t(i)];
        b.set_steps(tau, sig);
    };
    auto kslot = [](std::vector<double>& kk, int i) { return kk.data() + size_t(i) * NKKT; };
    b.set_base(x0, z0);
    std::vector<uint8_t> all(static_cast<size_t>(B), 1), active(static_cast<size_t>(B), 0);
    std::vector<double> kk(size_t(B) * NKKT, 0.0), k_last(static_cast<size_t>(B)), k_prev(static_cast<size_t>(B), kInf);
    std::vector<double> fps;  // fixed-point residual terms of the last evaluation, 3 per instance
    auto measure = [&](int i, double wi) {  // run_fast's restart measure for instance i at weight wi
        if (o.restart_metric == 1) {
            const double ei = eta[size_t(i)];
            const double* f = fps.data() + 3 * size_t(i);
            return std::sqrt(std::max(0.0, f[0] * wi / ei + f[1] / (ei * wi) + 2 * f[2]));
        }
        const double* k = kslot(kk, i);
        const double g = k[POBJ] - (k[DROW] + k[DCOL]);
        return std::sqrt(wi * k[PR2S] + k[DR2S] / wi + g * g);
    };
    std::vector<int64_t> n_since(static_cast<size_t>(B), 0), it_of(static_cast<size_t>(B), 0);
    std::vector<Rel> best(static_cast<size_t>(B), Rel{kInf, kInf, kInf});
    b.begin_round(all);
    push_steps();
    b.restart(all);
    b.eval(all, kk);
    b.fixed_points(fps);
    int live = 0;
    for (int i = 0; i < B; ++i) {
        best[size_t(i)] = rel_of(kslot(kk, i), *P[size_t(i)], o.row_rel);
        k_last[size_t(i)] = measure(i, w[size_t(i)]);
        if (best[size_t(i)].worst() <= o.tol) r[size_t(i)].status = "near_optimal";
        else active[size_t(i)] = 1, ++live;
    }
    int64_t it = 0;
    std::vector<uint8_t> confirm(static_cast<size_t>(B)), rs_move(static_cast<size_t>(B)), rs_any(static_cast<size_t>(B));
    while (live > 0) {
        const int steps = int(std::min<int64_t>(o.eval_freq, o.max_iter - it));
        if (steps <= 0) break;
        b.iterate(steps, active);
        it += steps;
        b.eval(active, kk);
        b.fixed_points(fps);
        std::fill(confirm.begin(), confirm.end(), 0), std::fill(rs_move.begin(), rs_move.end(), 0);
        std::fill(rs_any.begin(), rs_any.end(), 0);
        bool any_confirm = false, any_move = false, any_rs = false;
        const bool out_of_time = since(t0) > time_budget;
        for (int i = 0; i < B; ++i) {
            if (!active[size_t(i)]) continue;
            Outcome& q = r[size_t(i)];
            n_since[size_t(i)] += steps, it_of[size_t(i)] = it;
            const double* k = kslot(kk, i);
            const Rel rc = rel_of(k, *P[size_t(i)], o.row_rel);
            best[size_t(i)] = rc;
            if (!std::isfinite(rc.worst())) {
                q.status = "numerical_failure", active[size_t(i)] = 0, --live;
                continue;
            }
            if (rc.worst() <= o.tol) {
                confirm[size_t(i)] = 1, any_confirm = true;
                continue;
            }
            if (out_of_time) {
                q.status = "time_limit", active[size_t(i)] = 0, --live;
                continue;
            }
            if (it >= o.max_iter) {
                active[size_t(i)] = 0, --live;
                continue;
            }
            const double kc = measure(i, w[size_t(i)]);
            const bool restart = double(n_since[size_t(i)]) >= 0.36 * double(it) || kc <= 0.2 * k_last[size_t(i)] ||
                                 (kc <= 0.8 * k_last[size_t(i)] && kc > k_prev[size_t(i)]);
            k_prev[size_t(i)] = kc;
            if (restart) {
                const double dx = std::sqrt(k[DX2]), dz = std::sqrt(k[DZ2]);
                w[size_t(i)] = next_omega(w[size_t(i)], dx, dz, rc.p, std::max(rc.d, rc.g), w_start[size_t(i)],
                                          o.omega_rule);
                if (o.omega_rule == 4 && q.restarts == 0) w_start[size_t(i)] = w[size_t(i)];
                if (o.refine && q.refinements < o.max_refinements) rs_move[size_t(i)] = 1, any_move = true, q.refinements++;
                rs_any[size_t(i)] = 1, any_rs = true, q.restarts++;
            }
        }
        if (any_confirm) {  // FP64 confirmation: recentre at the point and evaluate there
            b.end_round(confirm);
            b.begin_round(confirm);
            push_steps();
            b.restart(confirm);
            b.eval(confirm, kk);
            b.fixed_points(fps);
            for (int i = 0; i < B; ++i) {
                if (!confirm[size_t(i)]) continue;
                best[size_t(i)] = rel_of(kslot(kk, i), *P[size_t(i)], o.row_rel);
                if (best[size_t(i)].worst() <= o.tol) {
                    r[size_t(i)].status = "near_optimal", active[size_t(i)] = 0, --live;
                    continue;
                }
                r[size_t(i)].refinements++;
                k_last[size_t(i)] = measure(i, w[size_t(i)]), k_prev[size_t(i)] = kInf, n_since[size_t(i)] = 0;
            }
        }
        if (any_rs) {
            if (any_move) {
                b.end_round(rs_move);
                b.begin_round(rs_move);
            }
            push_steps();
            b.restart(rs_any);
            for (int i = 0; i < B; ++i)
                if (rs_any[size_t(i)])
                    k_last[size_t(i)] = measure(i, w[size_t(i)]), k_prev[size_t(i)] = kInf, n_since[size_t(i)] = 0;
        }
        if ((o.verbose && (it / o.eval_freq) % 50 == 0) || progress_due(o, t_progress)) {
            double worst = 0;
            for (int i = 0; i < B; ++i)
                if (active[size_t(i)]) worst = std::max(worst, best[size_t(i)].worst());
            std::fprintf(stderr, "batch it %ld  live %d of %d  worst live rel KKT %.2e\n", long(it), live, B, worst);
        }
    }
    const double iterate_s = since(t0) - power_s;
    for (int i = 0; i < B; ++i) {
        Outcome& q = r[size_t(i)];
        q.iterations = it_of[size_t(i)];
        q.rel_primal = best[size_t(i)].p, q.rel_dual = best[size_t(i)].d, q.rel_gap = best[size_t(i)].g;
        q.omega = w[size_t(i)], q.from_average = false, q.iterate_s = iterate_s;
        q.x.resize(size_t(P[size_t(i)]->n)), q.z.resize(size_t(P[size_t(i)]->m));
        b.download(i, q.x.data(), q.z.data());
    }
    const KernelTimes kt = b.kernel_times();
    for (Outcome& q : r) q.kernels = kt;
    return r;
}

BatchSession::BatchSession(const std::vector<const LP*>& lps, const Options& o, int cap) : o_(o) {
    auto t = Clock::now();
    if (lps.empty()) throw std::invalid_argument("empty batch");
    lps_.reserve(lps.size());
    for (const LP* lp : lps) {
        if (!lp) throw std::invalid_argument("null batch instance");
        lps_.push_back(*lp);
    }
    P_.resize(lps_.size());
    parallel_for(int64_t(lps_.size()), 1, [&](int64_t i) {  // instances scale independently
        const LP& lp = lps_[size_t(i)];
        P_[size_t(i)] = build_scaled_csc(lp.m, lp.n, lp.colptr.data(), lp.rowidx.data(), lp.val.data(),
                                         lp.cost.data(), lp.col_lo.data(), lp.col_up.data(), lp.row_lo.data(),
                                         lp.row_up.data(), lp.maximize ? -1.0 : 1.0);
    });
    if (device_count() == 0)
        throw std::runtime_error("no HIP device visible: the PDHG solver runs on an AMD GPU only (there is no CPU path)");
    BatchOptions bo;
    bo.precision_bits = o.precision_bits, bo.eval_freq = o.eval_freq, bo.use_graph = o.graph, bo.rho = o.rho;
    bo.cap = cap;
    std::vector<const Scaled*> ps;
    for (const Scaled& P : P_) ps.push_back(&P);
    dev_ = make_batch_device(ps, bo);
    setup_s_ = since(t);
}

std::vector<Report> BatchSession::solve(const std::vector<const double*>& x0, const std::vector<const double*>& y0,
                                        const std::vector<double>& omega0) {
    auto t_call = Clock::now();
    const int B = count();
    std::vector<std::vector<double>> xs(static_cast<size_t>(B)), zs(static_cast<size_t>(B));
    std::vector<const double*> xw(static_cast<size_t>(B), nullptr), zw(static_cast<size_t>(B), nullptr);
    for (int i = 0; i < B; ++i) {
        const LP& lp = lps_[size_t(i)];
        const Scaled& P = P_[size_t(i)];
        if (size_t(i) < x0.size() && x0[size_t(i)]) {
            xs[size_t(i)].resize(size_t(lp.n));
            for (int j = 0; j < lp.n; ++j) xs[size_t(i)][size_t(j)] = x0[size_t(i)][j] / P.Dc[size_t(j)];
            xw[size_t(i)] = xs[size_t(i)].data();
        }
        if (size_t(i) < y0.size() && y0[size_t(i)]) {
            zs[size_t(i)].resize(size_t(lp.m));
            for (int k = 0; k < lp.m; ++k) zs[size_t(i)][size_t(k)] = y0[size_t(i)][k] / P.Dr[size_t(k)];
            zw[size_t(i)] = zs[size_t(i)].data();
        }
    }
    std::vector<uint8_t> warm(static_cast<size_t>(B), 0);
    std::vector<double> wk(static_cast<size_t>(B), 0.0), ck(static_cast<size_t>(B), 0.0);
    bool any_warm = false;
    for (int i = 0; i < B; ++i) warm[size_t(i)] = xw[size_t(i)] || zw[size_t(i)], any_warm = any_warm || warm[size_t(i)];
    if (any_warm && o_.warm_fallback) {  // relative KKT of the warm and of the cold start, all instances at once
        std::vector<uint8_t> all(static_cast<size_t>(B), 1);
        std::vector<double> kk;
        auto probe = [&](const std::vector<const double*>& bx, const std::vector<const double*>& bz,
                         std::vector<double>& dst) {
            dev_->set_base(bx, bz);
            dev_->begin_round(all);
            dev_->eval(all, kk);
            for (int i = 0; i < B; ++i) dst[size_t(i)] = rel_of(kk.data() + size_t(i) * NKKT, P_[size_t(i)], o_.row_rel).worst();
        };
        probe(xw, zw, wk);
        probe({}, {}, ck);
        for (int i = 0; i < B; ++i)
            if (warm[size_t(i)] && !(wk[size_t(i)] < ck[size_t(i)])) warm[size_t(i)] = 0, xw[size_t(i)] = zw[size_t(i)] = nullptr;
    }
    std::vector<double> om(static_cast<size_t>(B), 0.0);
    for (int i = 0; i < B; ++i)
        if (warm[size_t(i)] && size_t(i) < omega0.size()) om[size_t(i)] = omega0[size_t(i)];
    std::vector<const Scaled*> ps;
    for (const Scaled& P : P_) ps.push_back(&P);
    if (!o_.reuse_norm) norms_.clear();  // every solve runs the power iteration
    std::vector<Outcome> outs = run_batch(*dev_, ps, o_, o_.time_limit, xw, zw, om, &norms_);
    const double solve_s = since(t_call);
    std::vector<Report> reps(static_cast<size_t>(B));
    for (int i = 0; i < B; ++i) {
        Report& rep = reps[size_t(i)];
        rep.out = std::move(outs[size_t(i)]);
        rep.out.warm_used = warm[size_t(i)] != 0, rep.out.warm_kkt = wk[size_t(i)], rep.out.cold_kkt = ck[size_t(i)];
        rep.setup_s = setup_s_, rep.solve_s = solve_s, rep.total_s = setup_s_ + solve_s;
        rep.device_name = dev_->name(), rep.device_mem_bytes = dev_->mem_used(), rep.nnz = P_[size_t(i)].A.nnz();
        finish_report(lps_[size_t(i)], P_[size_t(i)], rep);
    }
    return reps;
}

void BatchSession::set_vectors(int i, const double* cost, const double* col_lo, const double* col_up,
                               const double* row_lo, const double* row_up) {
    if (i < 0 || i >= count()) throw std::invalid_argument("instance index out of range");
    LP& lp = lps_[size_t(i)];
    lp.cost.assign(cost, cost + lp.n), lp.col_lo.assign(col_lo, col_lo + lp.n), lp.col_up.assign(col_up, col_up + lp.n);
    lp.row_lo.assign(row_lo, row_lo + lp.m), lp.row_up.assign(row_up, row_up + lp.m);
    Scaled& P = P_[size_t(i)];
    scale_vectors(P, lp.cost.data(), lp.col_lo.data(), lp.col_up.data(), lp.row_lo.data(), lp.row_up.data(),
                  lp.maximize ? -1.0 : 1.0);
    dev_->set_vectors(i, P.c.data(), P.l.data(), P.u.data(), P.rl.data(), P.ru.data());
}

Report solve_lp(const LP& lp, const Options& o, Timings tm, Clock::time_point t_start, const double* x0_orig,
                const double* y0_orig) {
    Session s(lp, o, tm);
    Report r = s.solve(x0_orig, y0_orig);
    r.total_s = since(t_start);
    return r;
}

std::string report_json(const Report& r, const LP& lp, const Options& o) {
    // Keys and formats of gpu/pdhg.cu:883-895, then the keys added here.
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
        r.tm.scale, r.tm.init, r.tm.transfer, r.solve_s, q.power_s, r.total_s, r.device_name.c_str(), r.backend.c_str(), r.threads,
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
