// Original-model recheck (recheck.hpp).
#include "recheck.hpp"

#include <algorithm>
#include <cmath>

namespace tpdhg {

namespace {
// min of coef * v over v in [lo, up]; false when the minimising side is infinite (the term is -inf).
bool min_over(double coef, double lo, double up, long double& term) {
    term = 0;
    if (coef == 0) return true;
    const double v = coef > 0 ? lo : up;
    if (!std::isfinite(v)) return false;
    term = (long double)coef * v;
    return true;
}
}  // namespace

Recheck recheck(const LP& lp, const double* x, const double* y) {
    Recheck r;
    std::vector<double> act(static_cast<size_t>(lp.m), 0.0);
    long double obj = lp.obj_const;
    for (int j = 0; j < lp.n; ++j) {
        const double xj = x[j];
        obj += (long double)lp.cost[size_t(j)] * xj;
        for (int64_t k = lp.colptr[size_t(j)]; k < lp.colptr[size_t(j) + 1]; ++k)
            act[size_t(lp.rowidx[size_t(k)])] += lp.val[size_t(k)] * xj;
        const double lo = lp.col_lo[size_t(j)], up = lp.col_up[size_t(j)];
        r.bound_viol = std::max(r.bound_viol, std::max(xj < lo ? lo - xj : 0.0, xj > up ? xj - up : 0.0));
    }
    r.objective = double(obj);
    for (int i = 0; i < lp.m; ++i) {
        const double a = act[size_t(i)], lo = lp.row_lo[size_t(i)], up = lp.row_up[size_t(i)];
        double v = 0;
        if (a < lo) v = (lo - a) / (1 + std::fabs(lo));
        else if (a > up) v = (a - up) / (1 + std::fabs(up));
        r.row_viol_rel = std::max(r.row_viol_rel, v);
    }
    if (!y) return r;
    r.dual_checked = true;
    const double s = lp.maximize ? -1.0 : 1.0;  // minimisation form: s * cost, s * obj_const
    struct Try {
        long double loose;
        double infeas;
        bool finite;
    };
    Try best{0, 0, false};
    int best_sign = 0;
    for (int sg : {1, -1}) {
        Try t{(long double)(s * lp.obj_const), 0.0, true};
        for (int i = 0; i < lp.m; ++i) {
            const double yi = sg * y[i];
            long double term;
            if (min_over(yi, lp.row_lo[size_t(i)], lp.row_up[size_t(i)], term)) t.loose += term;
            else t.finite = false, t.infeas = std::max(t.infeas, std::fabs(yi) / (1 + std::fabs(yi)));
        }
        for (int j = 0; j < lp.n; ++j) {
            const double c = s * lp.cost[size_t(j)];
            long double aty = 0;
            for (int64_t k = lp.colptr[size_t(j)]; k < lp.colptr[size_t(j) + 1]; ++k)
                aty += (long double)lp.val[size_t(k)] * (sg * y[lp.rowidx[size_t(k)]]);
            const double d = double((long double)c - aty);
            long double term;
            if (min_over(d, lp.col_lo[size_t(j)], lp.col_up[size_t(j)], term)) t.loose += term;
            else t.finite = false, t.infeas = std::max(t.infeas, std::fabs(d) / (1 + std::fabs(c)));
        }
        if (best_sign == 0 || t.infeas < best.infeas) best = t, best_sign = sg;
    }
    const double obj_min = s * r.objective, loose_min = double(best.loose);
    r.dual_sign = best_sign;
    r.dual_infeas_rel = best.infeas;
    r.dual_obj_loose = s * loose_min;
    r.dual_bound_finite = best.finite;
    r.dual_bound = best.finite ? s * loose_min : -s * kInf;
    r.dual_gap_rel = (obj_min - loose_min) / std::max(1.0, std::fabs(obj_min));
    return r;
}

}  // namespace tpdhg
