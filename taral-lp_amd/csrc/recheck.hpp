// Original-model recheck of a returned point, independent of the solver's scaled iterates: the checks of
// gpu/vector_check.py:46-74 and benchmarks/orig_check.py:63-88 in C++ (no Python, no HiGHS needed).
#pragma once
#include "driver.hpp"

namespace tpdhg {

struct Recheck {
    double objective = 0;       // cost'x + obj_const on the original model
    double row_viol_rel = 0;    // max over rows of the violation / (1 + |the violated bound|)
    double bound_viol = 0;      // max over columns of the bound violation
    bool dual_checked = false;  // y was given
    // Weak-duality bound from the row multipliers (either sign, the one with the smaller infeasibility):
    //   LB(y) = sum_j min_{x_j in [l_j,u_j]} d_j x_j + sum_i min_{r_i in [lo_i,up_i]} y_i r_i + const, d = c - A'y
    // for the minimisation form (a maximisation is checked as min -c'x; bounds are reported in its own sense).
    bool dual_bound_finite = false;  // no term had a wrong-sign coefficient on an infinite side
    double dual_bound = 0;           // the bound (valid only when dual_bound_finite)
    double dual_obj_loose = 0;       // the same sum with the infinite terms dropped (not a bound)
    double dual_infeas_rel = 0;      // largest dropped term: |y_i|/(1+|y_i|) for rows, |d_j|/(1+|c_j|) for columns
    double dual_gap_rel = 0;         // (objective - loose) / max(1, |objective|), minimisation sense
    int dual_sign = 0;               // +1 or -1
};

// x: n values; y: m row multipliers or null. Long double accumulation for the objective and the bound.
Recheck recheck(const LP& lp, const double* x, const double* y);

}  // namespace tpdhg
