"""Independent original-space KKT check of a returned point (x, y), written from the reference's definitions
(gpu/pdhg.cu:12-14, 236-271, 791-796) with y = Dr z the row multipliers the solver returns (minimisation form,
reduced cost d = c + A^T y). Used by the GPU tests to verify every reported relative KKT figure."""
import numpy as np


def rel_kkt(md, x, y):
    """(rel_primal, rel_dual, rel_gap) of the point on the model md (gen.LPModel), in FP64."""
    sign = -1.0 if md.maximize else 1.0
    c = sign * md.cost
    ax = md.matvec(x)
    viol = np.maximum(md.row_lo - ax, 0) + np.maximum(ax - md.row_up, 0)
    lo, up = md.row_lo, md.row_up
    b = np.where(np.isfinite(lo) & np.isfinite(up), np.maximum(abs(lo), abs(up)),
                 np.where(np.isfinite(lo), abs(lo), np.where(np.isfinite(up), abs(up), 0.0)))
    bnorm, cnorm = np.linalg.norm(b), np.linalg.norm(c)
    drow = -np.sum(np.where((y > 0) & np.isfinite(up), y * np.where(np.isfinite(up), up, 0), 0)) \
        - np.sum(np.where((y < 0) & np.isfinite(lo), y * np.where(np.isfinite(lo), lo, 0), 0))
    d = c + md.rmatvec(y)
    cl, cu = md.col_lo, md.col_up
    dcol = np.sum(np.where((d > 0) & np.isfinite(cl), d * np.where(np.isfinite(cl), cl, 0), 0)) \
        + np.sum(np.where((d < 0) & np.isfinite(cu), d * np.where(np.isfinite(cu), cu, 0), 0))
    res = np.where((d > 0) & ~np.isfinite(cl), d, 0) + np.where((d < 0) & ~np.isfinite(cu), -d, 0)
    po, du = c @ x, drow + dcol
    return (np.linalg.norm(viol) / (1 + bnorm), np.linalg.norm(res) / (1 + cnorm),
            abs(po - du) / (1 + abs(po) + abs(du)))
