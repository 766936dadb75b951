"""ctypes wrapper of the CPU oracle (oracle/pdhg_oracle.c). TEST INFRASTRUCTURE ONLY.

Imported by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, where it is the oracle
(or the timed CPU baseline), never the thing measured on the GPU.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle_pdhg.so")
STATUS = {0: "near_optimal", 1: "iteration_limit", 2: "time_limit", 3: "numerical_failure"}

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)


class Scaled(C.Structure):
    _fields_ = [("m", C.c_int), ("n", C.c_int), ("nnz", C.c_int64),
                ("aptr", _i32), ("aidx", _i32), ("aval", _d),
                ("tptr", _i32), ("tidx", _i32), ("tval", _d),
                ("c", _d), ("l", _d), ("u", _d), ("rl", _d), ("ru", _d), ("Dr", _d), ("Dc", _d),
                ("bnorm", C.c_double), ("cnorm", C.c_double)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_long), ("restarts", C.c_int),
                ("from_average", C.c_int), ("rel_primal", C.c_double), ("rel_dual", C.c_double),
                ("rel_gap", C.c_double), ("norm_a", C.c_double), ("omega", C.c_double),
                ("power_s", C.c_double), ("iterate_s", C.c_double)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_build_scaled.argtypes = [C.c_int, C.c_int, _i64, _i32, _d, _d, _d, _d, _d, _d, C.c_double,
                                      C.POINTER(Scaled)]
        L.or_free_scaled.argtypes = [C.POINTER(Scaled)]
        L.or_spmv.argtypes = [C.c_int, _i32, _i32, _d, _d, _d]
        L.or_eval.argtypes = [C.POINTER(Scaled), C.c_int, C.c_double, _d, _d, _d, _d, _d, _d, _d]
        L.or_power_norm.argtypes = [C.POINTER(Scaled), C.c_int, C.c_int]
        L.or_power_norm.restype = C.c_double
        L.or_solve.argtypes = [C.POINTER(Scaled), C.c_int, C.c_double, C.c_long, C.c_int, C.c_double,
                               C.c_double, _d, _d, _d, _d, C.POINTER(Result)]
        L.or_block.argtypes = [C.POINTER(Scaled), C.c_int, C.c_double, C.c_double, _d, _d, C.c_int] + [_d] * 8
        L.or_block_mt.argtypes = [C.POINTER(Scaled), C.c_int, C.c_double, C.c_double, _d, _d, C.c_int, C.c_int] + [_d] * 8
        L.or_finish.argtypes = [C.POINTER(Scaled), C.c_int, _i64, _i32, _d, _d, C.c_double, _d, _d, _d, _d,
                                _d, _d, _d, _d]
        _lib = L
    return _lib


def _p(a, t=_d):
    return a.ctypes.data_as(t) if a is not None else None


class OracleLP:
    """Scaled problem plus the arrays it was built from (kept alive for the C side)."""

    def __init__(self, model):
        self.model = model
        self._keep = [np.ascontiguousarray(model.colptr, dtype=np.int64),
                      np.ascontiguousarray(model.rowidx, dtype=np.int32),
                      np.ascontiguousarray(model.val, dtype=np.float64)]
        self._keep += [np.ascontiguousarray(getattr(model, k), dtype=np.float64)
                       for k in ("cost", "col_lo", "col_up", "row_lo", "row_up")]
        self.S = Scaled()
        cp, ri, v, c, cl, cu, rl, ru = self._keep
        sign = -1.0 if model.maximize else 1.0
        lib().or_build_scaled(model.m, model.n, _p(cp, _i64), _p(ri, _i32), _p(v), _p(c), _p(cl), _p(cu),
                              _p(rl), _p(ru), sign, C.byref(self.S))

    def __del__(self):
        if _lib is not None and getattr(self, "S", None) is not None:
            _lib.or_free_scaled(C.byref(self.S))

    # views of the scaled problem (copies, so they outlive the C memory)
    def vec(self, name):
        k = self.S.m if name in ("rl", "ru", "Dr") else self.S.n
        return np.ctypeslib.as_array(getattr(self.S, name), shape=(k,)).copy() if k else np.zeros(0)

    def csr(self, which="A"):
        rows = self.S.m if which == "A" else self.S.n
        pre = "a" if which == "A" else "t"
        nnz = self.S.nnz
        ptr = np.ctypeslib.as_array(getattr(self.S, pre + "ptr"), shape=(rows + 1,)).copy()
        idx = np.ctypeslib.as_array(getattr(self.S, pre + "idx"), shape=(max(nnz, 1),))[:nnz].copy()
        val = np.ctypeslib.as_array(getattr(self.S, pre + "val"), shape=(max(nnz, 1),))[:nnz].copy()
        return ptr, idx, val

    def power_norm(self, fp32=False, iters=200):
        return lib().or_power_norm(C.byref(self.S), int(fp32), iters)

    def eval(self, Ax, z, x, ATz, xlast, zlast, f=1.0, fp32=False):
        out = np.zeros(10)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (Ax, z, x, ATz, xlast, zlast)]
        lib().or_eval(C.byref(self.S), int(fp32), f, *[_p(a) for a in arrs], _p(out))
        return out

    def block(self, tau, sigma, steps, x0=None, z0=None, fp32=False, threads=1):
        """start + iterate(steps) of the CPU backend (its iterations on `threads` threads, partitioned as the
        reference's Pool); dict of x, z, xs, zs, Ax, ATz, Axs, ATzs."""
        m, n = self.S.m, self.S.n
        out = {k: np.zeros(max(m if k in ("z", "zs", "Ax", "Axs") else n, 1))
               for k in ("x", "z", "xs", "zs", "Ax", "ATz", "Axs", "ATzs")}
        x0 = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
        z0 = None if z0 is None else np.ascontiguousarray(z0, dtype=np.float64)
        lib().or_block_mt(C.byref(self.S), int(fp32), float(tau), float(sigma), _p(x0), _p(z0), int(steps),
                          int(threads), *[_p(out[k]) for k in ("x", "z", "xs", "zs", "Ax", "ATz", "Axs", "ATzs")])
        return {k: v[:(m if k in ("z", "zs", "Ax", "Axs") else n)] for k, v in out.items()}

    def solve(self, fp32=False, tol=1e-4, max_iter=2_000_000_000, eval_freq=64, row_rel=0.0,
              time_limit=3600.0, x0=None, z0=None):
        m, n = self.S.m, self.S.n
        x = np.zeros(max(n, 1))
        z = np.zeros(max(m, 1))
        r = Result()
        x0 = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
        z0 = None if z0 is None else np.ascontiguousarray(z0, dtype=np.float64)
        lib().or_solve(C.byref(self.S), int(fp32), tol, max_iter, eval_freq, row_rel, time_limit, _p(x0), _p(z0),
                       _p(x), _p(z), C.byref(r))
        md = self.model
        obj, rv, bv = C.c_double(), C.c_double(), C.c_double()
        cp, ri, v, c, cl, cu, rl, ru = self._keep
        lib().or_finish(C.byref(self.S), n, _p(cp, _i64), _p(ri, _i32), _p(v), _p(c), float(md.obj_const),
                        _p(cl), _p(cu), _p(rl), _p(ru), _p(x), C.byref(obj), C.byref(rv), C.byref(bv))
        objective = obj.value
        return dict(status=STATUS[r.status], iterations=r.iterations, restarts=r.restarts,
                    from_average=bool(r.from_average), rel_primal=r.rel_primal, rel_dual=r.rel_dual,
                    rel_gap=r.rel_gap, norm_a=r.norm_a, omega=r.omega, power_s=r.power_s, iterate_s=r.iterate_s,
                    objective=objective, max_row_violation=rv.value, max_bound_violation=bv.value,
                    x=x[:n], z=z[:m])
