"""ctypes binding of libtaral_pdhg.so (include/taral_pdhg.h), the MI355X restarted-PDHG LP solver.

The Python side mirrors the reference's backend concept (gpu/pdhg.cu:313-514: init, power_norm, set_steps,
start, iterate, eval, restart, download) and its solve path (gpu/pdhg.cu:913-935), so tests read like the
reference's own harness. All compute runs in the HIP library; this module only moves arrays across the C ABI.
It raises if the library is missing, and the library raises (TPDHG_ERR_NO_DEVICE) when no GPU is visible.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "build", "libtaral_pdhg.so")
CLI_PATH = os.path.join(PKG_ROOT, "build", "pdhg")

OK, ERR_ARG, ERR_PARSE, ERR_UNSUPPORTED, ERR_DEVICE, ERR_NO_DEVICE, ERR_TIME = 0, -1, -2, -3, -4, -5, -6
STATUS = {0: "near_optimal", 1: "iteration_limit", 2: "time_limit", 3: "numerical_failure"}
KKT_NAMES = ["PR2S", "PR2O", "PR2R", "DROW", "DZ2", "DR2S", "DR2O", "POBJ", "DCOL", "DX2"]

# every symbol include/taral_pdhg.h declares
EXPORTS = [
    "tpdhg_last_error", "tpdhg_status_string", "tpdhg_device_count",
    "tpdhg_model_read_mps", "tpdhg_model_read_mps_text", "tpdhg_model_from_csc", "tpdhg_model_free",
    "tpdhg_model_get_info", "tpdhg_model_export", "tpdhg_model_name_of", "tpdhg_model_input_error",
    "tpdhg_scaled_build", "tpdhg_scaled_view", "tpdhg_scaled_free",
    "tpdhg_dev_create", "tpdhg_dev_power_norm", "tpdhg_dev_set_steps", "tpdhg_dev_start", "tpdhg_dev_iterate",
    "tpdhg_dev_eval", "tpdhg_dev_eval2", "tpdhg_dev_restart", "tpdhg_dev_download", "tpdhg_dev_spmv",
    "tpdhg_dev_info", "tpdhg_dev_set_profile", "tpdhg_dev_kernel_times", "tpdhg_dev_destroy",
    "tpdhg_fast_create", "tpdhg_fast_power_norm", "tpdhg_fast_set_base", "tpdhg_fast_begin_round",
    "tpdhg_fast_set_steps", "tpdhg_fast_restart", "tpdhg_fast_iterate", "tpdhg_fast_eval", "tpdhg_fast_fixed_point",
    "tpdhg_fast_end_round",
    "tpdhg_fast_download", "tpdhg_fast_destroy",
    "tpdhg_options_default", "tpdhg_solve", "tpdhg_session_create", "tpdhg_session_solve", "tpdhg_session_free",
    "tpdhg_session_set_vectors", "tpdhg_session_set_primal_weight",
    "tpdhg_result_json", "tpdhg_spmv_bench", "tpdhg_stream_bench", "tpdhg_spmv_noalloc", "tpdhg_spmv_noalloc_max_nnz",
    "tpdhg_recheck",
    "tpdhg_batch_create", "tpdhg_batch_get_info", "tpdhg_batch_set_vectors", "tpdhg_batch_solve", "tpdhg_batch_probe",
    "tpdhg_batch_free",
    "tpdhg_mlp_create", "tpdhg_mlp_param_count", "tpdhg_mlp_set_params", "tpdhg_mlp_get_params",
    "tpdhg_mlp_forward", "tpdhg_mlp_loss_grad", "tpdhg_mlp_train", "tpdhg_mlp_destroy",
]


class TpdhgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)


class ModelInfo(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("nnz", C.c_int64), ("maximize", C.c_int32),
                ("has_integers", C.c_int32), ("qobj_terms", C.c_int64), ("sentinel_bounds", C.c_int64),
                ("obj_const", C.c_double)]


class ScaledArrays(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("nnz", C.c_int64),
                ("a_ptr", _i32), ("a_idx", _i32), ("a_val", _d), ("at_ptr", _i32), ("at_idx", _i32), ("at_val", _d),
                ("c", _d), ("l", _d), ("u", _d), ("rl", _d), ("ru", _d), ("Dr", _d), ("Dc", _d),
                ("bnorm", C.c_double), ("cnorm", C.c_double)]


class Options(C.Structure):
    _fields_ = [("tol", C.c_double), ("time_limit", C.c_double), ("max_iter", C.c_int64),
                ("eval_freq", C.c_int32), ("row_rel", C.c_double), ("precision_bits", C.c_int32),
                ("refine", C.c_int32), ("inner_tol", C.c_double), ("max_refinements", C.c_int32),
                ("use_graph", C.c_int32), ("verbose", C.c_int32), ("profile_kernels", C.c_int32),
                ("halpern", C.c_int32), ("reflection", C.c_double), ("primal_weight", C.c_double),
                ("warm_fallback", C.c_int32), ("progress_interval", C.c_double), ("omega_rule", C.c_int32),
                ("reuse_norm", C.c_int32), ("cpu_threads", C.c_int32), ("restart_metric", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int64), ("restarts", C.c_int32),
                ("refinements", C.c_int32), ("from_average", C.c_int32), ("objective", C.c_double),
                ("rel_primal", C.c_double), ("rel_dual", C.c_double), ("rel_gap", C.c_double),
                ("max_row_violation", C.c_double), ("max_bound_violation", C.c_double),
                ("setup_s", C.c_double), ("parse_s", C.c_double), ("scale_s", C.c_double),
                ("device_init_s", C.c_double), ("transfer_s", C.c_double), ("solve_s", C.c_double),
                ("power_iter_s", C.c_double), ("total_s", C.c_double), ("norm_a", C.c_double),
                ("primal_weight", C.c_double), ("device_mem_bytes", C.c_int64), ("m", C.c_int32), ("n", C.c_int32),
                ("nnz", C.c_int64), ("row_kernel_ms", C.c_double), ("col_kernel_ms", C.c_double),
                ("row_launches", C.c_int64), ("col_launches", C.c_int64), ("device", C.c_char * 64),
                ("warm_start_used", C.c_int32), ("warm_kkt", C.c_double), ("cold_kkt", C.c_double),
                ("threads", C.c_int32)]


class RecheckResult(C.Structure):
    _fields_ = [("objective", C.c_double), ("row_viol_rel", C.c_double), ("bound_viol", C.c_double),
                ("dual_checked", C.c_int32), ("dual_bound_finite", C.c_int32), ("dual_sign", C.c_int32),
                ("dual_bound", C.c_double), ("dual_obj_loose", C.c_double), ("dual_infeas_rel", C.c_double),
                ("dual_gap_rel", C.c_double)]


class BatchInfo(C.Structure):
    _fields_ = [("count", C.c_int32), ("unique_matrices", C.c_int32), ("cap", C.c_int32),
                ("nnz_total", C.c_int64), ("nnz_stored", C.c_int64), ("device_mem_bytes", C.c_int64),
                ("setup_s", C.c_double), ("tune_ms", C.c_float * 4), ("persist_max", C.c_int32),
                ("persistent_blocks", C.c_int64), ("lockstep_blocks", C.c_int64), ("stage_switches", C.c_int64)]


_lib = None


def lib():
    """The loaded library; raises FileNotFoundError when it was not built (no fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is missing: build it with `make -C taral-lp_amd` "
                                    "(the solver has no Python or CPU fallback)")
        L = C.CDLL(LIB_PATH)
        L.tpdhg_last_error.restype = C.c_char_p
        L.tpdhg_status_string.restype = C.c_char_p
        L.tpdhg_model_name_of.restype = C.c_char_p
        L.tpdhg_model_input_error.restype = C.c_char_p
        L.tpdhg_model_read_mps.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        L.tpdhg_model_read_mps_text.argtypes = [C.c_char_p, C.c_int64, C.POINTER(C.c_void_p)]
        L.tpdhg_model_from_csc.argtypes = [C.c_int32, C.c_int32, _i64, _i32, _d, _d, _d, _d, _d, _d, C.c_double,
                                           C.c_int32, C.POINTER(C.c_void_p)]
        L.tpdhg_model_free.argtypes = [C.c_void_p]
        L.tpdhg_model_get_info.argtypes = [C.c_void_p, C.POINTER(ModelInfo)]
        L.tpdhg_model_export.argtypes = [C.c_void_p, _i64, _i32, _d, _d, _d, _d, _d, _d, C.POINTER(C.c_int8)]
        L.tpdhg_model_name_of.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.tpdhg_model_input_error.argtypes = [C.c_void_p]
        L.tpdhg_scaled_build.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        L.tpdhg_scaled_view.argtypes = [C.c_void_p, C.POINTER(ScaledArrays)]
        L.tpdhg_scaled_free.argtypes = [C.c_void_p]
        L.tpdhg_dev_create.argtypes = [C.POINTER(ScaledArrays), C.c_int32, C.c_int32, C.c_int32,
                                       C.POINTER(C.c_void_p)]
        L.tpdhg_dev_power_norm.argtypes = [C.c_void_p, C.c_int32, _d]
        L.tpdhg_dev_set_steps.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.tpdhg_dev_start.argtypes = [C.c_void_p, _d, _d]
        L.tpdhg_dev_iterate.argtypes = [C.c_void_p, C.c_int32]
        L.tpdhg_dev_eval.argtypes = [C.c_void_p, C.c_int32, C.c_double, _d]
        L.tpdhg_dev_eval2.argtypes = [C.c_void_p, C.c_double, _d, _d]
        L.tpdhg_dev_restart.argtypes = [C.c_void_p, C.c_int32, C.c_double]
        L.tpdhg_dev_download.argtypes = [C.c_void_p, C.c_int32, C.c_double, _d, _d]
        L.tpdhg_dev_spmv.argtypes = [C.c_void_p, C.c_int32, _d, _d]
        L.tpdhg_dev_info.argtypes = [C.c_void_p, _i64, C.c_char_p, C.c_int32]
        L.tpdhg_dev_set_profile.argtypes = [C.c_void_p, C.c_int32]
        L.tpdhg_dev_kernel_times.argtypes = [C.c_void_p, _d, _d, _i64, _i64]
        L.tpdhg_dev_destroy.argtypes = [C.c_void_p]
        L.tpdhg_fast_create.argtypes = [C.POINTER(ScaledArrays), C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                        C.POINTER(C.c_void_p)]
        L.tpdhg_fast_power_norm.argtypes = [C.c_void_p, C.c_int32, _d]
        L.tpdhg_fast_set_base.argtypes = [C.c_void_p, _d, _d]
        for nm in ("begin_round", "restart", "end_round", "destroy"):
            getattr(L, "tpdhg_fast_" + nm).argtypes = [C.c_void_p]
        L.tpdhg_fast_set_steps.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.tpdhg_fast_iterate.argtypes = [C.c_void_p, C.c_int32]
        L.tpdhg_fast_eval.argtypes = [C.c_void_p, _d]
        L.tpdhg_fast_fixed_point.argtypes = [C.c_void_p, _d]
        L.tpdhg_fast_download.argtypes = [C.c_void_p, _d, _d]
        L.tpdhg_options_default.argtypes = [C.POINTER(Options)]
        L.tpdhg_solve.argtypes = [C.c_void_p, C.POINTER(Options), _d, _d, C.POINTER(Result), _d, _d]
        L.tpdhg_session_create.argtypes = [C.c_void_p, C.POINTER(Options), C.POINTER(C.c_void_p)]
        L.tpdhg_session_solve.argtypes = [C.c_void_p, _d, _d, C.POINTER(Result), _d, _d]
        L.tpdhg_session_free.argtypes = [C.c_void_p]
        L.tpdhg_session_set_vectors.argtypes = [C.c_void_p, _d, _d, _d, _d, _d]
        L.tpdhg_session_set_primal_weight.argtypes = [C.c_void_p, C.c_double]
        L.tpdhg_result_json.argtypes = [C.c_void_p, C.POINTER(Options), C.POINTER(Result), C.c_char_p, C.c_int32]
        L.tpdhg_spmv_bench.argtypes = [C.c_int32, C.c_int32, C.c_int32, _i32, _i32, _d, _d, _d, C.c_int32, _d, _d]
        L.tpdhg_stream_bench.argtypes = [C.c_int64, C.c_int32, _d, _d]
        L.tpdhg_spmv_noalloc.argtypes = [C.c_int32, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p]
        L.tpdhg_spmv_noalloc_max_nnz.restype = C.c_int64
        L.tpdhg_recheck.argtypes = [C.c_void_p, _d, _d, C.POINTER(RecheckResult)]
        _pp = C.POINTER(C.c_void_p)
        L.tpdhg_batch_create.argtypes = [_pp, C.c_int32, C.POINTER(Options), C.c_int32, C.POINTER(C.c_void_p)]
        L.tpdhg_batch_get_info.argtypes = [C.c_void_p, C.POINTER(BatchInfo)]
        L.tpdhg_batch_set_vectors.argtypes = [C.c_void_p, C.c_int32, _d, _d, _d, _d, _d]
        L.tpdhg_batch_solve.argtypes = [C.c_void_p, _pp, _pp, _d, C.POINTER(Result), _pp, _pp]
        L.tpdhg_batch_probe.argtypes = [C.c_void_p, C.c_int32, _d, _d, _i64, _i64]
        L.tpdhg_batch_free.argtypes = [C.c_void_p]
        _f = C.POINTER(C.c_float)
        L.tpdhg_mlp_create.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)]
        L.tpdhg_mlp_param_count.argtypes = [C.c_void_p]
        L.tpdhg_mlp_param_count.restype = C.c_int64
        L.tpdhg_mlp_set_params.argtypes = [C.c_void_p, _f]
        L.tpdhg_mlp_get_params.argtypes = [C.c_void_p, _f]
        L.tpdhg_mlp_forward.argtypes = [C.c_void_p, _f, C.c_int32, _f]
        L.tpdhg_mlp_loss_grad.argtypes = [C.c_void_p, _f, _f, C.c_int32, _d, _f]
        L.tpdhg_mlp_train.argtypes = [C.c_void_p, _f, _f, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _d]
        L.tpdhg_mlp_destroy.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _check(code):
    if code != OK:
        raise TpdhgError(code, lib().tpdhg_last_error().decode(errors="replace"))


def _p(a, t=_d):
    return None if a is None else a.ctypes.data_as(t)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def device_count() -> int:
    return lib().tpdhg_device_count()


class Model:
    """The input contract: an LP read from MPS (read_mps, src/mps.cpp:388) or built from column-wise arrays."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    @classmethod
    def from_mps(cls, path: str) -> "Model":
        h = C.c_void_p()
        _check(lib().tpdhg_model_read_mps(os.fsencode(path), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_mps_text(cls, text: str) -> "Model":
        b = text.encode()
        h = C.c_void_p()
        _check(lib().tpdhg_model_read_mps_text(b, len(b), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_csc(cls, m, n, colptr, rowidx, val, cost, col_lo, col_up, row_lo, row_up, obj_const=0.0,
                 maximize=False) -> "Model":
        keep = [np.ascontiguousarray(colptr, dtype=np.int64), np.ascontiguousarray(rowidx, dtype=np.int32)]
        keep += [_f64(a) for a in (val, cost, col_lo, col_up, row_lo, row_up)]
        h = C.c_void_p()
        _check(lib().tpdhg_model_from_csc(int(m), int(n), _p(keep[0], _i64), _p(keep[1], _i32),
                                          *[_p(a) for a in keep[2:]], float(obj_const), int(bool(maximize)),
                                          C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_lp(cls, md) -> "Model":
        """From a gen.LPModel."""
        return cls.from_csc(md.m, md.n, md.colptr, md.rowidx, md.val, md.cost, md.col_lo, md.col_up, md.row_lo,
                            md.row_up, md.obj_const, md.maximize)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.tpdhg_model_free(self._h)
            self._h = C.c_void_p()

    def info(self) -> ModelInfo:
        i = ModelInfo()
        _check(lib().tpdhg_model_get_info(self._h, C.byref(i)))
        return i

    def arrays(self) -> dict:
        i = self.info()
        m, n, nnz = i.m, i.n, i.nnz
        out = dict(colptr=np.zeros(n + 1, np.int64), rowidx=np.zeros(max(nnz, 1), np.int32),
                   val=np.zeros(max(nnz, 1)), cost=np.zeros(max(n, 1)), col_lo=np.zeros(max(n, 1)),
                   col_up=np.zeros(max(n, 1)), row_lo=np.zeros(max(m, 1)), row_up=np.zeros(max(m, 1)),
                   is_int=np.zeros(max(n, 1), np.int8))
        _check(lib().tpdhg_model_export(self._h, _p(out["colptr"], _i64), _p(out["rowidx"], _i32), _p(out["val"]),
                                        _p(out["cost"]), _p(out["col_lo"]), _p(out["col_up"]), _p(out["row_lo"]),
                                        _p(out["row_up"]), out["is_int"].ctypes.data_as(C.POINTER(C.c_int8))))
        for k, size in (("rowidx", nnz), ("val", nnz), ("cost", n), ("col_lo", n), ("col_up", n), ("row_lo", m),
                        ("row_up", m), ("is_int", n)):
            out[k] = out[k][:size]
        return out

    def name_of(self, is_row: bool, i: int) -> str:
        return lib().tpdhg_model_name_of(self._h, int(bool(is_row)), int(i)).decode()

    def input_error(self):
        r = lib().tpdhg_model_input_error(self._h)
        return None if r is None else r.decode()


class Scaled:
    """build_scaled (gpu/pdhg.cu:80-133) on the host in FP64; arrays are copied out on access."""

    def __init__(self, model: Model):
        h = C.c_void_p()
        _check(lib().tpdhg_scaled_build(model._h, C.byref(h)))
        self._h = h
        self.view = ScaledArrays()
        _check(lib().tpdhg_scaled_view(self._h, C.byref(self.view)))

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.tpdhg_scaled_free(self._h)
            self._h = C.c_void_p()

    def vec(self, name):
        k = self.view.m if name in ("rl", "ru", "Dr") else self.view.n
        return np.ctypeslib.as_array(getattr(self.view, name), shape=(k,)).copy() if k else np.zeros(0)

    def csr(self, which="A"):
        rows = self.view.m if which == "A" else self.view.n
        pre = "a_" if which == "A" else "at_"
        nnz = self.view.nnz
        g = lambda k, s: np.ctypeslib.as_array(getattr(self.view, pre + k), shape=(s,)).copy()  # noqa: E731
        return g("ptr", rows + 1), (g("idx", nnz) if nnz else np.zeros(0, np.int32)), (g("val", nnz) if nnz else np.zeros(0))


class Device:
    """The device backend (Gpu<T>, gpu/pdhg.cu:313-514) on the current HIP device."""

    def __init__(self, scaled: Scaled, precision_bits=64, eval_freq=64, use_graph=True):
        self.scaled = scaled  # keeps the host arrays alive
        self.m, self.n = scaled.view.m, scaled.view.n
        h = C.c_void_p()
        _check(lib().tpdhg_dev_create(C.byref(scaled.view), int(precision_bits), int(eval_freq), int(bool(use_graph)),
                                      C.byref(h)))
        self._h = h

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.tpdhg_dev_destroy(self._h)
            self._h = C.c_void_p()

    def power_norm(self, iters=200) -> float:
        r = C.c_double()
        _check(lib().tpdhg_dev_power_norm(self._h, iters, C.byref(r)))
        return r.value

    def set_steps(self, tau, sigma):
        _check(lib().tpdhg_dev_set_steps(self._h, float(tau), float(sigma)))

    def start(self, x0=None, z0=None):
        x0, z0 = _f64(x0), _f64(z0)
        _check(lib().tpdhg_dev_start(self._h, _p(x0), _p(z0)))

    def iterate(self, steps):
        _check(lib().tpdhg_dev_iterate(self._h, int(steps)))

    def eval(self, avg=False, f=1.0):
        out = np.zeros(10)
        _check(lib().tpdhg_dev_eval(self._h, int(bool(avg)), float(f), _p(out)))
        return out

    def eval2(self, f=1.0):
        cur, avg = np.zeros(10), np.zeros(10)
        _check(lib().tpdhg_dev_eval2(self._h, float(f), _p(cur), _p(avg)))
        return cur, avg

    def restart(self, avg, f):
        _check(lib().tpdhg_dev_restart(self._h, int(bool(avg)), float(f)))

    def download(self, avg=False, f=1.0):
        x, z = np.zeros(max(self.n, 1)), np.zeros(max(self.m, 1))
        _check(lib().tpdhg_dev_download(self._h, int(bool(avg)), float(f), _p(x), _p(z)))
        return x[:self.n], z[:self.m]

    def spmv(self, v, transpose=False):
        v = _f64(v)
        y = np.zeros(max(self.n if transpose else self.m, 1))
        _check(lib().tpdhg_dev_spmv(self._h, int(bool(transpose)), _p(v), _p(y)))
        return y[:(self.n if transpose else self.m)]

    def info(self):
        mem = C.c_int64()
        buf = C.create_string_buffer(128)
        _check(lib().tpdhg_dev_info(self._h, C.byref(mem), buf, 128))
        return dict(mem_used=mem.value, name=buf.value.decode())

    def set_profile(self, on=True):
        _check(lib().tpdhg_dev_set_profile(self._h, int(bool(on))))

    def kernel_times(self):
        a, b = C.c_double(), C.c_double()
        na, nb = C.c_int64(), C.c_int64()
        _check(lib().tpdhg_dev_kernel_times(self._h, C.byref(a), C.byref(b), C.byref(na), C.byref(nb)))
        return dict(row_ms=a.value, col_ms=b.value, row_launches=na.value, col_launches=nb.value)


class FastDevice:
    """The fast-path backend (reflected Halpern PDHG around an FP64 base point) on the current HIP device."""

    def __init__(self, scaled: Scaled, precision_bits=32, eval_freq=64, use_graph=True, reflection=1.9):
        self.scaled = scaled
        self.m, self.n = scaled.view.m, scaled.view.n
        h = C.c_void_p()
        _check(lib().tpdhg_fast_create(C.byref(scaled.view), int(precision_bits), int(eval_freq),
                                       int(bool(use_graph)), float(reflection), C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.tpdhg_fast_destroy(self._h)
            self._h = C.c_void_p()

    def power_norm(self, iters=200):
        r = C.c_double()
        _check(lib().tpdhg_fast_power_norm(self._h, iters, C.byref(r)))
        return r.value

    def set_base(self, x_b=None, z_b=None):
        x_b, z_b = _f64(x_b), _f64(z_b)
        _check(lib().tpdhg_fast_set_base(self._h, _p(x_b), _p(z_b)))

    def begin_round(self):
        _check(lib().tpdhg_fast_begin_round(self._h))

    def set_steps(self, tau, sigma):
        _check(lib().tpdhg_fast_set_steps(self._h, float(tau), float(sigma)))

    def restart(self):
        _check(lib().tpdhg_fast_restart(self._h))

    def iterate(self, steps):
        _check(lib().tpdhg_fast_iterate(self._h, int(steps)))

    def eval(self):
        out = np.zeros(10)
        _check(lib().tpdhg_fast_eval(self._h, _p(out)))
        return out

    def fixed_point(self):
        """(||x - x+||^2, ||z - z+||^2, (z - z+)' A (x - x+)) of the inner point at the last eval()."""
        out = np.zeros(3)
        _check(lib().tpdhg_fast_fixed_point(self._h, _p(out)))
        return out

    def end_round(self):
        _check(lib().tpdhg_fast_end_round(self._h))

    def download(self):
        x, z = np.zeros(max(self.n, 1)), np.zeros(max(self.m, 1))
        _check(lib().tpdhg_fast_download(self._h, _p(x), _p(z)))
        return x[:self.n], z[:self.m]


def default_options(**kw) -> Options:
    o = Options()
    lib().tpdhg_options_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def _result_dict(model, o, r, x, y, m, n):
    out = {name: getattr(r, name) for name, _ in Result._fields_ if name != "device"}
    out["device"] = r.device.decode(errors="replace")
    out["status"] = STATUS[r.status]
    out["x"], out["y"] = x[:n], y[:m]
    buf = C.create_string_buffer(8192)
    _check(lib().tpdhg_result_json(model._h, C.byref(o), C.byref(r), buf, 8192))
    out["json"] = buf.value.decode()
    return out


class Session:
    """A problem resident on the GPU (scaled and uploaded once); each solve() is a full solve from a cold or
    warm start."""

    def __init__(self, model: Model, **opts):
        self.model = model
        self.opts = default_options(**opts)
        i = model.info()
        self.m, self.n = i.m, i.n
        h = C.c_void_p()
        _check(lib().tpdhg_session_create(model._h, C.byref(self.opts), C.byref(h)))
        self._h = h

    def solve(self, x0=None, y0=None, primal_weight=None) -> dict:
        """primal_weight: initial primal weight of this and later warm-started solves (0 = default)."""
        if primal_weight is not None:
            _check(lib().tpdhg_session_set_primal_weight(self._h, float(primal_weight)))
        x, y = np.zeros(max(self.n, 1)), np.zeros(max(self.m, 1))
        r = Result()
        x0, y0 = _f64(x0), _f64(y0)
        _check(lib().tpdhg_session_solve(self._h, _p(x0), _p(y0), C.byref(r), _p(x), _p(y)))
        return _result_dict(self.model, self.opts, r, x, y, self.m, self.n)

    def set_vectors(self, cost, col_lo, col_up, row_lo, row_up):
        """New cost and bounds (original space) for the same A (scenario families)."""
        keep = [_f64(v) for v in (cost, col_lo, col_up, row_lo, row_up)]
        _check(lib().tpdhg_session_set_vectors(self._h, *[_p(v) for v in keep]))

    def close(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.tpdhg_session_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()


def solve(model: Model, x0=None, y0=None, **opts) -> dict:
    """Scale, upload and solve on the GPU (the pdhg CLI path). Returns the result fields plus x (original space)
    and y = Dr z (row multipliers, minimisation form)."""
    o = default_options(**opts)
    i = model.info()
    x, y = np.zeros(max(i.n, 1)), np.zeros(max(i.m, 1))
    r = Result()
    x0, y0 = _f64(x0), _f64(y0)
    _check(lib().tpdhg_solve(model._h, C.byref(o), _p(x0), _p(y0), C.byref(r), _p(x), _p(y)))
    return _result_dict(model, o, r, x, y, i.m, i.n)


def recheck(model: "Model", x, y=None) -> dict:
    """Original-model recheck of a point (objective, violations, weak-duality bound from y); host only."""
    x = _f64(x)
    y = _f64(y)
    r = RecheckResult()
    _check(lib().tpdhg_recheck(model._h, _p(x), _p(y), C.byref(r)))
    out = {k: getattr(r, k) for k, _ in RecheckResult._fields_}
    if not r.dual_bound_finite:
        out["dual_bound"] = None
    return out


def _ptr_array(arrays):
    """(void* array for C, keep-alive list) from a list of float64 arrays or Nones."""
    keep = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in arrays]
    arr = (C.c_void_p * len(keep))(*[None if a is None else a.ctypes.data for a in keep])
    return arr, keep


class Batch:
    """Many LPs solved together on the GPU by the fast path (block-diagonal packing, per-instance scalars). Every
    instance takes the decisions a single solve with the same options takes; identical matrices are stored once."""

    def __init__(self, models, cap=0, **opts):
        opts.setdefault("halpern", 1)
        self.models = list(models)
        self.opts = default_options(**opts)
        self.dims = [(md.info().m, md.info().n) for md in self.models]
        arr = (C.c_void_p * len(self.models))(*[md._h.value for md in self.models])
        h = C.c_void_p()
        _check(lib().tpdhg_batch_create(arr, len(self.models), C.byref(self.opts), int(cap), C.byref(h)))
        self._h = h

    def __len__(self):
        return len(self.models)

    def info(self) -> dict:
        i = BatchInfo()
        _check(lib().tpdhg_batch_get_info(self._h, C.byref(i)))
        out = {k: getattr(i, k) for k, _ in BatchInfo._fields_ if k != "tune_ms"}
        out["tune_ms"] = list(i.tune_ms)
        return out

    def probe(self, iters=24) -> dict:
        """HIP-event times of `iters` all-live batched iterations (row and column launches apart)."""
        r, c = C.c_double(), C.c_double()
        nr, nc = C.c_int64(), C.c_int64()
        _check(lib().tpdhg_batch_probe(self._h, int(iters), C.byref(r), C.byref(c), C.byref(nr), C.byref(nc)))
        return dict(row_ms=r.value, col_ms=c.value, row_launches=nr.value, col_launches=nc.value)

    def set_vectors(self, i, cost, col_lo, col_up, row_lo, row_up):
        keep = [_f64(v) for v in (cost, col_lo, col_up, row_lo, row_up)]
        _check(lib().tpdhg_batch_set_vectors(self._h, int(i), *[_p(v) for v in keep]))

    def solve(self, x0=None, y0=None, primal_weight=None, points=True) -> list:
        B = len(self.models)
        res = (Result * B)()
        ax, kx = _ptr_array(x0 if x0 is not None else [None] * B)
        ay, ky = _ptr_array(y0 if y0 is not None else [None] * B)
        pw = None if primal_weight is None else np.ascontiguousarray(primal_weight, dtype=np.float64)
        xs = [np.zeros(max(n, 1)) for (m, n) in self.dims] if points else None
        ys = [np.zeros(max(m, 1)) for (m, n) in self.dims] if points else None
        px = (C.c_void_p * B)(*[a.ctypes.data for a in xs]) if points else None
        py = (C.c_void_p * B)(*[a.ctypes.data for a in ys]) if points else None
        _check(lib().tpdhg_batch_solve(self._h, ax if x0 is not None else None, ay if y0 is not None else None,
                                       _p(pw), res, px, py))
        out = []
        for i, (m, n) in enumerate(self.dims):
            x = xs[i] if points else np.zeros(0)
            y = ys[i] if points else np.zeros(0)
            out.append(_result_dict(self.models[i], self.opts, res[i], x, y, m, n))
        return out

    def close(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.tpdhg_batch_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_f32p = C.POINTER(C.c_float)


def _f32(a, cols):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != cols:
        raise ValueError(f"expected a 2-D array with {cols} columns, got shape {a.shape}")
    return a


class Mlp:
    """Warm-start network on the GPU: n_in -> h1 -> h2 -> n_out, ReLU, bf16 MFMA products, FP32 master weights.
    Parameters are one FP32 vector in the order W1 [n_in x h1], b1, W2 [h1 x h2], b2, W3 [h2 x n_out], b3."""

    def __init__(self, n_in, h1, h2, n_out, seed=0):
        self.shape = (int(n_in), int(h1), int(h2), int(n_out))
        h = C.c_void_p()
        _check(lib().tpdhg_mlp_create(*self.shape, int(seed), C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            lib().tpdhg_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def param_count(self) -> int:
        return int(lib().tpdhg_mlp_param_count(self._h))

    def get_params(self) -> np.ndarray:
        p = np.empty(self.param_count(), dtype=np.float32)
        _check(lib().tpdhg_mlp_get_params(self._h, p.ctypes.data_as(_f32p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.param_count():
            raise ValueError(f"expected {self.param_count()} parameters, got {p.size}")
        _check(lib().tpdhg_mlp_set_params(self._h, p.ctypes.data_as(_f32p)))

    def forward(self, x) -> np.ndarray:
        x = _f32(x, self.shape[0])
        y = np.empty((x.shape[0], self.shape[3]), dtype=np.float32)
        _check(lib().tpdhg_mlp_forward(self._h, x.ctypes.data_as(_f32p), x.shape[0], y.ctypes.data_as(_f32p)))
        return y

    def loss_grad(self, x, t):
        x, t = _f32(x, self.shape[0]), _f32(t, self.shape[3])
        if x.shape[0] != t.shape[0]:
            raise ValueError("x and t differ in rows")
        g = np.empty(self.param_count(), dtype=np.float32)
        loss = C.c_double()
        _check(lib().tpdhg_mlp_loss_grad(self._h, x.ctypes.data_as(_f32p), t.ctypes.data_as(_f32p), x.shape[0],
                                         C.byref(loss), g.ctypes.data_as(_f32p)))
        return loss.value, g

    def train(self, x, t, batch=256, epochs=10, lr=1e-3, seed=0) -> float:
        x, t = _f32(x, self.shape[0]), _f32(t, self.shape[3])
        if x.shape[0] != t.shape[0]:
            raise ValueError("x and t differ in rows")
        loss = C.c_double()
        _check(lib().tpdhg_mlp_train(self._h, x.ctypes.data_as(_f32p), t.ctypes.data_as(_f32p), x.shape[0],
                                     int(batch), int(epochs), float(lr), int(seed), C.byref(loss)))
        return loss.value


def spmv_bench(ptr, idx, val, v, precision_bits=32, reps=20):
    """y = M v with the solver's SpMV kernel; returns (y, mean_ms, median_ms) from HIP events."""
    ptr = np.ascontiguousarray(ptr, dtype=np.int32)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    val, v = _f64(val), _f64(v)
    rows = ptr.size - 1
    y = np.zeros(max(rows, 1))
    mean, med = C.c_double(), C.c_double()
    _check(lib().tpdhg_spmv_bench(int(precision_bits), rows, v.size, _p(ptr, _i32), _p(idx, _i32), _p(val), _p(v),
                                  _p(y), int(reps), C.byref(mean), C.byref(med)))
    return y[:rows], mean.value, med.value


def stream_bench(nbytes=1 << 30, reps=20) -> dict:
    """Median GB/s of the library's streaming-read and copy kernels over nbytes of HBM."""
    r, c = C.c_double(), C.c_double()
    _check(lib().tpdhg_stream_bench(int(nbytes), int(reps), C.byref(r), C.byref(c)))
    return dict(bytes=int(nbytes), read_gbs=r.value, copy_gbs=c.value)


def spmv_noalloc(precision_bits, rows, nnz, ptr, idx, val, v, y, stream=0):
    """The allocation-free SpMV on device pointers (integers: e.g. torch tensors' data_ptr()), enqueued on the
    HIP stream handle `stream` (0 = the default stream); returns without waiting."""
    _check(lib().tpdhg_spmv_noalloc(int(precision_bits), int(rows), int(nnz), ptr, idx, val, v, y,
                                    C.c_void_p(stream) if stream else None))
