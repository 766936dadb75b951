"""The C-ABI library and the pdhg CLI on the CPU: exported symbols, the input gate, exit codes.

No compute call needs a GPU here: the gate runs before any device work, and without a device the solve calls
report TPDHG_ERR_NO_DEVICE instead of falling back to a CPU path.
"""
import os
import re
import subprocess

import pytest

from taral_pdhg import ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MPS = os.path.join(ROOT, "tests", "golden", "mps")


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "taral_pdhg.h")).read()
    declared = set(re.findall(r"\b(tpdhg_[a-z0-9_]+)\s*\(", header))
    assert declared == set(ffi.EXPORTS)
    lib = ffi.lib()
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", ffi.LIB_PATH], capture_output=True, text=True).stdout
    for name in declared:
        assert re.search(rf"\bT {name}$", out, re.M), name


GATE_REJECT = {"gate_quadobj": ("quadratic",), "gate_qmatrix": ("quadratic",), "gate_qsection": ("quadratic",),
               "gate_maximize_qp": ("quadratic",), "gate_intorg": ("integer",), "gate_binary_bound": ("integer",),
               "gate_integer_lower": ("integer",), "gate_integer_upper": ("integer",),
               "gate_fixed_integer": ("integer",), "gate_miqp": ("quadratic", "integer")}


@pytest.mark.parametrize("case", sorted(GATE_REJECT))
@pytest.mark.parametrize("device", ["gpu", "cpu"])
@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("existing", [False, True])
def test_cli_rejects_unsupported_models(case, device, fp32, existing, tmp_path):
    """gpu/pdhg_input_tests.py:89-109: exit 2, stdout exactly the status line, the reason on stderr, and no
    output file created or modified."""
    outs = [tmp_path / "a.json", tmp_path / "a.sol", tmp_path / "a.dual"]
    for o in outs:
        if existing:
            o.write_text("sentinel\n")
    cmd = [ffi.CLI_PATH, os.path.join(MPS, case + ".mps"), "--device", device, "--json", str(outs[0]), "--sol",
           str(outs[1]), "--dual", str(outs[2])] + (["--fp32"] if fp32 else [])
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 2
    assert r.stdout.strip() == "status unsupported_model"
    assert "continuous LPs only" in r.stderr and all(w in r.stderr for w in GATE_REJECT[case])
    for o in outs:
        assert (o.read_text() == "sentinel\n") if existing else not o.exists()


def test_c_abi_gate_and_no_device():
    md = ffi.Model.from_mps(os.path.join(MPS, "gate_quadobj.mps"))
    with pytest.raises(ffi.TpdhgError) as e:
        ffi.solve(md)
    assert e.value.code == ffi.ERR_UNSUPPORTED and "quadratic" in str(e.value)
    if ffi.device_count() == 0:
        ok = ffi.Model.from_mps(os.path.join(MPS, "gate_continuous_lp.mps"))
        with pytest.raises(ffi.TpdhgError) as e:
            ffi.solve(ok)
        assert e.value.code == ffi.ERR_NO_DEVICE


def test_cli_parse_error_and_usage(tmp_path):
    r = subprocess.run([ffi.CLI_PATH, os.path.join(MPS, "err_bad_number.mps")], capture_output=True, text=True)
    assert r.returncode == 1 and r.stdout.startswith("status parse_error: bad number '1e'")
    r = subprocess.run([ffi.CLI_PATH], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([ffi.CLI_PATH, os.path.join(MPS, "gate_continuous_lp.mps"), "--device", "tpu"],
                       capture_output=True, text=True)
    assert r.returncode == 2
    # the fast path on the CPU backend: refused (it runs on the GPU only), no file written
    out = tmp_path / "x.json"
    r = subprocess.run([ffi.CLI_PATH, os.path.join(MPS, "gate_continuous_lp.mps"), "--device", "cpu", "--fast",
                        "--json", str(out)], capture_output=True, text=True)
    assert r.returncode == 2 and "GPU only" in r.stderr and not out.exists()
    # fast-path settings out of range: refused before the model is read or a device touched
    for flag, bad in (("--omega-rule", "5"), ("--restart-metric", "2")):
        r = subprocess.run([ffi.CLI_PATH, os.path.join(MPS, "gate_continuous_lp.mps"), "--device", "gpu", "--fast",
                            flag, bad, "--json", str(out)], capture_output=True, text=True)
        assert r.returncode == 2 and flag in r.stderr and not out.exists()


def test_model_from_csc_validates():
    import numpy as np
    good = dict(m=2, n=2, colptr=[0, 1, 2], rowidx=[0, 1], val=[1.0, 2.0], cost=[1.0, 1.0], col_lo=[0, 0],
                col_up=[1, 1], row_lo=[0, 0], row_up=[1, 1])
    assert ffi.Model.from_csc(**good).info().nnz == 2
    for bad in (dict(rowidx=[0, 2]), dict(colptr=[0, 2, 1]), dict(val=[1.0, np.nan])):
        with pytest.raises(ffi.TpdhgError) as e:
            ffi.Model.from_csc(**{**good, **bad})
        assert e.value.code == ffi.ERR_ARG


def test_session_refuses_without_device_or_bad_options():
    md = ffi.Model.from_mps(os.path.join(MPS, "gate_continuous_lp.mps"))
    with pytest.raises(ffi.TpdhgError) as e:
        ffi.Session(md, precision_bits=16)
    assert e.value.code == ffi.ERR_ARG
    with pytest.raises(ffi.TpdhgError) as e:
        ffi.Session(md, precision_bits=64, refine=1)
    assert e.value.code == ffi.ERR_ARG


def test_mlp_create_checks_widths_then_device():
    with pytest.raises(ffi.TpdhgError) as e:
        ffi.Mlp(30, 512, 512, 257)
    assert e.value.code == ffi.ERR_ARG and "multiples of 16" in str(e.value)
    if ffi.device_count() == 0:
        with pytest.raises(ffi.TpdhgError) as e:
            ffi.Mlp(32, 512, 512, 257)
        assert e.value.code == ffi.ERR_NO_DEVICE


def test_options_carry_warm_start_fields():
    o = ffi.default_options()
    assert o.primal_weight == 0 and o.warm_fallback == 0
    assert ffi.Result().warm_start_used == 0


def _passes_host_checks(call):
    """A valid setting passes the host-side checks: without a GPU the call then stops at the device check, with one
    it solves."""
    try:
        call()
    except ffi.TpdhgError as e:
        assert e.code == ffi.ERR_NO_DEVICE, e


def test_restart_metric_option():
    """The fast path restarts on the fixed-point residual by default; only 0 (weighted KKT) and 1 are accepted,
    refused on the host before any device work."""
    from taral_pdhg import gen
    assert ffi.default_options().restart_metric == 1 and ffi.default_options().omega_rule == 3
    md = ffi.Model.from_lp(gen.planted_lp(20, 40, 3, seed=0))
    for bad in (-1, 2):
        with pytest.raises(ffi.TpdhgError) as e:
            ffi.solve(md, halpern=1, restart_metric=bad)
        assert e.value.code == ffi.ERR_ARG and "restart_metric" in str(e.value)
        with pytest.raises(ffi.TpdhgError) as e:
            ffi.Batch([md], restart_metric=bad)
        assert e.value.code == ffi.ERR_ARG and "restart_metric" in str(e.value)
    _passes_host_checks(lambda: ffi.solve(md, halpern=1, restart_metric=0))


def test_omega_rule_option():
    """Weight guards 0-4 are accepted (4: the bound anchored at the first restart's weight), others refused on the
    host before any device work."""
    from taral_pdhg import gen
    md = ffi.Model.from_lp(gen.planted_lp(20, 40, 3, seed=0))
    for bad in (-1, 5):
        with pytest.raises(ffi.TpdhgError) as e:
            ffi.solve(md, halpern=1, omega_rule=bad)
        assert e.value.code == ffi.ERR_ARG and "omega_rule" in str(e.value)
    _passes_host_checks(lambda: ffi.solve(md, halpern=1, omega_rule=4))


def test_batch_create_validates_before_the_device():
    from taral_pdhg import gen
    md = ffi.Model.from_lp(gen.planted_lp(20, 40, 3, seed=0))
    with pytest.raises(ffi.TpdhgError) as e:
        ffi.Batch([md], halpern=0)
    assert e.value.code == ffi.ERR_ARG and "halpern" in str(e.value)
    with pytest.raises(ffi.TpdhgError) as e:
        ffi.Batch([md], cap=3000)
    assert e.value.code == ffi.ERR_ARG
    mip = ffi.Model.from_mps(os.path.join(os.path.dirname(__file__), "golden", "mps", "gate_intorg.mps"))
    with pytest.raises(ffi.TpdhgError) as e:
        ffi.Batch([md, mip])
    assert e.value.code == ffi.ERR_UNSUPPORTED
    if ffi.device_count() == 0:
        with pytest.raises(ffi.TpdhgError) as e:
            ffi.Batch([md, md])
        assert e.value.code == ffi.ERR_NO_DEVICE


def test_cli_batch_list_gates_before_the_device(tmp_path):
    lst = tmp_path / "list.txt"
    r = subprocess.run([ffi.CLI_PATH, "--batch", str(tmp_path / "missing.txt")], capture_output=True, text=True)
    assert r.returncode == 2
    lst.write_text("# comment\n\n" + os.path.join(MPS, "gate_continuous_lp.mps") + "\n"
                   + os.path.join(MPS, "gate_intorg.mps") + "\n")
    r = subprocess.run([ffi.CLI_PATH, "--batch", str(lst)], capture_output=True, text=True)
    assert r.returncode == 2 and r.stdout.strip() == "status unsupported_model" and "gate_intorg" in r.stderr
    lst.write_text(os.path.join(MPS, "err_bad_number.mps") + "\n")
    r = subprocess.run([ffi.CLI_PATH, "--batch", str(lst)], capture_output=True, text=True)
    assert r.returncode == 1 and r.stdout.startswith("status parse_error:")
    if ffi.device_count() == 0:
        lst.write_text(os.path.join(MPS, "gate_continuous_lp.mps") + "\n")
        r = subprocess.run([ffi.CLI_PATH, "--batch", str(lst)], capture_output=True, text=True)
        assert r.returncode == 1 and "no HIP device" in r.stderr


def _arrays(m, n, a, at):
    """A tpdhg_scaled_arrays over numpy arrays (kept alive in the returned tuple)."""
    import ctypes as C
    import numpy as np
    keep = [np.ascontiguousarray(x) for x in (*a, *at)] + [np.ones(max(n, 1)), np.ones(max(m, 1))]
    d = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))
    i = lambda x: x.ctypes.data_as(C.POINTER(C.c_int32))
    col, row = keep[6], keep[7]
    s = ffi.ScaledArrays(m=m, n=n, nnz=int(a[0][-1]), a_ptr=i(keep[0]), a_idx=i(keep[1]), a_val=d(keep[2]),
                         at_ptr=i(keep[3]), at_idx=i(keep[4]), at_val=d(keep[5]), c=d(col), l=d(col), u=d(col),
                         rl=d(row), ru=d(row), Dr=d(row), Dc=d(col), bnorm=1.0, cnorm=1.0)
    return s, keep


def test_raw_csr_entry_points_validate_before_the_device():
    """tpdhg_spmv_bench, tpdhg_dev_create and tpdhg_fast_create take a caller's CSR, which the kernels index
    without bounds checks: a malformed one is refused on the host (TPDHG_ERR_ARG) before any device work; a valid
    one gets as far as the device (TPDHG_ERR_NO_DEVICE here when no GPU is visible)."""
    import ctypes as C
    import numpy as np
    ptr, idx, val = np.array([0, 2, 3], np.int32), np.array([0, 2, 1], np.int32), np.array([1.0, 2.0, 3.0])
    v = np.ones(3)
    bad = {
        "row pointers must start at 0": (np.array([1, 2, 3], np.int32), idx, val),
        "row pointers must not decrease": (np.array([0, 3, 2], np.int32), idx, val),
        "column index out of range": (ptr, np.array([0, 3, 1], np.int32), val),
    }
    for msg, (p, i, x) in bad.items():
        with pytest.raises(ffi.TpdhgError) as e:
            ffi.spmv_bench(p, i, x, v)
        assert e.value.code == ffi.ERR_ARG and msg in str(e.value), msg
    with pytest.raises(ffi.TpdhgError) as e:
        ffi.spmv_bench(ptr, np.array([0, -1, 1], np.int32), val, v)
    assert e.value.code == ffi.ERR_ARG and "column index out of range" in str(e.value)
    # A: 2 x 3 with rows {0, 2} and {1}; A^T: 3 x 2
    a = (ptr, idx, val)
    at = (np.array([0, 1, 2, 3], np.int32), np.array([0, 1, 0], np.int32), np.array([1.0, 3.0, 2.0]))
    at_bad = (at[0], np.array([0, 2, 0], np.int32), at[2])      # a row index of A^T beyond m = 2
    at_short = (np.array([0, 1, 2, 2], np.int32), at[1], at[2])  # ends at 2 nonzeros, A has 3
    L = ffi.lib()
    for arr, msg in ((at_bad, "column index out of range"), (at_short, "nonzero count")):
        s, keep = _arrays(2, 3, a, arr)
        h = C.c_void_p()
        assert L.tpdhg_dev_create(C.byref(s), 64, 64, 1, C.byref(h)) == ffi.ERR_ARG
        assert msg in L.tpdhg_last_error().decode()
        assert L.tpdhg_fast_create(C.byref(s), 32, 64, 1, 1.9, C.byref(h)) == ffi.ERR_ARG
        assert msg in L.tpdhg_last_error().decode()
    if ffi.device_count() == 0:
        with pytest.raises(ffi.TpdhgError) as e:
            ffi.spmv_bench(ptr, idx, val, v)
        assert e.value.code == ffi.ERR_NO_DEVICE
        s, keep = _arrays(2, 3, a, at)
        h = C.c_void_p()
        assert L.tpdhg_dev_create(C.byref(s), 64, 64, 1, C.byref(h)) == ffi.ERR_NO_DEVICE


def test_handle_calls_refuse_null_outputs_and_bad_sizes():
    """Entry points that write through caller pointers or take sizes refuse null outputs and non-positive sizes
    with TPDHG_ERR_ARG before touching a handle or a device (no crash on a caller's mistake)."""
    import ctypes as C
    L = ffi.lib()
    nd = C.POINTER(C.c_double)()
    nf = C.POINTER(C.c_float)()
    assert L.tpdhg_dev_power_norm(None, 10, nd) == ffi.ERR_ARG
    assert L.tpdhg_dev_power_norm(None, 0, C.byref(C.c_double())) == ffi.ERR_ARG
    assert L.tpdhg_fast_power_norm(None, 10, nd) == ffi.ERR_ARG
    assert L.tpdhg_dev_eval(None, 0, 1.0, nd) == ffi.ERR_ARG
    assert L.tpdhg_dev_eval2(None, 1.0, nd, nd) == ffi.ERR_ARG
    assert L.tpdhg_fast_eval(None, nd) == ffi.ERR_ARG
    assert L.tpdhg_dev_spmv(None, 0, nd, nd) == ffi.ERR_ARG
    x = (C.c_float * 32)()
    assert L.tpdhg_mlp_forward(None, x, 0, x) == ffi.ERR_ARG
    assert L.tpdhg_mlp_loss_grad(None, x, x, 0, None, nf) == ffi.ERR_ARG
    assert L.tpdhg_mlp_train(None, x, x, 1, 1, 1, C.c_float(0.0), 0, None) == ffi.ERR_ARG
