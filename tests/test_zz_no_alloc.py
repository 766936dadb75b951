"""The allocation-free SpMV (tpdhg_spmv_noalloc) against the library's SpMV entry point (tpdhg_spmv_bench) on the
same CSR: FP64 and FP32, two calls in a row on one device (the second on a smaller matrix, so the static scratch
array holds the first call's partials when the second begins), rows spanning many work segments, empty rows, and
the size limit refused before any launch. Device arrays come from torch (plumbing only); both entry points are
called through the C ABI.

Tolerance: both results approximate the same exact product with a different summation order, so per row
|y1 - y2| <= 2 (k + 2) u sum_j |a_ij v_j| (u = 2^-53 FP64, 2^-24 FP32, k the row's nonzeros).
"""
import numpy as np
import pytest

from taral_pdhg import ffi

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _csr(lengths, cols, rng):
    lengths = np.asarray(lengths, dtype=np.int64)
    ptr = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int32)
    idx = rng.integers(0, cols, int(ptr[-1])).astype(np.int32)
    val = rng.standard_normal(int(ptr[-1]))
    return ptr, idx, val


def _noalloc(bits, ptr, idx, val, v):
    dt = torch.float64 if bits == 64 else torch.float32
    dev = torch.device("cuda:0")
    dp = torch.from_numpy(ptr).to(dev)
    di = torch.from_numpy(idx).to(dev)
    dval = torch.from_numpy(val).to(dev, dt)
    dv = torch.from_numpy(v).to(dev, dt)
    dy = torch.full((ptr.size - 1,), float("nan"), dtype=dt, device=dev)
    st = torch.cuda.current_stream()
    ffi.spmv_noalloc(bits, ptr.size - 1, int(ptr[-1]), dp.data_ptr(), di.data_ptr(), dval.data_ptr(), dv.data_ptr(),
                     dy.data_ptr(), st.cuda_stream)
    st.synchronize()
    return dy.double().cpu().numpy()


def _check(bits, ptr, idx, val, v):
    y = _noalloc(bits, ptr, idx, val, v)
    ref, _, _ = ffi.spmv_bench(ptr, idx, val, v, precision_bits=bits, reps=1)
    if bits == 32:
        val = val.astype(np.float32).astype(np.float64)
        v = v.astype(np.float32).astype(np.float64)
    mag = np.zeros(ptr.size - 1)
    np.add.at(mag, np.repeat(np.arange(ptr.size - 1), np.diff(ptr)), np.abs(val * v[idx]))
    u = 2.0 ** -53 if bits == 64 else 2.0 ** -24
    tol = 2 * (np.diff(ptr) + 2) * u * mag + 1e-300
    assert np.all(np.isfinite(y))
    assert np.all(np.abs(y - ref) <= tol)
    return y


@pytest.mark.parametrize("bits", [64, 32])
def test_noalloc_matches_spmv_entry_two_calls(bits):
    rng = np.random.default_rng(21)
    cols = 50_000
    lengths = np.concatenate([rng.integers(0, 40, 20000), [9000, 2048, 4096, 70000],
                              np.where(rng.random(5000) < 0.8, 0, rng.integers(1, 300, 5000))])
    ptr, idx, val = _csr(lengths, cols, rng)
    v = rng.standard_normal(cols)
    print(f"noalloc bits={bits} first call rows={ptr.size - 1} nnz={ptr[-1]}", flush=True)
    _check(bits, ptr, idx, val, v)
    # second call, smaller input: the scratch array still holds the first call's partials
    ptr2, idx2, val2 = _csr(np.concatenate([rng.integers(0, 12, 3000), [5000], [0, 0, 3]]), 4000, rng)
    v2 = rng.standard_normal(4000)
    print(f"noalloc bits={bits} second call rows={ptr2.size - 1} nnz={ptr2[-1]}", flush=True)
    _check(bits, ptr2, idx2, val2, v2)


def test_noalloc_empty_and_limits():
    rng = np.random.default_rng(3)
    ptr = np.zeros(17, np.int32)  # 16 empty rows
    y = _noalloc(64, ptr, np.zeros(1, np.int32), np.zeros(1), rng.standard_normal(4))
    assert np.array_equal(y, np.zeros(16))
    limit = ffi.lib().tpdhg_spmv_noalloc_max_nnz()
    assert limit == 2048 * 65536
    with pytest.raises(ffi.TpdhgError, match="scratch"):
        ffi.spmv_noalloc(64, 10, limit + 1, 1, 1, 1, 1, 1)
