"""HIP kernels against the CPU oracle, through the C ABI (libtaral_pdhg.so).

SpMV (both layouts: column-sorted blocks and row blocks, plain and XCD-slab with 8 and 16 slabs, long rows, the
largest blocks and workgroups): FP64 within
summation-order error of the FP64 product; FP32 within (k + 2) 2^-24 sum |a_k x_k| of the product of the
FP32-rounded inputs (k = row length). Backend operations (power iteration, start, one
64-step iteration block, fused current+average evaluation, restart, download) against the oracle's CPU
backend on the same scaled problem: FP64 to 1e-10 relative (summation order is the only difference).
"""
import numpy as np
import pytest

from oracle import oracle
from taral_pdhg import ffi, gen

pytestmark = pytest.mark.gpu


def _csr_from_rows(lengths, cols, rng):
    ptr = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int32)
    idx = np.concatenate([np.sort(rng.choice(cols, size=k, replace=False)) if k else np.zeros(0, np.int64)
                          for k in lengths]).astype(np.int32)
    val = rng.standard_normal(int(ptr[-1]))
    return ptr, idx, val


def _spmv_ref(ptr, idx, val, v):
    rows = ptr.size - 1
    y = np.zeros(rows)
    mag = np.zeros(rows)
    r = np.repeat(np.arange(rows), np.diff(ptr))
    np.add.at(y, r, val * v[idx])
    np.add.at(mag, r, np.abs(val * v[idx]))
    return y, mag


SPMV_CASES = {  # row lengths; gathered vectors of 2^19 entries or more take the XCD column-slab layout
    "poisson_rows": lambda rng: rng.poisson(25, 3000),
    "slab_poisson_rows": lambda rng: rng.poisson(40, 20000),
    "slab_long_rows": lambda rng: np.concatenate([rng.integers(0, 30, 3000), [30000], rng.integers(0, 30, 300)]),
    "empty_rows": lambda rng: np.where(rng.random(5000) < 0.7, 0, rng.integers(1, 9, 5000)),
    "long_rows": lambda rng: np.concatenate([rng.integers(0, 20, 500), [10240], rng.integers(0, 20, 500), [4097],
                                             [2048], [2049]]),
    "many_single_nnz_rows": lambda rng: np.ones(1000, np.int64),
    "one_row": lambda rng: np.array([300]),
    "all_empty": lambda rng: np.zeros(64, np.int64),
}


@pytest.mark.parametrize("layout", ["sorted", "unsorted", "slab"])
@pytest.mark.parametrize("bits", [64, 32])
@pytest.mark.parametrize("case", sorted(SPMV_CASES))
def test_spmv_parity(case, bits, layout, monkeypatch):
    # column-sorted blocks, row blocks, and sorted column slabs forced on every size ("slab")
    monkeypatch.setenv("TPDHG_LAYOUT", layout)
    rng = np.random.default_rng(11)
    lengths = SPMV_CASES[case](rng)
    cols = 600_000 if case.startswith("slab") else 12000
    ptr, idx, val = _csr_from_rows(lengths, cols, rng)
    v = rng.standard_normal(cols)
    print(f"spmv case={case} bits={bits} layout={layout} rows={ptr.size - 1} nnz={ptr[-1]}", flush=True)
    y, _, _ = ffi.spmv_bench(ptr, idx, val, v, precision_bits=bits, reps=2)
    if bits == 64:
        ref, mag = _spmv_ref(ptr, idx, val, v)
        tol = 4 * np.diff(ptr) * 2.0 ** -53 * mag + 1e-300
    else:
        v32, val32 = v.astype(np.float32).astype(np.float64), val.astype(np.float32).astype(np.float64)
        ref, mag = _spmv_ref(ptr, idx, val32, v32)
        tol = (np.diff(ptr) + 2) * 2.0 ** -24 * mag + 1e-30
    assert np.all(np.abs(y - ref) <= tol)


LARGE_CASES = {  # (rows, cols, mean row length): sizes that reach the largest blocks and workgroups
    # 16 column slabs (gathered vector > 2^20 entries), 16,384-nonzero blocks (FP32) on 1,024 threads, 8,192 (FP64)
    "slab16_big_blocks": (900_000, 1_200_000, 20),
    # plain sorted layout (< 2^19 columns) with 8,192-nonzero blocks on 512 threads
    "plain_big_blocks": (100_000, 400_000, 86),
}


@pytest.mark.parametrize("layout", ["sorted", "unsorted"])
@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("case", sorted(LARGE_CASES))
def test_spmv_parity_large(case, bits, layout, monkeypatch):
    monkeypatch.setenv("TPDHG_LAYOUT", layout)
    rows, cols, mean = LARGE_CASES[case]
    rng = np.random.default_rng(12)
    lengths = rng.poisson(mean, rows)
    ptr = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int32)
    idx = rng.integers(0, cols, int(ptr[-1])).astype(np.int32)  # repeated columns in a row are summed
    val = rng.standard_normal(int(ptr[-1]))
    v = rng.standard_normal(cols)
    print(f"spmv large case={case} bits={bits} layout={layout} rows={rows} cols={cols} nnz={ptr[-1]}", flush=True)
    y, _, _ = ffi.spmv_bench(ptr, idx, val, v, precision_bits=bits, reps=2)
    if bits == 64:
        ref, mag = _spmv_ref(ptr, idx, val, v)
        tol = 4 * np.diff(ptr) * 2.0 ** -53 * mag + 1e-300
    else:
        v32, val32 = v.astype(np.float32).astype(np.float64), val.astype(np.float32).astype(np.float64)
        ref, mag = _spmv_ref(ptr, idx, val32, v32)
        tol = (np.diff(ptr) + 2) * 2.0 ** -24 * mag + 1e-30
    assert np.all(np.abs(y - ref) <= tol)


@pytest.mark.parametrize("layout", ["sorted", "unsorted"])
@pytest.mark.parametrize("bits", [32, 64])
def test_spmv_parity_block_structured(bits, layout, monkeypatch):
    """A block-diagonal matrix with more than 2^19 columns plus a few rows spanning all blocks: most (row, slab)
    pairs are empty, so the layout is one piece (row blocks) instead of column slabs; its transpose likewise."""
    monkeypatch.setenv("TPDHG_LAYOUT", layout)
    rng = np.random.default_rng(13)
    blocks, rows_b, cols_b = 64, 1200, 10000
    rows, cols = blocks * rows_b + 8, blocks * cols_b
    lengths = np.concatenate([rng.integers(1, 25, blocks * rows_b), np.full(8, 3000)])
    ptr = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int32)
    blk = np.concatenate([np.repeat(np.arange(blocks * rows_b) // rows_b, lengths[:-8]), np.full(8 * 3000, -1)])
    idx = np.where(blk >= 0, blk * cols_b + rng.integers(0, cols_b, blk.size), rng.integers(0, cols, blk.size))
    idx = idx.astype(np.int32)
    val = rng.standard_normal(int(ptr[-1]))
    from scipy.sparse import csr_matrix
    T = csr_matrix((val, idx, ptr), shape=(rows, cols)).T.tocsr()  # repeated entries kept apart
    T.sort_indices()
    mats = {"A": (ptr, idx, val, cols),
            "AT": (T.indptr.astype(np.int32), T.indices.astype(np.int32), T.data.astype(np.float64), rows)}
    for name, (p_, i_, v_, n_) in mats.items():
        x = rng.standard_normal(n_)
        print(f"spmv block-structured {name} bits={bits} layout={layout} rows={p_.size - 1} cols={n_} nnz={p_[-1]}",
              flush=True)
        y, _, _ = ffi.spmv_bench(p_, i_, v_, x, precision_bits=bits, reps=2)
        if bits == 64:
            ref, mag = _spmv_ref(p_, i_, v_, x)
            tol = 4 * np.diff(p_) * 2.0 ** -53 * mag + 1e-300
        else:
            x32, v32 = x.astype(np.float32).astype(np.float64), v_.astype(np.float32).astype(np.float64)
            ref, mag = _spmv_ref(p_, i_, v32, x32)
            tol = (np.diff(p_) + 2) * 2.0 ** -24 * mag + 1e-30
        assert np.all(np.abs(y - ref) <= tol)


@pytest.fixture(scope="module")
def packing_problem():
    md, _ = gen.packing_mps(10000)
    model = ffi.Model.from_lp(md)
    return md, ffi.Scaled(model), oracle.OracleLP(md)


@pytest.fixture(scope="module")
def cfg0_problem(cfg0):
    return cfg0, ffi.Scaled(ffi.Model.from_lp(cfg0)), oracle.OracleLP(cfg0)


def _close(a, b, rel=1e-10):
    scale = max(1.0, float(np.max(np.abs(b))) if np.size(b) else 1.0)
    assert np.max(np.abs(np.asarray(a) - np.asarray(b))) <= rel * scale


@pytest.mark.parametrize("layout", ["auto", "sorted", "slab"])
@pytest.mark.parametrize("which", ["packing", "cfg0"])
def test_backend_block_fp64(which, layout, packing_problem, cfg0_problem, monkeypatch):
    """One 64-step block of the reference scheme on each layout: row blocks (the default at these sizes), column-
    sorted blocks (k_sorted with EpiRow / EpiCol) and sorted slabs (k_sorted_slab + k_combine_epi with EpiRow /
    EpiCol), each against the oracle's block."""
    if layout == "auto":
        monkeypatch.delenv("TPDHG_LAYOUT", raising=False)
    else:
        monkeypatch.setenv("TPDHG_LAYOUT", layout)
    md, S, O = packing_problem if which == "packing" else cfg0_problem
    print(f"backend block fp64 {which} layout={layout}", flush=True)
    dev = ffi.Device(S, precision_bits=64, eval_freq=64, use_graph=True)
    norm = dev.power_norm(200)
    assert norm == pytest.approx(O.power_norm(fp32=False), rel=1e-10)
    eta = 0.95 / norm
    tau, sigma = eta / 0.8, eta * 0.8
    dev.set_steps(tau, sigma)
    dev.start()
    want0 = O.block(tau, sigma, 0)
    cur, _ = dev.eval2(1.0)
    k_ref = O.eval(want0["Ax"], want0["z"], want0["x"], want0["ATz"], want0["x"], want0["z"])
    _close(cur, k_ref)
    dev.iterate(64)  # graph-captured block
    want = O.block(tau, sigma, 64)
    x, z = dev.download(False)
    xa, za = dev.download(True, 1 / 64)
    _close(x, want["x"]), _close(z, want["z"])
    _close(xa, want["xs"] / 64), _close(za, want["zs"] / 64)
    cur, avg = dev.eval2(1 / 64)
    _close(cur, O.eval(want["Ax"], want["z"], want["x"], want["ATz"], want0["x"], want0["z"]))
    _close(avg, O.eval(want["Axs"], want["zs"], want["xs"], want["ATzs"], want0["x"], want0["z"], f=1 / 64))
    # single evaluations agree with the fused one
    _close(dev.eval(False, 1.0), cur, rel=0), _close(dev.eval(True, 1 / 64), avg, rel=0)
    # a second block without the graph continues identically to the oracle's 128 steps
    dev.restart(True, 1 / 64)
    xr, zr = dev.download(False)
    _close(xr, want["xs"] / 64), _close(zr, want["zs"] / 64)
    _close(dev.spmv(xr), O.model.matvec(xr * O.vec("Dc")) * O.vec("Dr"), rel=1e-11)


def test_backend_warm_start_and_fp32(cfg0_problem):
    md, S, O = cfg0_problem
    rng = np.random.default_rng(4)
    x0, z0 = rng.random(md.n), rng.standard_normal(md.m)
    for bits in (64, 32):
        print(f"backend warm start bits={bits}", flush=True)
        dev = ffi.Device(S, precision_bits=bits, eval_freq=16, use_graph=False)
        dev.set_steps(0.3, 0.5)
        dev.start(x0, z0)
        dev.iterate(16)
        want = O.block(0.3, 0.5, 16, x0=x0, z0=z0, fp32=bits == 32)
        x, z = dev.download(False)
        _close(x, want["x"], rel=1e-10 if bits == 64 else 2e-4)
        _close(z, want["z"], rel=1e-10 if bits == 64 else 2e-4)


def test_graph_and_plain_launches_agree(packing_problem):
    md, S, O = packing_problem
    outs = []
    for graph in (True, False):
        dev = ffi.Device(S, precision_bits=64, eval_freq=64, use_graph=graph)
        dev.set_steps(0.9, 0.9)
        dev.start()
        for _ in range(3):
            dev.iterate(64)
            dev.eval2(1 / 64)
        outs.append(dev.download(True, 1 / 192))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


def test_kernel_event_profile(packing_problem):
    md, S, O = packing_problem
    dev = ffi.Device(S, precision_bits=32, eval_freq=64)
    dev.set_steps(0.9, 0.9)
    dev.start()
    dev.set_profile(True)
    dev.iterate(64)
    dev.eval2(1 / 64)
    t = dev.kernel_times()
    assert t["row_launches"] == 64 and t["col_launches"] == 64
    assert t["row_ms"] > 0 and t["col_ms"] > 0


def test_csr_matvec_counterpart():
    """python -m taral_pdhg.csr_matvec (gpu/csr_matvec.cu's shapes and fields) on its smallest size: the GPU product
    matches the CPU loop and the GB/s field follows from the byte count and the median time."""
    from taral_pdhg import csr_matvec
    for bits in (32, 64):
        r = csr_matvec.bench(1024, bits, repeats=3)
        assert r["max_rel_error"] <= (1e-5 if bits == 32 else 1e-13)
        assert r["gpu_gbs"] == pytest.approx(r["bytes"] / r["gpu_median_s"] / 1e9)
