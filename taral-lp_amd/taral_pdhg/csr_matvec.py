"""The SpMV microbenchmark of gpu/csr_matvec.cu, on this library's SpMV kernels.

Same shapes and output fields as the reference's (gpu/csr_matvec.cu:24-119): random n x n matrices with each entry
kept with probability 0.01 and values U(-1, 1), x ~ U(-1, 1), n = 1024 ... 16384, FP32 then FP64, median of 20
timed launches, the result checked against a CPU loop; here the GPU time is the median HIP-event time of the
library's SpMV kernel (tpdhg_spmv_bench, matrix resident on the device) and the CPU baseline is scipy's CSR
product on one host thread. Added fields: the algorithmic bytes of one product, B = (s + 4) nnz + 4 (n + 1) + s n +
s n (matrix once, x read once, y written once; s = 4 or 8 bytes), and the achieved GB/s, B / median time.
The matrices come from numpy's generator (seed 12345), not the reference's mt19937_64 stream: same distribution,
not the same entries.

    python -m taral_pdhg.csr_matvec [--sizes 1024,2048,...] [--repeats 20]
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np
import scipy.sparse as sp

from . import ffi

SIZES = (1024, 2048, 4096, 8192, 16384)


def matrix(n: int, density: float = 0.01, seed: int = 12345):
    """CSR (ptr, idx, val) of an n x n matrix with entries kept independently with probability `density`, values
    U(-1, 1), and x ~ U(-1, 1)."""
    rng = np.random.default_rng(seed + n)
    keep = rng.random((n, n)) < density
    rows, cols = np.nonzero(keep)
    val = rng.uniform(-1.0, 1.0, rows.size)
    ptr = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(np.bincount(rows, minlength=n), out=ptr[1:])
    x = rng.uniform(-1.0, 1.0, n)
    return ptr, cols.astype(np.int32), val, x


def bench(n: int, bits: int, repeats: int = 20) -> dict:
    ptr, idx, val, x = matrix(n)
    s = 4 if bits == 32 else 8
    dt = np.float32 if bits == 32 else np.float64
    y_gpu, _, med_ms = ffi.spmv_bench(ptr, idx, val, x, precision_bits=bits, reps=repeats)
    a = sp.csr_matrix((val.astype(dt), idx, ptr), shape=(n, n))
    xv = x.astype(dt)
    ts = []
    for _ in range(repeats):
        t = time.perf_counter()
        y_cpu = a @ xv
        ts.append(time.perf_counter() - t)
    cpu = float(np.median(ts))
    gpu = med_ms * 1e-3
    nnz = int(ptr[-1])
    nbytes = (s + 4) * nnz + 4 * (n + 1) + 2 * s * n
    scale = max(float(np.max(np.abs(y_cpu))), 1e-300)
    return {"n": n, "dtype": "float32" if bits == 32 else "float64", "nnz": nnz, "cpu_median_s": cpu,
            "gpu_median_s": gpu, "cpu_threads": 1, "speedup_compute": cpu / gpu,
            "max_rel_error": float(np.max(np.abs(y_gpu - y_cpu.astype(np.float64)))) / scale,
            "bytes": nbytes, "gpu_gbs": nbytes / gpu / 1e9}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--sizes", default=",".join(map(str, SIZES)))
    ap.add_argument("--repeats", type=int, default=20)
    a = ap.parse_args()
    sizes = [int(v) for v in a.sizes.split(",")]
    out = {"tool": "taral_pdhg.csr_matvec", "kernel": "the solver's LDS row-block / column-sorted SpMV (HIP, gfx950)",
           "density": 0.01, "repeats": a.repeats, "results": []}
    for bits in (32, 64):
        for n in sizes:
            out["results"].append(bench(n, bits, a.repeats))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
