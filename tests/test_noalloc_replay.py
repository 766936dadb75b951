"""CPU replay of the allocation-free SpMV's indexing (taral-lp_amd/csrc/noalloc.hip), no GPU: for CSR matrices
with rows spanning many 2,048-nonzero segments, runs of empty rows, rows ending exactly on segment edges and an
empty matrix, the replay follows the two kernels' rules (segment b covers nonzeros [2048 b, 2048 (b + 1)); its
first row's partial goes to slot 2b when the row began earlier, its last row's to slot 2b + 1 when the row goes
on; the fix-up adds slot 2 b0 + 1 and slots 2b of the later segments) and checks that every index stays inside
its array, that every slot the fix-up reads was written by the first kernel, that every row is written exactly
once, and that y equals the CSR product."""
import bisect

import numpy as np
import pytest

SEG = 2048
MAX_SEGS = 1 << 16


def replay(ptr, idx, val, v):
    rows, nnz = len(ptr) - 1, int(ptr[-1])
    assert nnz <= SEG * MAX_SEGS
    y = np.full(rows, np.nan)
    writes = np.zeros(rows, int)
    slots = {}
    nseg = (nnz + SEG - 1) // SEG
    for b in range(nseg):
        k0, k1 = b * SEG, min(b * SEG + SEG, nnz)
        rf = bisect.bisect_right(ptr, k0) - 1       # last r with ptr[r] <= k0, as row_of
        rl = bisect.bisect_right(ptr, k1 - 1) - 1
        assert 0 <= rf <= rl < rows and ptr[rf] <= k0 < ptr[rf + 1] and ptr[rl] <= k1 - 1 < ptr[rl + 1]
        assert 0 <= k0 < k1 <= nnz and k1 - k0 <= SEG
        sh = val[k0:k1] * v[idx[k0:k1]]
        for r in range(rf, rl + 1):
            e0, e1 = int(ptr[r]), int(ptr[r + 1])
            if e0 == e1:
                continue
            a, e = max(e0, k0) - k0, min(e1, k1) - k0
            assert 0 <= a <= e <= k1 - k0
            s = float(np.sum(sh[a:e]))
            if r == rf and e0 < k0:
                assert 2 * b not in slots
                slots[2 * b] = s
            elif r == rl and e1 > k1:
                assert 2 * b + 1 not in slots
                slots[2 * b + 1] = s
            else:
                y[r] = s
                writes[r] += 1
    assert all(0 <= k < 2 * MAX_SEGS for k in slots)
    for r in range(rows):
        e0, e1 = int(ptr[r]), int(ptr[r + 1])
        if e0 == e1:
            y[r] = 0.0
            writes[r] += 1
            continue
        b0, b1 = e0 // SEG, (e1 - 1) // SEG
        if b0 == b1:
            continue
        s = slots[2 * b0 + 1]                      # KeyError: a slot read that was never written
        for q in range(b0 + 1, b1 + 1):
            s += slots[2 * q]
        y[r] = s
        writes[r] += 1
    assert np.all(writes == 1)
    return y


def csr(lengths, cols, rng):
    lengths = np.asarray(lengths, dtype=np.int64)
    ptr = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    idx = rng.integers(0, cols, int(ptr[-1]))
    val = rng.standard_normal(int(ptr[-1]))
    return ptr, idx, val


CASES = {
    "short_rows": lambda rng: rng.integers(0, 30, 5000),
    "long_rows_many_segments": lambda rng: np.concatenate([rng.integers(0, 10, 300), [9000], [2048], [4096],
                                                          rng.integers(0, 10, 300), [70000], [1]]),
    "edges_on_row_boundaries": lambda rng: np.full(64, 512),          # segments start exactly at rows
    "runs_of_empty_rows": lambda rng: np.where(rng.random(20000) < 0.9, 0, rng.integers(1, 400, 20000)),
    "one_long_row": lambda rng: np.array([10 * SEG + 7]),
    "leading_and_trailing_empty": lambda rng: np.concatenate([np.zeros(100, int), rng.integers(1, 50, 300),
                                                               np.zeros(100, int)]),
    "all_empty": lambda rng: np.zeros(50, int),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_noalloc_indexing_replayed(case):
    rng = np.random.default_rng(5)
    ptr, idx, val = csr(CASES[case](rng), 3000, rng)
    v = rng.standard_normal(3000)
    y = replay(list(ptr), idx, val, v)
    ref = np.zeros(len(ptr) - 1)
    np.add.at(ref, np.repeat(np.arange(len(ptr) - 1), np.diff(ptr)), val * v[idx])
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)
