"""Shared helpers for the parity tests (the oracle is the checker)."""
from __future__ import annotations

import numpy as np

import oracle as O


def random_csr(rng, m, n, density=None, row_len=None, sorted_rows=True, with_vals=False, dup=False):
    """Random CSR with int64 ptr/int32 col; row lengths drawn from row_len (callable) or density."""
    ptr = [0]
    cols = []
    for _ in range(m):
        k = int(row_len(rng)) if row_len else int(rng.binomial(n, density))
        k = max(0, min(k, n if not dup else 4 * n))
        if n == 0:
            k = 0
        if dup:
            c = rng.integers(0, n, size=k)
        else:
            c = rng.choice(n, size=k, replace=False) if k else np.zeros(0, np.int64)
        if sorted_rows:
            c = np.sort(c)
        cols.append(c.astype(np.int32))
        ptr.append(ptr[-1] + k)
    col = np.concatenate(cols) if cols else np.zeros(0, np.int32)
    val = None
    if with_vals:
        val = rng.integers(0, 2**64, size=len(col), dtype=np.uint64)
    return O.CSR(m, n, np.asarray(ptr, np.int64), col.astype(np.int32), val)


def assert_same_csr(ptr, col, val, ref: O.CSR, col_map=None, what=""):
    """GPU (ptr, col, val) vs oracle CSR (local columns mapped through col_map when given)."""
    assert len(ptr) == ref.nrows + 1, what
    np.testing.assert_array_equal(ptr, ref.ptr, err_msg=f"{what}: row pointers differ")
    rc = ref.col if col_map is None else col_map[ref.col]
    np.testing.assert_array_equal(col, rc, err_msg=f"{what}: columns differ")
    np.testing.assert_array_equal(val, ref.values(), err_msg=f"{what}: values differ")


def type_bins(hg):
    return hg.bins
