"""ctypes front-end of the oracle (CPU restatement of the reference path).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py``, and only as the checker.  The product (libsmcm.so) never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


class OrcCSR(C.Structure):
    _fields_ = [
        ("nrows", C.c_int32),
        ("ncols", C.c_int32),
        ("nnz", C.c_int64),
        ("ptr", C.POINTER(C.c_int64)),
        ("col", C.POINTER(C.c_int32)),
        ("val", C.POINTER(C.c_uint64)),
    ]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
        L = C.CDLL(LIB_PATH)
        PC = C.POINTER(OrcCSR)
        L.orc_csr_free.argtypes = [PC]
        L.orc_csr_free.restype = None
        L.orc_typed_matrix.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                       C.c_int32, C.c_int32, C.c_void_p, C.c_int64, PC]
        L.orc_estimate.argtypes = [C.c_int32, C.POINTER(PC), C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]
        L.orc_dp.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_dp.restype = C.c_double
        L.orc_plan.argtypes = [C.c_int32, C.c_void_p, C.c_void_p]
        L.orc_plan.restype = None
        L.orc_spgemm.argtypes = [PC, PC, PC]
        L.orc_execute_plan.argtypes = [C.c_int32, C.POINTER(PC), C.c_void_p, PC]
        L.orc_row_digest.argtypes = [PC, C.c_void_p]
        L.orc_row_digest.restype = None
        L.orc_plan_digest.argtypes = [C.c_int32, C.POINTER(PC), C.c_void_p, C.c_int32, C.c_void_p,
                                      C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.orc_plan_digest.restype = C.c_int
        L.orc_products.argtypes = [PC, PC]
        L.orc_products.restype = C.c_int64
        L.orc_num_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int32]
        L.orc_set_threads.restype = None
        _lib = L
    return _lib


class CSR:
    """Owned copy of an oracle CSR as numpy arrays."""

    def __init__(self, nrows, ncols, ptr, col, val=None):
        self.nrows, self.ncols = int(nrows), int(ncols)
        self.ptr = np.ascontiguousarray(ptr, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = None if val is None else np.ascontiguousarray(val, dtype=np.uint64)

    @property
    def nnz(self) -> int:
        return int(self.ptr[-1])

    @classmethod
    def _from_struct(cls, s: OrcCSR) -> "CSR":
        nnz = int(s.nnz)
        ptr = np.ctypeslib.as_array(s.ptr, shape=(s.nrows + 1,)).copy()
        col = np.ctypeslib.as_array(s.col, shape=(nnz,)).copy() if nnz else np.zeros(0, np.int32)
        val = None
        if bool(s.val):
            val = np.ctypeslib.as_array(s.val, shape=(nnz,)).copy() if nnz else np.zeros(0, np.uint64)
        out = cls(s.nrows, s.ncols, ptr, col, val)
        lib().orc_csr_free(C.byref(s))
        return out

    def _as_struct(self) -> OrcCSR:
        s = OrcCSR()
        s.nrows, s.ncols, s.nnz = self.nrows, self.ncols, self.nnz
        s.ptr = self.ptr.ctypes.data_as(C.POINTER(C.c_int64))
        s.col = self.col.ctypes.data_as(C.POINTER(C.c_int32))
        s.val = self.val.ctypes.data_as(C.POINTER(C.c_uint64)) if self.val is not None else C.POINTER(C.c_uint64)()
        s._keep = self  # keep arrays alive
        return s

    def values(self) -> np.ndarray:
        return self.val if self.val is not None else np.ones(self.nnz, dtype=np.uint64)

    def to_dense(self) -> np.ndarray:
        d = np.zeros((self.nrows, self.ncols), dtype=np.uint64)
        rows = np.repeat(np.arange(self.nrows), np.diff(self.ptr))
        np.add.at(d, (rows, self.col), self.values())
        return d


def typed_matrix(g, ts: int, e: int, td: int, rows=None) -> CSR:
    s = OrcCSR()
    rp = None
    if rows is not None:
        rp = np.ascontiguousarray(rows, dtype=np.int32)
    rc = lib().orc_typed_matrix(g.n, g.vertexTypes.ctypes.data, g.adj_ptr.ctypes.data, g.adj_nbr.ctypes.data,
                                g.adj_et.ctypes.data, ts, e, td, None if rp is None else rp.ctypes.data,
                                0 if rp is None else len(rp), C.byref(s))
    if rc != 0:
        raise ValueError("bad row selection")
    return CSR._from_struct(s)


def chain_matrices(g, mp, src_rows=None) -> list[CSR]:
    """Matrices of the meta-path; src_rows (local ids of the start type) restricts M_0."""
    mats = []
    for i in range(mp.pathLen):
        rows = src_rows if i == 0 else None
        mats.append(typed_matrix(g, mp.vertexTypes[i], mp.edgeTypes[i], mp.vertexTypes[i + 1], rows))
    return mats


def _struct_array(mats):
    structs = [m._as_struct() for m in mats]
    arr = (C.POINTER(OrcCSR) * len(mats))(*[C.pointer(s) for s in structs])
    return arr, structs


def estimate(mats: list[CSR], coldims) -> tuple[np.ndarray, list[np.ndarray]]:
    L = len(mats)
    arr, keep = _struct_array(mats)
    cd = np.ascontiguousarray(coldims, dtype=np.int32)
    c = np.zeros(L * L, dtype=np.float64)
    R = [np.zeros(m.nrows * (L - i), dtype=np.float64) for i, m in enumerate(mats)]
    Rp = (C.c_void_p * L)(*[r.ctypes.data for r in R])
    lib().orc_estimate(L, arr, cd.ctypes.data, c.ctypes.data, Rp)
    return c.reshape(L, L), [r.reshape(m.nrows, L - i) for i, (r, m) in enumerate(zip(R, mats))]


def dp(c: np.ndarray) -> tuple[float, np.ndarray, np.ndarray, np.ndarray]:
    L = c.shape[0]
    cc = np.ascontiguousarray(c, dtype=np.float64).ravel()
    m = np.zeros(L * L, dtype=np.float64)
    s = np.zeros(L * L, dtype=np.int32)
    cost = lib().orc_dp(L, cc.ctypes.data, m.ctypes.data, s.ctypes.data)
    pairs = np.zeros(2 * max(L - 1, 1), dtype=np.int32)
    lib().orc_plan(L, s.ctypes.data, pairs.ctypes.data)
    return cost, m.reshape(L, L), s.reshape(L, L), pairs[: 2 * (L - 1)].reshape(-1, 2)


def spgemm(A: CSR, B: CSR) -> CSR:
    out = OrcCSR()
    a, b = A._as_struct(), B._as_struct()
    if lib().orc_spgemm(C.byref(a), C.byref(b), C.byref(out)) != 0:
        raise ValueError("shape mismatch")
    return CSR._from_struct(out)


def execute_plan(mats: list[CSR], pairs) -> CSR:
    L = len(mats)
    arr, keep = _struct_array(mats)
    pp = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).ravel())
    out = OrcCSR()
    if lib().orc_execute_plan(L, arr, pp.ctypes.data if pp.size else None, C.byref(out)) != 0:
        raise ValueError("plan execution failed")
    return CSR._from_struct(out)


def plan_digest(mats: list[CSR], pairs, block_rows: int = 1 << 14):
    """Per-row digests of the plan's result (L >= 2) without materialising it, row block by row
    block of M_0; returns (digests, nnz, intermediate products of the plan over all rows)."""
    L = len(mats)
    arr, keep = _struct_array(mats)
    pp = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).ravel())
    dig = np.zeros(mats[0].nrows, dtype=np.uint64)
    nnz, prods = C.c_int64(), C.c_int64()
    if lib().orc_plan_digest(L, arr, pp.ctypes.data, int(block_rows), dig.ctypes.data, C.byref(nnz),
                             C.byref(prods)) != 0:
        raise ValueError("plan digest failed")
    return dig, int(nnz.value), int(prods.value)


def plan_products(mats: list[CSR], pairs) -> int:
    """Intermediate products of a plan (sum over its SpGEMMs of sum_{(u,k) in A} |row_k(B)|),
    from the operands' structure: the product sizes come from the oracle's SpGEMM."""
    stack, total = [], 0
    steps = np.asarray(pairs, dtype=np.int32).reshape(-1, 2)
    for q, (k, l) in enumerate(steps):
        A = mats[k] if k >= 0 else stack.pop(-2 if l < 0 else -1)
        B = mats[l] if l >= 0 else stack.pop()
        total += products(A, B)
        if q < len(steps) - 1:  # the root's own result is not needed
            stack.append(spgemm(A, B))
    return total


def left_to_right(L: int) -> np.ndarray:
    """The plan (0,1), (-1,2), ..., (-1,L-1)."""
    return np.array([[0, 1]] + [[-1, q] for q in range(2, L)], dtype=np.int32)[: max(L - 1, 0)]


def products(A: CSR, B: CSR) -> int:
    a, b = A._as_struct(), B._as_struct()
    return int(lib().orc_products(C.byref(a), C.byref(b)))


def row_digest(Cm: CSR) -> np.ndarray:
    out = np.zeros(Cm.nrows, dtype=np.uint64)
    s = Cm._as_struct()
    lib().orc_row_digest(C.byref(s), out.ctypes.data)
    return out


def coldims(g, mp) -> np.ndarray:
    return np.array([g.dimensions[mp.vertexTypes[i + 1]] for i in range(mp.pathLen)], dtype=np.int32)


def run_chain(g, mp, src_rows=None, plan=None):
    """Full reference pipeline: matrices, estimator, DP, plan, product.  Returns
    (result CSR, c table, R list, pairs)."""
    mats = chain_matrices(g, mp, src_rows)
    L = mp.pathLen
    if L == 1:
        return mats[0], np.zeros((1, 1)), [], np.zeros((0, 2), np.int32)
    c, R = estimate(mats, coldims(g, mp))
    _, _, _, pairs = dp(c)
    if plan is not None:
        pairs = np.asarray(plan, dtype=np.int32).reshape(-1, 2)
    return execute_plan(mats, pairs), c, R, pairs


def set_threads(t: int) -> None:
    lib().orc_set_threads(int(t))


def num_threads() -> int:
    return int(lib().orc_num_threads())


# ---- product-cost estimate (no counterpart in the reference) ----------------------------
# Restates the library's column-density estimate of product counts with numpy (float64 sums in
# numpy's order, logarithms unquantised) so the GPU tables can be checked within a tolerance.

def products_estimate(mats: list[CSR], R: list[np.ndarray], coldims) -> np.ndarray:
    """P[i, k, j] = rows_i * sum_v rho_ik[v] * R[k+1][v, j-k-1] * coldims[j] for i <= k < j."""
    L = len(mats)
    P = np.zeros((L, L, L))
    for i in range(L - 1):
        rows_i = mats[i].nrows
        cnt = np.bincount(mats[i].col, minlength=mats[i].ncols).astype(np.float64)
        rho = np.minimum(1.0, cnt / rows_i) if rows_i > 0 else np.zeros(mats[i].ncols)
        for k in range(i, L - 1):
            if k > i:
                M = mats[k]
                rowid = np.repeat(np.arange(M.nrows), np.diff(M.ptr))
                with np.errstate(divide="ignore"):
                    lg = np.maximum(np.log1p(-rho), -64.0)
                S = np.zeros(M.ncols)
                np.add.at(S, M.col, lg[rowid])
                rho = -np.expm1(S)
            for j in range(k + 1, L):
                w = R[k + 1][:, j - k - 1] * float(coldims[j])
                P[i, k, j] = rows_i * float(np.dot(rho, w))
    return P


def split_bytes(c: np.ndarray, i: int, k: int, j: int, root: bool) -> float:
    """Estimated device bytes of the product X_ik * X_k+1,j: 28 per entry of an intermediate left
    operand (column, count and SpGEMM scratch), 12 per entry of an intermediate right operand and
    of the output below the root; base relations count 0."""
    left = c[i, k] if k > i else 0.0
    right = c[k + 1, j] if j > k + 1 else 0.0
    out = 0.0 if root else c[i, j]
    return 28.0 * left + 12.0 * right + 12.0 * out


def dp_products(c: np.ndarray, P: np.ndarray, budget: float = np.inf):
    """cost(i,j) = min_k cost(i,k) + cost(k+1,j) + P[i,k,j] (+ c[i,j] below the root), lowest k on
    ties; splits whose split_bytes exceed budget are excluded unless every plan needs one, then
    the DP runs without the bound.  Returns (cost table, split table, plan pairs)."""
    L = c.shape[0]
    m = np.zeros((L, L))
    s = np.zeros((L, L), np.int32)

    def run(bound):
        for l in range(1, L):
            for i in range(L - l):
                j = i + l
                best = np.inf
                for k in range(i, j):
                    if split_bytes(c, i, k, j, l == L - 1) > bound:
                        continue
                    v = m[i, k] + m[k + 1, j] + P[i, k, j]
                    if v < best:
                        best, s[i, j] = v, k
                m[i, j] = best + (c[i, j] if l < L - 1 else 0.0)
        return L < 2 or m[0, L - 1] < np.inf

    if not run(budget):
        run(np.inf)
    pairs = []

    def rec(i, j):
        if i == j:
            return i
        k = rec(i, s[i, j])
        q = rec(s[i, j] + 1, j)
        pairs.append((k, q))
        return -1

    if L > 1:
        rec(0, L - 1)
    return m, s, np.array(pairs, np.int32).reshape(-1, 2)
