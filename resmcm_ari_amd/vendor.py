"""rocSPARSE generic SpGEMM (binary64 values) on host CSR operands: a timing baseline for the
library's SpGEMM.  Not part of the product path."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _build

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = os.path.join(_build.LIBDIR, "libsmcm_vendor.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build()")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.smcm_vendor_spgemm.argtypes = [C.c_int32, C.c_int32, C.c_int32, vp, vp, vp, vp, vp, vp, C.c_int32, vp, vp,
                                         vp, vp, vp, C.c_char_p, C.c_int32]
        L.smcm_vendor_spgemm.restype = C.c_int
        _lib = L
    return _lib


def spgemm_f64(m, k, n, a_ptr, a_col, a_val, b_ptr, b_col, b_val, reps: int = 3, want_result: bool = True):
    """C = A * B with rocSPARSE; values as binary64 (None = all ones).  Returns (ms per product
    (nnz and compute stages), nnz, (ptr, col, val) or None)."""
    ap = np.ascontiguousarray(a_ptr, np.int64)
    ac = np.ascontiguousarray(a_col, np.int32)
    av = np.ones(len(ac)) if a_val is None else np.ascontiguousarray(a_val, np.float64)
    bp = np.ascontiguousarray(b_ptr, np.int64)
    bc = np.ascontiguousarray(b_col, np.int32)
    bv = np.ones(len(bc)) if b_val is None else np.ascontiguousarray(b_val, np.float64)
    ms = C.c_double()
    nnz = C.c_int64()
    err = C.create_string_buffer(256)
    cp = np.zeros(m + 1, np.int64) if want_result else None
    # two calls: the first sizes the result, the second returns it
    rc = lib().smcm_vendor_spgemm(m, k, n, ap.ctypes.data, ac.ctypes.data, av.ctypes.data, bp.ctypes.data,
                                  bc.ctypes.data, bv.ctypes.data, reps, C.byref(ms), C.byref(nnz),
                                  None, None, None, err, 256)
    if rc != 0:
        # -22: an invalid operand, rejected on the host before any device work
        raise (ValueError if rc == -22 else RuntimeError)("rocSPARSE SpGEMM failed: " + err.value.decode())
    if not want_result:
        return ms.value, nnz.value, None
    cc = np.zeros(max(nnz.value, 1), np.int32)
    cv = np.zeros(max(nnz.value, 1), np.float64)
    rc = lib().smcm_vendor_spgemm(m, k, n, ap.ctypes.data, ac.ctypes.data, av.ctypes.data, bp.ctypes.data,
                                  bc.ctypes.data, bv.ctypes.data, 1, C.byref(C.c_double()), C.byref(nnz),
                                  cp.ctypes.data, cc.ctypes.data, cv.ctypes.data, err, 256)
    if rc != 0:
        raise RuntimeError("rocSPARSE SpGEMM failed: " + err.value.decode())
    return ms.value, nnz.value, (cp, cc[: nnz.value], cv[: nnz.value])
