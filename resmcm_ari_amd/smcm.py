"""ctypes front-end of ``lib/libsmcm.so`` (the C ABI in include/smcm.h).

``build_pair`` keeps the reference entry's meaning (efficient/resmcm.cpp:774-779): a meta-path
over a resident graph in, the result matrix and the times ``es_time / dp_time / spgemm_time /
smcm_time`` out.  There is no CPU fallback: loading fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass

import numpy as np

from . import _build
from .hin import HeteroGraph, MetaPath

SMCM_OK = 0
SMCM_EINVAL = -22
SMCM_ENOMEM = -12
SMCM_EBUDGET = -28
SMCM_EDEVICE = -5
SMCM_EAGAIN = -11
SMCM_ECANCELED = -125

LIB_PATH = os.path.join(_build.LIBDIR, "libsmcm.so")

# every entry point declared in include/smcm.h
EXPORTS = [
    "smcm_ctx_create", "smcm_ctx_destroy", "smcm_set_option", "smcm_get_option", "smcm_last_error",
    "smcm_last_stats", "smcm_ctx_sync", "smcm_graph_upload", "smcm_graph_free", "smcm_graph_dims",
    "smcm_chain_run", "smcm_chain_run_plan", "smcm_chain_plan", "smcm_estimate", "smcm_estimate_noalloc",
    "smcm_estimator_rows",
    "smcm_result_nrows", "smcm_result_nnz", "smcm_result_copy", "smcm_result_copy_rows", "smcm_result_row_ids",
    "smcm_result_row_digest", "smcm_result_free", "smcm_spgemm_csr", "smcm_chain_run_blocked",
    "smcm_kernel_stats", "smcm_kernel_stats_reset", "smcm_copy_bandwidth", "smcm_chain_products_estimate",
    "smcm_batch_run", "smcm_server_create", "smcm_server_destroy", "smcm_server_submit", "smcm_server_poll",
    "smcm_server_stats", "smcm_server_replay", "smcm_batch_autotune", "smcm_cost_estimate", "smcm_server_cancel",
]

# DP cost providers (option plan_cost and Graph.cost_estimate)
PROVIDERS = {"metaac": 2, "mnc": 3, "dmap": 4, "lgraph": 5}


class Times(C.Structure):
    _fields_ = [("total", C.c_double), ("spgemm", C.c_double), ("dp", C.c_double), ("est", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("products", C.c_int64), ("nnz", C.c_int64), ("rows_tier", C.c_int64 * 4),
                ("n_spgemm", C.c_int32), ("pairs", C.c_int32 * 30), ("c_root", C.c_double),
                ("ref_nnz", C.c_int64)]


class BatchStats(C.Structure):
    _fields_ = [("chains", C.c_int32), ("jobs", C.c_int32), ("small_jobs", C.c_int32), ("large_jobs", C.c_int32),
                ("rerun_jobs", C.c_int32), ("groups", C.c_int32), ("nodes", C.c_int64),
                ("tiles", C.c_int64), ("products_est", C.c_int64), ("plan_s", C.c_double), ("exec_s", C.c_double),
                ("large_s", C.c_double), ("total_s", C.c_double), ("alone_jobs", C.c_int32), ("oom_retries", C.c_int32)]


class BatchTuning(C.Structure):
    _fields_ = [("baseline_s", C.c_double), ("best_s", C.c_double), ("small_products", C.c_int64),
                ("threads", C.c_int32), ("sort_cap", C.c_int32), ("tile_rows", C.c_int32), ("trials", C.c_int32),
                ("streams", C.c_int32), ("pad", C.c_int32)]


class ServerOpts(C.Structure):
    _fields_ = [("slots", C.c_int32), ("plan_cost", C.c_int32), ("budget_bytes", C.c_uint64),
                ("cache_bytes", C.c_uint64), ("cache_entry_max", C.c_uint64), ("block_bytes", C.c_uint64),
                ("large_slots", C.c_int32), ("large_cu_percent", C.c_int32), ("large_bytes", C.c_uint64)]


class QueryOut(C.Structure):
    _fields_ = [("nnz", C.c_int64), ("t_submit", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double),
                ("predicted_bytes", C.c_uint64), ("status", C.c_int32), ("cache_k", C.c_int32),
                ("blocks", C.c_int32), ("reruns", C.c_int32), ("error", C.c_char * 160)]


class SmcmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def load() -> C.CDLL:
    """Load libsmcm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, u64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_double
        sig = {
            "smcm_ctx_create": [C.c_int, u64, vp],
            "smcm_ctx_destroy": [vp],
            "smcm_set_option": [vp, C.c_char_p, i64],
            "smcm_get_option": [vp, C.c_char_p, vp],
            "smcm_last_stats": [vp, vp],
            "smcm_ctx_sync": [vp],
            "smcm_graph_upload": [vp, i32, vp, vp, vp, vp, vp],
            "smcm_graph_free": [vp],
            "smcm_graph_dims": [vp, vp, i32, vp],
            "smcm_chain_run": [vp, vp, vp, vp, i32, vp, i64, vp, vp],
            "smcm_chain_run_plan": [vp, vp, vp, vp, i32, vp, i64, vp, vp, vp],
            "smcm_chain_plan": [vp, vp, vp, vp, i32, vp, i64, vp, vp, vp, vp],
            "smcm_estimate": [vp, vp, vp, vp, i32, vp, i64, vp],
            "smcm_estimate_noalloc": [vp, vp, vp, vp, i32, vp],
            "smcm_cost_estimate": [vp, vp, vp, vp, i32, i32, i32, vp, vp, vp, vp],
            "smcm_chain_products_estimate": [vp, vp, vp, vp, i32, vp, i64, vp],
            "smcm_batch_run": [vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp],
            "smcm_batch_autotune": [vp, vp, i32, vp, vp, vp, vp, vp, i32, vp],
            "smcm_server_create": [vp, vp, i32, vp, vp],
            "smcm_server_destroy": [vp],
            "smcm_server_submit": [vp, i32, vp, vp, i32, vp, i64, vp, vp],
            "smcm_server_poll": [vp, i64, i32, vp],
            "smcm_server_cancel": [vp, i64],
            "smcm_server_stats": [vp, vp, vp, vp],
            "smcm_server_replay": [vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp],
            "smcm_estimator_rows": [vp, vp, vp, vp, i32, vp, i64, i32, vp, i64],
            "smcm_result_nrows": [vp],
            "smcm_result_nnz": [vp],
            "smcm_result_copy": [vp, vp, vp, vp],
            "smcm_result_copy_rows": [vp, i32, i32, vp, vp, vp],
            "smcm_result_row_ids": [vp, vp],
            "smcm_result_row_digest": [vp, vp],
            "smcm_result_free": [vp],
            "smcm_spgemm_csr": [vp, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp],
            "smcm_chain_run_blocked": [vp, vp, vp, vp, i32, vp, i64, u64, vp, vp, vp],
            "smcm_kernel_stats": [vp, i32, C.c_char_p, i32, vp, vp, vp],
            "smcm_kernel_stats_reset": [vp],
            "smcm_copy_bandwidth": [vp, u64, i32, vp],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.smcm_result_nrows.restype = C.c_int32
        L.smcm_result_nnz.restype = C.c_int64
        L.smcm_last_error.restype = C.c_char_p
        L.smcm_last_error.argtypes = []
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != SMCM_OK:
        raise SmcmError(rc, load().smcm_last_error().decode())


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


class Context:
    """One device context (own HIP stream, stream-ordered memory pool)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self._h = C.c_void_p()
        self._children = weakref.WeakSet()  # graphs and results; closed before the context
        _check(load().smcm_ctx_create(device, hbm_budget, C.byref(self._h)))

    def set_option(self, key: str, value: int) -> None:
        _check(load().smcm_set_option(self._h, key.encode(), int(value)))

    def get_option(self, key: str) -> int:
        v = C.c_int64()
        _check(load().smcm_get_option(self._h, key.encode(), C.byref(v)))
        return int(v.value)

    def last_stats(self) -> dict:
        s = Stats()
        _check(load().smcm_last_stats(self._h, C.byref(s)))
        n = max(s.n_spgemm, 0)
        return {"products": int(s.products), "nnz": int(s.nnz), "rows_tier": list(s.rows_tier),
                "n_spgemm": int(s.n_spgemm), "pairs": [int(x) for x in s.pairs[: 2 * n]], "c_root": float(s.c_root),
                "ref_nnz": int(s.ref_nnz)}

    def sync(self) -> None:
        _check(load().smcm_ctx_sync(self._h))

    def kernel_stats(self) -> dict:
        """{kernel name: (total ms, launches, algorithmic bytes)} since the last reset."""
        out = {}
        i = 0
        name = C.create_string_buffer(128)
        ms, cnt, by = C.c_double(), C.c_int64(), C.c_double()
        while load().smcm_kernel_stats(self._h, i, name, 128, C.byref(ms), C.byref(cnt), C.byref(by)) == SMCM_OK:
            out[name.value.decode()] = (ms.value, int(cnt.value), by.value)
            i += 1
        return out

    def copy_bandwidth(self, nbytes: int = 1 << 32, iters: int = 10) -> float:
        """Measured HBM copy bandwidth (GB/s, read + write) of a plain copy kernel."""
        g = C.c_double()
        _check(load().smcm_copy_bandwidth(self._h, int(nbytes), int(iters), C.byref(g)))
        return float(g.value)

    def kernel_stats_reset(self) -> None:
        _check(load().smcm_kernel_stats_reset(self._h))

    def close(self) -> None:
        if self._h:
            kids = list(self._children)
            # servers first (they use the graphs), then results and graphs
            for ch in sorted(kids, key=lambda c: 0 if isinstance(c, Server) else 1):
                ch.free() if isinstance(ch, Result) else ch.close()
            _check(load().smcm_ctx_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def spgemm_csr(self, m, k, n, a_ptr, a_col, a_val, b_ptr, b_col, b_val) -> "Result":
        ap = np.ascontiguousarray(a_ptr, np.int64)
        ac = np.ascontiguousarray(a_col, np.int32)
        av = None if a_val is None else np.ascontiguousarray(a_val, np.uint64)
        bp = np.ascontiguousarray(b_ptr, np.int64)
        bc = np.ascontiguousarray(b_col, np.int32)
        bv = None if b_val is None else np.ascontiguousarray(b_val, np.uint64)
        h = C.c_void_p()
        _check(load().smcm_spgemm_csr(self._h, m, k, n, _ptr(ap), _ptr(ac), _ptr(av), _ptr(bp), _ptr(bc), _ptr(bv),
                                      C.byref(h)))
        return Result(h, self)


class Graph:
    """A HeteroGraph resident in HBM (typed relation matrices are built on the device and
    cached per meta-path edge)."""

    def __init__(self, ctx: Context, hg: HeteroGraph):
        self.ctx = ctx
        self.hg = hg
        self._h = C.c_void_p()
        ctx._children.add(self)
        _check(load().smcm_graph_upload(ctx._h, hg.n, _ptr(hg.vertexTypes), _ptr(hg.adj_ptr), _ptr(hg.adj_nbr),
                                        _ptr(hg.adj_et), C.byref(self._h)))

    def close(self) -> None:
        if self._h and self.ctx._h:
            load().smcm_graph_free(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _path(mp: MetaPath):
        vt = np.ascontiguousarray(mp.vertexTypes, np.int32)
        et = np.ascontiguousarray(mp.edgeTypes if mp.pathLen else [0], np.int32)
        return vt, et

    @staticmethod
    def _src(src_rows):
        if src_rows is None:
            return None, 0
        s = np.ascontiguousarray(src_rows, np.int32)
        return s, len(s)

    def run(self, mp: MetaPath, src_rows=None, plan=None) -> tuple["Result", dict]:
        vt, et = self._path(mp)
        s, ns = self._src(src_rows)
        h = C.c_void_p()
        t = Times()
        if plan is None:
            _check(load().smcm_chain_run(self.ctx._h, self._h, _ptr(vt), _ptr(et), mp.pathLen, _ptr(s), ns,
                                         C.byref(h), C.byref(t)))
        else:
            pp = np.ascontiguousarray(np.asarray(plan, np.int32).ravel())
            _check(load().smcm_chain_run_plan(self.ctx._h, self._h, _ptr(vt), _ptr(et), mp.pathLen, _ptr(s), ns,
                                              _ptr(pp), C.byref(h), C.byref(t)))
        times = {"total": t.total, "spgemm": t.spgemm, "dp": t.dp, "est": t.est}
        return Result(h, self.ctx), times

    def run_blocked(self, mp: MetaPath, block_bytes: int, src_rows=None, digest: bool = False):
        """Row-blocked run: returns (nnz, per-row digests or None, times)."""
        vt, et = self._path(mp)
        s, ns = self._src(src_rows)
        nrows = ns if s is not None else int(self.hg.dimensions[mp.vertexTypes[0]])
        dig = np.zeros(max(nrows, 1), np.uint64) if digest else None
        nnz = C.c_int64()
        t = Times()
        _check(load().smcm_chain_run_blocked(self.ctx._h, self._h, _ptr(vt), _ptr(et), mp.pathLen, _ptr(s), ns,
                                             int(block_bytes), C.byref(nnz), _ptr(dig), C.byref(t)))
        times = {"total": t.total, "spgemm": t.spgemm, "dp": t.dp, "est": t.est}
        return int(nnz.value), (dig[:nrows] if digest else None), times

    def plan(self, mp: MetaPath, src_rows=None):
        L = mp.pathLen
        vt, et = self._path(mp)
        s, ns = self._src(src_rows)
        c = np.zeros(L * L, np.float64)
        m = np.zeros(L * L, np.float64)
        st = np.zeros(L * L, np.int32)
        pairs = np.zeros(max(2 * (L - 1), 1), np.int32)
        _check(load().smcm_chain_plan(self.ctx._h, self._h, _ptr(vt), _ptr(et), L, _ptr(s), ns, _ptr(c), _ptr(m),
                                      _ptr(st), _ptr(pairs)))
        return c.reshape(L, L), m.reshape(L, L), st.reshape(L, L), pairs[: 2 * (L - 1)].reshape(-1, 2)

    def estimate(self, mp: MetaPath, src_rows=None) -> float:
        vt, et = self._path(mp)
        s, ns = self._src(src_rows)
        x = C.c_double()
        _check(load().smcm_estimate(self.ctx._h, self._h, _ptr(vt), _ptr(et), mp.pathLen, _ptr(s), ns, C.byref(x)))
        return float(x.value)

    def estimate_noalloc(self, mp: MetaPath, d_out: int) -> None:
        """Allocation-free estimate: c_[0][L-1] written to the device double at address d_out
        (queued on the context's stream; sync the context before reading it)."""
        vt, et = self._path(mp)
        _check(load().smcm_estimate_noalloc(self.ctx._h, self._h, _ptr(vt), _ptr(et), mp.pathLen,
                                            C.c_void_p(int(d_out))))

    def cost_estimate(self, mp: MetaPath, provider, accuracy: bool = False):
        """One of the reference's other DP cost providers ("metaac", "mnc", "dmap", "lgraph" or
        2-5) over the whole chain: (estimate, c, m, s) - the number the reference's accuracy
        program prints (accuracy=True) or the planner's tables (accuracy=False)."""
        L = mp.pathLen
        vt, et = self._path(mp)
        pid = PROVIDERS[provider] if isinstance(provider, str) else int(provider)
        est = C.c_double()
        c = np.zeros(L * L, np.float64)
        m = np.zeros(L * L, np.float64)
        st = np.zeros(L * L, np.int32)
        _check(load().smcm_cost_estimate(self.ctx._h, self._h, _ptr(vt), _ptr(et), L, pid, 1 if accuracy else 0,
                                         C.byref(est), _ptr(c), _ptr(m), _ptr(st)))
        return float(est.value), c.reshape(L, L), m.reshape(L, L), st.reshape(L, L)

    def products_estimate(self, mp: MetaPath, src_rows=None) -> np.ndarray:
        """Estimated product count of every split: P[i, k, j] for X_ik * X_{k+1,j}."""
        L = mp.pathLen
        vt, et = self._path(mp)
        s, ns = self._src(src_rows)
        out = np.zeros(max(L, 1) ** 3, np.float64)
        _check(load().smcm_chain_products_estimate(self.ctx._h, self._h, _ptr(vt), _ptr(et), L, _ptr(s), ns,
                                                   _ptr(out)))
        return out.reshape(L, L, L)

    def run_batch(self, paths, src_rows=None, digests: bool = True, results: bool = False):
        """Batched form: paths[i] (MetaPath) from src_rows[i] (global ids; None = all rows of the
        start type for every chain).  Returns (nnz per chain, per-chain digest arrays or None,
        result handles or None, stats dict)."""
        n = len(paths)
        plen = np.array([mp.pathLen for mp in paths], np.int32)
        vts = np.concatenate([np.asarray(mp.vertexTypes, np.int32) for mp in paths]) if n else np.zeros(1, np.int32)
        ets = np.concatenate([np.asarray(mp.edgeTypes, np.int32) for mp in paths]) if n else np.zeros(1, np.int32)
        if src_rows is None:
            soff, srows = None, None
            nrows = [int(self.hg.dimensions[mp.vertexTypes[0]]) for mp in paths]
        else:
            parts = [np.ascontiguousarray(r, np.int32) for r in src_rows]
            nrows = [len(r) for r in parts]
            soff = np.zeros(n + 1, np.int64)
            soff[1:] = np.cumsum(nrows)
            srows = np.concatenate(parts) if parts and soff[-1] > 0 else np.zeros(1, np.int32)
        nnz = np.zeros(max(n, 1), np.int64)
        dig = np.zeros(max(int(sum(nrows)), 1), np.uint64) if digests else None
        hs = (C.c_void_p * max(n, 1))() if results else None
        st = BatchStats()
        _check(load().smcm_batch_run(self.ctx._h, self._h, n, _ptr(plen), _ptr(vts), _ptr(ets),
                                     None if soff is None else _ptr(soff), None if srows is None else _ptr(srows),
                                     _ptr(nnz), None if dig is None else _ptr(dig), hs, C.byref(st)))
        out_dig = None
        if digests:
            out_dig, o = [], 0
            for r in nrows:
                out_dig.append(dig[o:o + r])
                o += r
        res = [Result(C.c_void_p(hs[i]), self.ctx) for i in range(n)] if results else None
        stats = {k: getattr(st, k) for k, _ in BatchStats._fields_}
        return nnz[:n], out_dig, res, stats

    def autotune_batch(self, paths, src_rows=None, reps: int = 1) -> dict:
        """Autotuner of the batched form on a sample of chains; leaves the best settings on the
        context and returns them."""
        n = len(paths)
        plen = np.array([mp.pathLen for mp in paths], np.int32)
        vts = np.concatenate([np.asarray(mp.vertexTypes, np.int32) for mp in paths])
        ets = np.concatenate([np.asarray(mp.edgeTypes, np.int32) for mp in paths])
        soff = srows = None
        if src_rows is not None:
            parts = [np.ascontiguousarray(r, np.int32) for r in src_rows]
            soff = np.zeros(n + 1, np.int64)
            soff[1:] = np.cumsum([len(p) for p in parts])
            srows = np.concatenate(parts) if soff[-1] > 0 else np.zeros(1, np.int32)
        t = BatchTuning()
        _check(load().smcm_batch_autotune(self.ctx._h, self._h, n, _ptr(plen), _ptr(vts), _ptr(ets), _ptr(soff),
                                          _ptr(srows), reps, C.byref(t)))
        return {k: getattr(t, k) for k, _ in BatchTuning._fields_}

    def estimator_rows(self, mp: MetaPath, i: int, src_rows=None) -> np.ndarray:
        L = mp.pathLen
        vt, et = self._path(mp)
        s, ns = self._src(src_rows)
        nrows = ns if (i == 0 and s is not None) else int(self.hg.dimensions[mp.vertexTypes[i]])
        out = np.zeros(max(nrows * (L - i), 1), np.float64)
        _check(load().smcm_estimator_rows(self.ctx._h, self._h, _ptr(vt), _ptr(et), L, _ptr(s), ns, i, _ptr(out),
                                          out.size))
        return out[: nrows * (L - i)].reshape(nrows, L - i)


class Result:
    def __init__(self, h, ctx: Context):
        self._h = h
        self.ctx = ctx
        ctx._children.add(self)

    @property
    def nrows(self) -> int:
        return int(load().smcm_result_nrows(self._h))

    @property
    def nnz(self) -> int:
        return int(load().smcm_result_nnz(self._h))

    def to_host(self):
        n, nnz = self.nrows, self.nnz
        ptr = np.zeros(n + 1, np.int64)
        col = np.zeros(max(nnz, 1), np.int32)
        val = np.zeros(max(nnz, 1), np.uint64)
        _check(load().smcm_result_copy(self._h, _ptr(ptr), _ptr(col), _ptr(val)))
        return ptr, col[:nnz], val[:nnz]

    def rows(self, b: int, e: int):
        ptr = np.zeros(e - b + 1, np.int64)
        _check(load().smcm_result_copy_rows(self._h, b, e, _ptr(ptr), None, None))
        cnt = int(ptr[-1])
        col = np.zeros(max(cnt, 1), np.int32)
        val = np.zeros(max(cnt, 1), np.uint64)
        _check(load().smcm_result_copy_rows(self._h, b, e, _ptr(ptr), _ptr(col), _ptr(val)))
        return ptr, col[:cnt], val[:cnt]

    def row_ids(self) -> np.ndarray:
        out = np.zeros(max(self.nrows, 1), np.int32)
        _check(load().smcm_result_row_ids(self._h, _ptr(out)))
        return out[: self.nrows]

    def row_digest(self) -> np.ndarray:
        out = np.zeros(max(self.nrows, 1), np.uint64)
        _check(load().smcm_result_row_digest(self._h, _ptr(out)))
        return out[: self.nrows]

    def free(self) -> None:
        if self._h and self.ctx._h:
            load().smcm_result_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class PairTimes:
    es_time: float
    dp_time: float
    spgemm_time: float
    smcm_time: float


def build_pair(graph: Graph, path: MetaPath, src_rows=None) -> tuple[Result, PairTimes]:
    """The reference's build_pair/SSP_dynamic (efficient/resmcm.cpp:653-779): returns the chain
    product and the four timings."""
    r, t = graph.run(path, src_rows)
    return r, PairTimes(es_time=t["est"], dp_time=t["dp"], spgemm_time=t["spgemm"], smcm_time=t["total"])


class Server:
    """Serving form: worker slots over graphs resident in HBM (see smcm_server_create)."""

    def __init__(self, ctx: Context, graphs: list, slots: int = 6, plan_cost: int = 1, budget: int = 200 << 30,
                 cache: int = 16 << 30, cache_entry_max: int = 2 << 30, block: int = 32 << 30,
                 large_slots: int | None = None, large_bytes: int = 4 << 30, large_cu_percent: int = 50):
        self.ctx = ctx
        self.graphs = list(graphs)
        arr = (C.c_void_p * len(graphs))(*[G._h for G in graphs])
        if large_slots is None:
            large_slots = min(2, slots // 2) if slots > 1 else 0
        o = ServerOpts(slots, plan_cost, budget, cache, cache_entry_max, block, large_slots, large_cu_percent,
                       large_bytes)
        h = C.c_void_p()
        _check(load().smcm_server_create(ctx._h, arr, len(graphs), C.byref(o), C.byref(h)))
        self._h = h
        self._pending = {}
        ctx._children.add(self)

    def close(self) -> None:
        if getattr(self, "_h", None):
            _check(load().smcm_server_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, graph: int, path, src_rows=None):
        """Queues one query; returns (ticket, per-row digest buffer filled when it finishes)."""
        vts = np.ascontiguousarray(path.vertexTypes, np.int32)
        ets = np.ascontiguousarray(path.edgeTypes, np.int32)
        if src_rows is None:
            src, n = None, int(self.graphs[graph].hg.dimensions[path.vertexTypes[0]])
        else:
            src = np.ascontiguousarray(src_rows, np.int32)
            n = len(src)
        dig = np.zeros(max(n, 1), np.uint64)
        t = C.c_int64()
        _check(load().smcm_server_submit(self._h, graph, _ptr(vts), _ptr(ets), path.pathLen, _ptr(src),
                                         0 if src is None else len(src), _ptr(dig), C.byref(t)))
        self._pending[t.value] = (dig, src)  # both stay alive until the query is polled
        return t.value, dig[:n]

    def poll(self, ticket: int, wait: bool = True):
        """The query's result dict, or None while it runs (wait=False)."""
        q = QueryOut()
        rc = load().smcm_server_poll(self._h, ticket, 1 if wait else 0, C.byref(q))
        if rc == SMCM_EAGAIN:
            return None
        _check(rc)
        self._pending.pop(ticket, None)
        d = {k: getattr(q, k) for k, _ in QueryOut._fields_}
        d["error"] = d["error"].decode(errors="replace")
        return d

    def cancel(self, ticket: int) -> None:
        _check(load().smcm_server_cancel(self._h, ticket))

    def stats(self) -> dict:
        hits, ins, byts = C.c_int64(), C.c_int64(), C.c_uint64()
        _check(load().smcm_server_stats(self._h, C.byref(hits), C.byref(ins), C.byref(byts)))
        return {"cache_hits": hits.value, "cache_inserts": ins.value, "cache_bytes": byts.value}

    def replay(self, queries, arrivals):
        """queries: list of (graph index, MetaPath, global source rows or None); arrivals: seconds.
        Returns (list of per-query dicts, list of per-row digest arrays)."""
        n = len(queries)
        gi = np.array([q[0] for q in queries], np.int32)
        plen = np.array([q[1].pathLen for q in queries], np.int32)
        vts = np.concatenate([np.asarray(q[1].vertexTypes, np.int32) for q in queries])
        ets = np.concatenate([np.asarray(q[1].edgeTypes, np.int32) for q in queries])
        if any(q[2] is None for q in queries):
            if not all(q[2] is None for q in queries):
                raise ValueError("either every query has source rows or none")
            soff, srows = None, None
            nrows = [int(self.graphs[q[0]].hg.dimensions[q[1].vertexTypes[0]]) for q in queries]
        else:
            parts = [np.ascontiguousarray(q[2], np.int32) for q in queries]
            nrows = [len(p) for p in parts]
            soff = np.zeros(n + 1, np.int64)
            soff[1:] = np.cumsum(nrows)
            srows = np.concatenate(parts) if soff[-1] > 0 else np.zeros(1, np.int32)
        arr = np.ascontiguousarray(arrivals, np.float64)
        dig = np.zeros(max(int(sum(nrows)), 1), np.uint64)
        outs = (QueryOut * max(n, 1))()
        _check(load().smcm_server_replay(self._h, n, _ptr(gi), _ptr(plen), _ptr(vts), _ptr(ets),
                                         None if soff is None else _ptr(soff), None if srows is None else _ptr(srows),
                                         _ptr(arr), _ptr(dig), outs))
        res, digs, o = [], [], 0
        for i in range(n):
            q = outs[i]
            d = {k: getattr(q, k) for k, _ in QueryOut._fields_}
            d["error"] = d["error"].decode(errors="replace")
            res.append(d)
            digs.append(dig[o:o + nrows[i]])
            o += nrows[i]
        return res, digs
