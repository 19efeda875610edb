"""Host-side HIN model mirroring the reference's loader classes.

``HeteroGraph`` (efficient/resmcm.cpp:181-240), ``MetaPath`` (:243-336) and
``read_metapathesN`` (:339-357) keep the reference's names, argument meaning and error
behaviour; the storage is a flat adjacency CSR owned by ``libsmcm_hin.so``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _build

HIN_MAX_TYPES = 32


class _HinGraph(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("m", C.c_int64),
        ("ntypes", C.c_int32),
        ("vtype", C.POINTER(C.c_int32)),
        ("adj_ptr", C.POINTER(C.c_int64)),
        ("adj_nbr", C.POINTER(C.c_int32)),
        ("adj_et", C.POINTER(C.c_int32)),
        ("dims", C.c_int32 * HIN_MAX_TYPES),
        ("type_offset", C.c_int32 * HIN_MAX_TYPES),
    ]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = os.path.join(_build.LIBDIR, "libsmcm_hin.so")
        if not os.path.exists(path):
            _build.build_hin()
        L = C.CDLL(path)
        P = C.POINTER(_HinGraph)
        L.hin_gen_bib.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                  C.c_double, C.c_int32, C.c_double, C.c_uint64, C.POINTER(P)]
        L.hin_gen_loc.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                  C.c_double, C.c_double, C.c_uint64, C.POINTER(P)]
        L.hin_from_arrays.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(P)]
        L.hin_write_text.argtypes = [P, C.c_char_p, C.c_char_p, C.c_char_p]
        L.hin_read_text.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(P)]
        L.hin_free.argtypes = [P]
        L.hin_free.restype = None
        L.hin_parse_metapath.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]
        L.hin_metapath_to_string.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_char_p, C.c_int32]
        L.hin_enumerate_metapaths.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32,
                                              C.c_int32, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]
        L.hin_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise RuntimeError(lib().hin_last_error().decode())


def _arr(ptr, n, dtype) -> np.ndarray:
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class HeteroGraph:
    """Typed graph with the reference's fields: ``n``, ``m``, ``max_vt``, ``vertexTypes``,
    ``dimensions`` (vertices per type) and ``bins`` (ascending ids per type).  The adjacency
    is ``adj_ptr``/``adj_nbr``/``adj_et`` (file order, edge types resolved)."""

    def __init__(self, graphFile: str | None = None, vertexFile: str | None = None,
                 edgeFile: str | None = None, *, _handle=None):
        if _handle is None:
            h = C.POINTER(_HinGraph)()
            _check(lib().hin_read_text(graphFile.encode(), vertexFile.encode(), edgeFile.encode(), C.byref(h)))
            _handle = h
        g = _handle.contents
        self.n = int(g.n)
        self.m = int(g.m)
        self.max_vt = int(g.ntypes) - 1
        self.vertexTypes = _arr(g.vtype, self.n, np.int32)
        self.adj_ptr = _arr(g.adj_ptr, self.n + 1, np.int64)
        self.adj_nbr = _arr(g.adj_nbr, self.m, np.int32)
        self.adj_et = _arr(g.adj_et, self.m, np.int32)
        self.dimensions = np.array([g.dims[t] for t in range(self.max_vt + 1)], dtype=np.int32)
        self.type_offset = np.array([g.type_offset[t] for t in range(self.max_vt + 1)], dtype=np.int32)
        lib().hin_free(_handle)
        self._bins = None

    @property
    def bins(self) -> list[np.ndarray]:
        if self._bins is None:
            order = np.argsort(self.vertexTypes, kind="stable")
            splits = np.cumsum(self.dimensions)[:-1]
            self._bins = [b.astype(np.int32) for b in np.split(order, splits)]
        return self._bins

    def local_ids(self) -> np.ndarray:
        loc = np.empty(self.n, dtype=np.int32)
        for b in self.bins:
            loc[b] = np.arange(len(b), dtype=np.int32)
        return loc

    @classmethod
    def from_arrays(cls, vtype, adj_ptr, adj_nbr, adj_et) -> "HeteroGraph":
        vt = np.ascontiguousarray(vtype, dtype=np.int32)
        ap = np.ascontiguousarray(adj_ptr, dtype=np.int64)
        an = np.ascontiguousarray(adj_nbr, dtype=np.int32)
        ae = np.ascontiguousarray(adj_et, dtype=np.int32)
        h = C.POINTER(_HinGraph)()
        _check(lib().hin_from_arrays(len(vt), vt.ctypes.data, ap.ctypes.data, an.ctypes.data, ae.ctypes.data,
                                     C.byref(h)))
        return cls(_handle=h)

    @classmethod
    def generate_bib(cls, nA, nP, nV, nT, lam_auth, s_auth, s_venue, terms_per_paper, s_term, seed) -> "HeteroGraph":
        h = C.POINTER(_HinGraph)()
        _check(lib().hin_gen_bib(nA, nP, nV, nT, lam_auth, s_auth, s_venue, terms_per_paper, s_term, seed,
                                 C.byref(h)))
        return cls(_handle=h)

    @classmethod
    def generate_loc(cls, nU, nL, nC, nY, nS, geo_p, s_checkin, s_city, seed) -> "HeteroGraph":
        h = C.POINTER(_HinGraph)()
        _check(lib().hin_gen_loc(nU, nL, nC, nY, nS, geo_p, s_checkin, s_city, seed, C.byref(h)))
        return cls(_handle=h)

    def write_text(self, graphFile: str, vertexFile: str, edgeFile: str) -> None:
        h = self._to_handle()
        try:
            _check(lib().hin_write_text(h, graphFile.encode(), vertexFile.encode(), edgeFile.encode()))
        finally:
            lib().hin_free(h)

    def _to_handle(self):
        h = C.POINTER(_HinGraph)()
        _check(lib().hin_from_arrays(self.n, self.vertexTypes.ctypes.data, self.adj_ptr.ctypes.data,
                                     self.adj_nbr.ctypes.data, self.adj_et.ctypes.data, C.byref(h)))
        return h


class MetaPath:
    """Meta-path ``vt-et-vt-...-vt``; ``pathLen`` is the number of matrices.  Raises
    ``ValueError("the meta-path is incorrect")`` where the reference throws."""

    def __init__(self, metaPathStr=None, *, vertexTypes=None, edgeTypes=None):
        if metaPathStr is not None:
            cap = 4096
            vt = (C.c_int32 * cap)()
            et = (C.c_int32 * cap)()
            L = C.c_int32(-1)
            rc = lib().hin_parse_metapath(metaPathStr.encode(), vt, et, cap, C.byref(L))
            if rc != 0:
                raise ValueError(lib().hin_last_error().decode())
            self.pathLen = int(L.value)
            self.vertexTypes = [int(vt[i]) for i in range(self.pathLen + 1)]
            self.edgeTypes = [int(et[i]) for i in range(self.pathLen)]
        else:
            if len(vertexTypes) != len(edgeTypes) + 1:
                raise ValueError("the meta-path is incorrect")
            self.vertexTypes = [int(x) for x in vertexTypes]
            self.edgeTypes = [int(x) for x in edgeTypes]
            self.pathLen = len(self.edgeTypes)

    def toString(self) -> str:
        s = ""
        for i in range(self.pathLen):
            s += f"{self.vertexTypes[i]}-{self.edgeTypes[i]}-"
        return s + str(self.vertexTypes[self.pathLen])

    __str__ = toString

    def __repr__(self) -> str:
        return f"MetaPath({self.toString()!r})"


def read_metapathesN(file: str) -> list[MetaPath]:
    """One meta-path per line (first whitespace token); stops at the first malformed line,
    as the reference does (resmcm.cpp:339-357)."""
    paths = []
    with open(file) as f:
        for line in f:
            tok = line.split()
            s = tok[0] if tok else ""
            try:
                paths.append(MetaPath(s))
            except ValueError:
                break
    return paths


# schemas of the reference's meta-path generators: start types in order, then relations
# (src type, edge type, dst type) in each type's neighbour order (tool/generate_DBLP.cpp:325-358,
# tool/generate_FourSquare.cpp:335-366, tool/generate_IMDB.cpp:304-328)
SCHEMAS = {
    "DBLP": ([0, 1, 2, 3], [(0, 0, 1), (0, 1, 2), (0, 2, 3), (1, 3, 0), (2, 4, 0), (3, 5, 0)]),
    "FourSquare": ([0, 1, 2, 3, 4], [(0, 0, 1), (0, 2, 2), (0, 4, 3), (0, 6, 4), (1, 1, 0), (2, 3, 0), (3, 5, 0),
                                     (4, 7, 0)]),
    "IMDB": ([0, 1, 2, 3], [(0, 0, 1), (0, 2, 2), (0, 4, 3), (1, 1, 0), (2, 3, 0), (3, 5, 0)]),
}


def enumerate_metapaths(schema, L: int) -> list[MetaPath]:
    """Every meta-path with L edges over a schema (name in SCHEMAS or (starts, relations)), in the
    order of the reference's generators."""
    starts, rels = SCHEMAS[schema] if isinstance(schema, str) else schema
    src = np.array([r[0] for r in rels], np.int32)
    et = np.array([r[1] for r in rels], np.int32)
    dst = np.array([r[2] for r in rels], np.int32)
    st = np.array(starts, np.int32)
    cnt = C.c_int64()
    _check(lib().hin_enumerate_metapaths(src.ctypes.data, et.ctypes.data, dst.ctypes.data, len(rels), st.ctypes.data,
                                         len(st), L, None, None, 0, C.byref(cnt)))
    n = cnt.value
    vts = np.zeros(max(n * (L + 1), 1), np.int32)
    ets = np.zeros(max(n * L, 1), np.int32)
    _check(lib().hin_enumerate_metapaths(src.ctypes.data, et.ctypes.data, dst.ctypes.data, len(rels), st.ctypes.data,
                                         len(st), L, vts.ctypes.data, ets.ctypes.data, n, C.byref(cnt)))
    return [MetaPath(vertexTypes=vts[i * (L + 1):(i + 1) * (L + 1)].tolist(), edgeTypes=ets[i * L:(i + 1) * L].tolist())
            for i in range(n)]
