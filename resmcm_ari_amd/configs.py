"""Seeded synthetic workloads (graphs, meta-paths, chain batches, query streams).

Type and edge-type numbering follow the reference's meta-path generators so that meta-path
strings read the same: bibliographic graphs use the DBLP schema (paper 0, author 1, venue 2,
term 3; P->A 0, P->V 1, P->T 2, A->P 3, V->P 4, T->P 5; tool/generate_DBLP.cpp:334-352) and
location graphs the FourSquare schema (venue 0, city 1, category 2, user 3, slot 4; V->Y 0,
Y->V 1, V->C 2, C->V 3, V->U 4, U->V 5, V->S 6, S->V 7; tool/generate_FourSquare.cpp:335-359).

Seeds are ``0x5EED0000 + k`` for workload k (1-based, list order).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .hin import HeteroGraph, MetaPath

SEED_BASE = 0x5EED0000

# bibliographic schema
P, A, V, T = 0, 1, 2, 3
BIB_EDGE = {(P, A): 0, (P, V): 1, (P, T): 2, (A, P): 3, (V, P): 4, (T, P): 5}
# location schema
LV, LY, LC, LU, LS = 0, 1, 2, 3, 4
LOC_EDGE = {(LV, LY): 0, (LY, LV): 1, (LV, LC): 2, (LC, LV): 3, (LV, LU): 4, (LU, LV): 5,
            (LV, LS): 6, (LS, LV): 7}

BIB_NAMES = {"P": P, "A": A, "V": V, "T": T}
LOC_NAMES = {"L": LV, "Y": LY, "C": LC, "U": LU, "S": LS}


def metapath_from_letters(letters: str, schema: str) -> MetaPath:
    """``"A-P-V-P-A"`` -> MetaPath over the named schema ("bib" or "loc")."""
    names = BIB_NAMES if schema == "bib" else LOC_NAMES
    edges = BIB_EDGE if schema == "bib" else LOC_EDGE
    vts = [names[c] for c in letters.split("-")]
    ets = [edges[(vts[i], vts[i + 1])] for i in range(len(vts) - 1)]
    return MetaPath(vertexTypes=vts, edgeTypes=ets)


@dataclass(frozen=True)
class BibSpec:
    nA: int
    nP: int
    nV: int
    nT: int
    lam_auth: float
    s_auth: float = 1.1
    s_venue: float = 0.8
    terms_per_paper: int = 8
    s_term: float = 1.0
    seed: int = SEED_BASE

    def build(self) -> HeteroGraph:
        return HeteroGraph.generate_bib(self.nA, self.nP, self.nV, self.nT, self.lam_auth, self.s_auth,
                                        self.s_venue, self.terms_per_paper, self.s_term, self.seed)


@dataclass(frozen=True)
class LocSpec:
    nU: int
    nL: int
    nC: int
    nY: int
    nS: int
    geo_p: float = 0.02
    s_checkin: float = 1.2
    s_city: float = 1.0
    seed: int = SEED_BASE

    def build(self) -> HeteroGraph:
        return HeteroGraph.generate_loc(self.nU, self.nL, self.nC, self.nY, self.nS, self.geo_p,
                                        self.s_checkin, self.s_city, self.seed)


# Workload 1: small bibliographic graph (A=20k, P=30k, V=200; T=20k added, drawn from its own
# stream so that A/P/V relations are those of the three-type graph).
CFG1 = BibSpec(nA=20_000, nP=30_000, nV=200, nT=20_000, lam_auth=2.0, seed=SEED_BASE + 1)
# Workload 2: large bibliographic graph.
CFG2 = BibSpec(nA=500_000, nP=800_000, nV=5_000, nT=20_000, lam_auth=3.0, seed=SEED_BASE + 2)
# Workload 3: location graph for long chains.
CFG3 = LocSpec(nU=100_000, nL=200_000, nC=500, nY=1_000, nS=168, seed=SEED_BASE + 3)

CFG1_PATHS = ["A-P-V-P-A"]
CFG2_PATHS = ["A-P-V-P-A", "A-P-T-P-A"]
# long location meta-paths, L = 8-12 matrices
CFG3_PATHS = ["U-L-C-L-Y-L-U-L-S-L-U", "U-L-Y-L-S-L-C-L-U", "U-L-S-L-C-L-U-L-Y",
              "U-L-C-L-U-L-Y-L-S-L-U-L-U"]


def tiny_bib(seed: int = 7, nA: int = 60, nP: int = 80, nV: int = 6, nT: int = 30) -> BibSpec:
    return BibSpec(nA=nA, nP=nP, nV=nV, nT=nT, lam_auth=2.0, terms_per_paper=3, seed=seed)


def tiny_loc(seed: int = 11) -> LocSpec:
    return LocSpec(nU=50, nL=90, nC=7, nY=9, nS=12, geo_p=0.2, seed=seed)


BIB_SCHEMA = {P: [A, V, T], A: [P], V: [P], T: [P]}
LOC_SCHEMA = {LV: [LY, LC, LU, LS], LY: [LV], LC: [LV], LU: [LV], LS: [LV]}


def random_walk_path(rng: np.random.Generator, L: int, schema: str, start: int | None = None) -> MetaPath:
    """Random schema-valid meta-path with L matrices."""
    sch = BIB_SCHEMA if schema == "bib" else LOC_SCHEMA
    edges = BIB_EDGE if schema == "bib" else LOC_EDGE
    t = int(rng.choice(sorted(sch))) if start is None else start
    vts = [t]
    for _ in range(L):
        t = int(rng.choice(sch[t]))
        vts.append(t)
    ets = [edges[(vts[i], vts[i + 1])] for i in range(L)]
    return MetaPath(vertexTypes=vts, edgeTypes=ets)


def log_uniform_int(rng: np.random.Generator, lo: int, hi: int) -> int:
    return int(min(hi, max(lo, math.floor(math.exp(rng.uniform(math.log(lo), math.log(hi + 1)))))))


@dataclass(frozen=True)
class ChainQuery:
    path: MetaPath
    src_rows: np.ndarray  # ascending local ids of the start type


def cfg4_batch(dims, n_chains: int = 4096, seed: int = SEED_BASE + 4, lmin: int = 2, lmax: int = 6,
               smin: int = 16, smax: int = 65_536) -> list[ChainQuery]:
    """Workload 4: random bibliographic meta-paths with lmin..lmax matrices, each from a source
    subset of log-uniform size smin..smax (capped at the start type's size), rows drawn without
    replacement and sorted."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_chains):
        L = int(rng.integers(lmin, lmax + 1))
        mp = random_walk_path(rng, L, "bib")
        n = log_uniform_int(rng, smin, smax)
        d0 = int(dims[mp.vertexTypes[0]])
        rows = np.sort(rng.choice(d0, size=min(n, d0), replace=False)).astype(np.int32)
        out.append(ChainQuery(mp, rows))
    return out


@dataclass(frozen=True)
class ServedQuery:
    graph: int             # 0: workload-1 bibliographic graph, 1: workload-3 location graph
    path: MetaPath
    src_rows: np.ndarray   # ascending local ids of the start type
    arrival: float         # seconds from the start of the stream
    parent: int            # index of the earlier query whose prefix (and sources) it shares, -1: none


def cfg5_stream(dims_bib, dims_loc, n_queries: int = 512, qps: float = 20.0, seed: int = SEED_BASE + 5,
                lmin: int = 2, lmax: int = 10, smin: int = 1, smax: int = 100_000,
                share: float = 0.7) -> list[ServedQuery]:
    """Workload 5: Poisson arrivals at `qps`; each query picks a graph, a meta-path with lmin..lmax
    matrices and a source subset of log-uniform size smin..smax (capped at the start type's size).
    With probability `share` it extends a prefix of an earlier query on the same graph (same
    sources, same first k matrices, 1 <= k < its own length)."""
    rng = np.random.default_rng(seed)
    out: list[ServedQuery] = []
    t = 0.0
    for q in range(n_queries):
        t += float(rng.exponential(1.0 / qps))
        L = int(rng.integers(lmin, lmax + 1))
        parent = -1
        if out and rng.random() < share:
            parent = int(rng.integers(0, len(out)))
            p = out[parent]
            schema = "bib" if p.graph == 0 else "loc"
            k = int(rng.integers(1, min(p.path.pathLen, L - 1) + 1)) if L > 1 else 1
            k = min(k, L - 1) if L > 1 else 1
            tail = random_walk_path(rng, L - k, schema, start=p.path.vertexTypes[k])
            vts = list(p.path.vertexTypes[: k + 1]) + list(tail.vertexTypes[1:])
            ets = list(p.path.edgeTypes[:k]) + list(tail.edgeTypes)
            out.append(ServedQuery(p.graph, MetaPath(vertexTypes=vts, edgeTypes=ets), p.src_rows, t, parent))
            continue
        gidx = int(rng.integers(0, 2))
        schema = "bib" if gidx == 0 else "loc"
        dims = dims_bib if gidx == 0 else dims_loc
        mp = random_walk_path(rng, L, schema)
        d0 = int(dims[mp.vertexTypes[0]])
        n = log_uniform_int(rng, smin, smax)
        rows = np.sort(rng.choice(d0, size=min(n, d0), replace=False)).astype(np.int32)
        out.append(ServedQuery(gidx, mp, rows, t, -1))
    return out
