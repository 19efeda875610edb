"""CPU restatement of the reference's alternative DP cost providers (MetaAC, MNC, DMap, LGraph).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's cost providers.

Every provider works in the reference's global index space: the typed matrix M_i of a meta-path
is an n x n matrix (n = all vertices of the graph) whose rows are the vertices of type t_i and
whose columns are global vertex ids.  ``GMat`` gives that view of an oracle typed matrix (local
row r is global vertex bins[t_i][r], local column c is bins[t_{i+1}][c]).

Floating point follows the reference: ``rate`` is ``double`` in meta.cpp and dm.cpp and ``long
double`` (x87 80-bit, numpy's ``longdouble`` on x86-64) in mnc.cpp and lg.cpp, sums are
sequential where the reference's are, and the LGraph ranks come from the seeded generator
``lgraph_ranks`` (the reference draws them from an unseeded ``mt19937``, lg.cpp setN).
"""
from __future__ import annotations

import math

import numpy as np

# the sparse cost model's weights (efficient/meta.cpp:44-46, identical in dm.cpp, mnc.cpp, lg.cpp)
ALPH, BT, GMM = -130.0, -16.0, 83.0
LD = np.longdouble

PROVIDERS = {"metaac": 2, "mnc": 3, "dmap": 4, "lgraph": 5}


class GMat:
    """Global-index view of a typed matrix: row_g[r] is the global id of local row r, col_g the
    global column id of every entry (CSR order)."""

    def __init__(self, M, row_bins, col_bins):
        self.ptr = np.asarray(M.ptr, np.int64)
        self.row_g = np.asarray(row_bins, np.int64)
        self.col_g = np.asarray(col_bins, np.int64)[np.asarray(M.col, np.int64)]
        self.lens = np.diff(self.ptr)
        self.rows_of_entries = np.repeat(self.row_g, self.lens)


def global_mats(g, mp, mats):
    bins = {}
    vt = np.asarray(g.vertexTypes)
    for t in set(mp.vertexTypes):
        bins[t] = np.flatnonzero(vt == t)
    return [GMat(mats[i], bins[mp.vertexTypes[i]], bins[mp.vertexTypes[i + 1]]) for i in range(len(mats))], bins


def _plan(s, L):
    pairs = []

    def rec(i, j):
        if i == j:
            return i
        k = rec(i, int(s[i, j]))
        l = rec(int(s[i, j]) + 1, j)
        pairs.append((k, l))
        return -1

    if L > 1:
        rec(0, L - 1)
    return pairs


# ------------------------------------------------------------------------------------------
# MetaAC: one density per matrix, uniform model.  efficient/meta.cpp:455-460 (getSparsity),
# :489-501 (product density and cost), :505-566 (DP); accuracy/meta.cpp:486-549 (unweighted
# cost, the chain's density as the estimate).
def metaac_density(G: GMat, n: int) -> float:
    # res += (rate)len / n / n, sequentially in row order
    terms = G.lens.astype(np.float64) / n / n
    return float(np.cumsum(terms)[-1]) if len(terms) else 0.0


def _metaac_cost(a, b, c, n, weighted):
    tA = n * float(a * n)
    tB = (((float(n) * n) * a) * n) * b
    tC = float(n * c) * n
    if weighted:
        return ALPH * tA + BT * tB + GMM * tC
    return tA + tB + tC


def metaac(gm, n, accuracy=False):
    L = len(gm)
    m = np.zeros((L, L))
    s = np.zeros((L, L), np.int32)
    sp = np.zeros((L, L))
    dens = [metaac_density(G, n) for G in gm]
    for l in range(1, L):
        for i in range(L - l):
            j = i + l
            m[i, j] = np.finfo(np.float64).max
            for k in range(i, j):
                a = sp[i, k] if sp[i, k] != 0 else dens[i]
                b = sp[k + 1, j] if sp[k + 1, j] != 0 else dens[j]
                r = 1.0 - math.pow(1.0 - a * b, n)
                cost = m[i, k] + m[k + 1, j] + _metaac_cost(a, b, r, n, not accuracy)
                if cost < m[i, j]:
                    m[i, j] = cost
                    s[i, j] = k
                    sp[i, j] = r
    # sparsities[len - 1] (accuracy/meta.cpp:546): entry (0, L-1), never set when L = 1
    est = float(sp[0, L - 1]) * n * n
    return {"m": m, "s": s, "c": sp, "pairs": _plan(s, L), "estimate": est}


# ------------------------------------------------------------------------------------------
# MNC sketches.  efficient/mnc.cpp: MNCSketches (:313-548), BoolMatrix::buildSketches
# (:612-660), DP (:739-800); accuracy/mnc.cpp:761-825 (cost = dot of the sketches).
def _wrap32(x):
    x = np.asarray(x, np.int64)
    return ((x + (1 << 31)) & 0xFFFFFFFF) - (1 << 31)


class Sketch:
    def __init__(self, n):
        self.h_r = np.zeros(n, np.int64)
        self.h_c = np.zeros(n, np.int64)
        self.h_er = None
        self.h_ec = None
        self.rN1 = 0
        self.cN1 = 0
        self.es_nnz = LD(0)

    def finish(self):
        self.h_r_nnz = int(np.count_nonzero(self.h_r))
        self.h_c_nnz = int(np.count_nonzero(self.h_c))
        self.max_h_r = int(self.h_r.max()) if len(self.h_r) else 0
        self.max_h_c = int(self.h_c.max()) if len(self.h_c) else 0


def mnc_leaf(G: GMat, n: int) -> Sketch:
    sk = Sketch(n)
    np.add.at(sk.h_r, G.rows_of_entries, 1)
    np.add.at(sk.h_c, G.col_g, 1)
    sk.finish()
    sk.rN1 = int(np.count_nonzero(sk.h_r == 1))
    sk.cN1 = int(np.count_nonzero(sk.h_c == 1))
    sk.h_er = np.zeros(n, np.int64)
    sk.h_ec = np.zeros(n, np.int64)
    # extended sketches; the column test of h_ec reads the column sketch at the ROW's id
    # (mnc.cpp:648-653)
    np.add.at(sk.h_er, G.rows_of_entries[sk.h_c[G.col_g] == 1], 1)
    np.add.at(sk.h_ec, G.col_g[sk.h_c[G.rows_of_entries] == 1], 1)
    return sk


def _union_recurrence(lsp):
    # sp = sp + lsp - sp * lsp, sequentially (long double)
    sp = LD(0)
    for x in lsp:
        sp = sp + x - sp * x
    return sp


def mnc_sparsity(a: Sketch, b: Sketch, n: int) -> LD:
    """MNCSketches::getSparsity(a, b, m = n, n = n, l = n), mnc.cpp:399-441."""
    if a.max_h_r <= 1 or b.max_h_c <= 1:
        nnz = LD(int(_wrap32(a.h_c * b.h_r).sum()))
    elif a.h_ec is not None and b.h_er is not None:
        mn_out = LD(int(_wrap32((a.h_r_nnz - a.rN1) * (b.h_c_nnz - b.cN1))))
        h1 = a.h_ec
        h2 = a.h_er  # the reference reads a's extended row sketch here (mnc.cpp:411)
        nnz = LD(int((h1 * b.h_r).sum())) + LD(int(((a.h_c - h1) * h2).sum()))
        with np.errstate(divide="ignore", invalid="ignore"):
            lsp = (a.h_c - h1).astype(LD) * (b.h_r - h2).astype(LD) / mn_out
        nnz = nnz + _union_recurrence(lsp) * mn_out
    else:
        p = LD(a.h_r_nnz) * LD(b.h_c_nnz)
        with np.errstate(divide="ignore", invalid="ignore"):
            lsp = a.h_c.astype(LD) * b.h_r.astype(LD) / p
        nnz = _union_recurrence(lsp) * p
    larger = LD(int(_wrap32(int(np.count_nonzero(a.h_r > n // 2)) * int(np.count_nonzero(b.h_c > n // 2)))))
    nnz = larger if nnz < larger else nnz
    return (nnz / LD(n)) * LD(1) / LD(n)


def _propagate_vec(v, sparsity_d, n):
    s = LD(int(v.sum()))
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        x = (v.astype(np.float64) * sparsity_d * float(n) * float(n)).astype(LD) / s
        # round half away from zero (roundl), then the int conversion; a NaN or out-of-range
        # value converts to INT_MIN as on x86 (an empty sketch gives 0 / 0)
        ax = np.abs(x)
        r = np.floor(ax)
        r = r + (ax - r >= LD(0.5))
        r = np.sign(x) * r
        bad = ~np.isfinite(r) | (r >= LD(2.0 ** 31)) | (r < LD(-(2.0 ** 31)))
        return np.where(bad, -(1 << 31), r).astype(np.int64)


def mnc_propagate(a: Sketch, b: Sketch, n: int) -> Sketch:
    sp = mnc_sparsity(a, b, n)
    out = Sketch(n)
    out.es_nnz = sp * LD(n) * LD(n)
    sd = float(sp)  # propagate*Sketch take the sparsity as double (mnc.cpp:544-545)
    out.h_r = _propagate_vec(a.h_r, sd, n)
    out.h_c = _propagate_vec(b.h_c, sd, n)
    out.finish()
    return out


def mnc(gm, n, accuracy=False):
    L = len(gm)
    m = np.zeros((L, L), LD)
    s = np.zeros((L, L), np.int32)
    E = {}
    for i in range(L):
        E[(i, i)] = mnc_leaf(gm[i], n)
    for l in range(1, L):
        for i in range(L - l):
            j = i + l
            m[i, j] = np.finfo(LD).max
            for k in range(i, j):
                a, b = E[(i, k)], E[(k + 1, j)]
                if accuracy:
                    tmp = LD(int(_wrap32(a.h_c * b.h_r).sum()))
                    prop = None
                else:
                    prop = mnc_propagate(a, b, n)
                    tmp = prop.es_nnz
                cost = m[i, k] + m[k + 1, j] + tmp
                if cost < m[i, j]:
                    m[i, j] = cost
                    s[i, j] = k
                    E[(i, j)] = prop if prop is not None else mnc_propagate(a, b, n)
    est = E[(0, L - 1)].es_nnz if L > 1 else LD(0)
    return {"m": m, "s": s, "pairs": _plan(s, L), "estimate": est}


# ------------------------------------------------------------------------------------------
# DMap: 256 x 256 density blocks.  efficient/dm.cpp: DenseBlock (:52-108), buildDenseBlock
# (:115-138), dense_map_estimator (:539-561), DP (:563-620).
DMAP_B = 256


def dmap_leaf(G: GMat, n: int) -> np.ndarray:
    bn = -(-n // DMAP_B)
    cnt = np.zeros((bn, bn), np.float64)
    np.add.at(cnt, (G.rows_of_entries // DMAP_B, G.col_g // DMAP_B), 1.0)
    return cnt / DMAP_B / DMAP_B


def dmap_product(left: np.ndarray, right: np.ndarray) -> np.ndarray:
    bn = left.shape[0]
    res = np.zeros_like(left)
    for i in range(bn):
        for k in range(bn):
            sp1 = left[i, k]
            if sp1 == 0:
                continue
            row = right[k]
            nz = row != 0
            if not nz.any():
                continue
            t1 = 1.0 - np.power(1.0 - sp1 * row[nz], float(DMAP_B))
            t2 = res[i, nz]
            res[i, nz] = t1 + t2 - t1 * t2
    return res


def dmap_density(D: np.ndarray) -> float:
    bn = D.shape[0]
    return float(D.sum()) / bn / bn


def dmap(gm, n):
    L = len(gm)
    leaves = [dmap_leaf(G, n) for G in gm]
    c = np.zeros((L, L))
    maps = {}
    for i in range(L - 1, -1, -1):
        maps[(i, i)] = leaves[i]
        c[i, i] = dmap_density(leaves[i])
        for j in range(i + 1, L):
            maps[(i, j)] = dmap_product(leaves[i], maps[(i + 1, j)])
            c[i, j] = dmap_density(maps[(i, j)])
    m = np.zeros((L, L))
    s = np.zeros((L, L), np.int32)
    for l in range(1, L):
        for i in range(L - l):
            j = i + l
            m[i, j] = np.finfo(np.float64).max
            for k in range(i, j):
                tmp = ALPH * c[i, k] + BT / n * c[i, k] * c[k + 1, j] + GMM * c[i, j]
                cost = m[i, k] + m[k + 1, j] + tmp
                if cost < m[i, j]:
                    m[i, j] = cost
                    s[i, j] = k
    return {"m": m, "s": s, "c": c, "pairs": _plan(s, L), "estimate": float(c[0, L - 1]) * n * n}


# ------------------------------------------------------------------------------------------
# LGraph: 32 exponential ranks per vertex, min-propagated along the chain.  efficient/lg.cpp:
# LayeredGraph (:52-132), buildLayeredGraph (:137-215), layered_graph_estimator (:657-681), DP
# (:683-740).  accuracy/lg.cpp:477-520, :586-652 (row and column nodes = all vertices of the
# types; the chain estimate only).
LG_R = 32
_M64 = (1 << 64) - 1


def lgraph_ranks(seed: int, layer: int, u: np.ndarray) -> np.ndarray:
    """Ranks of vertices u (global ids) for leaf `layer`: 32 Exp(1) draws each, from a
    splitmix64 hash of (seed, layer, u, idx)."""
    u = np.asarray(u, np.uint64)
    idx = np.arange(LG_R, dtype=np.uint64)
    key = (np.uint64(layer) << np.uint64(58)) ^ (u[:, None] << np.uint64(5)) ^ idx[None, :]
    with np.errstate(over="ignore"):
        z = np.uint64(seed & _M64) + (key + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    uni = ((z >> np.uint64(11)).astype(np.float64) + 1.0) * (2.0 ** -53)
    return -np.log(uni)


def lgraph_col_nodes(G: GMat, n: int) -> np.ndarray:
    """The column node list of buildLayeredGraph (lg.cpp:150-166): every entry's column is
    pushed unless its flag is set, and the flag set is the ROW's."""
    flag = np.zeros(n, bool)
    out = []
    order = np.argsort(G.row_g, kind="stable")
    for r in order:
        u = int(G.row_g[r])
        for v in G.col_g[G.ptr[r]:G.ptr[r + 1]]:
            v = int(v)
            if not flag[v]:
                out.append(v)
                flag[u] = True
    return np.asarray(out, np.int64)


def _lg_estimate(R, nodes):
    sums = R[nodes].astype(LD).sum(axis=1)
    res = LD(0)
    for x in sums:
        res += LD(LG_R - 1) / x
    return res


def lgraph(gm, n, seed, accuracy=False, bins=None, types=None):
    L = len(gm)
    c = np.zeros((L, L), LD)
    BIG = np.finfo(np.float64).max
    cur = None
    for i in range(L):
        G = gm[i]
        R = np.full((n, LG_R), BIG)
        if accuracy:
            rows = bins[types[i]]
            cols = bins[types[i + 1]]
        else:
            rows = np.unique(G.rows_of_entries)
            cols = lgraph_col_nodes(G, n)
        R[rows] = lgraph_ranks(seed, i, rows)
        new = np.full((n, LG_R), BIG)
        np.minimum.at(new, G.col_g, R[G.rows_of_entries])
        ucols = np.unique(cols)
        R[ucols] = new[ucols]
        if not accuracy:
            c[i, i] = _lg_estimate(R, cols)
        lefts = R
        for j in range(i + 1, L):
            Gj = gm[j]
            nodes = bins[types[j + 1]] if accuracy else lgraph_col_nodes(Gj, n)
            Rj = np.full((n, LG_R), BIG)
            np.minimum.at(Rj, Gj.col_g, lefts[Gj.rows_of_entries])
            if not accuracy:
                # only the column-node list is propagated (lg.cpp:671-679)
                keep = np.zeros(n, bool)
                keep[nodes] = True
                Rj[~keep] = BIG
            c[i, j] = _lg_estimate(Rj, nodes)
            lefts = Rj
    m = np.zeros((L, L), LD)
    s = np.zeros((L, L), np.int32)
    for l in range(1, L):
        for i in range(L - l):
            j = i + l
            m[i, j] = np.finfo(LD).max
            for k in range(i, j):
                tmp = c[i, k] + c[k + 1, j] + c[i, j]
                cost = m[i, k] + m[k + 1, j] + tmp
                if cost < m[i, j]:
                    m[i, j] = cost
                    s[i, j] = k
    return {"m": m, "s": s, "c": c, "pairs": _plan(s, L), "estimate": c[0, L - 1]}
