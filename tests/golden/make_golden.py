#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the reference's own programs.

The reference (efficient/resmcm.cpp, accuracy/res.cpp) is compiled from its sources by
oracle/Makefile into oracle/_ref/ (never copied into this repository).  This script, run in the
build container where the reference tree exists, writes seeded graphs in the reference's text
format, runs
  * oracle/_ref/ref_dump  (a harness around the reference's DynamicOptimizer / SSP_dynamic, one
                           OpenMP thread so that c_ij sums are the sequential ones),
  * oracle/_ref/resmcm    (the reference program itself: 6-column output lines),
  * oracle/_ref/res       (the reference estimator program: "mp estimate" lines),
and stores their outputs as fixtures (inputs are regenerated by the seeded generators):
  golden_<graph>.npz      c tables, DP split tables, chain orders, r_ij vectors (tiny graphs) or
                          their SHA-256 (larger graphs), result structure (full for tiny graphs,
                          per-row widths + structure digests for larger graphs)
  golden_<graph>_resmcm.txt / golden_<graph>_res.txt   the reference programs' output lines
"""
from __future__ import annotations

import hashlib
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from resmcm_ari_amd import configs  # noqa: E402
from resmcm_ari_amd.hin import HeteroGraph, MetaPath  # noqa: E402

REF_DIR = os.path.join(ROOT, "oracle", "_ref")


def edge_case_graph() -> HeteroGraph:
    """Hand-made HIN with duplicate entries, unsorted adjacency, isolated vertices and an edge
    type that never occurs between some types (DBLP type/edge numbering)."""
    P, A, V, T = 0, 1, 2, 3
    vt = [P, P, P, P, A, A, A, V, V, T, T, A, P]
    n = len(vt)
    adj = {v: [] for v in range(n)}

    def link(u, w, e_uw, e_wu, times=1):
        for _ in range(times):
            adj[u].append((w, e_uw))
            adj[w].append((u, e_wu))

    link(0, 5, 0, 3)
    link(0, 4, 0, 3)            # paper 0 authors listed out of order: 5 then 4
    link(1, 4, 0, 3, times=2)   # duplicate authorship entry
    link(1, 6, 0, 3)
    link(2, 6, 0, 3)
    link(3, 11, 0, 3)
    link(0, 8, 1, 4)
    link(1, 7, 1, 4)
    link(2, 8, 1, 4)
    link(3, 7, 1, 4)
    link(0, 10, 2, 5)
    link(2, 9, 2, 5)
    link(2, 10, 2, 5)
    # vertex 12 (a paper) is isolated; vertex 3's term list is empty
    ptr = [0]
    nbr, et = [], []
    for v in range(n):
        for w, e in adj[v]:
            nbr.append(w)
            et.append(e)
        ptr.append(len(nbr))
    return HeteroGraph.from_arrays(np.array(vt), np.array(ptr), np.array(nbr), np.array(et))


def letters(ps, schema):
    return [configs.metapath_from_letters(p, schema) for p in ps]


BIB_TINY_PATHS = ["A-P-A", "P-V-P", "A-P-V-P-A", "A-P-T-P-A", "P-A-P-T-P", "V-P-A-P-V", "T-P-A-P-V-P-T",
                  "A-P-A-P-A-P-A", "P-T-P"]
LOC_TINY_PATHS = ["U-L-U", "U-L-C-L-U", "U-L-C-L-Y-L-U-L-S-L-U", "L-Y-L-C-L-S-L", "C-L-U-L-Y"]
EDGE_PATHS = ["A-P-A", "A-P-V-P-A", "P-T-P", "A-P-T-P-A", "V-P-T", "P-A", "A-P-A-P-A"]
CFG1_PATHS = ["A-P-V-P-A", "P-V-P-A", "V-P-A-P-V", "A-P-V"]

GRAPHS = {
    "bib_tiny": (lambda: configs.tiny_bib(seed=7).build(), "bib", BIB_TINY_PATHS, True),
    "bib_tiny2": (lambda: configs.tiny_bib(seed=21, nA=150, nP=200, nV=9, nT=40).build(), "bib", BIB_TINY_PATHS, True),
    "loc_tiny": (lambda: configs.tiny_loc(seed=11).build(), "loc", LOC_TINY_PATHS, True),
    "edge": (edge_case_graph, "bib", EDGE_PATHS, True),
    "cfg1": (lambda: configs.CFG1.build(), "bib", CFG1_PATHS, False),
}


def read_dump(path):
    """Parse the binary records written by oracle/ref_harness.cpp."""
    recs = []
    with open(path, "rb") as f:
        data = f.read()
    off = 0

    def take(fmt, count=1):
        nonlocal off
        size = np.dtype(fmt).itemsize * count
        arr = np.frombuffer(data, dtype=np.dtype(fmt), count=count, offset=off)
        off += size
        return arr

    while off < len(data):
        L = int(take("<i4")[0])
        vts = take("<i4", L + 1).copy()
        ets = take("<i4", L).copy()
        c = take("<f8", L * L).reshape(L, L).copy()
        s = take("<i4", L * L).reshape(L, L).copy()
        npairs = int(take("<i4")[0])
        pairs = take("<i4", 2 * npairs).reshape(npairs, 2).copy()
        r = {}
        for i in range(L):
            for j in range(i, L):
                nr = int(take("<i8")[0])
                r[(i, j)] = take("<f8", nr).copy()
        nr = int(take("<i8")[0])
        lens = np.zeros(nr, np.int64)
        cols = []
        for q in range(nr):
            ln = int(take("<i8")[0])
            lens[q] = ln
            cols.append(take("<i4", ln).copy())
        recs.append(dict(L=L, vts=vts, ets=ets, c=c, s=s, pairs=pairs, r=r, lens=lens, cols=cols))
    return recs


def structure_digest(lens, cols_concat):
    """Per-row digest of sorted global column ids (order-free sum of mixed ids + width)."""
    from numpy import uint64 as U

    def mix(z):
        z = (z ^ (z >> U(30))) * U(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> U(27))) * U(0x94D049BB133111EB)
        return z ^ (z >> U(31))

    with np.errstate(over="ignore"):
        h = mix(cols_concat.astype(np.uint64) ^ U(0x9E3779B97F4A7C15))
        rows = np.repeat(np.arange(len(lens)), lens)
        out = np.zeros(len(lens), np.uint64)
        np.add.at(out, rows, h)
        out += mix(lens.astype(np.uint64) ^ U(0x243F6A8885A308D3))
    return out


def main():
    for exe in ("ref_dump", "resmcm", "res"):
        if not os.path.exists(os.path.join(REF_DIR, exe)):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
            break
    only = set(sys.argv[1:])
    for name, (build, schema, paths, full) in GRAPHS.items():
        if only and name not in only:
            continue
        hg = build()
        mps = letters(paths, schema)
        with tempfile.TemporaryDirectory() as td:
            gf, vf, ef, mf = (os.path.join(td, x) for x in ("graph.txt", "vertex.txt", "edge.txt", "mp.txt"))
            hg.write_text(gf, vf, ef)
            with open(mf, "w") as f:
                for mp in mps:
                    f.write(mp.toString() + "\n")
            dump = os.path.join(td, "dump.bin")
            subprocess.run([os.path.join(REF_DIR, "ref_dump"), "1", gf, vf, ef, mf, dump], check=True,
                           stdout=subprocess.DEVNULL)
            recs = read_dump(dump)
            out_resmcm = os.path.join(td, "resmcm.txt")
            subprocess.run([os.path.join(REF_DIR, "resmcm"), "1", "1", gf, vf, ef, mf, out_resmcm], check=True,
                           stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            out_res = os.path.join(td, "res.txt")
            subprocess.run([os.path.join(REF_DIR, "res"), "1", "1", gf, vf, ef, mf, out_res], check=True,
                           stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            resmcm_lines = open(out_resmcm).read()
            res_lines = open(out_res).read()
        arrays = {"paths": np.array([mp.toString() for mp in mps])}
        for k, rec in enumerate(recs):
            pre = f"p{k}_"
            arrays[pre + "c"] = rec["c"]
            arrays[pre + "s"] = rec["s"]
            arrays[pre + "pairs"] = rec["pairs"]
            arrays[pre + "lens"] = rec["lens"]
            cols = np.concatenate(rec["cols"]) if rec["cols"] else np.zeros(0, np.int32)
            if full:
                arrays[pre + "cols"] = cols.astype(np.int32)
                for (i, j), v in rec["r"].items():
                    arrays[pre + f"r_{i}_{j}"] = v
            else:
                arrays[pre + "digest"] = structure_digest(rec["lens"], cols)
                for (i, j), v in rec["r"].items():
                    arrays[pre + f"rsha_{i}_{j}"] = np.frombuffer(hashlib.sha256(v.tobytes()).digest(), np.uint8)
                    arrays[pre + f"rhead_{i}_{j}"] = v[:64]
        np.savez_compressed(os.path.join(HERE, f"golden_{name}.npz"), **arrays)
        # drop the time columns of the reference program's lines: keep "mp nnz"
        with open(os.path.join(HERE, f"golden_{name}_resmcm.txt"), "w") as f:
            for line in resmcm_lines.splitlines():
                parts = line.split()
                f.write(f"{parts[0]} {parts[1]}\n")
        with open(os.path.join(HERE, f"golden_{name}_res.txt"), "w") as f:
            f.write(res_lines)
        print(f"{name}: {len(recs)} meta-paths", flush=True)


if __name__ == "__main__":
    main()
