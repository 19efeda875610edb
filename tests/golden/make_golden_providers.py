#!/usr/bin/env python3
"""Writes the golden fixtures of the reference's alternative cost providers.

The reference's programs are compiled from their sources by oracle/Makefile into oracle/_ref/
(never copied into this repository): the estimate programs accuracy/meta.cpp, mnc.cpp and dm.cpp
(`meta_acc`, `mnc_acc`, `dm_acc`: "mp estimate" lines) and the planner harness
oracle/provider_harness.cpp around efficient/meta.cpp, mnc.cpp and dm.cpp (`meta_dump`,
`mnc_dump`, `dm_dump`: the DP tables m_ and s_ and the planner's return value).  LGraph
(lg.cpp) draws its ranks from an unseeded generator, so it has no fixture.

This script, run in the build container where the reference tree exists, writes the seeded
graphs of make_golden.py in the reference's text format, runs those programs with one thread and
stores:
  golden_<graph>_providers.npz   per provider and meta-path: m, s, ret
  golden_<graph>_<prov>_acc.txt  the estimate programs' output lines (6 significant digits)
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_golden import GRAPHS, REF_DIR, ROOT, letters  # noqa: E402

PROVIDERS = ("meta", "mnc", "dm")
PROVIDER_GRAPHS = ("bib_tiny", "bib_tiny2", "loc_tiny", "edge", "cfg1")


def read_tables(path):
    data = open(path, "rb").read()
    off = 0
    out = []
    while off < len(data):
        L = int(np.frombuffer(data, "<i4", 1, off)[0])
        off += 4
        m = np.frombuffer(data, "<f8", L * L, off).reshape(L, L).copy()
        off += 8 * L * L
        s = np.frombuffer(data, "<i4", L * L, off).reshape(L, L).copy()
        off += 4 * L * L
        ret = float(np.frombuffer(data, "<f8", 1, off)[0])
        off += 8
        out.append((m, s, ret))
    return out


def main():
    needed = [f"{p}_{k}" for p in PROVIDERS for k in ("acc", "dump")]
    if any(not os.path.exists(os.path.join(REF_DIR, x)) for x in needed):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    only = set(sys.argv[1:])
    for name in PROVIDER_GRAPHS:
        if only and name not in only:
            continue
        build, schema, paths, _ = GRAPHS[name]
        hg = build()
        mps = letters(paths, schema)
        arrays = {"paths": np.array([mp.toString() for mp in mps])}
        with tempfile.TemporaryDirectory() as td:
            gf, vf, ef, mf = (os.path.join(td, x) for x in ("graph.txt", "vertex.txt", "edge.txt", "mp.txt"))
            hg.write_text(gf, vf, ef)
            with open(mf, "w") as f:
                for mp in mps:
                    f.write(mp.toString() + "\n")
            for prov in PROVIDERS:
                dump = os.path.join(td, f"{prov}.bin")
                subprocess.run([os.path.join(REF_DIR, f"{prov}_dump"), "1", gf, vf, ef, mf, dump], check=True,
                               stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
                for k, (m, s, ret) in enumerate(read_tables(dump)):
                    arrays[f"{prov}_p{k}_m"] = m
                    arrays[f"{prov}_p{k}_s"] = s
                    arrays[f"{prov}_p{k}_ret"] = np.array(ret)
                acc = os.path.join(td, f"{prov}_acc.txt")
                subprocess.run([os.path.join(REF_DIR, f"{prov}_acc"), "1", "1", gf, vf, ef, mf, acc], check=True,
                               stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
                with open(os.path.join(HERE, f"golden_{name}_{prov}_acc.txt"), "w") as f:
                    f.write(open(acc).read())
        np.savez_compressed(os.path.join(HERE, f"golden_{name}_providers.npz"), **arrays)
        print(f"{name}: {len(mps)} meta-paths", flush=True)


if __name__ == "__main__":
    main()
