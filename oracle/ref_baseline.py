"""CPU baseline from the reference's own program (oracle/_ref/resmcm, built from efficient/resmcm.cpp
by oracle/Makefile) on a bounded sample of a benchmark workload.

TEST INFRASTRUCTURE ONLY: used by bench.py's CPU-baseline leg and by the calibration recipe below,
never by the product.  The reference always computes every row of the start type
(efficient/resmcm.cpp:716, :859-868), so a sample of source rows is given to it as the same graph
without the start edges of the rows outside the sample: their rows of M_0 are empty, every other
relation is unchanged.  Its output line's smcm_time column (:883) is the timed region; the
intermediate products of its plan (printed to stdout, :713) are counted by the oracle on the same
graph.

    python oracle/ref_baseline.py [--threads T] [--frac F]   # calibration: reference vs oracle
"""
from __future__ import annotations

import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF = os.path.join(HERE, "_ref", "resmcm")
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)


def threads_available() -> int:
    """The CPU share: OMP_NUM_THREADS when set, else the affinity mask."""
    v = os.environ.get("OMP_NUM_THREADS")
    if v and v.isdigit() and int(v) > 0:
        return int(v)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def sampled_graph(hg, start_type: int, start_edge: int, rows):
    """hg without the entries of edge type start_edge leaving vertices of start_type outside rows
    (local ids): the reference's M_0 then has rows only for the sample."""
    from resmcm_ari_amd.hin import HeteroGraph
    keep = np.zeros(hg.n, bool)
    keep[hg.bins[start_type][rows]] = True
    src = np.repeat(np.arange(hg.n), np.diff(hg.adj_ptr))
    drop = (hg.adj_et == start_edge) & (hg.vertexTypes[src] == start_type) & ~keep[src]
    kept = ~drop
    ptr = np.zeros(hg.n + 1, np.int64)
    ptr[1:] = np.cumsum(np.bincount(src[kept], minlength=hg.n))
    return HeteroGraph.from_arrays(hg.vertexTypes, ptr, hg.adj_nbr[kept], hg.adj_et[kept])


def run_reference(hg, metapaths, threads: int, workdir: str | None = None):
    """Runs the reference program on hg for the meta-path strings; returns per path
    (nnz, total s, spgemm s, dp s, est s, plan pairs)."""
    if not os.path.exists(REF):
        raise RuntimeError(f"{REF} is missing: build it with make -C oracle ref")
    with tempfile.TemporaryDirectory(dir=workdir) as d:
        files = [os.path.join(d, x) for x in ("graph.txt", "vertex.txt", "edge.txt")]
        hg.write_text(*files)
        mpf, out = os.path.join(d, "mp.txt"), os.path.join(d, "out.txt")
        with open(mpf, "w") as f:
            f.write("".join(p + "\n" for p in metapaths))
        env = dict(os.environ, OMP_PROC_BIND="spread", OMP_NUM_THREADS=str(threads))
        p = subprocess.run([REF, str(threads), "1", *files, mpf, out], capture_output=True, text=True, env=env)
        if p.returncode != 0:
            raise RuntimeError(f"reference program failed ({p.returncode}): {p.stderr[-500:]}")
        # the chain order printed per meta-path (efficient/resmcm.cpp:713), one "(k, l)" per line
        plans, cur = [], []
        for line in p.stdout.splitlines():
            m = re.fullmatch(r"\((-?\d+), (-?\d+)\)", line.strip())
            if m:
                cur.append((int(m.group(1)), int(m.group(2))))
            elif cur:
                plans.append(cur)
                cur = []
        if cur:
            plans.append(cur)
        rows = [line.split() for line in open(out) if line.strip()]
    res = []
    for i, r in enumerate(rows):
        res.append((int(r[1]), float(r[2]), float(r[3]), float(r[4]), float(r[5]), plans[i] if i < len(plans) else []))
    return res


def calibrate(threads: int, frac: float) -> None:
    """Reference program vs the oracle (the CPU restatement with uint64 counts and sorted rows) at
    equal thread counts, on the config-1 graph and on a sample of the config-2 authors."""
    import oracle as O
    from resmcm_ari_amd import configs
    O.set_threads(threads)
    print(f"threads {threads}, CPU {cpu_model()}")
    cases = [("config 1", configs.CFG1.build(), None), ("config 2", configs.CFG2.build(), frac)]
    for name, hg, f in cases:
        if f is not None:
            rng = np.random.default_rng(1234)
            nA = int(hg.dimensions[configs.A])
            rows = np.sort(rng.choice(nA, size=int(nA * f), replace=False))
            hg = sampled_graph(hg, configs.A, configs.BIB_EDGE[(configs.A, configs.P)], rows)
        for letters in ("A-P-V-P-A", "A-P-T-P-A"):
            mp = configs.metapath_from_letters(letters, "bib")
            (nnz, tot, spg, dp, est, plan), = run_reference(hg, [mp.toString()], threads)
            mats = O.chain_matrices(hg, mp)
            t0 = time.perf_counter()
            c, _ = O.estimate(mats, O.coldims(hg, mp))
            _, _, _, pairs = O.dp(c)
            dig, onnz, prods = O.plan_digest(mats, pairs, 1 << 14)
            t_orc = time.perf_counter() - t0
            same_plan = [tuple(x) for x in pairs.tolist()] == plan
            print(f"{name} {letters}: reference {tot:.3f} s (spgemm {spg:.3f}, est {est:.3f}), nnz {nnz}; "
                  f"oracle {t_orc:.3f} s, nnz {onnz}; products {prods:.3e}; same plan {same_plan}; "
                  f"reference/oracle time {tot / t_orc:.2f}")


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=threads_available())
    ap.add_argument("--frac", type=float, default=0.02)
    a = ap.parse_args()
    calibrate(a.threads, a.frac)
