#!/usr/bin/env python3
"""Benchmark: sparse matrix chain products on one MI355X.

Workload (the largest single-GPU configuration): seeded bibliographic graph with A=500k authors,
P=800k papers, V=5k venues, T=20k terms (1+Poisson(3) authors per paper ~ Zipf(1.1), one venue
~ Zipf(0.8), 8 distinct terms ~ Zipf(1.0)); one step runs the meta-paths A-P-V-P-A and A-P-T-P-A
through the full path (typed matrices resident in HBM; estimator, chain-order DP and every
SpGEMM of the plan).  A-P-T-P-A's result (~4.5e10 entries, ~540 GB) exceeds HBM, so the root
product is materialised in row blocks of at most --block-gb GB; each block is complete in HBM
before its memory is reused (the reference's driver likewise keeps only the nnz of a result).

value = intermediate products per second (sum over the plan's SpGEMMs; the plan is the
reference DP's, which the tests check).  roofline = algorithmic bytes / HIP-event duration of the
dominant kernel over the timed region; cpu_baseline = the oracle (CPU restatement of the
reference path, OpenMP) on a fixed sample of source rows of the same chains.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from resmcm_ari_amd import configs, smcm  # noqa: E402

PEAK_HBM_GBPS = 8000.0
CHAINS = ["A-P-V-P-A", "A-P-T-P-A"]
WORKLOAD_TAG = "bib_A500k_P800k_V5k_T20k.APVPA_APTPA"


def profiled_traffic(kernel: str):
    """HBM bytes per launch of `kernel` on this workload, from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (profiles/<kernel>.<workload>.md, written by resmcm_ari_amd.rocprof_summary);
    None when there is no such summary."""
    path = os.path.join(ROOT, "profiles", f"{kernel}.{WORKLOAD_TAG}.md")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch:" and parts[1] == kernel:
                return float(parts[2]), os.path.relpath(path, ROOT)
    return None, None


def cpu_baseline(hg, sample_frac: float, seed: int = 1234) -> dict:
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # test infrastructure: the checker and the CPU baseline only

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    O.set_threads(cores)
    rng = np.random.default_rng(seed)
    nA = int(hg.dimensions[configs.A])
    rows = np.sort(rng.choice(nA, size=max(1, int(nA * sample_frac)), replace=False)).astype(np.int32)
    total_products, total_s = 0, 0.0
    for p in CHAINS:
        mp = configs.metapath_from_letters(p, "bib")
        mats = O.chain_matrices(hg, mp, rows)  # typed matrices are built before the timer, as in the reference
        t0 = time.perf_counter()
        c, _ = O.estimate(mats, O.coldims(hg, mp))
        _, _, _, pairs = O.dp(c)
        res, prods = run_plan(O, mats, pairs)
        total_s += time.perf_counter() - t0
        total_products += prods
        del res
    return {"value": total_products / total_s, "unit": "products/s", "cores": cores, "kind": "port",
            "sample": f"{len(rows)} of {nA} source authors ({sample_frac:.2%}, seed {seed}) for "
                      f"{' and '.join(CHAINS)}; {total_products:.3e} products in {total_s:.2f} s"}


def run_plan(O, mats, pairs):
    """The reference's operand stack (efficient/resmcm.cpp:709-760) over the oracle's SpGEMM;
    returns the product and its intermediate-product count."""
    stack, total = [], 0
    for k, l in pairs:
        A = mats[k] if k >= 0 else stack.pop(-2 if l < 0 else -1)
        B = mats[l] if l >= 0 else stack.pop()
        total += O.products(A, B)
        stack.append(O.spgemm(A, B))
    return stack[0], total


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--block-gb", type=float, default=96.0)
    ap.add_argument("--cpu-sample", type=float, default=0.02, help="fraction of source rows for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--chains", default=",".join(CHAINS))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by source rows, see DESIGN.md)", file=sys.stderr)
        return 2
    chains = args.chains.split(",")
    block_bytes = int(args.block_gb * (1 << 30))

    t_gen = time.perf_counter()
    hg = configs.CFG2.build()
    print(f"[bench] graph n={hg.n} m={hg.m} built in {time.perf_counter() - t_gen:.1f} s", file=sys.stderr, flush=True)
    ctx = smcm.Context(0)
    G = smcm.Graph(ctx, hg)
    mps = [configs.metapath_from_letters(p, "bib") for p in chains]

    def step():
        prods, nnz = 0, 0
        for mp in mps:
            n, _, _ = G.run_blocked(mp, block_bytes)
            st = ctx.last_stats()
            prods += st["products"]
            nnz += n
        return prods, nnz

    for w in range(args.warmup):
        step()
        print(f"[bench] warmup {w} done", file=sys.stderr, flush=True)
    ctx.set_option("kernel_timing", 1)
    ctx.kernel_stats_reset()
    ctx.sync()
    t0 = time.perf_counter()
    prods_step, nnz_step = 0, 0
    for k in range(args.steps):
        ts = time.perf_counter()
        prods_step, nnz_step = step()
        print(f"[bench] step {k} done in {time.perf_counter() - ts:.3f} s", file=sys.stderr, flush=True)
    ctx.sync()
    dt = time.perf_counter() - t0
    ks = ctx.kernel_stats()
    ctx.set_option("kernel_timing", 0)

    dom = max(ks.items(), key=lambda kv: kv[1][0])
    name, (ms, cnt, byts) = dom
    achieved = (byts / cnt) / (ms / cnt * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS,
                "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": None,
                "launches": cnt, "avg_launch_ms": round(ms / cnt, 3),
                "algorithmic_bytes_per_launch": round(byts / cnt)}
    kernels = {k: {"ms": round(v[0], 3), "launches": v[1], "GBps": round(v[2] / v[0] / 1e6, 1) if v[0] else 0.0,
                   "bytes_per_launch": round(v[2] / v[1]) if v[1] else 0}
               for k, v in sorted(ks.items(), key=lambda kv: -kv[1][0])}

    if chains == CHAINS:
        traffic, src = profiled_traffic(name)
        if traffic:
            roofline["traffic"] = round(traffic)
            roofline["traffic_source"] = src

    copy_gbps = ctx.copy_bandwidth(1 << 32, 10)
    roofline["measured_copy_GBps"] = round(copy_gbps, 1)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(hg, args.cpu_sample)

    value = prods_step * args.steps / dt
    out = {
        "metric": "chain products/s (bit-exact); served queries/s + p50/p99 vs plain GPU",
        "value": value,
        "unit": "intermediate products/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded generator)",
        "config": {"workload": "bibliographic graph A=500k P=800k V=5k T=20k (seed 0x5EED0002); "
                               "chains " + " + ".join(chains) + " per step, plain per-chain path",
                   "products_per_step": prods_step, "result_nnz_per_step": nnz_step,
                   "chains_per_s": len(chains) * args.steps / dt,
                   "block_gb": args.block_gb, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(out))
    G.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
