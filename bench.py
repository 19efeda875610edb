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
dominant kernel over the timed region; cpu_baseline = the reference's own resmcm program (built
from efficient/resmcm.cpp) on a fixed sample of source authors of the same chains, every thread of
the host's share.

Beside it, on the same GPU in the same run:
  config1  the 20k/30k/200 graph's A-P-V-P-A chain: SpGEMM kernels' algorithmic bytes / time;
  config3  the long location meta-paths over all 100k users (results of 1e10 entries streamed in
           row blocks), reference plan and product plan;
  vendor   rocSPARSE's SpGEMM (binary64 values) on the final products of A-P-V-P-A, configs 1 and 2;
  batched  4096 chains (L 2-6, log-uniform sources 16..65536) over that graph with T=20k:
           smcm_batch_run vs the per-chain path, queries/s, every result's digests compared;
  served   a Poisson stream over that graph and the 100k-user location graph (L 2-10, sources
           log-uniform 1..100k, 70% sharing a prefix): the server vs one-at-a-time execution,
           queries/s and p50/p99 latency, every result's digests compared.
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
    """The reference's own program (efficient/resmcm.cpp built from source into oracle/_ref) on a
    sample of the source authors of both chains, on every thread of the host's share; its products
    are counted by the oracle for the plan the program printed.  Without that build, the oracle's
    restatement is timed instead (kind "port")."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # test infrastructure: the checker and the CPU baseline only
    import ref_baseline as RB

    threads = RB.threads_available()
    rng = np.random.default_rng(seed)
    nA = int(hg.dimensions[configs.A])
    rows = np.sort(rng.choice(nA, size=max(1, int(nA * sample_frac)), replace=False)).astype(np.int32)
    sample = f"{len(rows)} of {nA} source authors ({sample_frac:.2%}, seed {seed}) for {' and '.join(CHAINS)}"
    if os.path.exists(RB.REF):
        sub = RB.sampled_graph(hg, configs.A, configs.BIB_EDGE[(configs.A, configs.P)], rows)
        mps = [configs.metapath_from_letters(p, "bib") for p in CHAINS]
        res = RB.run_reference(sub, [mp.toString() for mp in mps], threads)
        O.set_threads(threads)
        total_products, total_s = 0, 0.0
        for mp, (nnz, tot, spg, dp, est, plan) in zip(mps, res):
            total_products += O.plan_products(O.chain_matrices(sub, mp), np.asarray(plan, np.int32))
            total_s += tot
        return {"value": total_products / total_s, "unit": "products/s", "cores": threads, "kind": "reference",
                "sample": f"{sample}; the reference resmcm program (efficient/resmcm.cpp, built from source) on the "
                          f"graph without the other authors' A->P edges; {total_products:.3e} products of its plans "
                          f"in {total_s:.2f} s (its smcm_time)",
                "cpu_model": RB.cpu_model(), "omp_proc_bind": "spread", "omp_num_threads": threads}
    O.set_threads(threads)
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
    return {"value": total_products / total_s, "unit": "products/s", "cores": threads, "kind": "port",
            "sample": f"{sample}; the oracle's restatement; {total_products:.3e} products in {total_s:.2f} s",
            "cpu_model": RB.cpu_model(), "omp_num_threads": threads}


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


def bench_config1(hg1, reps: int = 20) -> dict:
    """A-P-V-P-A over all 20k authors of the workload-1 graph: SpGEMM kernels' algorithmic bytes
    per second of kernel time (HIP events)."""
    ctx = smcm.Context(0)
    G = smcm.Graph(ctx, hg1)
    mp = configs.metapath_from_letters("A-P-V-P-A", "bib")
    for _ in range(3):
        r, _ = G.run(mp)
        r.free()
    # whole chains without kernel timing (the timer synchronises after every timed kernel group)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        r, _ = G.run(mp)
        r.free()
    ctx.sync()
    dt = time.perf_counter() - t0
    ctx.set_option("kernel_timing", 1)
    ctx.kernel_stats_reset()
    prods = 0
    for _ in range(reps):
        r, _ = G.run(mp)
        prods += ctx.last_stats()["products"]
        r.free()
    ctx.sync()
    ks = ctx.kernel_stats()
    ms = sum(v[0] for k, v in ks.items() if k.startswith("spgemm."))
    byts = sum(v[2] for k, v in ks.items() if k.startswith("spgemm."))
    ctx.set_option("kernel_timing", 0)
    copy_gbps = ctx.copy_bandwidth(1 << 32, 2)  # also the counter calibration of a profiled run
    G.close()
    ctx.close()
    gbps = byts / (ms * 1e-3) / 1e9 if ms else 0.0
    return {"workload": "A-P-V-P-A over A=20k P=30k V=200 (seed 0x5EED0001), all rows",
            "chain_ms": round(dt / reps * 1e3, 3), "spgemm_kernel_ms": round(ms / reps, 3),
            "products_per_chain": prods // reps, "spgemm_GBps": round(gbps, 1),
            "spgemm_frac_of_peak": round(gbps / PEAK_HBM_GBPS, 4), "measured_copy_GBps": round(copy_gbps, 1),
            "kernels": {k: {"ms": round(v[0] / reps, 4), "launches": v[1] // reps,
                            "GBps": round(v[2] / v[0] / 1e6, 1) if v[0] else 0.0,
                            "bytes_per_launch": round(v[2] / v[1]) if v[1] else 0}
                        for k, v in sorted(ks.items(), key=lambda kv: -kv[1][0])}}


def bench_config3(block_gb: float = 32.0) -> dict:
    """Workload 3: the long location meta-paths over all 100k users, roots streamed in row blocks;
    the reference plan (drop-in) and the product plan, same results."""
    hg3 = configs.CFG3.build()
    ctx = smcm.Context(0)
    G = smcm.Graph(ctx, hg3)
    out = {"workload": "U=100k L=200k C=500 Y=1k S=168 (seed 0x5EED0003), all users, roots in "
                       f"{block_gb:g} GB row blocks", "chains": []}
    tot = {0: [0.0, 0], 1: [0.0, 0]}
    for p in configs.CFG3_PATHS:
        mp = configs.metapath_from_letters(p, "loc")
        row = {"path": p}
        digs = []
        for pc in (0, 1):
            ctx.set_option("plan_cost", pc)
            t0 = time.perf_counter()
            n, d, _ = G.run_blocked(mp, int(block_gb * (1 << 30)), digest=True)
            dt = time.perf_counter() - t0
            st = ctx.last_stats()
            digs.append(d)
            key = "reference_plan" if pc == 0 else "product_plan"
            row[key] = {"seconds": round(dt, 3), "products": st["products"], "nnz": n}
            tot[pc][0] += dt
            tot[pc][1] += st["products"]
            print(f"[bench] config 3 {p} plan_cost={pc}: {dt:.2f} s", file=sys.stderr, flush=True)
        row["same_result"] = bool(np.array_equal(digs[0], digs[1]))
        out["chains"].append(row)
    ctx.set_option("plan_cost", 0)
    out["reference_plan_products_per_s"] = round(tot[0][1] / tot[0][0], 1)
    out["product_plan_seconds"] = round(tot[1][0], 3)
    out["reference_plan_seconds"] = round(tot[0][0], 3)
    G.close()
    ctx.close()
    return out


def bench_vendor(hg1, hg2=None) -> dict:
    """The final products of A-P-V-P-A (config 1, and config 2 when given): this library's SpGEMM
    kernels (HIP-event time, uint64 counts) against rocSPARSE's generic SpGEMM (binary64 values),
    same operands; the structures of both results are compared."""
    from resmcm_ari_amd import vendor

    out = {"note": "rocSPARSE times its nnz and compute stages; ours the analysis, symbolic and numeric kernels "
                   "(allocations excluded in both)", "cases": []}
    ctx = smcm.Context(0)
    for name, hg, reps in (("config1", hg1, 5), ("config2", hg2, 2)):
        if hg is None:
            continue
        G = smcm.Graph(ctx, hg)
        # operands: X = A-P-V (left) and Y = V-P-A (right), as the DP's plan builds them
        X, _ = G.run(configs.metapath_from_letters("A-P-V", "bib"))
        Y, _ = G.run(configs.metapath_from_letters("V-P-A", "bib"))
        xp, xc, xv = X.to_host()
        yp, yc, yv = Y.to_host()
        off = hg.type_offset
        xc = (xc - off[configs.V]).astype(np.int32)  # type-local columns
        yc = (yc - off[configs.A]).astype(np.int32)
        X.free()
        Y.free()
        m, k, n = len(xp) - 1, len(yp) - 1, int(hg.dimensions[configs.A])
        ctx.set_option("kernel_timing", 1)
        ms_ours = []
        ours = None
        for r in range(reps + 1):
            ctx.kernel_stats_reset()
            res = ctx.spgemm_csr(m, k, n, xp, xc, xv, yp, yc, yv)
            ks = ctx.kernel_stats()
            if r > 0:
                ms_ours.append(sum(v[0] for kk, v in ks.items() if kk.startswith("spgemm.")))
            if r == reps and name == "config1":
                ours = res.to_host()
            nnz_ours = res.nnz
            res.free()
        ctx.set_option("kernel_timing", 0)
        ms_v, nnz_v, vres = vendor.spgemm_f64(m, k, n, xp, xc, xv.astype(np.float64), yp, yc,
                                              yv.astype(np.float64), reps=reps, want_result=(name == "config1"))
        same = nnz_v == nnz_ours
        if vres is not None and ours is not None:
            vp, vc, vv = vres
            gp, gc, gv = ours
            # rocSPARSE rows may come unsorted: compare each row's column set
            same = same and np.array_equal(vp, gp)
            if same:
                order = np.lexsort((vc, np.repeat(np.arange(m), np.diff(vp))))
                same = np.array_equal(vc[order], gc.astype(np.int32))  # both type-local
        t = float(np.median(ms_ours))
        out["cases"].append({"case": f"{name} A-P-V * V-P-A ({m}x{k} * {k}x{n})", "nnz": nnz_ours,
                             "ours_ms": round(t, 3), "rocsparse_ms": round(ms_v, 3),
                             "speedup": round(ms_v / t, 2) if t else None, "same_structure": bool(same)})
        print(f"[bench] vendor {name}: ours {t:.3f} ms, rocSPARSE {ms_v:.3f} ms", file=sys.stderr, flush=True)
        G.close()
    ctx.close()
    return out


def bench_batched(hg1, n_chains: int = 4096, ref_sample: int = 1024) -> dict:
    """Workload 4: the batched form against the per-chain path on the same chains: with the
    product-aware plans the batch uses (all chains, every result's digests compared) and with the
    reference DP's plans (the drop-in; a sample of the first chains, it is slower)."""
    ctx = smcm.Context(0)
    G = smcm.Graph(ctx, hg1)
    batch = configs.cfg4_batch(hg1.dimensions, n_chains)
    off = hg1.type_offset
    paths = [q.path for q in batch]
    srcs = [(q.src_rows + off[q.path.vertexTypes[0]]).astype(np.int32) for q in batch]
    G.run_batch(paths[:32], srcs[:32])  # typed matrices, code objects
    tune = G.autotune_batch(paths[:256], srcs[:256])  # on a sample, before the timed run
    print(f"[bench] batched: autotuned in {tune['trials']} trials", file=sys.stderr, flush=True)
    # one untimed run of the whole batch: the tuner hands its trials' device blocks back, and the
    # timed run then measures the steady state of a serving process (blocks mapped and reused),
    # as the per-chain runs below do after their first chains
    G.run_batch(paths, srcs)
    print("[bench] batched: warm-up run done", file=sys.stderr, flush=True)
    ctx.sync()
    t0 = time.perf_counter()
    nnz, dig, _, st = G.run_batch(paths, srcs, digests=True)
    tb = time.perf_counter() - t0
    print(f"[bench] batched: {n_chains} chains in {tb:.2f} s", file=sys.stderr, flush=True)
    per_chain = {}
    bad = 0
    for pc, count in ((1, n_chains), (0, min(ref_sample, n_chains))):
        ctx.set_option("plan_cost", pc)
        ctx.sync()
        t0 = time.perf_counter()
        for i in range(count):
            r, _ = G.run(paths[i], srcs[i])
            d = r.row_digest()
            if r.nnz != nnz[i] or not np.array_equal(d, dig[i]):
                bad += 1
            r.free()
            if i % 1024 == 1023:
                print(f"[bench] per-chain path, plan_cost={pc}: {i + 1} chains, {time.perf_counter() - t0:.1f} s",
                      file=sys.stderr, flush=True)
        per_chain[pc] = (count, time.perf_counter() - t0)
    ctx.set_option("plan_cost", 0)
    G.close()
    ctx.close()
    (n1, t1), (n0, t0s) = per_chain[1], per_chain[0]
    qps, qps1, qps0 = n_chains / tb, n1 / t1, n0 / t0s
    return {"workload": f"{n_chains} chains, L 2-6 over A/P/V/T, sources log-uniform 16..65536 "
                        f"(seed 0x5EED0004), graph A=20k P=30k V=200 T=20k",
            "queries_per_s": round(qps, 2), "seconds": round(tb, 3),
            "product_plan_queries_per_s": round(qps1, 2), "product_plan_seconds": round(t1, 3),
            "speedup_vs_product_plan": round(qps / qps1, 2),
            "plain_queries_per_s": round(qps0, 2), "plain_sample_chains": n0, "plain_sample_seconds": round(t0s, 3),
            "speedup": round(qps / qps0, 2),
            "mismatching_chains": bad, "jobs": st["jobs"], "executor_jobs": st["small_jobs"],
            "tiered_jobs": st["large_jobs"], "alone_jobs": st["alone_jobs"], "oom_retries": st["oom_retries"], "plan_s": round(st["plan_s"], 3), "executor_s": round(st["exec_s"], 3),
            "tiered_s": round(st["large_s"], 3),
            "autotuned": {k: tune[k] for k in ("streams", "small_products", "threads", "sort_cap", "tile_rows",
                                               "trials")},
            "result_nnz": int(nnz.sum()),
            "note": "product_plan = the per-chain path with the batch's product-aware plans (plan_cost 1), all "
                    "chains; plain = the per-chain path with the reference DP's plans (the drop-in), on the "
                    f"first {n0} chains; every per-chain result's digests are compared with the batch's; the "
                    "batch is timed on its second run after tuning (the first, untimed, maps its device blocks)"}


def served_queries(hg1, hg3, n_queries: int, qps: float):
    hgs = [hg1, hg3]
    stream = configs.cfg5_stream(hg1.dimensions, hg3.dimensions, n_queries=n_queries, qps=qps)
    queries = [(q.graph, q.path, (q.src_rows + hgs[q.graph].type_offset[q.path.vertexTypes[0]]).astype(np.int32))
               for q in stream]
    return queries, np.array([q.arrival for q in stream])


def latencies(outs, arr):
    lat = np.array([o["t_done"] - a for o, a in zip(outs, arr)])
    span = max(o["t_done"] for o in outs) - float(np.min(arr))
    return lat, span


def bench_served(hg1, n_queries: int = 128, qps: float = 5.0) -> dict:
    """Workload 5: the server against one-at-a-time execution on the same Poisson stream, and the
    server's capacity (every query submitted at once)."""
    hg3 = configs.CFG3.build()
    ctx = smcm.Context(0)
    G = [smcm.Graph(ctx, hg1), smcm.Graph(ctx, hg3)]
    queries, arr = served_queries(hg1, hg3, n_queries, qps)
    res = {}
    for name, kw, at in (("served", dict(plan_cost=1), arr),
                         ("plain", dict(slots=1, plan_cost=0, cache=0), arr),
                         ("saturated", dict(plan_cost=1), np.zeros(n_queries))):
        srv = smcm.Server(ctx, G, **kw)
        outs, digs = srv.replay(queries, at)
        stats = srv.stats()
        srv.close()
        lat, span = latencies(outs, at)
        res[name] = (outs, digs, lat, span, stats)
        print(f"[bench] {name}: {n_queries} queries, span {span:.1f} s", file=sys.stderr, flush=True)
    (o1, d1, l1, s1, st1), (o0, d0, l0, s0, _), (o2, d2, l2, s2, _) = res["served"], res["plain"], res["saturated"]
    bad = sum(1 for i in range(n_queries)
              if o1[i]["status"] or o0[i]["status"] or o2[i]["status"] or o1[i]["nnz"] != o0[i]["nnz"]
              or not np.array_equal(d1[i], d0[i]) or not np.array_equal(d2[i], d0[i]))
    for G_ in G:
        G_.close()
    ctx.close()
    return {"workload": f"{n_queries} queries, Poisson {qps} /s, L 2-10, sources log-uniform 1..100k, 70% sharing "
                        "a prefix (seed 0x5EED0005); graphs A=20k P=30k V=200 T=20k and U=100k L=200k C=500 Y=1k "
                        "S=168 resident",
            "queries_per_s": round(n_queries / s1, 3), "p50_ms": round(float(np.percentile(l1, 50)) * 1e3, 1),
            "p99_ms": round(float(np.percentile(l1, 99)) * 1e3, 1),
            "plain_queries_per_s": round(n_queries / s0, 3),
            "plain_p50_ms": round(float(np.percentile(l0, 50)) * 1e3, 1),
            "plain_p99_ms": round(float(np.percentile(l0, 99)) * 1e3, 1),
            "capacity_queries_per_s": round(n_queries / s2, 3),
            "capacity_p50_ms": round(float(np.percentile(l2, 50)) * 1e3, 1),
            "mismatching_queries": bad, "cache_hits": st1["cache_hits"],
            "server": "6 slots (2 for large queries, their persistent kernels sized for 50% of the CUs), products "
                      "plan, 200 GiB admission budget, 16 GiB prefix cache",
            "plain": "one query at a time, reference DP plan, no cache",
            "capacity": "the same queries submitted at once to the same server configuration"}


def bench_served_sweep(hg1, n_queries: int = 512, rates=(5.0, 10.0, 20.0)) -> dict:
    """Serving capacity: n_queries of workload 5 at several arrival rates and all at once, each
    query's latency in the mix against its own time when queries run one at a time through the
    same server configuration (one slot; consecutive completions), results compared."""
    hg3 = configs.CFG3.build()
    ctx = smcm.Context(0)
    G = [smcm.Graph(ctx, hg1), smcm.Graph(ctx, hg3)]
    queries, _ = served_queries(hg1, hg3, n_queries, rates[0])
    srv = smcm.Server(ctx, G, slots=1, large_slots=0, plan_cost=1)
    alone, dig0 = [], []
    for c in range(0, n_queries, 64):  # chunks of one server (its cache kept), progress per chunk
        outs, digs = srv.replay(queries[c:c + 64], np.zeros(len(queries[c:c + 64])))
        done = np.array([o["t_done"] for o in outs])
        alone.extend(np.diff(np.concatenate([[0.0], done])))  # FIFO on one slot: completion gaps
        dig0.extend(digs)
        print(f"[bench] served sweep: one at a time, {c + len(outs)} queries", file=sys.stderr, flush=True)
    srv.close()
    alone = np.array(alone)
    out = {"workload": f"{n_queries} queries of workload 5 (seed 0x5EED0005), same queries at every rate",
           "alone_seconds": round(float(alone.sum()), 3), "runs": []}
    for rate in list(rates) + [None]:
        if rate is None:
            q2, arr = queries, np.zeros(n_queries)
        else:
            q2, arr = served_queries(hg1, hg3, n_queries, rate)
            assert all(a[1].vertexTypes == b[1].vertexTypes and np.array_equal(a[2], b[2]) for a, b in zip(q2, queries))
        srv = smcm.Server(ctx, G, plan_cost=1)
        o, d = srv.replay(q2, arr)
        hits = srv.stats()["cache_hits"]
        srv.close()
        lat, span = latencies(o, arr)
        bad = sum(1 for i in range(n_queries) if o[i]["status"] or not np.array_equal(d[i], dig0[i]))
        slow = lat / np.maximum(alone, 1e-4)
        worst = np.argsort(-slow)[:5]
        out["runs"].append({
            "offered_qps": rate if rate is not None else "all at once", "queries_per_s": round(n_queries / span, 3),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 1), "p90_ms": round(float(np.percentile(lat, 90)) * 1e3, 1),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 1), "max_ms": round(float(lat.max()) * 1e3, 1),
            "slowdown_p50": round(float(np.percentile(slow, 50)), 2), "slowdown_p90": round(float(np.percentile(slow, 90)), 2),
            "slowdown_max": round(float(slow.max()), 2), "cache_hits": hits, "mismatching_queries": bad,
            "worst": [{"query": int(i), "path": queries[i][1].toString(), "graph": int(queries[i][0]),
                       "sources": int(len(queries[i][2])), "latency_ms": round(float(lat[i]) * 1e3, 1),
                       "alone_ms": round(float(alone[i]) * 1e3, 1)} for i in worst]})
        print(f"[bench] served sweep: rate {rate}: span {span:.1f} s", file=sys.stderr, flush=True)
    for G_ in G:
        G_.close()
    ctx.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--block-gb", type=float, default=96.0)
    ap.add_argument("--cpu-sample", type=float, default=0.06,
                    help="fraction of source rows for the CPU baseline (about 10 s of CPU work)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--chains", default=",".join(CHAINS))
    ap.add_argument("--no-side", action="store_true", help="skip the config-1, batched and served measurements")
    ap.add_argument("--batch-chains", type=int, default=4096)
    ap.add_argument("--serve-queries", type=int, default=128)
    ap.add_argument("--serve-qps", type=float, default=5.0)
    ap.add_argument("--only", choices=["config1", "config3", "batched", "served", "served_sweep", "vendor"],
                    help="run one side measurement alone and print its JSON (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the library runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by source rows: each GPU runs the chain on a slice of start vertices)", file=sys.stderr)
        return 2
    if args.only:
        hg1 = configs.CFG1.build() if args.only != "config3" else None
        res = {"config1": lambda: bench_config1(hg1), "config3": lambda: bench_config3(),
               "batched": lambda: bench_batched(hg1, args.batch_chains),
               "vendor": lambda: bench_vendor(hg1, configs.CFG2.build()),
               "served": lambda: bench_served(hg1, args.serve_queries, args.serve_qps),
               "served_sweep": lambda: bench_served_sweep(hg1)}[args.only]()
        print(json.dumps({args.only: res}))
        return 0
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
    G.close()
    ctx.close()

    side = {}
    if not args.no_side:
        hg1 = configs.CFG1.build()
        side["config1"] = bench_config1(hg1)
        side["config3"] = bench_config3()
        side["vendor"] = bench_vendor(hg1, hg)
        side["batched"] = bench_batched(hg1, args.batch_chains)
        side["served"] = bench_served(hg1, args.serve_queries, args.serve_qps)

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
        **side,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
