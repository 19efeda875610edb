"""Short text summaries of rocprofv3 CSV output for the kernels of this library.

    python -m resmcm_ari_amd.rocprof_summary --trace DIR [--fetch DIR --write DIR] [--bench FILE]
        [--kernel NAME] [--command TEXT ...] [--title TEXT]

--trace   directory of `rocprofv3 --kernel-trace --stats --output-format csv` (its *_kernel_stats.csv)
--fetch   directory of `rocprofv3 --pmc FETCH_SIZE --output-format csv` (its *_counter_collection.csv)
--write   the same for WRITE_SIZE
--bench   a file holding the JSON line bench.py printed in the traced run (HIP-event durations and the
          algorithmic bytes per launch of each timed kernel)
--kernel  the timed kernel name the traffic line is written for (default: bench's roofline kernel)
--sq      directory of `rocprofv3 --pmc SQ_WAVES SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_INSTS_LDS
          SQ_LDS_BANK_CONFLICT SQ_LDS_IDX_ACTIVE SQ_INSTS_VMEM_RD SQ_INSTS_VALU GRBM_GUI_ACTIVE`: LDS
          bank conflicts, occupancy and instruction mix per kernel

Counter reductions (as rocprofv3 reports one value per dispatch): GRBM_GUI_ACTIVE is summed over
the 8 XCDs, so a dispatch's cycles are GRBM_GUI_ACTIVE / 8; SQ_WAVE_CYCLES counts resident waves in
units of 4 cycles (rocprofv3's OccupancyPercent = 400 * SQ_WAVE_CYCLES / cycles / CUs / 32); bank
conflicts per LDS access cycle are SQ_LDS_BANK_CONFLICT / (SQ_LDS_IDX_ACTIVE - SQ_LDS_BANK_CONFLICT)
(rocprofv3's LdsBankConflict).

Traffic calibration: the copy kernel (`copy_kernel`, 16-byte loads and stores, launched by
`smcm_copy_bandwidth` with a known byte count) runs in the same process; the ratio of its known
bytes to what FETCH_SIZE / WRITE_SIZE report for it is the scale applied to every other kernel.
The last line of the output, `traffic_bytes_per_launch: <kernel> <bytes>`, is what bench.py reads.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import re
import sys
from collections import defaultdict

# timed-kernel name -> pattern on the HIP kernel name (template arguments included)
TIMED = {
    "spgemm.numeric.window": r"window_(tier|major)_kernel<",
    "spgemm.numeric.bitmap": r"bitmap_tier_kernel<true",
    "spgemm.numeric.bitmap.hbm": None,  # the same kernel as spgemm.numeric.bitmap (values in HBM)
    "spgemm.symbolic.bitmap": r"bitmap_tier_kernel<false",
    "spgemm.symbolic.window": r"window_symbolic_kernel",
    "spgemm.numeric.wavebitmap": r"wave_bitmap_kernel<true",
    "spgemm.symbolic.wavebitmap": r"wave_bitmap_kernel<false",
    "batch.executor": r"batch_kernel<",
    "spgemm.numeric.sort": r"sort_tier_kernel<true",
    "spgemm.symbolic.sort": r"sort_tier_kernel<false",
    "spgemm.numeric.wave": r"wave_tier_kernel<true",
    "spgemm.symbolic.wave": r"wave_tier_kernel<false",
    "spgemm.analysis": r"^entry_len_kernel$",
    "spgemm.dense.accumulate": r"^dense_(scatter|count)_kernel",
    "spgemm.dense.compact": r"^dense_compact_kernel$",
    "spgemm.symbolic.bbits": r"^bbits_(kernel|symbolic_kernel|heavy_kernel)$",
    "spgemm.numeric.copy": r"^copy_(rows|tier)_kernel<",
    "spgemm.symbolic.copy": r"^copy_width_kernel$",
    "estimator.pass": r"est_(light|heavy)_kernel<",
    "copy": r"^copy_kernel$",
}


def short_name(full: str) -> str:
    s = full.strip()
    if s.startswith("void "):
        s = s[5:]
    s = s.replace("smcm::(anonymous namespace)::", "").replace("smcm::", "")
    depth, out = 0, []
    for ch in s:  # drop the parameter list, keep template arguments
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            break
        out.append(ch)
    return "".join(out)


def one_csv(d: str, suffix: str) -> str:
    files = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not files:
        raise SystemExit(f"no *{suffix} under {d}")
    return files[0]


def kernel_stats(d: str):
    rows = []
    with open(one_csv(d, "_kernel_stats.csv")) as f:
        for r in csv.DictReader(f):
            rows.append((short_name(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"]),
                         float(r["Percentage"])))
    return rows


def counter_per_kernel(d: str, counter: str):
    """kernel short name -> (dispatches, summed counter value)"""
    acc = defaultdict(lambda: [0, 0.0])
    with open(one_csv(d, "_counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            a = acc[short_name(r["Kernel_Name"])]
            a[0] += 1
            a[1] += float(r["Counter_Value"])
    return acc


SQ_COUNTERS = ["SQ_WAVES", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES", "SQ_INSTS_LDS", "SQ_LDS_BANK_CONFLICT",
               "SQ_LDS_IDX_ACTIVE", "SQ_INSTS_VMEM_RD", "SQ_INSTS_VALU", "GRBM_GUI_ACTIVE"]
XCDS, CUS, MAX_WAVES_PER_CU = 8, 256, 32


def sq_table(d: str, top: int):
    """kernel -> per-dispatch averages and derived LDS / occupancy figures, heaviest first."""
    acc = defaultdict(lambda: defaultdict(float))
    disp = defaultdict(set)
    with open(one_csv(d, "_counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            k = short_name(r["Kernel_Name"])
            acc[k][r["Counter_Name"]] += float(r["Counter_Value"])
            disp[k].add(r["Dispatch_Id"])
    rows = []
    for k, c in acc.items():
        n = len(disp[k])
        cycles = c["GRBM_GUI_ACTIVE"] / XCDS
        if cycles <= 0 or c["SQ_WAVES"] <= 0:
            continue
        waves_per_cu = 4.0 * c["SQ_WAVE_CYCLES"] / cycles / CUS
        lds_active = c["SQ_LDS_IDX_ACTIVE"] - c["SQ_LDS_BANK_CONFLICT"]
        rows.append({"kernel": k, "dispatches": n, "cycles": cycles / n, "waves": c["SQ_WAVES"] / n,
                     "occupancy_pct": 100.0 * waves_per_cu / MAX_WAVES_PER_CU, "waves_per_cu": waves_per_cu,
                     "lds_per_wave": c["SQ_INSTS_LDS"] / c["SQ_WAVES"],
                     "vmem_rd_per_wave": c["SQ_INSTS_VMEM_RD"] / c["SQ_WAVES"],
                     "valu_per_wave": c["SQ_INSTS_VALU"] / c["SQ_WAVES"],
                     "conflicts_per_access": c["SQ_LDS_BANK_CONFLICT"] / lds_active if lds_active > 0 else 0.0,
                     "conflict_stall_pct": 100.0 * c["SQ_LDS_BANK_CONFLICT"] / cycles / CUS,
                     "wave_cycles": c["SQ_WAVE_CYCLES"]})
    rows.sort(key=lambda r: -r["wave_cycles"])
    return rows[:top]


def match(pattern: str, names):
    rx = re.compile(pattern)
    return [n for n in names if rx.search(n)]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--bench")
    ap.add_argument("--sq")
    ap.add_argument("--kernel")
    ap.add_argument("--copy-bytes", type=float, default=float(1 << 32),
                    help="bytes read (and written) by one copy_kernel launch")
    ap.add_argument("--command", action="append", default=[])
    ap.add_argument("--title", default="rocprofv3 summary")
    ap.add_argument("--top", type=int, default=12)
    a = ap.parse_args(argv)

    bench = None
    if a.bench:
        with open(a.bench) as f:
            for line in f:
                line = line.strip()
                if line.startswith("{"):
                    d = json.loads(line)
                    # the contract line, or one side measurement (bench.py --only X)
                    bench = d if "metric" in d else next(iter(d.values()))
    kernel = a.kernel or (bench["roofline"]["kernel"] if bench and "roofline" in bench else "spgemm.numeric.window")

    out = [f"# {a.title}", ""]
    if a.command:
        out += ["Commands (from the repository root, one MI355X):", "", "```"] + a.command + ["```", ""]

    ks = kernel_stats(a.trace)
    total = sum(r[2] for r in ks)
    out += [f"## Kernel trace (`--kernel-trace --stats`), top {a.top} of {len(ks)} kernels by total time", "",
            "| kernel | calls | total ms | avg ms | % |", "|---|---:|---:|---:|---:|"]
    for n, c, t, avg, pct in sorted(ks, key=lambda r: -r[2])[: a.top]:
        out.append(f"| `{n}` | {c} | {t / 1e6:.3f} | {avg / 1e6:.3f} | {pct:.2f} |")
    out += ["", f"All kernels: {total / 1e6:.3f} ms.", ""]

    # a timed launch may span several dispatches (the window tier's second, 64-bit launch for rows
    # whose counts may reach 2^32): with the contract line's steps and warm-up, the trace's timed
    # launches are counted and its figures are per timed launch, else per dispatch
    per_launch = {}  # timed kernel -> dispatches per timed launch in the traced run
    if bench:
        steps, warm = bench.get("steps"), bench.get("warmup")
        out += ["## HIP-event timing in the same run (bench.py `kernels`) vs the trace", "",
                "| timed kernel | launches | HIP-event avg ms | trace avg ms | dispatches per launch | algorithmic GB/launch |",
                "|---|---:|---:|---:|---:|---:|"]
        for tname, v in bench.get("kernels", {}).items():
            pat = TIMED.get(tname)
            names = match(pat, [r[0] for r in ks]) if pat else []
            calls = sum(r[1] for r in ks if r[0] in names)
            traced = v["launches"] * (steps + warm) / steps if steps and warm is not None and v["launches"] else 0
            per_launch[tname] = calls / traced if traced else 1.0
            div = traced if traced else calls
            tr = sum(r[2] for r in ks if r[0] in names) / div / 1e6 if div else float("nan")
            ev = v["ms"] / v["launches"] if v["launches"] else float("nan")  # per launch
            gb = v.get("bytes_per_launch", 0) / 1e9
            out.append(f"| {tname} | {v['launches']} | {ev:.3f} | {tr:.3f} | {per_launch[tname]:.2f} | {gb:.3f} |")
        out += ["", "The trace also covers the warm-up step; estimator.pass spans several kernels (light and heavy "
                "rows, partial and final sums), so its trace column is a per-kernel average.", ""]

    if a.sq:
        out += ["## LDS, occupancy and instruction mix (`--pmc` " + " ".join(SQ_COUNTERS) + ")", "",
                "Per dispatch: cycles = GRBM_GUI_ACTIVE / 8; occupancy = resident waves per CU "
                "(4 * SQ_WAVE_CYCLES / cycles / 256) as a share of 32; bank conflicts = conflict cycles per "
                "conflict-free LDS cycle; stall = conflict cycles as a share of all CU cycles.", "",
                "| kernel | dispatches | Mcycles/dispatch | waves/dispatch | occupancy % (waves/CU) | LDS instr/wave | "
                "VMEM rd/wave | VALU/wave | bank conflicts/access | conflict stall % |",
                "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
        for r in sq_table(a.sq, a.top):
            out.append(f"| `{r['kernel']}` | {r['dispatches']} | {r['cycles'] / 1e6:.3f} | {r['waves']:.0f} | "
                       f"{r['occupancy_pct']:.1f} ({r['waves_per_cu']:.1f}) | {r['lds_per_wave']:.0f} | "
                       f"{r['vmem_rd_per_wave']:.0f} | {r['valu_per_wave']:.0f} | {r['conflicts_per_access']:.3f} | "
                       f"{r['conflict_stall_pct']:.2f} |")
        out.append("")

    traffic_line = None
    if a.fetch and a.write:
        fe = counter_per_kernel(a.fetch, "FETCH_SIZE")
        wr = counter_per_kernel(a.write, "WRITE_SIZE")
        cname = match(TIMED["copy"], fe.keys())
        if not cname:
            raise SystemExit("the copy kernel is missing from the FETCH_SIZE run: no calibration")
        cf, cw = fe[cname[0]], wr[cname[0]]
        f_scale = a.copy_bytes / (cf[1] / cf[0] * 1024.0)
        w_scale = a.copy_bytes / (cw[1] / cw[0] * 1024.0)
        out += ["## HBM traffic (`--pmc FETCH_SIZE`, `--pmc WRITE_SIZE`, KiB per dispatch)", "",
                f"Calibration with `copy_kernel` ({a.copy_bytes / 2**30:.0f} GiB read + written per launch): "
                f"FETCH_SIZE reads {cf[1] / cf[0] / 2**20:.3f} GiB -> scale {f_scale:.3f}; "
                f"WRITE_SIZE reads {cw[1] / cw[0] / 2**20:.3f} GiB -> scale {w_scale:.3f}.", "",
                "| kernel | dispatches | FETCH GB/dispatch (scaled) | WRITE GB/dispatch (scaled) |",
                "|---|---:|---:|---:|"]
        for n in sorted(fe, key=lambda k: -fe[k][1])[: a.top]:
            d, fv = fe[n]
            wv = wr.get(n, [d, 0.0])[1]
            out.append(f"| `{n}` | {d} | {fv / d * 1024 * f_scale / 1e9:.3f} | {wv / d * 1024 * w_scale / 1e9:.3f} |")
        out.append("")
        pat = TIMED.get(kernel)
        names = match(pat, fe.keys()) if pat else []
        if names:
            d = sum(fe[n][0] for n in names) / per_launch.get(kernel, 1.0)  # timed launches
            fb = sum(fe[n][1] for n in names) * 1024 * f_scale
            wb = sum(wr[n][1] for n in names if n in wr) * 1024 * w_scale
            per = (fb + wb) / d
            alg = None
            if bench and kernel in bench.get("kernels", {}):
                v = bench["kernels"][kernel]
                alg = v.get("bytes_per_launch") or None
            out += [f"{kernel} ({', '.join('`' + n + '`' for n in names)}): {d:.0f} launches, "
                    f"{fb / d / 1e9:.3f} GB fetched + {wb / d / 1e9:.3f} GB written = {per / 1e9:.3f} GB per launch"
                    + (f"; algorithmic {alg / 1e9:.3f} GB per launch (ratio {per / alg:.2f})." if alg else "."), ""]
            traffic_line = f"traffic_bytes_per_launch: {kernel} {per:.0f}"
    if traffic_line:
        out.append(traffic_line)
    print("\n".join(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
