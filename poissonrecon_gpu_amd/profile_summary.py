"""Summaries of rocprofv3 output for the benchmark's dominant kernel.

Usage (after the rocprofv3 runs of ``bench.py``; see the command lines in the
committed summaries):

    python -m poissonrecon_gpu_amd.profile_summary traffic \\
        --fetch DIR/pmc_fetch/run_counter_collection.csv \\
        --write DIR/pmc_write/run_counter_collection.csv --workload superquadrics_2M_d10

    python -m poissonrecon_gpu_amd.profile_summary trace --stats DIR/prof/bench_kernel_stats.csv \\
        --trace DIR/prof/bench_kernel_trace.csv

``traffic`` turns FETCH_SIZE / WRITE_SIZE (KiB per dispatch) into HBM bytes per
launch of the finest-depth CG SpMV.  Each counter is scaled by what it reads for
the 512 MiB device-to-device copy that ``bench.py --calibrate-copy`` issues at
the end (a copy kernel of known size: 512 MiB read and 512 MiB written).  The
finest-depth SpMV dispatches are the ``k_cg_spmv`` launches after the last
``k_cg_init`` (depths are solved coarse to fine), less the launches that found
the solve finished (their counters read near zero).
"""
from __future__ import annotations

import argparse
import csv
import json
import statistics
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
COPY_KIB = 512 * 1024


def _rows(path):
    with open(path) as f:
        rows = list(csv.DictReader(f))
    rows.sort(key=lambda r: int(r.get("Dispatch_Id", 0)))
    return rows


def finest_spmv_kib(path):
    """(per-launch KiB of the working finest-depth SpMV launches, calibration copy KiB)."""
    rows = _rows(path)
    inits = [i for i, r in enumerate(rows) if "k_cg_init" in r["Kernel_Name"]]
    if not inits:
        raise SystemExit(f"{path}: no k_cg_init dispatch")
    tail = [r for r in rows[inits[-1]:] if "k_cg_spmv" in r["Kernel_Name"]]
    vals = [float(r["Counter_Value"]) for r in tail]
    top = max(vals)
    work = [v for v in vals if v > 0.25 * top]
    copies = [float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"]]
    return work, max(copies)


def traffic(args):
    fw, fcopy = finest_spmv_kib(args.fetch)
    ww, wcopy = finest_spmv_kib(args.write)
    f_scale, w_scale = COPY_KIB / fcopy, COPY_KIB / wcopy
    fetch = statistics.mean(fw) * 1024 * f_scale
    write = statistics.mean(ww) * 1024 * w_scale
    out = {
        "kernel": "k_cg_spmv (finest depth)",
        "workload": args.workload,
        "launches": len(fw),
        "fetch_kib_raw_mean": round(statistics.mean(fw), 1),
        "write_kib_raw_mean": round(statistics.mean(ww), 1),
        "copy_calibration": {"bytes": COPY_KIB * 1024, "fetch_kib_read": fcopy, "write_kib_read": wcopy,
                             "fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4)},
        "hbm_fetch_bytes_per_launch": round(fetch),
        "hbm_write_bytes_per_launch": round(write),
        "hbm_bytes_per_launch": round(fetch + write),
    }
    dst = Path(args.out) if args.out else ROOT / "profiles" / f"traffic_k_cg_spmv_{args.workload}.json"
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out))


def trace(args):
    """Top kernels by total time, and per-dispatch statistics of the finest-depth CG."""
    with open(args.stats) as f:
        st = list(csv.DictReader(f))
    st.sort(key=lambda r: -float(r["TotalDurationNs"]))
    print("| kernel | calls | total ms | avg us | % |")
    print("|---|---|---|---|---|")
    for r in st[: args.top]:
        name = r["Name"].replace("prg::(anonymous namespace)::", "").split("(")[0]
        print(f"| {name} | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
              f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} |")
    if args.trace:
        rows = _rows(args.trace)
        rows.sort(key=lambda r: int(r["Start_Timestamp"]))
        inits = [i for i, r in enumerate(rows) if "k_cg_init" in r["Kernel_Name"]]
        tail = rows[inits[-1]:]
        for kname in ("k_cg_spmv", "k_cg_update"):
            d = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in tail if kname in r["Kernel_Name"]]
            top = max(d)
            work = [x for x in d if x > 0.25 * top]
            print(f"\nfinest depth {kname}: {len(work)} working dispatches, mean {statistics.mean(work):.1f} us, "
                  f"min {min(work):.1f} us, max {max(work):.1f} us ({len(d) - len(work)} idle dispatches)")


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("traffic")
    a.add_argument("--fetch", required=True)
    a.add_argument("--write", required=True)
    a.add_argument("--workload", default="superquadrics_2M_d10")
    a.add_argument("--out")
    b = sub.add_parser("trace")
    b.add_argument("--stats", required=True)
    b.add_argument("--trace")
    b.add_argument("--top", type=int, default=30)
    args = ap.parse_args()
    traffic(args) if args.cmd == "traffic" else trace(args)


if __name__ == "__main__":
    main()
