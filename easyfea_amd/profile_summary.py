"""Text summaries of rocprofv3 output (the files kept under ``profiles/``).

``python -m easyfea_amd.profile_summary stats --stats S_kernel_stats.csv
[--trace S_kernel_trace.csv --kernel k_pcg2_step1 --harness-launches 500]``
    Markdown table of the kernels by total time; with a trace, the named
    kernel's launches split into those of the solves and the last
    ``--harness-launches`` ones (the timing harness ``bench.py`` runs after
    the solves), each with its average duration.

``python -m easyfea_amd.profile_summary traffic --fetch F.csv --write W.csv
--kernel k_pcg2_step1 --copy-bytes N``
    HBM bytes per launch of the named kernel from two ``--pmc`` runs
    (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by what the
    counter reads for the device-to-device copy of ``N`` bytes each way that
    the same program runs last (``__amd_rocclr_copyBuffer``).  Launches that
    returned at once (a converged solve's tail: under 10% of the median
    fetch) are left out.  Prints JSON.
"""

from __future__ import annotations

import argparse
import csv
import json
import sys

import numpy as np


def _short(name: str) -> str:
    n = name.split("(")[0]
    return n.replace("void ", "").replace("efea::", "")


def stats_table(path: str, top: int = 15) -> str:
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = ["| kernel | calls | total ms | avg us | % of kernel time |", "|---|---|---|---|---|"]
    for r in rows[:top]:
        out.append(f"| `{_short(r['Name'])}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.2f} | "
                   f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['Percentage']):.1f} |")
    return "\n".join(out)


def trace_split(path: str, kernel: str, harness: int) -> dict:
    """Average durations (us) of ``kernel``'s launches: all, those of the
    solves (all but the last ``harness``) with and without the launches that
    returned at once, and the harness's."""
    d = []
    for r in csv.DictReader(open(path)):
        if kernel in r["Kernel_Name"]:
            d.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]) - int(r["Start_Timestamp"])))
    d.sort()
    dur = np.array([x[1] for x in d], dtype=float) / 1e3
    solve, bench = dur[:-harness] if harness else dur, dur[-harness:] if harness else dur[:0]
    med = float(np.median(solve)) if solve.size else 0.0
    working = solve[solve >= 0.5 * med] if solve.size else solve
    return {
        "launches": int(dur.size),
        "avg_us_all": float(dur.mean()) if dur.size else None,
        "solve_launches": int(solve.size),
        "solve_avg_us": float(solve.mean()) if solve.size else None,
        "solve_working_launches": int(working.size),
        "solve_working_avg_us": float(working.mean()) if working.size else None,
        "harness_launches": int(bench.size),
        "harness_avg_us": float(bench.mean()) if bench.size else None,
    }


def _counter(path: str, name_part: str):
    vals = []
    for r in csv.DictReader(open(path)):
        if name_part in r["Kernel_Name"]:
            vals.append((r["Kernel_Name"], float(r["Counter_Value"])))
    return vals


def traffic(fetch_csv: str, write_csv: str, kernel: str, copy_bytes: int) -> dict:
    cf = _counter(fetch_csv, "__amd_rocclr_copyBuffer")
    cw = _counter(write_csv, "__amd_rocclr_copyBuffer")
    # the calibration copy is the largest copy of the run
    f_copy = max(v for _, v in cf)
    w_copy = max(v for _, v in cw)
    fetch_scale = copy_bytes / 1024 / f_copy
    write_scale = copy_bytes / 1024 / w_copy
    kf = _counter(fetch_csv, kernel)
    kw = _counter(write_csv, kernel)
    name = kf[0][0]
    f = np.array([v for _, v in kf])
    w = np.array([v for _, v in kw])
    keep_f = f >= 0.1 * np.median(f)
    keep_w = w >= 0.1 * np.median(w)
    fk, wk = float(f[keep_f].mean()), float(w[keep_w].mean())
    return {
        "kernel": name.split("(")[0],
        "dispatches": int(f.size),
        "dispatches_used": [int(keep_f.sum()), int(keep_w.sum())],
        "fetch_kib": fk,
        "write_kib": wk,
        "fetch_scale": fetch_scale,
        "write_scale": write_scale,
        "calibration": f"device-to-device copy of {copy_bytes} bytes each way: FETCH_SIZE {f_copy:.1f} KiB, "
                       f"WRITE_SIZE {w_copy:.1f} KiB",
        "hbm_bytes_per_launch": (fk * fetch_scale + wk * write_scale) * 1024,
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("--stats", required=True)
    s.add_argument("--top", type=int, default=15)
    s.add_argument("--trace")
    s.add_argument("--kernel")
    s.add_argument("--harness-launches", type=int, default=0)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        print(stats_table(args.stats, args.top))
        if args.trace and args.kernel:
            print()
            print(json.dumps(trace_split(args.trace, args.kernel, args.harness_launches)))
    else:
        print(json.dumps(traffic(args.fetch, args.write, args.kernel, args.copy_bytes), indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
