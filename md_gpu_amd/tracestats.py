"""Summaries of rocprofv3 kernel traces of the MD engine.

    python -m md_gpu_amd.tracestats <prefix>_kernel_trace.csv [--window N]
    python -m md_gpu_amd.tracestats --pmc <prefix>_counter_collection.csv

The step kernel and the list-build chain are launched every step / every
check interval but exit at once when a replica has nothing to do (list stale,
target reached, list still valid).  rocprofv3 --stats averages those idle
launches in; this tool reports real launches (longer than --idle-us) and idle
ones separately, per kernel, over the longest stretch of back-to-back
launches (the timed region) or over the whole trace.
"""
from __future__ import annotations

import argparse
import collections
import csv
import statistics


def _name(r):
    n = r["Kernel_Name"]
    n = n.split("(")[0]
    return n[5:] if n.startswith("void ") else n


def load_trace(path):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    return [(_name(r), int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in rows]


def longest_busy_window(rows, gap_ns=2_000_000):
    """Indices [a, b) of the longest stretch without a gap above gap_ns."""
    best, a = (0, 0), 0
    for k in range(1, len(rows) + 1):
        if k == len(rows) or rows[k][1] - rows[k - 1][2] > gap_ns:
            if k - a > best[1] - best[0]:
                best = (a, k)
            a = k
    return best


def summarize(rows, idle_us):
    per = collections.defaultdict(list)
    for n, s, e in rows:
        per[n].append((e - s) / 1e3)
    out = []
    for n, d in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        real = [x for x in d if x > idle_us]
        idle = [x for x in d if x <= idle_us]
        out.append(dict(kernel=n, calls=len(d), total_us=sum(d), mean_us=statistics.fmean(d),
                        real=len(real), real_mean_us=statistics.fmean(real) if real else 0.0,
                        real_median_us=statistics.median(real) if real else 0.0,
                        idle=len(idle), idle_mean_us=statistics.fmean(idle) if idle else 0.0))
    return out


def pmc_summary(path, idle_us):
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(path)):
        dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        if dur <= idle_us:
            continue
        per[_name(r)][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: statistics.median(v) for c, v in cs.items()} for k, cs in per.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("path")
    ap.add_argument("--idle-us", type=float, default=10.0)
    ap.add_argument("--whole", action="store_true", help="whole trace instead of the longest busy stretch")
    ap.add_argument("--pmc", action="store_true", help="path is a counter_collection.csv")
    args = ap.parse_args()
    if args.pmc:
        for k, cs in pmc_summary(args.path, args.idle_us).items():
            print(f"{k:36s} " + "  ".join(f"{c}={v:.6g}" for c, v in sorted(cs.items())))
        return
    rows = load_trace(args.path)
    a, b = (0, len(rows)) if args.whole else longest_busy_window(rows)
    win = rows[a:b]
    wall = (win[-1][2] - win[0][1]) / 1e3 if win else 0.0
    print(f"launches {len(win)} (of {len(rows)}), wall {wall:.1f} us, idle threshold {args.idle_us} us")
    print(f"{'kernel':36s} {'calls':>7s} {'total_us':>11s} {'mean_us':>9s} {'real':>6s} {'real_mean':>10s} "
          f"{'real_med':>9s} {'idle':>6s} {'idle_mean':>9s}")
    for r in summarize(win, args.idle_us):
        print(f"{r['kernel']:36s} {r['calls']:7d} {r['total_us']:11.1f} {r['mean_us']:9.2f} {r['real']:6d} "
              f"{r['real_mean_us']:10.2f} {r['real_median_us']:9.2f} {r['idle']:6d} {r['idle_mean_us']:9.2f}")


if __name__ == "__main__":
    main()
