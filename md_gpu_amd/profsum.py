"""Summaries of rocprofv3 output for the step and list-build kernels.

    python -m md_gpu_amd.profsum --pmc DIR [DIR ...] [--trace DIR] [--calib DIR ...]
                                 [--kernel SUBSTR] [--workload NAME] [--json OUT]

--pmc      directories of `rocprofv3 --pmc ... --output-format csv` runs (one
           counter group each); every *counter_collection.csv below them is read
--trace    a `rocprofv3 --kernel-trace --stats` directory (kernel_stats.csv)
--calib    counter runs of `python -m md_gpu_amd.profrun --system copy`: the
           largest `__amd_rocclr_copyBuffer` dispatch moves a known number of
           bytes (--calib-bytes, 64 MiB by default) each way, and FETCH_SIZE /
           WRITE_SIZE read for it give the scale from counter KiB to bytes
--kernel   substring of the kernel whose traffic goes into --json
--only     print the tables of these kernels only (substrings)

Printed: a markdown table per kernel of the per-dispatch means of every
counter, the derived HBM bytes, FP32 flop executed (64 lanes x (ADD + MUL +
2 FMA + TRANS) wave instructions) and VALU busy (SQ_ACTIVE_INST_VALU, in
units of 4 cycles per SIMD, x 4 over SIMDs x cycles; GRBM_GUI_ACTIVE is summed
over the 8 XCDs of the device and is divided by 8 first; ~1 means the VALU
issues every cycle, and small excesses over 1 come from the two counters being
read in different runs), plus the trace statistics.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sys
from collections import defaultdict

N_XCD = 8
N_SIMD = 256 * 4


def _csvs(d, suffix):
    return sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))


def short(name: str) -> str:
    return name.replace("(StepArgs)", "").replace("(BuildArgs)", "").replace("void ", "").strip()


def read_counters(dirs):
    """{kernel: {counter: [values per dispatch]}} plus {kernel: [durations ns]}"""
    vals = defaultdict(lambda: defaultdict(list))
    dur = defaultdict(list)
    for d in dirs:
        for path in _csvs(d, "counter_collection.csv"):
            with open(path) as f:
                for r in csv.DictReader(f):
                    k = r["Kernel_Name"]
                    vals[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
                    if r["Counter_Name"] in ("FETCH_SIZE", "SQ_WAVES"):
                        dur[k].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    return vals, dur


def calibration(dirs, nbytes):
    """(fetch scale, write scale): bytes per counter byte for a copy of nbytes."""
    best = {}
    for d in dirs:
        for path in _csvs(d, "counter_collection.csv"):
            with open(path) as f:
                for r in csv.DictReader(f):
                    if "copyBuffer" not in r["Kernel_Name"]:
                        continue
                    g = int(r["Grid_Size"])
                    c = r["Counter_Name"]
                    if c not in best or g > best[c][0]:
                        best[c] = (g, [])
                    if g == best[c][0]:
                        best[c][1].append(float(r["Counter_Value"]))
    sc = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        if c in best:
            kib = sum(best[c][1]) / len(best[c][1])
            sc[c] = nbytes / (kib * 1024.0)
    return sc.get("FETCH_SIZE"), sc.get("WRITE_SIZE")


def derive(cnt, fscale, wscale):
    m = {c: sum(v) / len(v) for c, v in cnt.items()}
    out = {}
    if "FETCH_SIZE" in m and fscale:
        out["fetch_bytes"] = m["FETCH_SIZE"] * 1024.0 * fscale
    if "WRITE_SIZE" in m and wscale:
        out["write_bytes"] = m["WRITE_SIZE"] * 1024.0 * wscale
    if "fetch_bytes" in out and "write_bytes" in out:
        out["bytes"] = out["fetch_bytes"] + out["write_bytes"]
    f32 = ("SQ_INSTS_VALU_ADD_F32", "SQ_INSTS_VALU_MUL_F32", "SQ_INSTS_VALU_FMA_F32", "SQ_INSTS_VALU_TRANS_F32")
    if all(c in m for c in f32):
        out["fp32_flop"] = 64.0 * (m[f32[0]] + m[f32[1]] + 2.0 * m[f32[2]] + m[f32[3]])
    if "SQ_ACTIVE_INST_VALU" in m and m.get("GRBM_GUI_ACTIVE"):
        out["valu_busy"] = m["SQ_ACTIVE_INST_VALU"] * 4.0 / (N_SIMD * m["GRBM_GUI_ACTIVE"] / N_XCD)
    if "SQ_INSTS_LDS" in m and "SQ_LDS_BANK_CONFLICT" in m:
        out["lds_conflict_cycles_per_lds_inst"] = m["SQ_LDS_BANK_CONFLICT"] / max(m["SQ_INSTS_LDS"], 1.0)
    return m, out


def read_trace(d):
    rows = []
    for path in _csvs(d, "kernel_stats.csv"):
        with open(path) as f:
            rows += list(csv.DictReader(f))
    return rows


def fmt(x):
    if isinstance(x, float):
        if x == 0 or 1e-3 <= abs(x) < 1e7:
            return f"{x:.4g}"
        return f"{x:.4e}"
    return str(x)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--pmc", nargs="*", default=[])
    ap.add_argument("--trace", default=None)
    ap.add_argument("--calib", nargs="*", default=[])
    ap.add_argument("--calib-bytes", type=int, default=64 << 20)
    ap.add_argument("--kernel", default="k_step<")
    ap.add_argument("--workload", default=None)
    ap.add_argument("--launch", default=None)
    ap.add_argument("--json", default=None)
    ap.add_argument("--min-dispatches", type=int, default=2)
    ap.add_argument("--only", nargs="*", default=None, help="print only kernels whose name contains one of these")
    args = ap.parse_args(argv)

    fscale, wscale = calibration(args.calib, args.calib_bytes) if args.calib else (None, None)
    vals, dur = read_counters(args.pmc)
    print(f"counter scale from a {args.calib_bytes} B device copy: FETCH_SIZE x {fmt(fscale)}, "
          f"WRITE_SIZE x {fmt(wscale)} (bytes per counted byte)\n")
    chosen = None
    for k in sorted(vals, key=lambda k: -sum(dur.get(k, [0]))):
        ndisp = max(len(v) for v in vals[k].values())
        if ndisp < args.min_dispatches or "__amd_rocclr" in k:
            continue
        if args.only and not any(o in k for o in args.only) and args.kernel not in k:
            continue
        m, d = derive(vals[k], fscale, wscale)
        us = sum(dur[k]) / len(dur[k]) / 1e3 if dur.get(k) else None
        print(f"### {short(k)}  ({ndisp} dispatches; mean duration under counters {fmt(us)} us)\n")
        print("| counter | mean per dispatch |\n|---|---|")
        for c in sorted(m):
            print(f"| {c} | {fmt(m[c])} |")
        for c in sorted(d):
            print(f"| *{c}* | {fmt(d[c])} |")
        if us and "bytes" in d:
            print(f"| *hbm_GBps* | {fmt(d['bytes'] / (us * 1e-6) / 1e9)} |")
        print()
        if args.kernel in k and (chosen is None or ndisp > chosen[1]):
            chosen = (k, ndisp, m, d, us)
    if args.trace:
        rows = read_trace(args.trace)
        if rows:
            print("### kernel trace (rocprofv3 --kernel-trace --stats)\n")
            print("| kernel | calls | mean us | min us | max us | share |\n|---|---|---|---|---|---|")
            for r in rows:
                print(f"| {short(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                      f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | {r['Percentage']}% |")
            print()
    if args.json:
        if chosen is None or "bytes" not in chosen[3]:
            print(f"no FETCH_SIZE/WRITE_SIZE data for a kernel matching {args.kernel!r}", file=sys.stderr)
            return 1
        k, ndisp, m, d, us = chosen
        out = {"kernel": short(k), "workload": args.workload, "launch": args.launch,
               "dispatches": ndisp, "bytes_per_launch": d["bytes"], "fetch_bytes": d["fetch_bytes"],
               "write_bytes": d["write_bytes"], "fetch_scale": fscale, "write_scale": wscale,
               "duration_us_under_counters": us}
        for c in ("fp32_flop", "valu_busy", "lds_conflict_cycles_per_lds_inst"):
            if c in d:
                out[c] = d[c]
        with open(args.json, "w") as f:
            json.dump(out, f, indent=1)
            f.write("\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
