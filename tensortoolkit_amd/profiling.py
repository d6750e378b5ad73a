"""Summaries of rocprofv3 output for the contraction kernels.

``python -m tensortoolkit_amd.profiling stats DIR`` prints per-kernel call
counts and durations from a ``--kernel-trace --stats`` CSV run;
``trace DIR --kernel REGEX`` groups one kernel's dispatches by grid size
(one group per contraction of the workload).

``python -m tensortoolkit_amd.profiling pmc --fetch DIR --write DIR
--kernel REGEX --label NAME --workload W --copy-bytes N`` turns two
``--pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` CSV runs into HBM bytes per
dispatch of one kernel.  Both counters are in KiB per dispatch; each is
scaled by what the same counter reads for the device-to-device copy of
``N`` bytes that the measured program ran first (``bench.py
--calibration-copy-mb``), which reads and writes exactly ``N`` bytes.  The
kernel's dispatches are those whose name matches REGEX with the largest
grid, i.e. the dominant contraction of the workload.
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import os
import re
import sys


def _short(name: str) -> str:
    name = name.replace("(anonymous namespace)::", "").replace("void ", "")
    return name.split("(")[0]


def _csv(directory: str, suffix: str) -> str:
    hits = sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True))
    if not hits:
        raise FileNotFoundError(f"no *{suffix} under {directory}")
    return hits[0]


def read_dispatches(directory: str):
    """Dispatches of a --pmc run: list of dicts with kernel, grid, counters."""
    rows = list(csv.DictReader(open(_csv(directory, "counter_collection.csv"))))
    disp = collections.OrderedDict()
    for r in rows:
        d = disp.setdefault(int(r["Dispatch_Id"]), {"kernel": _short(r["Kernel_Name"]),
                                                     "grid": int(r["Grid_Size"]),
                                                     "wg": int(r["Workgroup_Size"]),
                                                     "counters": collections.defaultdict(float)})
        d["counters"][r["Counter_Name"]] += float(r["Counter_Value"])
    return list(disp.values())


def kernel_stats(directory: str) -> str:
    rows = list(csv.DictReader(open(_csv(directory, "kernel_stats.csv"))))
    out = [f"{'kernel':64s} {'calls':>6s} {'total_us':>11s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s} {'pct':>6s}"]
    for r in rows:
        out.append(f"{_short(r['Name'])[:64]:64s} {int(r['Calls']):6d} {float(r['TotalDurationNs']) / 1e3:11.1f} "
                   f"{float(r['AverageNs']) / 1e3:9.2f} {float(r['MinNs']) / 1e3:9.2f} "
                   f"{float(r['MaxNs']) / 1e3:9.2f} {float(r['Percentage']):6.2f}")
    return "\n".join(out)


def kernel_trace_by_grid(directory: str, kernel_regex: str) -> str:
    """Durations of the dispatches of one kernel from a --kernel-trace CSV,
    grouped by grid (workgroups): one contraction's launch per group."""
    pat = re.compile(kernel_regex)
    rows = list(csv.DictReader(open(_csv(directory, "kernel_trace.csv"))))
    groups = collections.defaultdict(list)
    for r in rows:
        name = _short(r["Kernel_Name"])
        if not pat.search(name):
            continue
        wgs = int(r["Grid_Size_X"]) // max(1, int(r["Workgroup_Size_X"]))
        groups[(name, wgs)].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    out = [f"{'kernel':44s} {'workgroups':>10s} {'calls':>6s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s}"]
    for (name, wgs), v in sorted(groups.items(), key=lambda kv: -kv[0][1]):
        out.append(f"{name[:44]:44s} {wgs:10d} {len(v):6d} {sum(v) / len(v):9.2f} {min(v):9.2f} {max(v):9.2f}")
    return "\n".join(out)


def _calibration(dispatches, counter: str, copy_bytes: int) -> float:
    copies = [d for d in dispatches if "copyBuffer" in d["kernel"]]
    if not copies:
        raise RuntimeError("no device-to-device copy dispatch to calibrate against")
    big = max(copies, key=lambda d: d["counters"][counter])
    return copy_bytes / (big["counters"][counter] * 1024.0)


def pmc_summary(fetch_dir: str, write_dir: str, kernel_regex: str, label: str, workload: str,
                copy_bytes: int, command: str) -> str:
    pat = re.compile(kernel_regex)
    lines = []
    per = {}
    for counter, directory in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        ds = read_dispatches(directory)
        factor = _calibration(ds, counter, copy_bytes)
        mine = [d for d in ds if pat.search(d["kernel"])]
        if not mine:
            raise RuntimeError(f"no dispatch of {kernel_regex} in {directory}")
        grid = max(d["grid"] for d in mine)
        sel = [d for d in mine if d["grid"] == grid]
        kib = sum(d["counters"][counter] for d in sel) / len(sel)
        per[counter] = kib * 1024.0 * factor
        lines.append(f"{counter.lower()}_kib_per_dispatch: {kib:.1f}")
        lines.append(f"{counter.lower()}_calibration_bytes_per_kib: {factor * 1024.0:.1f}")
        lines.append(f"{counter.lower()}_dispatches_averaged: {len(sel)} (grid {grid // sel[0]['wg']} workgroups)")
        lines.append(f"{counter.lower()}_bytes_per_dispatch: {per[counter]:.0f}")
    head = [f"kernel: {label}", f"workload: {workload}", f"kernel_regex: {kernel_regex}",
            f"calibration_copy_bytes: {copy_bytes}",
            f"traffic_bytes_per_dispatch: {per['FETCH_SIZE'] + per['WRITE_SIZE']:.0f}"]
    return "\n".join(head + lines + ["command: " + command])


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    t = sub.add_parser("trace")
    t.add_argument("dir")
    t.add_argument("--kernel", required=True, help="regex on the short kernel name")
    p = sub.add_parser("pmc")
    p.add_argument("--fetch", required=True)
    p.add_argument("--write", required=True)
    p.add_argument("--kernel", required=True, help="regex on the short kernel name")
    p.add_argument("--label", required=True, help="kernel name as the kernel probe reports it")
    p.add_argument("--workload", required=True)
    p.add_argument("--copy-bytes", type=int, required=True)
    p.add_argument("--command", default="")
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        print(kernel_stats(a.dir))
    elif a.cmd == "trace":
        print(kernel_trace_by_grid(a.dir, a.kernel))
    else:
        print(pmc_summary(a.fetch, a.write, a.kernel, a.label, a.workload, a.copy_bytes, a.command))
    return 0


if __name__ == "__main__":
    sys.exit(main())
