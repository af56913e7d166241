"""Summaries of rocprofv3 CSV output for this library's kernels.

    python -m fastimagingpipe_amd.rocprof_summary stats DIR [--last N]
        Per-kernel call count and average / total duration from the
        ``*kernel_stats.csv`` that ``rocprofv3 --kernel-trace --stats -f csv``
        writes under DIR.  With --last N, from the ``*kernel_trace.csv``
        instead, over each kernel's last N dispatches only (the timed steps of
        a run, without its warm-up), with their median and spread.

    python -m fastimagingpipe_amd.rocprof_summary traffic FETCH_DIR WRITE_DIR \
            --copy-bytes N [--json OUT]
        HBM bytes per dispatch from two ``rocprofv3 --pmc`` runs (FETCH_SIZE,
        then WRITE_SIZE; both in KiB per dispatch).  Each counter is scaled by
        what it reads for the device-to-device copy kernel of N bytes that
        the profiled program runs, so a plain copy of N bytes counts N bytes
        read and N bytes written.

    python -m fastimagingpipe_amd.rocprof_summary counters DIR [DIR ...]
        Average per dispatch of every counter in the ``rocprofv3 --pmc`` runs
        under the DIRs (one pass each), per pipeline stage, and ratios of
        those counters as the hardware reports them.
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

# Kernel name pattern -> pipeline stage (the names the library's per-stage
# event timing reports).
STAGES = [
    (r"k_reset", "reset"), (r"k_bin", "bin"), (r"k_scan", "scan"), (r"k_scatter", "scatter"),
    (r"k_stage1", "grid_colfft"), (r"k_stage2", "row_c2r"), (r"k_splat<\d+, true>", "splat_trigger"),
    (r"k_splat", "splat"), (r"k_trigger", "trigger"),
]
COPY_PATTERN = r"copyBuffer|Copy"


def stage_of(name: str) -> str | None:
    for pat, st in STAGES:
        if re.search(pat, name):
            return st
    return None


def _col(row: dict, *names):
    low = {k.lower(): v for k, v in row.items()}
    for n in names:
        if n.lower() in low:
            return low[n.lower()]
    raise KeyError(names)


def _find(directory: str, suffix: str) -> list[str]:
    files = glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True)
    if not files:
        raise FileNotFoundError(f"no *{suffix} under {directory}")
    return sorted(files)


def short(name: str, width: int = 70) -> str:
    name = re.sub(r"\s+", " ", name)
    return name if len(name) <= width else name[: width - 3] + "..."


def kernel_stats(directory: str) -> list[dict]:
    rows = []
    for path in _find(directory, "kernel_stats.csv"):
        with open(path) as f:
            for r in csv.DictReader(f):
                rows.append({"name": _col(r, "Name", "KernelName"),
                             "calls": int(_col(r, "Calls")),
                             "total_ns": float(_col(r, "TotalDurationNs")),
                             "avg_ns": float(_col(r, "AverageNs")),
                             "pct": float(_col(r, "Percentage"))})
    rows.sort(key=lambda r: -r["total_ns"])
    return rows


def counter_values(directory: str, counter: str) -> dict[str, list[float]]:
    """Counter value of every dispatch, per kernel name."""
    vals = defaultdict(dict)
    for path in _find(directory, "counter_collection.csv"):
        with open(path) as f:
            for r in csv.DictReader(f):
                if _col(r, "Counter_Name") != counter:
                    continue
                key = _col(r, "Dispatch_Id", "Correlation_Id")
                name = _col(r, "Kernel_Name")
                vals[name][key] = vals[name].get(key, 0.0) + float(_col(r, "Counter_Value"))
    return {k: list(v.values()) for k, v in vals.items()}


def counter_table(directories: list[str]) -> dict[str, dict[str, float]]:
    """Per stage: average per dispatch of every counter found under the directories."""
    sums = defaultdict(lambda: defaultdict(dict))   # stage -> counter -> dispatch -> value
    for d in directories:
        for path in _find(d, "counter_collection.csv"):
            with open(path) as f:
                for r in csv.DictReader(f):
                    st = stage_of(_col(r, "Kernel_Name"))
                    if st is None:
                        continue
                    cn = _col(r, "Counter_Name")
                    key = (d, _col(r, "Dispatch_Id", "Correlation_Id"))
                    sums[st][cn][key] = sums[st][cn].get(key, 0.0) + float(_col(r, "Counter_Value"))
    return {st: {cn: sum(v.values()) / len(v) for cn, v in cs.items()} for st, cs in sums.items()}


# Derived ratios: name -> (numerator, denominator).
RATIOS = [
    ("VALU issue / wave cycles", "SQ_ACTIVE_INST_VALU", "SQ_WAVE_CYCLES"),
    ("LDS issue / wave cycles", "SQ_ACTIVE_INST_LDS", "SQ_WAVE_CYCLES"),
    ("LDS wait / wave cycles", "SQ_WAIT_INST_LDS", "SQ_WAVE_CYCLES"),
    ("any wait / wave cycles", "SQ_WAIT_ANY", "SQ_WAVE_CYCLES"),
    ("LDS bank-conflict / LDS active", "SQ_LDS_BANK_CONFLICT", "SQ_LDS_IDX_ACTIVE"),
    ("LDS instr / VALU instr", "SQ_INSTS_LDS", "SQ_INSTS_VALU"),
]


def traffic(fetch_dir: str, write_dir: str, copy_bytes: int) -> dict:
    """copy_bytes: size of ONE calibration copy; the largest copy dispatch is taken
    as that copy (smaller copy dispatches are the program's input uploads)."""
    fetch = counter_values(fetch_dir, "FETCH_SIZE")
    write = counter_values(write_dir, "WRITE_SIZE")

    def copy_factor(table, what):
        cands = [max(v) for k, v in table.items() if re.search(COPY_PATTERN, k)]
        if not cands:
            raise RuntimeError(f"no copy kernel in the {what} run")
        kib = max(cands)
        return (copy_bytes / 1024.0) / kib, kib

    ff, fk = copy_factor(fetch, "FETCH_SIZE")
    fw, wk = copy_factor(write, "WRITE_SIZE")
    out = {"calibration": {"copy_bytes": copy_bytes, "fetch_kib": fk, "write_kib": wk,
                           "fetch_factor": ff, "write_factor": fw},
           "kernels": {}}
    for name in sorted(set(fetch) | set(write)):
        st = stage_of(name)
        if st is None:
            continue
        fv, wv = fetch.get(name, [0.0]), write.get(name, [0.0])
        f_b = sum(fv) / len(fv) * 1024.0 * ff
        w_b = sum(wv) / len(wv) * 1024.0 * fw
        rec = {"kernel": name, "fetch_bytes": f_b, "write_bytes": w_b, "bytes": f_b + w_b,
               "dispatches": len(fv)}
        prev = out["kernels"].get(st)
        if prev is None or rec["dispatches"] > prev["dispatches"]:
            out["kernels"][st] = rec
    return out


def kernel_trace_last(directory: str, last: int) -> list[dict]:
    """Per kernel name: the last `last` dispatches of the kernel trace (in
    dispatch order), their mean, median, min and max duration."""
    per = defaultdict(list)
    for path in _find(directory, "kernel_trace.csv"):
        with open(path) as f:
            for r in csv.DictReader(f):
                name = _col(r, "Kernel_Name", "KernelName", "Name")
                start, end = int(_col(r, "Start_Timestamp")), int(_col(r, "End_Timestamp"))
                per[name].append((int(_col(r, "Dispatch_Id", "Correlation_Id")), end - start))
    rows = []
    for name, v in per.items():
        v.sort()
        d = sorted(x for _, x in v[-last:])
        n = len(d)
        rows.append({"name": name, "calls": n, "total_ns": float(sum(d)), "avg_ns": sum(d) / n,
                     "median_ns": float(d[n // 2] if n % 2 else 0.5 * (d[n // 2 - 1] + d[n // 2])),
                     "min_ns": float(d[0]), "max_ns": float(d[-1])})
    tot = sum(r["total_ns"] for r in rows) or 1.0
    for r in rows:
        r["pct"] = 100.0 * r["total_ns"] / tot
    rows.sort(key=lambda r: -r["total_ns"])
    return rows


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("dir")
    a.add_argument("--last", type=int, default=0)
    b = sub.add_parser("traffic")
    b.add_argument("fetch_dir")
    b.add_argument("write_dir")
    b.add_argument("--copy-bytes", type=int, required=True)
    b.add_argument("--json")
    c = sub.add_parser("counters")
    c.add_argument("dirs", nargs="+")
    args = ap.parse_args(argv)
    if args.cmd == "counters":
        tab = counter_table(args.dirs)
        order = [st for _, st in STAGES if st in tab]
        order = list(dict.fromkeys(order))
        names = sorted({cn for st in tab for cn in tab[st]})
        print(f"{'counter (avg per dispatch)':<32}" + "".join(f"{st:>14}" for st in order))
        for cn in names:
            print(f"{cn:<32}" + "".join(f"{tab[st].get(cn, float('nan')):>14.4g}" for st in order))
        print()
        for label, num, den in RATIOS:
            if all(num in tab[st] and den in tab[st] for st in order):
                print(f"{label:<32}" + "".join(
                    f"{(tab[st][num] / tab[st][den] if tab[st][den] else float('nan')):>14.3f}" for st in order))
        return 0
    if args.cmd == "stats" and args.last > 0:
        rows = kernel_trace_last(args.dir, args.last)
        print(f"{'kernel':<72} {'stage':<13} {'calls':>6} {'avg_us':>10} {'median_us':>10} {'min_us':>8} "
              f"{'max_us':>8} {'pct':>6}")
        for r in rows:
            print(f"{short(r['name']):<72} {stage_of(r['name']) or '-':<13} {r['calls']:>6} "
                  f"{r['avg_ns'] / 1e3:>10.2f} {r['median_ns'] / 1e3:>10.2f} {r['min_ns'] / 1e3:>8.2f} "
                  f"{r['max_ns'] / 1e3:>8.2f} {r['pct']:>6.2f}")
        return 0
    if args.cmd == "stats":
        rows = kernel_stats(args.dir)
        print(f"{'kernel':<72} {'stage':<13} {'calls':>6} {'avg_us':>10} {'total_ms':>10} {'pct':>6}")
        for r in rows:
            print(f"{short(r['name']):<72} {stage_of(r['name']) or '-':<13} {r['calls']:>6} "
                  f"{r['avg_ns'] / 1e3:>10.2f} {r['total_ns'] / 1e6:>10.3f} {r['pct']:>6.2f}")
        return 0
    t = traffic(args.fetch_dir, args.write_dir, args.copy_bytes)
    c = t["calibration"]
    print(f"copy calibration: {c['copy_bytes']} B copied -> FETCH_SIZE {c['fetch_kib']:.0f} KiB "
          f"(x{c['fetch_factor']:.4f}), WRITE_SIZE {c['write_kib']:.0f} KiB (x{c['write_factor']:.4f})")
    print(f"{'stage':<13} {'dispatches':>10} {'fetch_MB':>10} {'write_MB':>10} {'total_MB':>10}  kernel")
    for st, r in t["kernels"].items():
        print(f"{st:<13} {r['dispatches']:>10} {r['fetch_bytes'] / 1e6:>10.2f} {r['write_bytes'] / 1e6:>10.2f} "
              f"{r['bytes'] / 1e6:>10.2f}  {short(r['kernel'], 50)}")
    if args.json:
        with open(args.json, "w") as f:
            json.dump(t, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
