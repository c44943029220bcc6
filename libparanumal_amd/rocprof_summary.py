"""Summarise rocprofv3 CSV output into the short text files kept beside the
benchmarks.

    python -m libparanumal_amd.rocprof_summary stats  <dir> [--top 12] [--cmd "..."]
    python -m libparanumal_amd.rocprof_summary kernel <dir> --kernel <substring> [--grid G] [--cmd "..."]
    python -m libparanumal_amd.rocprof_summary kernels <dir> --match <sub1,sub2,...> [--cmd "..."]
        (substrings separated by "|" instead when they contain commas)
    python -m libparanumal_amd.rocprof_summary traffic --fetch <dir> --write <dir>
           --kernel <substring> --copy-bytes <B> [--copy-kernel <substring>] [--grid G]
           [--cmd-fetch ...] [--cmd-write ...] [--json out.json]

`stats` reads `*_kernel_stats.csv` (from `rocprofv3 --kernel-trace --stats
--output-format csv`). `traffic` reads `*_counter_collection.csv` from two
`--pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs. Both counters are in KiB per
dispatch. Each is calibrated against a plain device-to-device copy of known
size in the same run (by default the largest dispatch of the library's HBM
copy probe, hbm_copy_kernel, which `bench.py --counters` runs on 1 GiB): the
factor is (copy bytes) / (counter KiB x 1024 of the copy dispatch). The
kernel's HBM bytes per dispatch are its counter times that factor.
"""
import argparse
import csv
import glob
import json
import os
import sys


def _one(pattern):
    hits = sorted(glob.glob(pattern))
    if not hits:
        raise SystemExit(f"no file matches {pattern}")
    return hits[0]


def kernel_stats(d):
    rows = list(csv.DictReader(open(_one(os.path.join(d, "*kernel_stats.csv")))))
    for r in rows:
        for k in ("Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs"):
            r[k] = float(r[k])
    return rows


def short(name, n=96):
    return name if len(name) <= n else name[: n - 3] + "..."


def stats_text(d, top=12, cmd=None):
    rows = kernel_stats(d)
    out = []
    if cmd:
        out.append(f"command: {cmd}")
    out.append(f"{'calls':>6} {'avg us':>10} {'min us':>10} {'max us':>10} {'total %':>8}  kernel")
    for r in rows[:top]:
        out.append(f"{int(r['Calls']):6d} {r['AverageNs'] / 1e3:10.1f} {r['MinNs'] / 1e3:10.1f} "
                   f"{r['MaxNs'] / 1e3:10.1f} {r['Percentage']:8.2f}  {short(r['Name'])}")
    return "\n".join(out)


def kernel_durations(d, kernel, grid=None):
    """Durations (ns) of the dispatches of one kernel from *kernel_trace.csv in
    start order, optionally only those with Grid_Size_X == grid."""
    out = []
    for r in csv.DictReader(open(_one(os.path.join(d, "*kernel_trace.csv")))):
        if kernel not in r["Kernel_Name"]:
            continue
        if grid is not None and int(r["Grid_Size_X"]) != grid:
            continue
        out.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]) - int(r["Start_Timestamp"])))
    if not out:
        raise SystemExit(f"no dispatch of {kernel!r} with grid {grid}")
    return [dur for _, dur in sorted(out)]


def _line(label, v):
    v = sorted(v)
    return (f"{label}: dispatches {len(v)}  mean {sum(v) / len(v) / 1e3:.1f} us  median {v[len(v) // 2] / 1e3:.1f} us  "
            f"min {v[0] / 1e3:.1f} us  max {v[-1] / 1e3:.1f} us")


def kernel_text(d, kernel, grid=None, cmd=None, first=None, count=None):
    v = kernel_durations(d, kernel, grid)
    out = []
    if cmd:
        out.append(f"command: {cmd}")
    out.append(f"kernel: {kernel}" + (f"  grid size: {grid}" if grid is not None else ""))
    out.append(_line("all", v))
    if first is not None:
        w = v[first:first + (count or len(v))]
        out.append(_line(f"dispatches {first}..{first + len(w) - 1} (in start order)", w))
    return "\n".join(out)


def kernels_text(d, subs, cmd=None):
    """Dispatch count, median and min duration per (kernel, grid) for every
    kernel whose name contains one of the substrings."""
    groups = {}
    for r in csv.DictReader(open(_one(os.path.join(d, "*kernel_trace.csv")))):
        name = r["Kernel_Name"]
        if not any(s in name for s in subs):
            continue
        key = (name.split("(")[0].replace("void ", "").replace("libp_amd::", ""), int(r["Grid_Size_X"]))
        groups.setdefault(key, []).append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    out = []
    if cmd:
        out.append(f"command: {cmd}")
    out.append(f"{'dispatches':>10} {'median us':>10} {'min us':>8}  kernel (grid size x)")
    for (name, grid), v in sorted(groups.items()):
        v.sort()
        out.append(f"{len(v):10d} {v[len(v) // 2] / 1e3:10.1f} {v[0] / 1e3:8.1f}  {name} ({grid})")
    return "\n".join(out)


def counters(d, counter, grid=None):
    """{kernel name: [value per dispatch]} for one counter (only dispatches of
    grid size `grid` when given)."""
    f = _one(os.path.join(d, "*counter_collection.csv"))
    agg = {}
    for r in csv.DictReader(open(f)):
        if r.get("Counter_Name") != counter:
            continue
        if grid is not None and int(r.get("Grid_Size", r.get("Grid_Size_X", -1))) != grid:
            continue
        agg.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    if not agg:
        raise SystemExit(f"{f}: no {counter} rows")
    return agg


def _pick(agg, sub):
    hits = [k for k in agg if sub in k]
    if not hits:
        raise SystemExit(f"no kernel matching {sub!r}; have {list(agg)[:8]}")
    return hits[0], agg[hits[0]]


def _median(v):
    v = sorted(v)
    return v[len(v) // 2]


def traffic(fetch_dir, write_dir, kernel, copy_bytes, copy_kernel="hbm_copy_kernel", grid=None):
    res = {"kernel": None, "copy_bytes": copy_bytes}
    total = 0.0
    for tag, d, counter in (("fetch", fetch_dir, "FETCH_SIZE"), ("write", write_dir, "WRITE_SIZE")):
        name, vals = _pick(counters(d, counter, grid), kernel)
        _, cvals = _pick(counters(d, counter), copy_kernel)
        copy_kib = max(cvals)  # the large calibration copy is the biggest copy dispatch
        factor = copy_bytes / (copy_kib * 1024.0)
        kib = _median(vals)
        res["kernel"] = name
        res[tag] = {"counter": counter, "kernel_kib_per_dispatch": kib, "dispatches": len(vals),
                    "copy_kib_per_dispatch": copy_kib, "factor": factor, "bytes": kib * 1024.0 * factor}
        total += kib * 1024.0 * factor
    res["traffic_bytes"] = total
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="mode", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--top", type=int, default=12)
    s.add_argument("--cmd")
    k = sub.add_parser("kernel")
    k.add_argument("dir")
    k.add_argument("--kernel", required=True)
    k.add_argument("--grid", type=int)
    k.add_argument("--first", type=int, help="also summarise the dispatches from this index (start order)")
    k.add_argument("--count", type=int)
    k.add_argument("--cmd")
    ks = sub.add_parser("kernels")
    ks.add_argument("dir")
    ks.add_argument("--match", required=True)
    ks.add_argument("--cmd")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--copy-kernel", default="hbm_copy_kernel")
    t.add_argument("--grid", type=int)
    t.add_argument("--cmd-fetch")
    t.add_argument("--cmd-write")
    t.add_argument("--json")
    a = ap.parse_args(argv)
    if a.mode == "stats":
        print(stats_text(a.dir, a.top, a.cmd))
        return 0
    if a.mode == "kernel":
        print(kernel_text(a.dir, a.kernel, a.grid, a.cmd, a.first, a.count))
        return 0
    if a.mode == "kernels":
        print(kernels_text(a.dir, a.match.split("|" if "|" in a.match else ","), a.cmd))
        return 0
    r = traffic(a.fetch, a.write, a.kernel, a.copy_bytes, a.copy_kernel, a.grid)
    r["commands"] = {"fetch": a.cmd_fetch, "write": a.cmd_write}
    if a.json:
        with open(a.json, "w") as f:
            json.dump(r, f, indent=1)
    print(f"kernel: {short(r['kernel'])}")
    for tag in ("fetch", "write"):
        x = r[tag]
        print(f"{x['counter']}: {x['kernel_kib_per_dispatch']:.0f} KiB/dispatch x factor {x['factor']:.3f} "
              f"(copy of {copy_bytes_str(r['copy_bytes'])}: {x['copy_kib_per_dispatch']:.0f} KiB) "
              f"= {x['bytes'] / 1e9:.4f} GB")
    print(f"HBM traffic per dispatch: {r['traffic_bytes'] / 1e9:.4f} GB")
    return 0


def copy_bytes_str(b):
    return f"{b / 2**30:.2f} GiB"


if __name__ == "__main__":
    sys.exit(main())
