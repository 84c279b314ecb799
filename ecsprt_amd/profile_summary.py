"""Condense rocprofv3 CSV output into short text summaries.

    python -m ecsprt_amd.profile_summary stats <rocprof dir> <out.txt> --command "<command>"
    python -m ecsprt_amd.profile_summary pmc <rocprof dir> <out.txt> --command "<command>"

`stats` reads ``*_kernel_stats.csv`` (rocprofv3 --kernel-trace --stats) and writes the
kernels ranked by total time with calls and average duration; when the kernel trace is
present it adds the launches of one time step (between the last two launches of the
step's first kernel).  `pmc` reads ``*_counter_collection.csv`` (rocprofv3 --pmc) and
writes the mean counter value per dispatch of each kernel.  With ``--drop`` the bulky
CSVs are deleted afterwards.
"""
import argparse
import collections
import csv
import glob
import os


def short(name, width=72):
    n = name.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "").replace("ecsprt::", "")
    return n if len(n) <= width else n[: width - 3] + "..."


def one(path_glob):
    hits = sorted(glob.glob(path_glob, recursive=True))
    return hits[0] if hits else None


def stats(d, out, command, top, step_marker, drop, by_grid=""):
    path = one(os.path.join(d, "**", "*kernel_stats.csv"))
    if path is None:
        raise SystemExit(f"no kernel_stats.csv under {d}")
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    lines = [f"command: {command}", f"source: rocprofv3 kernel statistics ({os.path.basename(path)})",
             f"kernel time total: {total / 1e6:.1f} ms over {sum(int(r['Calls']) for r in rows)} launches", "",
             f"{'kernel':72s} {'calls':>7s} {'total_ms':>10s} {'avg_us':>10s} {'min_us':>9s} {'max_us':>9s} {'pct':>6s}"]
    for r in rows[:top]:
        lines.append(f"{short(r['Name']):72s} {int(r['Calls']):7d} {float(r['TotalDurationNs']) / 1e6:10.2f} "
                     f"{float(r['AverageNs']) / 1e3:10.2f} {float(r['MinNs']) / 1e3:9.2f} "
                     f"{float(r['MaxNs']) / 1e3:9.2f} {float(r['Percentage']):6.2f}")
    trace = one(os.path.join(d, "**", "*kernel_trace.csv"))
    if trace and by_grid:
        lines += [""] + grid_lines(trace, by_grid)
    if trace and step_marker:
        lines += [""] + step_lines(trace, step_marker)
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")
    if drop:
        for p in glob.glob(os.path.join(d, "**", "*.csv"), recursive=True):
            os.remove(p)


def grid_lines(trace, prefixes):
    """Launches of the kernels whose names start with one of `prefixes`, grouped by launch
    shape (one workload's launches of a kernel share it), in order of first launch."""
    rows = list(csv.DictReader(open(trace)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    groups = collections.OrderedDict()
    for r in rows:
        k = short(r["Kernel_Name"], 400)
        if not any(k.startswith(p) for p in prefixes.split(";")):
            continue
        grid = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
        block = int(r["Workgroup_Size_X"]) * int(r["Workgroup_Size_Y"]) * int(r["Workgroup_Size_Z"])
        g = groups.setdefault((short(r["Kernel_Name"]), grid // max(block, 1), block), [])
        g.append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    out = ["launches by shape (workgroups x threads), in order of first launch:",
           f"{'kernel':60s} {'workgroups':>10s} {'threads':>7s} {'calls':>6s} {'avg_us':>10s} {'min_us':>9s} {'max_us':>9s}"]
    for (k, wg, b), v in groups.items():
        out.append(f"{k[:60]:60s} {wg:10d} {b:7d} {len(v):6d} {sum(v) / len(v):10.2f} {min(v):9.2f} {max(v):9.2f}")
    return out


def step_lines(trace, marker):
    rows = list(csv.DictReader(open(trace)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    out = []
    for m in marker.split(";"):
        idx = [i for i, r in enumerate(rows) if short(r["Kernel_Name"], 400).startswith(m)]
        if len(idx) < 2:
            continue
        seg = rows[idx[-2]: idx[-1]]
        by = collections.OrderedDict()
        busy = 0.0
        for r in seg:
            us = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
            k = short(r["Kernel_Name"])
            c = by.setdefault(k, [0, 0.0])
            c[0] += 1
            c[1] += us
            busy += us
        span = (int(seg[-1]["End_Timestamp"]) - int(seg[0]["Start_Timestamp"])) / 1e3
        out.append(f"one step (from the last-but-one launch of {m}): {len(seg)} launches, "
                   f"span {span:.1f} us, kernel time {busy:.1f} us")
        out.append(f"{'kernel':72s} {'n':>5s} {'total_us':>10s}")
        for k, (n, us) in sorted(by.items(), key=lambda kv: -kv[1][1]):
            out.append(f"{k:72s} {n:5d} {us:10.1f}")
        out.append("")
    return out


def pmc(d, out, command, drop):
    path = one(os.path.join(d, "**", "*counter_collection.csv"))
    if path is None:
        raise SystemExit(f"no counter_collection.csv under {d}")
    vals = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        vals[(short(r["Kernel_Name"]), r["Counter_Name"])].append(float(r["Counter_Value"]))
    lines = [f"command: {command}", f"source: rocprofv3 counter collection ({os.path.basename(path)})", "",
             f"{'kernel':72s} {'counter':>12s} {'dispatches':>10s} {'mean_per_dispatch':>18s}"]
    for (k, c), v in sorted(vals.items(), key=lambda kv: -sum(kv[1]) / len(kv[1])):
        lines.append(f"{k:72s} {c:>12s} {len(v):10d} {sum(v) / len(v):18.1f}")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")
    if drop:
        for p in glob.glob(os.path.join(d, "**", "*.csv"), recursive=True):
            os.remove(p)


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("mode", choices=["stats", "pmc"])
    ap.add_argument("dir")
    ap.add_argument("out")
    ap.add_argument("--command", required=True, help="the command that produced the profile")
    ap.add_argument("--top", type=int, default=25)
    ap.add_argument("--step-marker", default="", help="kernel-name prefixes that start a step, separated by ';'")
    ap.add_argument("--by-grid", default="", help="kernel-name prefixes to list per launch shape, separated by ';'")
    ap.add_argument("--drop", action="store_true", help="delete the CSVs after summarising")
    a = ap.parse_args()
    if a.mode == "stats":
        stats(a.dir, a.out, a.command, a.top, a.step_marker, a.drop, a.by_grid)
    else:
        pmc(a.dir, a.out, a.command, a.drop)


if __name__ == "__main__":
    main()
