"""Summarise a rocprofv3 CSV output directory as a short markdown report.

usage: python -m qcd_amd.profile_summary DIR --command "rocprofv3 ... -- python3 ..." [--top 12]

Reads, when present, `*_kernel_stats.csv` (per-kernel totals), `*_kernel_trace.csv`
(per-dispatch times; grouped by kernel name and grid size, because the autotuner's
candidates share kernel names but differ in grid size) and `*_counter_collection.csv`
(PMC counters per dispatch, averaged per kernel and grid).  For FETCH_SIZE and
WRITE_SIZE the plain FP64 copy (`k_lincomb<0>`, 192 B read + 192 B written per
site of the same lattice) calibrates the counter: factor = true bytes / counter
bytes, and calibrated bytes per lattice site use --sites.
"""
import argparse
import csv
import glob
import os
import re
import sys
from collections import defaultdict

COPY_KERNEL = "k_lincomb<0>"
COPY_BYTES_PER_SITE = {"FETCH_SIZE": 192, "WRITE_SIZE": 192}


def short(name, width=90):
    name = re.sub(r"\(.*\)$", "", name).replace("void ", "").replace("qcdhip::", "")
    return name if len(name) <= width else name[:width - 3] + "..."


def rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def find(d, suffix):
    hits = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    return hits[0] if hits else None


def stats_table(path, top):
    out = ["| kernel | calls | average µs | total % |", "|---|---|---|---|"]
    for r in rows(path)[:top]:
        out.append(f"| `{short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['Percentage']):.1f} |")
    return out


def trace_groups(path):
    """(kernel, grid) -> list of durations in ns, in dispatch order."""
    g = defaultdict(list)
    for r in sorted(rows(path), key=lambda r: int(r["Dispatch_Id"])):
        grid = (int(r["Grid_Size_X"]), int(r["Grid_Size_Y"]), int(r["Grid_Size_Z"]))
        g[(r["Kernel_Name"], grid)].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    return g


def trace_table(path, top):
    g = trace_groups(path)
    items = sorted(g.items(), key=lambda kv: -sum(kv[1]))[:top]
    out = ["| kernel | grid | dispatches | mean µs | median µs |", "|---|---|---|---|---|"]
    for (name, grid), d in items:
        s = sorted(d)
        out.append(f"| `{short(name)}` | {grid[0]}x{grid[1]}x{grid[2]} | {len(d)} | {sum(d) / len(d) / 1e3:.1f} | "
                   f"{s[len(s) // 2] / 1e3:.1f} |")
    return out


def counter_means(path):
    """(kernel, grid, counter) -> mean value per dispatch; plus the copy kernel's grid."""
    acc = defaultdict(list)
    for r in rows(path):
        acc[(r["Kernel_Name"], int(r["Grid_Size"]), r["Counter_Name"])].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}, {k: len(v) for k, v in acc.items()}


def calibration(means, counter, sites):
    """true bytes / counter bytes of the plain copy, or None."""
    for (name, grid, c), v in means.items():
        if c == counter and COPY_KERNEL in name and v > 0:
            return COPY_BYTES_PER_SITE[counter] * sites / (v * 1024.0)
    return None


def counter_table(path, sites, top):
    means, counts = counter_means(path)
    out = []
    for counter in sorted({k[2] for k in means}):
        f = calibration(means, counter, sites) if (counter in COPY_BYTES_PER_SITE and sites) else None
        head = f"**{counter}** (KiB per dispatch"
        if f:
            head += f"; copy-kernel calibration factor {f:.3f} = true bytes / counter bytes"
        out += ["", head + ")", "", "| kernel | grid | dispatches | counter KiB | calibrated B/site |",
                "|---|---|---|---|---|"]
        items = sorted(((k, v) for k, v in means.items() if k[2] == counter), key=lambda kv: -kv[1])[:top]
        for (name, grid, _), v in items:
            per_site = f"{v * 1024.0 * f / sites:.1f}" if f else "-"
            out.append(f"| `{short(name)}` | {grid} | {counts[(name, grid, counter)]} | {v:.1f} | {per_site} |")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("dir")
    ap.add_argument("--command", required=True, help="the full command that produced DIR")
    ap.add_argument("--title", default="rocprofv3 summary")
    ap.add_argument("--sites", type=int, default=0, help="lattice sites of the profiled run, for B/site")
    ap.add_argument("--top", type=int, default=12)
    a = ap.parse_args(argv)
    out = [f"# {a.title}", "", "Command:", "", "```", a.command, "```"]
    p = find(a.dir, "_kernel_stats.csv")
    if p:
        out += ["", "## Kernel statistics (rocprofv3 --stats)", ""] + stats_table(p, a.top)
    p = find(a.dir, "_kernel_trace.csv")
    if p:
        out += ["", "## Dispatches grouped by kernel and grid (from the kernel trace)", ""] + trace_table(p, a.top)
    p = find(a.dir, "_counter_collection.csv")
    if p:
        out += ["", "## Counters"] + counter_table(p, a.sites, a.top)
    print("\n".join(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
