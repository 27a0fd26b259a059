"""Summarise a rocprofv3 kernel trace into a per-kernel table.

Usage: python -m hierarchicaldd_amd.profile_summary <rocprof output dir> [--out file] [--cmd "..."]

Reads the rocpd SQLite database (or a *kernel_stats.csv) that
`rocprofv3 --kernel-trace --stats` writes and prints calls, total and average
duration and share per kernel, with the command that produced the trace.
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import sqlite3
import sys


def short(name: str) -> str:
    name = name.replace("(anonymous namespace)::", "")
    return name.split("(")[0]


def from_db(path):
    con = sqlite3.connect(path)
    rows = con.execute("select name, count(*), sum(duration), avg(duration) from kernels group by name").fetchall()
    con.close()
    return [(short(n), int(c), float(t) / 1e3, float(a) / 1e3) for n, c, t, a in rows]  # ns -> us


def from_csv(path):
    out = []
    with open(path) as f:
        for r in csv.DictReader(f):
            out.append((short(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]) / 1e3, float(r["AverageNs"]) / 1e3))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("dir")
    ap.add_argument("--out")
    ap.add_argument("--cmd", default="")
    a = ap.parse_args()
    dbs = glob.glob(os.path.join(a.dir, "**", "*_results.db"), recursive=True)
    csvs = glob.glob(os.path.join(a.dir, "**", "*kernel_stats.csv"), recursive=True)
    rows = from_db(dbs[0]) if dbs else from_csv(csvs[0])
    agg = {}
    for n, c, t, _ in rows:
        c0, t0 = agg.get(n, (0, 0.0))
        agg[n] = (c0 + c, t0 + t)
    total = sum(t for _, t in agg.values())
    lines = []
    if a.cmd:
        lines.append(f"command: {a.cmd}")
    lines.append(f"total kernel time: {total / 1e3:.1f} ms over {sum(c for c, _ in agg.values())} launches")
    lines.append(f"{'kernel':48s} {'calls':>9s} {'total ms':>10s} {'avg us':>9s} {'share':>7s}")
    for n, (c, t) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        lines.append(f"{n[:48]:48s} {c:9d} {t / 1e3:10.2f} {t / c:9.2f} {100 * t / total:6.1f}%")
    text = "\n".join(lines) + "\n"
    if a.out:
        with open(a.out, "w") as f:
            f.write(text)
    sys.stdout.write(text)


if __name__ == "__main__":
    main()
