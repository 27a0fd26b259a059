"""Summarise a rocprofv3 kernel trace into a per-kernel table.

Usage: python -m hierarchicaldd_amd.profile_summary <rocprof output dir> [--out file] [--cmd "..."]

Reads the rocpd SQLite database (or a *kernel_stats.csv) that
`rocprofv3 --kernel-trace --stats` writes and prints calls, total and average
duration and share per kernel, with the command that produced the trace.
With --by-grid (database input only) each kernel is split by launch grid size,
which separates the launches of one kernel on differently sized levels. With
--from KERNEL (database input only) only the dispatches from the first launch
of a kernel whose name contains KERNEL on are counted (e.g. k_laplace_pcg: the
solves, without the set-up).
With --pmc the input is the *counter_collection.csv of a `rocprofv3 --pmc`
run (written with --output-format csv): the mean value of each counter per
kernel and grid size is printed. With --pmc --per-solve N --copy-bytes B
(the run of `bench.py --plan-only`: calibration copies of B bytes read and B
written, then N solves of one triangular-solve plan) the counter is also
summed over every dispatch after the last k_copy and divided by N, and the
copy's mean value is printed beside B, so that the plan's HBM bytes per solve
are the sum scaled by B over the copy's value. With --per-solve N on a kernel
trace of the same run, the kernel time and the span (first start to last end)
per solve after the last k_copy are printed, with each kernel's share.
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


def grid_column(con):
    cols = [r[1] for r in con.execute("pragma table_info(kernels)").fetchall()]
    for c in cols:
        if "grid" in c.lower() and c.lower().endswith("x"):
            return c
    raise SystemExit(f"no grid-size column in the kernels view (columns: {cols})")


def start_column(con):
    cols = [r[1] for r in con.execute("pragma table_info(kernels)").fetchall()]
    for c in cols:
        if c.lower() == "start":
            return c
    raise SystemExit(f"no start column in the kernels view (columns: {cols})")


def from_db(path, by_grid=False, first=None):
    con = sqlite3.connect(path)
    where = ""
    if first:
        st = start_column(con)
        t0 = con.execute(f"select min({st}) from kernels where name like ?", (f"%{first}%",)).fetchone()[0]
        if t0 is None:
            raise SystemExit(f"no dispatch of a kernel named like {first}")
        where = f" where {st} >= {int(t0)}"
    if by_grid:
        g = grid_column(con)
        q = f"select name, {g}, count(*), sum(duration), avg(duration) from kernels{where} group by name, {g}"
        rows = [(f"{short(n)} [grid {gx}]", c, t, a) for n, gx, c, t, a in con.execute(q).fetchall()]
    else:
        rows = [(short(n), c, t, a) for n, c, t, a in
                con.execute(f"select name, count(*), sum(duration), avg(duration) from kernels{where} group by name").fetchall()]
    con.close()
    return [(n, int(c), float(t) / 1e3, float(a) / 1e3) for n, c, t, a in rows]  # ns -> us


def trace_per_solve(path, n_solves):
    """Kernel trace of `bench.py --plan-only`: the dispatches after the last k_copy are the N
    solves of one plan. Per solve: the summed kernel time, the span from the first dispatch's
    start to the last one's end (what HIP events around the solve measure, launch gaps
    included), and each kernel's dispatches and time."""
    con = sqlite3.connect(path)
    st = start_column(con)
    rows = con.execute(f"select name, {st}, duration from kernels order by {st}").fetchall()
    con.close()
    copies = [i for i, (n, t, d) in enumerate(rows) if short(n).endswith("k_copy")]
    if not copies:
        raise SystemExit("no k_copy dispatch: --per-solve needs the calibration copies before the solves")
    after = rows[copies[-1] + 1:]
    if not after:
        raise SystemExit("no dispatch after the last k_copy")
    total = sum(float(d) for _, _, d in after)
    span = max(float(t) + float(d) for _, t, d in after) - float(after[0][1])
    kernels = {}
    for n, _, d in after:
        k = kernels.setdefault(short(n), [0, 0.0])
        k[0] += 1
        k[1] += float(d)
    lines = [f"solves: {n_solves}, dispatches after the last copy: {len(after)}",
             f"kernel time per solve: {total / n_solves / 1e3:.1f} us",
             f"span per solve (first start to last end, launch gaps included): {span / n_solves / 1e3:.1f} us",
             f"{'kernel':48s} {'dispatches/solve':>16s} {'us/solve':>10s}"]
    for n, (c, t) in sorted(kernels.items(), key=lambda kv: -kv[1][1]):
        lines.append(f"{n[:48]:48s} {c / n_solves:16.1f} {t / n_solves / 1e3:10.1f}")
    return lines


def from_csv(path):
    out = []
    with open(path) as f:
        for r in csv.DictReader(f):
            out.append((short(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]) / 1e3, float(r["AverageNs"]) / 1e3))
    return out


def pmc_rows(path):
    acc = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            key = (short(r["Kernel_Name"]), int(r["Grid_Size"]), r["Counter_Name"])
            c, t = acc.get(key, (0, 0.0))
            acc[key] = (c + 1, t + float(r["Counter_Value"]))
    return acc


def per_solve_lines(files, n_solves, copy_bytes):
    rows = []
    for fpath in files:
        with open(fpath) as f:
            for r in csv.DictReader(f):
                rows.append((int(r["Dispatch_Id"]), short(r["Kernel_Name"]), r["Counter_Name"], float(r["Counter_Value"])))
    rows.sort()
    copies = [i for i, (d, n, c, v) in enumerate(rows) if n.endswith("k_copy")]
    if not copies:
        raise SystemExit("no k_copy dispatch: --per-solve needs the calibration copies before the solves")
    last = copies[-1]
    copy_vals = [v for d, n, c, v in rows if n.endswith("k_copy")]
    after = rows[last + 1:]
    total = sum(v for d, n, c, v in after)
    kernels = {}
    for d, n, c, v in after:
        k = kernels.setdefault(n, [0, 0.0])
        k[0] += 1
        k[1] += v
    lines = [f"per-solve KiB: {total / n_solves:.1f}", f"copy KiB: {sum(copy_vals) / len(copy_vals):.1f}", f"copy bytes: {copy_bytes:.0f}",
             f"solves: {n_solves}, dispatches after the last copy: {len(after)}",
             f"{'kernel':48s} {'dispatches/solve':>16s} {'KiB/solve':>12s}"]
    for n, (c, t) in sorted(kernels.items(), key=lambda kv: -kv[1][1]):
        lines.append(f"{n[:48]:48s} {c / n_solves:16.1f} {t / n_solves:12.1f}")
    return lines


def main_pmc(a):
    files = glob.glob(os.path.join(a.dir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {a.dir}")
    if a.per_solve:
        return ([f"command: {a.cmd}"] if a.cmd else []) + per_solve_lines(files, a.per_solve, a.copy_bytes)
    acc = {}
    for fpath in files:
        for k, (c, t) in pmc_rows(fpath).items():
            c0, t0 = acc.get(k, (0, 0.0))
            acc[k] = (c0 + c, t0 + t)
    lines = [f"command: {a.cmd}"] if a.cmd else []
    lines.append(f"{'kernel':48s} {'grid':>10s} {'counter':>12s} {'dispatches':>10s} {'mean value':>14s}")
    for (n, g, cn), (c, t) in sorted(acc.items(), key=lambda kv: -kv[1][1] / kv[1][0]):
        lines.append(f"{n[:48]:48s} {g:10d} {cn:>12s} {c:10d} {t / c:14.1f}")
    return lines


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("dir")
    ap.add_argument("--out")
    ap.add_argument("--cmd", default="")
    ap.add_argument("--by-grid", action="store_true")
    ap.add_argument("--top", type=int, default=0, help="print only the N largest rows")
    ap.add_argument("--pmc", action="store_true")
    ap.add_argument("--from", dest="first", default=None, help="count dispatches from the first launch of this kernel on")
    ap.add_argument("--per-solve", type=int, default=0,
                    help="counter sum (with --pmc) or kernel time and span (kernel trace) per solve after the last k_copy")
    ap.add_argument("--copy-bytes", type=float, default=0.0, help="with --per-solve: bytes read (and written) by one calibration copy")
    a = ap.parse_args()
    if a.pmc:
        text = "\n".join(main_pmc(a)) + "\n"
        if a.out:
            with open(a.out, "w") as f:
                f.write(text)
        sys.stdout.write(text)
        return
    dbs = glob.glob(os.path.join(a.dir, "**", "*_results.db"), recursive=True)
    csvs = glob.glob(os.path.join(a.dir, "**", "*kernel_stats.csv"), recursive=True)
    if a.first and not dbs:
        raise SystemExit("--from needs the rocpd database output")
    if a.per_solve:
        if not dbs:
            raise SystemExit("--per-solve on a kernel trace needs the rocpd database output")
        text = "\n".join(([f"command: {a.cmd}"] if a.cmd else []) + trace_per_solve(dbs[0], a.per_solve)) + "\n"
        if a.out:
            with open(a.out, "w") as f:
                f.write(text)
        sys.stdout.write(text)
        return
    rows = from_db(dbs[0], a.by_grid, a.first) if dbs else from_csv(csvs[0])
    agg = {}
    for n, c, t, _ in rows:
        c0, t0 = agg.get(n, (0, 0.0))
        agg[n] = (c0 + c, t0 + t)
    total = sum(t for _, t in agg.values())
    lines = []
    if a.cmd:
        lines.append(f"command: {a.cmd}")
    if a.first:
        lines.append(f"dispatches from the first {a.first} on")
    lines.append(f"total kernel time: {total / 1e3:.1f} ms over {sum(c for c, _ in agg.values())} launches")
    lines.append(f"{'kernel':64s} {'calls':>9s} {'total ms':>10s} {'avg us':>9s} {'share':>7s}")
    ranked = sorted(agg.items(), key=lambda kv: -kv[1][1])
    for n, (c, t) in ranked[:a.top] if a.top else ranked:
        lines.append(f"{n[:64]:64s} {c:9d} {t / 1e3:10.2f} {t / c:9.2f} {100 * t / total:6.1f}%")
    text = "\n".join(lines) + "\n"
    if a.out:
        with open(a.out, "w") as f:
            f.write(text)
    sys.stdout.write(text)


if __name__ == "__main__":
    main()
