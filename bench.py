#!/usr/bin/env python3
"""Benchmark: DD-PCG solve of the 256^3 sphere domain on one MI355X.

A step is one full preconditioned-CG solve from x = 0 to max|r| <= 1e-6 max|b|
of the 256^3-cell sphere-obstacle problem (radius 64 at (128,128,128), delta
1.0 at (64,64,64), sub_size 32, vlevels 4, sigma levels 3, the reference's
default smoothing counts), with every SPGrid page, flag and interface-solver
plan resident in HBM before the timed region.

Printed JSON line: value = solves/s (1 / time-to-tolerance); the roofline
object covers the level-0 damped-Jacobi sweep kernel, timed with HIP events
on the context stream; "sigma_solves" gives the same for one exact solve of
each triangular-solve plan of the interface solver (time per solve, bytes of
the factor entries read); cpu_baseline is the oracle (the CPU restatement of the
reference dd_pcg) on a bounded sample: the complete 64^3 problem (setup +
solve), set beside the GPU's rate on the same 64^3 problem.

"fp16_storage" repeats the 256^3 solve with the V-cycle vectors stored in FP16
(FP32 arithmetic) and reports its time, iterations and level-0 sweep rate.

The other two parts of the metric are added to the same line:
  "batched": uneven independent problems (edge 64-256, 1-4 sphere obstacles,
      delta right-hand side, sub_size 32; --batch-count of them, 64 by default)
      solved one at a time on the GPU, then all at once in
      three forms: every problem's own context iterating at once on its own
      stream, one batched context (every stage one launch for all problems),
      and four batched contexts of similar-size problems iterating at once;
      solves/s of each (solve phase only: host setup is the same for all and is
      reported separately), the best form's speedup, and each batched result's
      difference from its solve alone;
  "served": a request stream (edge log-uniform 32-384, up to 40% blocked by
      spheres, N(0,1) right-hand side, tolerance 1e-6) with Poisson arrivals at
      the rates of --serve-rates, through the serving loop (host setup workers;
      continuous batching: units of the requests ready together iterate side by
      side on the GPU, and a unit admitted while others are mid-solve joins them
      at their next iteration); p50/p99 latency (submit to result, host setup
      included) and solves/s per rate, beside the plain path (each request set
      up and solved alone, one after another).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")
METRIC = "256³ time-to-1e-6 rel-res; batched solves/s; p50/p99 served latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def memory_in_use():
    """Device memory in use (hipMemGetInfo) and this process's resident host memory, in GiB."""
    import ctypes as C
    free, total = C.c_size_t(), C.c_size_t()
    dev = float("nan")
    try:
        if C.CDLL("libamdhip64.so").hipMemGetInfo(C.byref(free), C.byref(total)) == 0:
            dev = (total.value - free.value) / 2**30
    except OSError:
        pass
    rss = float("nan")
    with open("/proc/self/status") as f:
        for line in f:
            if line.startswith("VmRSS:"):
                rss = int(line.split()[1]) / 2**20
    return f"device {dev:.1f} GiB in use, host RSS {rss:.1f} GiB"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="sphere256", choices=["sphere256", "sphere64", "sparse512"],
                    help="sparse512: the 512^3 union-of-spheres domain, FP32 and FP16-storage V-cycles at 1e-6 and 1e-8")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--max-iterations", type=int, default=500)
    ap.add_argument("--batch-count", type=int, default=64, help="problems of the batched leg (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=32, help="requests per rate of the served leg (0: skip)")
    ap.add_argument("--serve-rates", default="8", help="comma-separated Poisson arrival rates (requests/s) of the served leg")
    ap.add_argument("--serve-nmax", type=int, default=256,
                    help="largest request edge of the served leg (384 for the full stream: its host set-up alone takes "
                         "about 6 minutes per pass over 32 requests, so the default run stops at 256)")
    ap.add_argument("--sweep-only", action="store_true",
                    help="set up, then run only level-0 sweeps and a plain copy of known size (for counter runs)")
    ap.add_argument("--plan-only", default=None, metavar="LEVEL,PLAN",
                    help="set up, then run 10 plain copies of known size and 1 + 10 solves of one triangular-solve "
                         "plan (PLAN: Aii, Arr or bottom), for counter runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding of subdomains is a design note, not an implementation)", file=sys.stderr)
        sys.exit(2)

    from hierarchicaldd_amd import domain, hdd

    if args.config == "sparse512":
        print(json.dumps(sparse_leg(hdd)), flush=True)
        return
    t0 = time.time()
    dom, cfg = domain.config1() if args.config == "sphere256" else domain.config0()
    log(f"[bench] domain {dom.meta.get('workload')} pages={dom.n_pages} generated in {time.time() - t0:.1f}s")
    t0 = time.time()
    solver = hdd.DDSolver(dom, cfg)
    st = solver.stats()
    log(f"[bench] setup {time.time() - t0:.1f}s (factorisation {st['factor_s']:.1f}s, interface-solver plans "
        f"{st['sigma_device_bytes'] / 2**30:.2f} GiB, bottom rho {st['bottom_rho']:.4f}, direct bottom {bool(st['bottom_direct'])}, "
        f"fused heights {st['fused_heights']:.0f} of which {st['wave_heights']:.0f} a wavefront per run)")
    tol = 1e-6 * float(np.abs(dom.b).max())

    if args.plan_only:
        lvl, plan = args.plan_only.split(",")
        copy_ms, copy_bytes = solver.profile_copy(10)
        ms, by = solver.profile_solve(int(lvl), plan, 10)
        log(f"[bench] calibration copy: {copy_bytes / 2:.0f} bytes read and {copy_bytes / 2:.0f} written per launch")
        log(f"[bench] level {lvl} {plan} solve: {ms * 1e3:.1f} us, {by / 1e6:.1f} MB algorithmic, 11 solves after the copies")
        return

    if args.sweep_only:
        sweep_ms, sweep_bytes = solver.profile_sweep(20)
        log(f"[bench] level-0 sweep: {sweep_ms * 1e3:.1f} us/launch, {sweep_bytes / 1e6:.1f} MB algorithmic")
        copy_ms, copy_bytes = solver.profile_copy(20)
        log(f"[bench] calibration copy: {copy_ms * 1e3:.1f} us/launch, {copy_bytes / 2:.0f} bytes read and {copy_bytes / 2:.0f} written")
        solver.set_fp16(True)
        h_ms, h_bytes = solver.profile_sweep(20)
        log(f"[bench] level-0 sweep, FP16 storage: {h_ms * 1e3:.1f} us/launch, {h_bytes / 1e6:.1f} MB algorithmic")
        return

    for w in range(args.warmup):
        it, hist, secs = solver.solve(tol, args.max_iterations)
        log(f"[bench] warmup {w}: {it} iterations, {secs:.3f}s")
    times, iters = [], []
    t_start = time.perf_counter()
    for k in range(args.steps):
        it, hist, secs = solver.solve(tol, args.max_iterations)
        times.append(secs)
        iters.append(it)
        log(f"[bench] step {k}: {it} iterations, {secs:.3f}s, final max|r| {hist[-1]:.3e}")
    wall = time.perf_counter() - t_start
    ms_per_step = 1000.0 * wall / args.steps

    # the level-0 smoothing sweep (the V-cycle kernel), HIP events on the context stream
    sweep_ms, sweep_bytes = solver.profile_sweep(100)
    sweep_gbs = sweep_bytes / (sweep_ms * 1e-3) / 1e9
    log(f"[bench] level-0 sweep: {sweep_ms * 1e3:.1f} us/launch, {sweep_bytes / 1e6:.1f} MB algorithmic, {sweep_gbs:.0f} GB/s")
    copy_ms, copy_bytes = solver.profile_copy(50)
    copy_gbs = copy_bytes / (copy_ms * 1e-3) / 1e9
    log(f"[bench] plain copy of one level-0 channel: {copy_ms * 1e3:.1f} us, {copy_gbs:.0f} GB/s (read + write)")
    # the same sweep on a level-0 grid larger than the 256 MB Infinity Cache: the 512^3 sparse
    # domain (30% of its 4^3 blocks Active), timed on its page set alone (no solver set-up)
    sweep512 = None
    if args.config == "sphere256":
        t0 = time.time()
        dom512, cfg512 = domain.sparse_spheres(target_coverage=0.30)
        ms512, by512 = hdd.profile_sweep_grid(dom512, cfg512, 100)
        gbs512 = by512 / (ms512 * 1e-3) / 1e9
        sweep512 = {"kernel": "k_smooth (level-0 damped-Jacobi sweep, FP32)", "workload": dom512.meta.get("workload"),
                    "pages": dom512.n_pages, "algorithmic_bytes": by512, "us_per_launch": ms512 * 1e3, "achieved": gbs512,
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs512 / HBM_PEAK_GBS}
        log(f"[bench] 512^3 sparse level-0 sweep ({dom512.n_pages} pages, {by512 / 1e6:.0f} MB): {ms512 * 1e3:.1f} us/launch, "
            f"{gbs512:.0f} GB/s, frac {gbs512 / HBM_PEAK_GBS:.2f} ({time.time() - t0:.1f}s with the domain)")
        del dom512

    # exact solves of the interface solver's triangular-solve plans: mean time per solve with HIP
    # events, algorithmic bytes = 4 B per factor entry per pass, two passes (subdomains that share
    # one factor counted each); solves per PCG iteration from the sigma V-cycle's schedule
    sigma_solves = []
    n_sigma, _ = solver.sigma_info()
    plans = [(l, w) for l in range(n_sigma) for w in ("Aii", "Arr")] + ([(n_sigma - 1, "bottom")] if st["bottom_direct"] else [])
    ms_per_it = 1000.0 * float(np.mean(times)) / max(1, int(np.median(iters)))
    for l, w in plans:
        ms, by = solver.profile_solve(l, w, 20)
        gbs = by / (ms * 1e-3) / 1e9
        per_it = solves_per_iteration(l, w, n_sigma, cfg, bool(st["bottom_direct"]))
        sigma_solves.append({"level": l, "plan": w, "us_per_solve": ms * 1e3, "algorithmic_MB": by / 1e6,
                             "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS, "solves_per_iteration": per_it,
                             "share_of_iteration": per_it * ms / ms_per_it})
        log(f"[bench] sigma level {l} {w} solve: {ms * 1e3:.0f} us, {by / 1e6:.0f} MB algorithmic, {gbs:.0f} GB/s, "
            f"{per_it} per iteration ({100 * per_it * ms / ms_per_it:.1f}% of it)")
    dom_plan = max(sigma_solves, key=lambda r: r["share_of_iteration"])
    dofs = dof_updates_per_iteration(solver, cfg)

    # FP16-storage V-cycles (FP32 arithmetic) on the same problem
    solver.set_fp16(True)
    solver.solve(tol, args.max_iterations)
    t16, it16 = [], []
    for k in range(max(1, min(args.steps, 2))):
        it, hist, secs = solver.solve(tol, args.max_iterations)
        t16.append(secs)
        it16.append(it)
    h_ms, h_bytes = solver.profile_sweep(100)
    solver.set_fp16(False)
    fp16 = {"time_to_tolerance_s": float(np.mean(t16)), "solves_per_s": 1.0 / float(np.mean(t16)), "iterations": int(np.median(it16)),
            "sweep_us": h_ms * 1e3, "sweep_achieved_GBs": h_bytes / (h_ms * 1e-3) / 1e9,
            "sweep_frac": h_bytes / (h_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "sweep_bytes_per_slot": 10}
    fp16["traffic"] = counter_traffic(dom.n_pages, "k_smooth<__half>") if args.config == "sphere256" else None
    log(f"[bench] fp16-storage V-cycles: {fp16['iterations']} iterations, {fp16['time_to_tolerance_s']:.3f}s; "
        f"level-0 sweep {h_ms * 1e3:.1f} us ({fp16['sweep_achieved_GBs']:.0f} GB/s algorithmic at 10 B/slot)")
    solver.close()
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(hdd)
    batched = batched_leg(hdd, args.batch_count) if args.batch_count > 0 else None
    served = served_leg(hdd, args.serve_requests, [float(r) for r in args.serve_rates.split(",")], args.serve_nmax) \
        if args.serve_requests > 0 else None

    out = {
        "metric": METRIC,
        "value": 1.0 / (np.mean(times)),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 dot products)",
        "data": "synthetic",
        "config": {"workload": dom.meta.get("workload"), "cells": "256^3", "pages": dom.n_pages,
                   "sub_size": cfg.sub_size, "vlevels": cfg.vlevels, "sigma_levels": cfg.sigma_levels,
                   "tolerance": "max|r| <= 1e-6 max|b|", "iterations": int(np.median(iters)),
                   "time_to_tolerance_s": float(np.mean(times)), "setup_s": st["setup_s"], "factor_s": st["factor_s"]},
        "roofline": {"kernel": f"level-{dom_plan['level']} {dom_plan['plan']} triangular-solve plan (k_tri_* / k_top_* chain, "
                               f"{100 * dom_plan['share_of_iteration']:.0f}% of a PCG iteration)",
                     "bound": "hbm", "achieved": dom_plan["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom_plan["achieved_GBs"] / HBM_PEAK_GBS,
                     "traffic": plan_counter_traffic(dom_plan["level"], dom_plan["plan"]) if args.config == "sphere256" else None,
                     "algorithmic_bytes": dom_plan["algorithmic_MB"] * 1e6, "us_per_solve": dom_plan["us_per_solve"]},
        "vcycle_sweep": {"kernel": "k_smooth (level-0 damped-Jacobi sweep)", "achieved": sweep_gbs, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": sweep_gbs / HBM_PEAK_GBS, "us_per_launch": sweep_ms * 1e3,
                         "algorithmic_bytes": sweep_bytes,
                         "traffic": counter_traffic(dom.n_pages) if args.config == "sphere256" else None},
        "vcycle_sweep_512": sweep512,
        "stream_copy_GBs": copy_gbs,
        "dof_updates_per_s": dofs * int(np.median(iters)) / float(np.mean(times)),
        "dof_updates_per_iteration": dofs,
        "sigma_solves": sigma_solves,
        "cpu_baseline": cpu,
        "fp16_storage": fp16,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(out), flush=True)


def solves_per_iteration(level, plan, n_sigma, cfg, bottom_direct):
    """Exact solves of one plan per PCG iteration (one preconditioner apply, sigma_mg V-cycles):
    above the bottom, sigma_smoothing smooths down and up (each one Aii and one Arr solve) and one
    residual (one Aii solve); the bottom's 200 smooths are one combined solve when it is direct."""
    s, mg = cfg.sigma_smoothing, cfg.sigma_mg
    if plan == "bottom":
        return mg if bottom_direct else 0
    if level < n_sigma - 1:
        return mg * (2 * s + 1 if plan == "Aii" else 2 * s)
    return 0 if bottom_direct else mg * 200


def dof_updates_per_iteration(solver, cfg):
    """Damped-Jacobi corrections of single Active cells per PCG iteration (one preconditioner apply =
    2 * nVcycles V-cycles): on each level above the bottom, boundary_smoothing sweeps of the
    Active&Boundary cells before and after interior_smoothing sweeps of the other Active cells, going
    down and coming up; 200 sweeps of every Active cell at the bottom (LIB/SPGrid_V_Cycle_Helper.h:156-192)."""
    from hierarchicaldd_amd import domain
    per_vcycle = 0
    for l in range(cfg.vlevels):
        fl = solver.level(l)[1]
        act = (fl & domain.ACTIVE) != 0
        bnd = int((act & ((fl & domain.BOUNDARY) != 0)).sum())
        inner = int(act.sum()) - bnd
        if l < cfg.vlevels - 1:
            per_vcycle += 2 * (2 * cfg.boundary_smoothing * bnd + cfg.interior_smoothing * inner)
        else:
            per_vcycle += 200 * int(act.sum())
    return 2 * cfg.nVcycles * per_vcycle


def plan_counter_traffic(level, plan):
    """HBM bytes of one solve of a triangular-solve plan from the committed rocprofv3 counter
    summaries (profiles/tri_plan_sphere256_l<level>_<plan>_{fetch,write}_size.txt): FETCH_SIZE and
    WRITE_SIZE summed over every kernel dispatch of the measured solves, divided by the solves, each
    scaled by what the counter reads for the calibration copy of the same run. None if absent."""
    base = os.path.join(PROFILES, f"tri_plan_sphere256_l{level}_{plan.lower()}")
    vals = {}
    for counter, suffix in (("FETCH_SIZE", "fetch_size"), ("WRITE_SIZE", "write_size")):
        path = f"{base}_{suffix}.txt"
        if not os.path.exists(path):
            return None
        for line in open(path):
            if line.startswith("per-solve KiB") or line.startswith("copy KiB") or line.startswith("copy bytes"):
                k, v = line.split(":")
                vals[(counter, k.strip())] = float(v)
    try:
        f = vals[("FETCH_SIZE", "per-solve KiB")] * vals[("FETCH_SIZE", "copy bytes")] / (1024.0 * vals[("FETCH_SIZE", "copy KiB")])
        w = vals[("WRITE_SIZE", "per-solve KiB")] * vals[("WRITE_SIZE", "copy bytes")] / (1024.0 * vals[("WRITE_SIZE", "copy KiB")])
    except KeyError:
        return None
    return 1024.0 * (f + w)


def sweep_grid_threads(pages):
    """Launch grid (threads) of the level-0 sweep: eight pages per 512-thread workgroup, XCD-chunked."""
    groups = (pages + 7) // 8
    return 8 * ((groups + 7) // 8) * 512


def counter_traffic(pages, kernel="k_smooth<float>"):
    """HBM bytes of one level-0 sweep from the committed rocprofv3 counter summaries.

    FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of the sweep kernel at its
    level-0 grid, each scaled by what the same counter reads for the
    calibration copy kernel of the same runs, which reads and writes
    (pages + 1) * 256 floats per launch (hdd_profile_copy). None when the
    summaries are absent or were taken on another page count.
    """
    def read(name, counter):
        path = os.path.join(PROFILES, name)
        if not os.path.exists(path):
            return None
        vals = {}
        for line in open(path):
            f = line.split()
            if len(f) >= 5 and f[-3] == counter and f[-4].isdigit():
                vals.setdefault(" ".join(f[:-4]).split("::")[-1], {})[int(f[-4])] = float(f[-1])
        return vals
    fetch = read("k_smooth_sphere256_fetch_size.txt", "FETCH_SIZE")
    write = read("k_smooth_sphere256_write_size.txt", "WRITE_SIZE")
    if not fetch or not write or kernel not in fetch or kernel not in write or "k_copy" not in fetch:
        return None
    copy_kib = (pages + 1) * 256 * 4 / 1024.0
    copy_grid = max(fetch["k_copy"], key=lambda g: fetch["k_copy"][g])
    f_scale = copy_kib / fetch["k_copy"][copy_grid]
    w_scale = copy_kib / write["k_copy"][copy_grid]
    grid = sweep_grid_threads(pages)
    if grid not in fetch[kernel] or grid not in write[kernel]:
        return None
    return 1024.0 * (fetch[kernel][grid] * f_scale + write[kernel][grid] * w_scale)


def true_residual(dom, x):
    """max|b - A x| / max|b| in float64 with the 7-point operator of the input flags
    (face f of an Active cell contributes x_c - x_n; Dirichlet neighbours are zero)."""
    from hierarchicaldd_amd import domain
    fl = dom.dense(dom.flags)
    xd = dom.dense(np.asarray(x).reshape(dom.n_pages, 256)).astype(np.float64)
    bd = dom.dense(dom.b).astype(np.float64)
    act = (fl & domain.ACTIVE) != 0
    ax = np.zeros_like(xd)
    xp = np.pad(xd, 1)
    X, Y, Z = xd.shape
    nbrs = [xp[0:X, 1:Y + 1, 1:Z + 1], xp[2:X + 2, 1:Y + 1, 1:Z + 1], xp[1:X + 1, 0:Y, 1:Z + 1],
            xp[1:X + 1, 2:Y + 2, 1:Z + 1], xp[1:X + 1, 1:Y + 1, 0:Z], xp[1:X + 1, 1:Y + 1, 2:Z + 2]]
    for f, nb in enumerate(nbrs):
        on = (fl & np.uint32(domain.FACE_MX << f)) != 0
        ax += np.where(on, xd - nb, 0.0)
    r = np.where(act, bd - ax, 0.0)
    return float(np.abs(r).max() / np.abs(bd).max())


def sparse_leg(hdd):
    """The 512^3 union-of-spheres domain: FP32 and FP16-storage V-cycles at 1e-6 and 1e-8."""
    from hierarchicaldd_amd import domain
    t0 = time.time()
    dom, cfg = domain.sparse_spheres(target_coverage=0.30)
    meta = {k: v for k, v in dom.meta.items() if k != "spheres"}
    log(f"[sparse] {meta} pages={dom.n_pages} generated in {time.time() - t0:.1f}s")
    t0 = time.time()
    h = hdd.DDSolver(dom, cfg, host_only=True)      # pages, flags, interface operators
    log(f"[sparse] host setup {time.time() - t0:.1f}s")
    t1 = time.time()
    g = hdd.DDBatch(from_solvers=[h])               # factors, plans, device upload
    h.close()
    st = g.stats()
    log(f"[sparse] factors and device setup {time.time() - t1:.1f}s (interface dofs {st['interface_dofs']:.0f}, "
        f"plans {st['sigma_device_bytes'] / 2**30:.2f} GiB)")
    st["setup_s"] = time.time() - t0
    bmax = float(np.abs(dom.b).max())
    rows = []
    for fp16 in (False, True):
        g.set_fp16(fp16)
        for tol in (1e-6, 1e-8):
            for outer in ("fp32", "fp64"):
                g.set_outer_fp64(outer == "fp64")
                it, hist, secs = g.solve(tol * bmax, 2000)
                x = g.x_f64() if outer == "fp64" else g.get(hdd.CH_X)
                res = true_residual(dom, x)
                row = {"storage": "fp16" if fp16 else "fp32", "outer": outer, "tolerance": tol, "iterations": it, "seconds": secs,
                       "final_max_r_rel": float(hist[-1] / bmax), "true_residual_rel_fp64": res}
                log(f"[sparse] {row}")
                rows.append(row)
    g.set_outer_fp64(False)
    ms, by = g.profile_sweep(50)
    g.set_fp16(False)
    ms32, by32 = g.profile_sweep(50)
    g.close()
    return {"metric": METRIC, "config": {**meta, "pages": dom.n_pages, "sub_size": cfg.sub_size,
                                         "vlevels": cfg.vlevels}, "setup_s": st["setup_s"], "solves": rows,
            "sweep_fp32": {"us": ms32 * 1e3, "GBs": by32 / (ms32 * 1e-3) / 1e9}, "sweep_fp16": {"us": ms * 1e3, "GBs": by / (ms * 1e-3) / 1e9}}


def batched_leg(hdd, count, seed=0):
    """Uneven problems one at a time on the GPU, then all of them in one batched context."""
    from hierarchicaldd_amd import domain
    t0 = time.time()
    probs = domain.batch_problems(count, seed=seed)
    log(f"[batched] {count} problems generated in {time.time() - t0:.1f}s "
        f"(edges {sorted(d.meta['n'] for d, _ in probs)}, {sum(d.n_pages for d, _ in probs)} pages)")
    # set-up of every problem alone (host setup, factors, device plans), four at a time: the
    # library releases the interpreter lock, and set-up is not part of either solve rate
    from concurrent.futures import ThreadPoolExecutor
    t0 = time.time()
    share = max(1, len(os.sched_getaffinity(0)) // 4)

    def setup_alone(p):
        hdd.set_host_threads(share)  # four set-ups side by side, each on its share of the host
        return hdd.DDSolver(*p)

    with ThreadPoolExecutor(max_workers=4) as ex:
        singles = list(ex.map(setup_alone, probs))
    setup_wall = time.time() - t0
    plain_s, setup_s, iters, xs, alone_s = 0.0, 0.0, [], [], []
    for (dom, cfg), g in zip(probs, singles):
        setup_s += g.setup_seconds
        it, hist, secs = g.solve(1e-6 * float(np.abs(dom.b).max()), 1000)
        plain_s += secs
        alone_s.append(secs)
        iters.append(it)
        xs.append(g.get(hdd.CH_X))
    by_edge = {}
    for (dom, _), secs in zip(probs, alone_s):
        by_edge.setdefault(dom.meta["n"], []).append(secs)
    log("[batched] plain solve seconds by edge: " + ", ".join(f"{n}: {len(v)} x {np.mean(v):.3f}s = {sum(v):.2f}s"
                                                            for n, v in sorted(by_edge.items())))
    log(f"[batched] set-up of {count} problems alone: {setup_wall:.1f}s wall, four at a time; {memory_in_use()}")
    log(f"[batched] plain: {count} solves in {plain_s:.3f}s ({count / plain_s:.2f} solves/s), host setup {setup_s:.1f}s")
    tol = [1e-6 * float(np.abs(d.b).max()) for d, _ in probs]
    # the same contexts, every one iterating at once on its own stream (hdd_batches_pcg_solve)
    hdd.solve_concurrently(singles, tol, 1000)  # warm-up
    t0 = time.perf_counter()
    res_c = hdd.solve_concurrently(singles, tol, 1000)
    conc_s = time.perf_counter() - t0
    log(f"[batched] {count} contexts iterating at once: {conc_s:.3f}s ({count / conc_s:.2f} solves/s, x {plain_s / conc_s:.2f} the plain path)")
    # four batched contexts of problems of similar size iterating at once, then one batched context
    # (every stage one launch for all problems); device memory holds at most two copies of the set
    order = sorted(range(count), key=lambda p: probs[p][0].n_pages)
    groups = [sorted(order[k * count // 4:(k + 1) * count // 4]) for k in range(4)] if count >= 8 else [list(range(count))]
    t0 = time.time()
    bts = [hdd.DDBatch(from_solvers=[singles[p] for p in grp]) for grp in groups]
    build_s = time.time() - t0
    log(f"[batched] size-grouped batches built: {memory_in_use()}")
    for g in singles:
        g.close()
    gtol = [tol[p] for grp in groups for p in grp]
    hdd.solve_concurrently(bts, gtol, 1000)  # warm-up
    t0 = time.perf_counter()
    res_g = hdd.solve_concurrently(bts, gtol, 1000)
    grouped_s = time.perf_counter() - t0
    git = [r[0] for r in res_g]
    log(f"[batched] {len(bts)} size-grouped batches at once: {grouped_s:.3f}s ({count / grouped_s:.2f} solves/s, x {plain_s / grouped_s:.2f})")
    bt = hdd.DDBatch(from_solvers=bts)
    log(f"[batched] one batch built: {memory_in_use()}")
    for b in bts:
        b.close()
    perm = [p for grp in groups for p in grp]  # problem order inside bt
    bt.solve_all([tol[p] for p in perm], 1000)  # warm-up
    t0 = time.perf_counter()
    res = bt.solve_all([tol[p] for p in perm], 1000)
    batch_s = time.perf_counter() - t0
    bit = [0] * count
    for k, p in enumerate(perm):
        bit[p] = res[k][0]
    xdiff = max(float(np.abs(bt.x(k) - xs[p]).max() / max(np.abs(xs[p]).max(), 1e-30)) for k, p in enumerate(perm))
    bt.close()
    log(f"[batched] one batch: {count} solves in {batch_s:.3f}s ({count / batch_s:.2f} solves/s, x {plain_s / batch_s:.2f} the plain path), "
        f"iterations alone {iters} batched {bit}, max x rel diff {xdiff:.2e}")
    forms = {"one_batch": batch_s, "contexts_at_once": conc_s, "size_grouped_batches_at_once": grouped_s}
    best = min(forms, key=forms.get)
    it_c = [r[0] for r in res_c]
    return {"problems": count, "seed": seed, "plain_solves_per_s": count / plain_s, "batched_solves_per_s": count / forms[best],
            "batched_form": best, "speedup": plain_s / forms[best],
            "forms_solves_per_s": {k: count / v for k, v in forms.items()},
            "plain_solve_s": plain_s, "batched_solve_s": forms[best], "host_setup_s": setup_s, "setup_wall_s": setup_wall,
            "batch_device_build_s": build_s,
            "max_iteration_diff": {"one_batch": int(max(abs(a - b) for a, b in zip(iters, bit))),
                                   "size_grouped_batches_at_once": int(max(abs(iters[p] - git[k]) for k, p in enumerate(perm))),
                                   "contexts_at_once": int(max(abs(a - b) for a, b in zip(iters, it_c)))},
            "iterations_off_by_more_than_one": {"one_batch": int(sum(abs(a - b) > 1 for a, b in zip(iters, bit))),
                                                "size_grouped_batches_at_once": int(sum(abs(iters[p] - git[k]) > 1 for k, p in enumerate(perm)))},
            "contexts_at_once_iterations_equal": bool(it_c == iters),
            "max_x_rel_diff_vs_alone": xdiff, "all_converged": bool(min(bit) >= 0 and min(it_c) >= 0 and min(git) >= 0)}


def served_leg(hdd, n_req, rates, nmax, seed=1):
    """Poisson request stream through the serving loop, and the same requests set up and solved one by one."""
    from hierarchicaldd_amd import domain
    rng = np.random.default_rng(seed)
    t0 = time.time()
    reqs = [domain.served_request(rng, nmax=nmax) for _ in range(n_req)]
    log(f"[served] {n_req} requests generated in {time.time() - t0:.1f}s (edges {[d.meta['n'] for d, _ in reqs]})")
    tols = [1e-6 * float(np.abs(d.b).max()) for d, _ in reqs]
    # plain path: each request set up and solved alone, one after another
    t0 = time.perf_counter()
    plain_lat = []
    for k, ((dom, cfg), tol) in enumerate(zip(reqs, tols)):
        t1 = time.perf_counter()
        g = hdd.DDSolver(dom, cfg)
        it, _, secs = g.solve(tol, 1000)
        g.close()
        plain_lat.append(time.perf_counter() - t1)
        log(f"[served] plain request {k} (edge {dom.meta['n']}): {it} iterations, solve {secs:.3f}s, with setup {plain_lat[-1]:.2f}s")
    plain_s = time.perf_counter() - t0
    log(f"[served] plain: {n_req} requests in {plain_s:.1f}s ({n_req / plain_s:.3f} solves/s)")
    out = {"requests": n_req, "seed": seed, "edge_max": nmax, "plain_solves_per_s": n_req / plain_s,
           "plain_p50_s": float(np.percentile(plain_lat, 50)), "rates": []}
    for lam in rates:
        arr = np.cumsum(rng.exponential(1.0 / lam, size=n_req))
        srv = hdd.Server(setup_threads=4, groups=4, max_batch=16)
        t0 = time.perf_counter()
        ids = []
        for (dom, cfg), tol, ta in zip(reqs, tols, arr):
            dt = ta - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(dom, cfg, tol, 1000))
        reps = []
        for k, i in enumerate(ids):
            reps.append(srv.wait(i)[2])
            log(f"[served] rate {lam}/s request {k}: latency {reps[-1]['latency_s']:.2f}s (unit of {reps[-1]['batch_size']}, "
                f"{reps[-1]['concurrent']} units on the device when it started)")
        span = time.perf_counter() - t0
        srv.close()
        lat = [r["latency_s"] for r in reps]
        ok = all(r["iterations"] >= 0 for r in reps)
        row = {"rate_per_s": lam, "solves_per_s": n_req / span, "p50_latency_s": float(np.percentile(lat, 50)),
               "p99_latency_s": float(np.percentile(lat, 99)), "mean_setup_s": float(np.mean([r["setup_s"] for r in reps])),
               "mean_queue_s": float(np.mean([r["queue_s"] for r in reps])), "mean_solve_s": float(np.mean([r["solve_s"] for r in reps])),
               "max_group": int(max(r["batch_size"] for r in reps)),
               "max_concurrent_units": int(max(r["concurrent"] for r in reps)),
               "joined_mid_solve": int(sum(r["concurrent"] >= 2 for r in reps)), "all_converged": ok}
        log(f"[served] rate {lam}/s: {row}")
        out["rates"].append(row)
    return out


def cpu_info():
    """CPU model, sockets and physical cores from lscpu (None where lscpu is absent)."""
    import subprocess
    info = {"model": None, "physical_cores": None, "logical_cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    kv = {}
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    info["model"] = kv.get("Model name")
    try:
        info["physical_cores"] = int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
    except (KeyError, ValueError):
        pass
    return info


def cpu_baseline(hdd):
    """The oracle (the CPU restatement of the reference dd_pcg) on a bounded sample with the
    256^3 workload's DD parameters (sub_size 32, vlevels 4) on the 64^3 sphere domain, with the
    host's OpenMP threads and with one thread, beside the GPU on the same sample."""
    from hierarchicaldd_amd import domain
    from oracle import pyoracle
    pyoracle.build()
    dom, cfg = domain.sphere_case(64, 32, 4)
    tol = 1e-6 * float(np.abs(dom.b).max())
    info = cpu_info()
    threads = pyoracle.lib().ora_max_threads()
    runs = {}
    for nt in (threads, 1):
        prev = pyoracle.set_threads(nt)
        try:
            t0 = time.perf_counter()
            o = pyoracle.Oracle(dom, cfg)
            setup = time.perf_counter() - t0
            it, hist, secs = o.solve(tol, 500)
            del o
        finally:
            pyoracle.set_threads(prev)
        runs[nt] = {"threads": nt, "setup_s": setup, "solve_s": secs, "iterations": it}
        log(f"[bench] cpu baseline ({nt} threads): setup {setup:.1f}s, solve {secs:.2f}s ({it} iterations)")
    g = hdd.DDSolver(dom, cfg)
    g.solve(tol, 500)
    git, gh, gsecs = g.solve(tol, 500)
    g_setup = g.setup_seconds
    g.close()
    full = runs[threads]
    log(f"[bench] gpu on the same sample: solve {gsecs:.4f}s ({git} iterations), setup {g_setup:.1f}s")
    return {"value": 1.0 / full["solve_s"], "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"oracle dd_pcg (CPU restatement of the reference) on the 64^3 sphere domain with the 256^3 "
                      f"workload's DD parameters (sub_size 32, vlevels 4, 8 subdomains): {full['iterations']} iterations, "
                      f"solve {full['solve_s']:.2f}s after {full['setup_s']:.1f}s setup on {threads} threads; one thread: "
                      f"solve {runs[1]['solve_s']:.2f}s, setup {runs[1]['setup_s']:.1f}s; the GPU solves the same problem "
                      f"in {gsecs:.4f}s ({git} iterations)",
            "cpu_model": info["model"], "physical_cores": info["physical_cores"], "logical_cpus": info["logical_cpus"],
            "runs": list(runs.values()), "gpu_same_sample_solves_per_s": 1.0 / gsecs, "gpu_same_sample_setup_s": g_setup}


if __name__ == "__main__":
    main()
