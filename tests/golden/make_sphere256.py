"""Writes the oracle's answer for the 256^3 sphere workload (sphere radius 64 at
(128,128,128), delta 1.0 at (64,64,64), sub_size 32, vlevels 4, sigma levels 3,
the reference's default smoothing counts) to tests/golden/sphere256_sub32.json.

The oracle is the CPU restatement of the reference dd_pcg (oracle/), run once in
double-precision sigma mode. On an 8-core host the set-up takes about 22 minutes
and each PCG iteration about 35 seconds, so the GPU tests compare against this
fixture instead of running the oracle.

Recorded: the PCG iteration count and max|r| history at the stopping rule
max|r| <= 1e-6 max|b|; checksums of x (sum, sum of squares, max|x|) and x at a
fixed sample of cells; and one preconditioner apply z = M^{-1} r for a seeded
random r (the same generator as tests/gpu_common.random_rhs, seed 11), as
checksums and the same sample of cells.

Usage: OMP_NUM_THREADS=8 python tests/golden/make_sphere256.py
       OMP_NUM_THREADS=8 python tests/golden/make_sphere256.py --single
The second form runs the oracle with single-precision sigma solves (the
rounding of the reference's single-precision PARDISO, iparm[27] = 1) and writes
its iteration count and history to tests/golden/sphere256_sub32_single.json.
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sphere256_sub32.json")
SAMPLE_SEED = 5
SAMPLE_COUNT = 4096
APPLY_SEED = 11


def sample_cells(flags, seed=SAMPLE_SEED, count=SAMPLE_COUNT):
    """Fixed sample of Active|Interface cells as (page, cell) index pairs into [n_pages, 256] arrays."""
    from hierarchicaldd_amd import domain
    cells = np.flatnonzero((flags.reshape(-1) & (domain.ACTIVE | domain.INTERFACE)) != 0)
    rng = np.random.default_rng(seed)
    pick = np.sort(rng.choice(cells, size=min(count, cells.size), replace=False))
    return pick


def checksums(v):
    v = np.asarray(v, dtype=np.float64).reshape(-1)
    return {"sum": float(v.sum()), "sum_sq": float((v * v).sum()), "max_abs": float(np.abs(v).max())}


def main_single():
    from hierarchicaldd_amd import domain
    from oracle import pyoracle
    dom, cfg = domain.config1()
    t0 = time.time()
    o = pyoracle.Oracle(dom, cfg, sigma_single=True)
    t_setup = time.time() - t0
    print(f"oracle (single-precision sigma solves) setup {t_setup:.1f}s", flush=True)
    tol = 1e-6 * float(np.abs(dom.b).max())
    it, hist, secs = o.solve(tol, 500)
    print(f"oracle solve: {it} iterations, {secs:.1f}s", flush=True)
    out = {"workload": "sphere256_sub32",
           "producer": "tests/golden/make_sphere256.py --single (oracle, single-precision sigma solves)",
           "n_pages": int(dom.n_pages), "tolerance": tol, "iterations": int(it), "history": [float(h) for h in hist],
           "oracle_seconds": {"setup": t_setup, "solve": secs, "threads": int(os.environ.get("OMP_NUM_THREADS", "0") or 0)}}
    path = OUT.replace(".json", "_single.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", path)


def main():
    from hierarchicaldd_amd import domain
    from oracle import pyoracle
    from tests.gpu_common import random_rhs
    dom, cfg = domain.config1()
    t0 = time.time()
    o = pyoracle.Oracle(dom, cfg)
    t_setup = time.time() - t0
    print(f"oracle setup {t_setup:.1f}s", flush=True)
    lv0 = o.level(0)[1]
    pick = sample_cells(lv0)
    r = random_rhs(lv0, APPLY_SEED)
    t1 = time.time()
    z = o.apply_preconditioner(r)
    t_apply = time.time() - t1
    print(f"oracle apply {t_apply:.1f}s", flush=True)
    tol = 1e-6 * float(np.abs(dom.b).max())
    it, hist, secs = o.solve(tol, 500)
    print(f"oracle solve: {it} iterations, {secs:.1f}s", flush=True)
    x = o.x()
    out = {
        "workload": "sphere256_sub32",
        "producer": "tests/golden/make_sphere256.py (oracle, double-precision sigma solves)",
        "n_pages": int(dom.n_pages),
        "tolerance": tol,
        "iterations": int(it),
        "history": [float(h) for h in hist],
        "x": checksums(x),
        "sample_seed": SAMPLE_SEED,
        "sample_cells": [int(c) for c in pick],
        "x_sample": [float(v) for v in x.reshape(-1)[pick]],
        "apply_seed": APPLY_SEED,
        "z": checksums(z),
        "z_sample": [float(v) for v in z.reshape(-1)[pick]],
        "oracle_seconds": {"setup": t_setup, "apply": t_apply, "solve": secs, "threads": int(os.environ.get("OMP_NUM_THREADS", "0") or 0)},
    }
    with open(OUT, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", OUT)


if __name__ == "__main__":
    if "--single" in sys.argv[1:]:
        main_single()
    else:
        main()
