"""The 256^3 sphere workload (sub_size 32, vlevels 4) on the GPU against the
oracle's answer recorded in tests/golden/sphere256_sub32.json (written by
tests/golden/make_sphere256.py: about an hour of CPU, so it is not rerun here).

Bars as in tests/test_gpu_pcg.py: with one refinement step after every sigma
solve (exact solves) the iteration count is within +-1 of the oracle's and the
max|r| history within 1e-5 relative; in the default single-precision mode the
iteration count is within +-1 and the history within 5e-4. One preconditioner
apply on the fixture's seeded residual and the solution x are compared on a
fixed sample of cells and by checksums.
"""
import json
import os

import numpy as np
import pytest

from hierarchicaldd_amd import domain
from tests.gpu_common import random_rhs

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sphere256_sub32.json")


@pytest.fixture(scope="module")
def case():
    from hierarchicaldd_amd import hdd
    if not os.path.exists(FIXTURE):
        pytest.skip("tests/golden/sphere256_sub32.json has not been generated")
    with open(FIXTURE) as f:
        gold = json.load(f)
    dom, cfg = domain.config1()
    assert dom.n_pages == gold["n_pages"]
    print(f"[sphere256_sub32] pages={dom.n_pages}: setting up", flush=True)
    g = hdd.DDSolver(dom, cfg)
    yield dom, gold, g
    g.close()


def _hist_rel(h, ref):
    n = min(len(h), len(ref))
    return float((np.abs(h[:n] - ref[:n]) / np.maximum(ref[:n], 1e-30)).max())


def _sample_rel(v, gold_sample, cells):
    a = np.asarray(v, dtype=np.float64).reshape(-1)[np.asarray(cells)]
    b = np.asarray(gold_sample, dtype=np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_preconditioner_apply(case):
    dom, gold, g = case
    r = random_rhs(g.level(0)[1], gold["apply_seed"])
    z = g.apply_preconditioner(r)
    e1 = _sample_rel(z, gold["z_sample"], gold["sample_cells"])
    g.set_sigma_refinement(1)
    try:
        zr = g.apply_preconditioner(r)
    finally:
        g.set_sigma_refinement(0)
    e2 = _sample_rel(zr, gold["z_sample"], gold["sample_cells"])
    s2 = abs(float(np.sum(zr, dtype=np.float64)) - gold["z"]["sum"]) / abs(gold["z"]["sum"])
    print(f"[sphere256_sub32] apply sample rel err: single {e1:.2e}, refined {e2:.2e}; refined sum rel {s2:.2e}")
    assert e1 < 2e-4
    assert e2 < 1e-5


def test_solve_matches_oracle_history(case):
    dom, gold, g = case
    tol = gold["tolerance"]
    ref = np.asarray(gold["history"])
    it, h, secs = g.solve(tol, 500)
    rel = _hist_rel(h, ref)
    xe = _sample_rel(g.get(3), gold["x_sample"], gold["sample_cells"])
    print(f"[sphere256_sub32] oracle {gold['iterations']} iterations; gpu {it} in {secs:.3f}s, history rel {rel:.2e}, x sample rel {xe:.2e}")
    assert it > 0 and abs(it - gold["iterations"]) <= 1
    assert h[-1] <= tol
    assert rel < 5e-4
    g.set_sigma_refinement(1)
    try:
        it_r, h_r, secs_r = g.solve(tol, 500)
        xr = g.get(3)
    finally:
        g.set_sigma_refinement(0)
    rel_r = _hist_rel(h_r, ref)
    xe_r = _sample_rel(xr, gold["x_sample"], gold["sample_cells"])
    print(f"[sphere256_sub32] refined: gpu {it_r} in {secs_r:.3f}s, history rel {rel_r:.2e}, x sample rel {xe_r:.2e}")
    assert it_r > 0 and abs(it_r - gold["iterations"]) <= 1
    assert rel_r < 1e-5
