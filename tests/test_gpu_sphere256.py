"""The 256^3 sphere workload (sub_size 32, vlevels 4) on the GPU against the
oracle's answer recorded in tests/golden/sphere256_sub32.json (written by
tests/golden/make_sphere256.py: about an hour of CPU, so it is not rerun here).

* One preconditioner apply on the fixture's seeded residual, on a fixed sample
  of cells: within 2e-4 in the default mode and 1e-5 with refined (exact) sigma
  solves.
* Reference-exact mode (refined sigma solves, the reference's 200 literal
  bottom smooths): the oracle's iteration count exactly, and its residual
  history bit for bit over the first iterations.
* Default mode: the oracle's iteration count +-1, and the solution meets the
  tolerance in float64 (max|b - A x| <= 10 tol from the input flags).

At this size the history after the first dozen iterations is not reproducible
at the 1e-5 level by any solver that does not round every operation as the
oracle does: a one-ulp difference (measured: at iteration 5 in the
reference-exact mode, from one sigma solve) grows by up to 70x per iteration
from iteration 12 on, as differences between the oracle's own single- and
double-precision sigma solves do: with the single-precision rounding of the
reference's PARDISO the oracle takes 64 iterations instead of 62, its history
within 4e-4 relative of the double one over the first 12 iterations and up
to 140% apart after (tests/golden/sphere256_sub32_single.json,
tests/test_oracle_pinning.py::test_sphere256_single_precision_oracle_fixture). The smaller cases of tests/test_gpu_pcg.py, where the
reference-exact mode reproduces the oracle's whole history bit for bit, carry
the 1e-5 history bar.
"""
import json
import os

import numpy as np
import pytest

from hierarchicaldd_amd import domain
from tests.gpu_common import random_rhs
from tests.reference_ops import face_residual

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sphere256_sub32.json")


@pytest.fixture(scope="module")
def case():
    if not os.path.exists(FIXTURE):
        pytest.skip("tests/golden/sphere256_sub32.json has not been generated")
    with open(FIXTURE) as f:
        gold = json.load(f)
    dom, cfg = domain.config1()
    assert dom.n_pages == gold["n_pages"]
    return dom, cfg, gold


def _sample_rel(v, gold_sample, cells):
    a = np.asarray(v, dtype=np.float64).reshape(-1)[np.asarray(cells)]
    b = np.asarray(gold_sample, dtype=np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


def _solver(dom, cfg, bottom=None):
    from hierarchicaldd_amd import hdd
    if bottom:
        os.environ["HDD_SIGMA_BOTTOM"] = bottom
    try:
        print(f"[sphere256_sub32] pages={dom.n_pages}: setting up (sigma bottom {bottom or 'auto'})", flush=True)
        return hdd.DDSolver(dom, cfg)
    finally:
        os.environ.pop("HDD_SIGMA_BOTTOM", None)


def test_preconditioner_apply(case):
    dom, cfg, gold = case
    g = _solver(dom, cfg)
    try:
        r = random_rhs(g.level(0)[1], gold["apply_seed"])
        z = g.apply_preconditioner(r)
        e1 = _sample_rel(z, gold["z_sample"], gold["sample_cells"])
        g.set_sigma_refinement(1)
        zr = g.apply_preconditioner(r)
    finally:
        g.close()
    e2 = _sample_rel(zr, gold["z_sample"], gold["sample_cells"])
    print(f"[sphere256_sub32] apply sample rel err: single {e1:.2e}, refined {e2:.2e}")
    assert e1 < 2e-4
    assert e2 < 1e-5


def test_solve_reference_exact_mode(case):
    dom, cfg, gold = case
    tol = gold["tolerance"]
    ref = np.asarray(gold["history"])
    g = _solver(dom, cfg, bottom="loop")
    try:
        g.set_sigma_refinement(1)
        it, h, secs = g.solve(tol, 500)
    finally:
        g.close()
    n = min(len(h), len(ref))
    rel = np.abs(h[:n] - ref[:n]) / ref[:n]
    same = int(np.argmax(rel > 0)) if (rel > 0).any() else n
    print(f"[sphere256_sub32] reference-exact mode: oracle {gold['iterations']} iterations, gpu {it} in {secs:.2f}s; "
          f"history identical over the first {same} iterations, max rel {rel.max():.2e}")
    assert it == gold["iterations"]
    assert h[-1] <= tol
    assert same >= 3


def test_solve_default_mode(case):
    dom, cfg, gold = case
    tol = gold["tolerance"]
    ref = np.asarray(gold["history"])
    g = _solver(dom, cfg)
    try:
        it, h, secs = g.solve(tol, 500)
        x = g.get(3)
    finally:
        g.close()
    xe = _sample_rel(x, gold["x_sample"], gold["sample_cells"])
    res = float(np.abs(face_residual(dom.page_offsets, dom.flags, x.astype(np.float64), dom.b.astype(np.float64))).max())
    n = min(len(h), len(ref))
    nrm = float(np.abs(h[:n] - ref[:n]).max() / ref[0])
    print(f"[sphere256_sub32] default mode: oracle {gold['iterations']} iterations; gpu {it} in {secs:.3f}s, history within "
          f"{nrm:.2e} of max|r0|, x sample rel {xe:.2e}, true residual {res / tol:.2f} tol")
    assert it > 0 and abs(it - gold["iterations"]) <= 1
    assert h[-1] <= tol
    assert res <= 10 * tol
