"""FP16-storage V-cycles (FP32 arithmetic) against the FP32 path and an FP64 oracle.

The preconditioner with FP16 storage is a perturbation of the reference's
(V-cycle values rounded to 11 significant bits), so the bar is on convergence
and on the error of the converged solution, stated per case:
  * one application of the preconditioner within 5e-3 relative (max-norm) of
    the FP32 one;
  * PCG meets the same stopping rule (max|r| <= tol * max|b|, tol = 1e-6 and
    1e-8) in at most max(2, 25%) more iterations than the FP32 path;
  * error of the solution against the oracle solved in double precision to
    1e-12: at most 10x the FP32 path's error, or 1e-6, whichever is larger.
"""
import numpy as np
import pytest

from hierarchicaldd_amd import domain
from tests.gpu_common import random_rhs, rel_err

pytestmark = pytest.mark.gpu

CASES = {
    "box32_sub16": lambda: domain.small_box(32, 16, 3, seed=1),
    "sphere64_sub16": domain.config0,
    "box48_sub16": lambda: domain.small_box(48, 16, 3, seed=2),
    # the 512^3 sparse workload's geometry (zero Dirichlet outside a union of spheres, N(0,1) rhs) at 96^3
    "sparse96_sub16": lambda: domain.sparse_spheres(96, 6, 6, 18, seed=2, sub=16, vlevels=3),
}


@pytest.fixture(scope="module", params=list(CASES))
def case(request, oracle_lib):
    from hierarchicaldd_amd import hdd
    dom, cfg = CASES[request.param]()
    bmax = float(np.abs(dom.b).max())
    o = oracle_lib.Oracle(dom, cfg, double=True)
    it_o, _, _ = o.solve(1e-12 * bmax, 1000)
    x_ref = o.x()
    print(f"[{request.param}] pages={dom.n_pages} fp64 oracle {it_o} iterations to 1e-12", flush=True)
    g = hdd.DDSolver(dom, cfg)
    yield request.param, dom, g, x_ref, bmax
    g.close()


def test_apply_fp16_close_to_fp32(case):
    name, dom, g, _, _ = case
    r = random_rhs(g.level(0)[1], 3)
    g.set_fp16(False)
    z32 = g.apply_preconditioner(r)
    g.set_fp16(True)
    z16 = g.apply_preconditioner(r)
    g.set_fp16(False)
    print(f"[{name}] apply fp16 vs fp32 rel diff {rel_err(z16, z32):.2e}")
    assert np.isfinite(z16).all()
    assert rel_err(z16, z32) < 5e-3


@pytest.mark.parametrize("tol", [1e-6, 1e-8])
def test_pcg_fp16_converges(case, tol):
    from hierarchicaldd_amd import hdd
    name, dom, g, x_ref, bmax = case
    g.set_fp16(False)
    it32, h32, _ = g.solve(tol * bmax, 1000)
    x32 = g.get(hdd.CH_X)
    g.set_fp16(True)
    it16, h16, _ = g.solve(tol * bmax, 1000)
    x16 = g.get(hdd.CH_X)
    g.set_fp16(False)
    e32, e16 = rel_err(x32, x_ref), rel_err(x16, x_ref)
    print(f"[{name} tol {tol:g}] fp32 {it32} iterations, error vs fp64 {e32:.2e}; fp16 storage {it16} iterations, error {e16:.2e}")
    assert it32 > 0 and it16 > 0
    assert h16[-1] <= tol * bmax
    assert it16 <= it32 + max(2, int(0.25 * it32))
    assert e16 <= max(10 * e32, 1e-6)
