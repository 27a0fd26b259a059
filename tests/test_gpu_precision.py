"""FP64 outer PCG (hdd_set_pcg_precision): FP64 rho/alpha/beta, the iterate in
FP64, and true-residual replacement. A solve that stops must have
max|b - L x| <= tolerance for its FP64 iterate (computed here in float64 from
the flags, independently of the library), at 1e-6 and 1e-8 relative. The
error of x against the FP64 oracle (the CPU restatement in double precision,
solved to 1e-12) is printed for both precisions and bounded."""
import numpy as np
import pytest

from hierarchicaldd_amd import domain
from tests.reference_ops import face_residual

pytestmark = pytest.mark.gpu

CASES = {
    "box32_sub16": lambda: domain.small_box(32, 16, 3, seed=1),
    "sphere64_sub16": domain.config0,
    "sparse96_sub16": lambda: domain.sparse_spheres(96, 6, 6, 18, seed=2, sub=16, vlevels=3),
}


@pytest.fixture(scope="module", params=list(CASES))
def case(request, oracle_lib):
    from hierarchicaldd_amd import hdd
    dom, cfg = CASES[request.param]()
    print(f"[case {request.param}] pages={dom.n_pages}", flush=True)
    g = hdd.DDSolver(dom, cfg)
    o = oracle_lib.Oracle(dom, cfg, double=True)
    bmax = float(np.abs(dom.b).max())
    it_o, _, _ = o.solve(1e-12 * bmax, 1000)
    assert it_o > 0
    xs = o.x()
    yield request.param, dom, g, xs
    g.close()


@pytest.mark.parametrize("rel_tol", [1e-6, 1e-8])
def test_fp64_outer_meets_true_residual(case, rel_tol):
    name, dom, g, xs = case
    off, fl = g.level(0)
    b = g.get(4).astype(np.float64)
    tol = rel_tol * float(np.abs(b).max())
    it_f, h_f, _ = g.solve(tol, 500)
    x_f = g.get(3).astype(np.float64)
    res_f = float(np.abs(face_residual(off, fl, x_f, b)).max())
    g.set_outer_fp64(True, 25)
    try:
        it_d, h_d, secs = g.solve(tol, 500)
        x_d = g.x_f64()
        x_dc = g.get(3)
    finally:
        g.set_outer_fp64(False)
    res_d = float(np.abs(face_residual(off, fl, x_d, b)).max())
    scale = float(np.abs(xs).max())
    err_f = float(np.abs(x_f - xs).max()) / scale
    err_d = float(np.abs(x_d - xs).max()) / scale
    print(f"[{name} tol {rel_tol:g}] float outer: {it_f} it, true res {res_f / tol:.3f} tol, err {err_f:.2e}; "
          f"FP64 outer: {it_d} it, true res {res_d / tol:.3f} tol, err {err_d:.2e}")
    assert it_d > 0
    assert res_d <= tol * (1 + 1e-6)
    assert np.array_equal(x_dc, x_d.astype(np.float32))
    # the oracle's x solves to 1e-12: x errors are bounded by the conditioning times the residual
    assert err_d < 1e-2 and err_f < 1e-2
