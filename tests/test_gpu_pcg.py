"""End-to-end DD-PCG on the GPU against the oracle's CPU dd_pcg.

Two accuracy modes of the interface solver, each with its bar:

* reference-exact mode: every sigma solve followed by one step of iterative
  refinement (exact solves, as the oracle's double-precision factors give) and
  the bottom of the sigma V-cycle run as the reference's 200 literal smooths
  (HDD_SIGMA_BOTTOM=loop). The residual history (max|r| per iteration) must
  match the oracle's within 1e-5 relative at every iteration, and the
  iteration count exactly; measured: identical to the last bit on every case.
* the default (fast) mode: FP32 factors, one solve each (the precision of the
  reference's single-precision PARDISO) and the bottom as one exact solve plus
  a low-rank correction (its 200 smooths evaluated in closed form). Iteration
  count within +-1, and the history within 1e-6 of the initial max|r| at every
  iteration. A per-iteration relative bar does not apply to single-precision
  solves: rounding the oracle's own sigma solves to float moves its late
  iterations by up to 41% on the sub_size-32 cases
  (tests/test_oracle_pinning.py::test_single_precision_sigma_solves_move_histories),
  while staying within 1e-6 of max|r0|.

Both stop at max|r| <= 1e-6 max|b|. Cases include the 64^3 workload itself and
the 256^3 workload's DD parameters (sub_size 32, vlevels 4) at 64^3 and 96^3.
"""
import numpy as np
import pytest

from hierarchicaldd_amd import domain
from tests.gpu_common import random_rhs, rel_err

pytestmark = pytest.mark.gpu


def _refgen_via_disk():
    """The reference 3D generator's domain at 64^3, written to and read back from
    the reference's SPGrid directory format (what its dd_pcg reads)."""
    import tempfile
    from hierarchicaldd_amd import spgrid_io
    dom, cfg = domain.reference_domain_3d(64, delta=(16, 16, 16))
    with tempfile.TemporaryDirectory() as d:
        spgrid_io.write_domain(d, dom)
        back = spgrid_io.read_domain(d, dom.size, meta=dom.meta)
    return back, cfg


CASES = {
    "box32_sub16": lambda: domain.small_box(32, 16, 3, seed=1),
    "refgen64_sub16": _refgen_via_disk,
    "sphere64_sub16": domain.config0,
    # the sparse-spheres geometry (zero Dirichlet outside a union of spheres, N(0,1) rhs) at 96^3
    "sparse96_sub16": lambda: domain.sparse_spheres(96, 6, 6, 18, seed=2, sub=16, vlevels=3),
    # sub_size 32 (adaptive strides up to 8), vlevels 4: the 256^3 workload's parameters
    "box64_sub32": lambda: domain.small_box(64, 32, 4, seed=3),
    "sphere96_sub32": lambda: domain.sphere_case(96, 32, 4, delta=(32, 32, 32)),
}

HISTORY_BAR_EXACT = 1e-5         # reference-exact mode, relative per iteration
HISTORY_BAR_NORMALISED = 1e-6    # default mode, |h_gpu - h_oracle| / max|r0|


@pytest.fixture(scope="module", params=list(CASES))
def pair(request, oracle_lib):
    from hierarchicaldd_amd import hdd
    dom, cfg = CASES[request.param]()
    print(f"[case {request.param}] pages={dom.n_pages}", flush=True)
    o = oracle_lib.Oracle(dom, cfg)
    o.solution = o.solve(1e-6 * float(np.abs(dom.b).max()), 200)[:2]  # (iterations, history), solved once
    g = hdd.DDSolver(dom, cfg)
    yield request.param, dom, cfg, o, g
    g.close()


def _history_rel(h_g, h_o):
    n = min(len(h_o), len(h_g))
    return float((np.abs(h_g[:n] - h_o[:n]) / np.maximum(h_o[:n], 1e-30)).max())


def _history_normalised(h_g, h_o):
    n = min(len(h_o), len(h_g))
    return float(np.abs(h_g[:n] - h_o[:n]).max() / h_o[0])


def test_apply_preconditioner(pair):
    name, dom, cfg, o, g = pair
    r = random_rhs(g.level(0)[1], 7)
    zo = o.apply_preconditioner(r)
    zg = g.apply_preconditioner(r)
    assert np.isfinite(zg).all()
    e1 = rel_err(zg, zo)
    g.set_sigma_refinement(1)
    try:
        zr = g.apply_preconditioner(r)
    finally:
        g.set_sigma_refinement(0)
    e2 = rel_err(zr, zo)
    print(f"[{name}] preconditioner rel err: single {e1:.2e}, refined {e2:.2e}")
    assert e1 < 2e-4, e1
    assert e2 < 1e-5, e2


def test_pcg_matches_oracle(pair):
    """Default mode: iterations +-1, history within 1e-6 of max|r0|."""
    name, dom, cfg, o, g = pair
    tol = 1e-6 * float(np.abs(dom.b).max())
    it_o, h_o = o.solution
    it_g, h_g, secs = g.solve(tol, 200)
    rel, nrm = _history_rel(h_g, h_o), _history_normalised(h_g, h_o)
    print(f"[{name}] oracle iters {it_o} gpu iters {it_g} gpu seconds {secs:.4f} history max rel diff {rel:.3e}, "
          f"normalised by max|r0| {nrm:.3e}")
    assert it_o > 0 and it_g > 0
    assert abs(it_g - it_o) <= 1
    assert h_g[-1] <= tol
    assert nrm < HISTORY_BAR_NORMALISED


def test_pcg_reference_exact_mode(pair):
    """Exact sigma solves and the reference's 200 literal bottom smooths: the oracle's history."""
    import os
    from hierarchicaldd_amd import hdd
    name, dom, cfg, o, _ = pair
    tol = 1e-6 * float(np.abs(dom.b).max())
    it_o, h_o = o.solution
    os.environ["HDD_SIGMA_BOTTOM"] = "loop"
    try:
        g = hdd.DDSolver(dom, cfg)
    finally:
        os.environ.pop("HDD_SIGMA_BOTTOM", None)
    try:
        g.set_sigma_refinement(1)
        it_r, h_r, secs_r = g.solve(tol, 200)
        # the inner products in the reference's serial summation order change nothing here
        g.set_serial_dots(True)
        it_s, h_s, _ = g.solve(tol, 200)
    finally:
        g.close()
    rel_r, rel_s = _history_rel(h_r, h_o), _history_rel(h_s, h_o)
    print(f"[{name}] reference-exact mode: gpu iters {it_r} seconds {secs_r:.4f} history max rel diff {rel_r:.3e}, "
          f"bit-identical {np.array_equal(h_r, h_o)}; serial-order inner products: {it_s} iters, max rel diff {rel_s:.3e}")
    assert it_r == it_o and it_s == it_o
    assert h_r[-1] <= tol
    assert rel_r < HISTORY_BAR_EXACT and rel_s < HISTORY_BAR_EXACT
