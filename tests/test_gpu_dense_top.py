"""Dense tops of the triangular-solve plans against the per-height solve and the oracle.

The top heights of a tall elimination tree (S, ancestor-closed) are solved as
two dense triangular matvecs with Linv = L_SS^{-1} (formed in double, stored in
float), instead of one forward and one backward step per height with the
explicit inverses of the diagonal blocks. Both are float evaluations of the
same exact solve, so the bars are the existing parity bars:
  * sigma smooths of every level within 2e-4 relative (max-norm) of the
    oracle's double Cholesky solves (the sigma parity bar);
  * PCG with dense tops on the PCG parity cases: iterations within one of the
    oracle's, residual history within 1e-4 relative of the oracle's (the PCG
    parity bar), solution within 1e-5 of the per-height plans' solution;
  * a second build gives bit-identical results (fixed-order sums).
(box48_sub16 is a sigma case only: its last residuals sit at the float noise
floor, 2.6e-3 from the oracle's with or without dense tops.)
HDD_DENSE_TOP=0 turns the tops off; HDD_DENSE_TOP=256 gives them to every
factor of at least 256 rows, so the subdomain batches get them too.
"""
import os

import numpy as np
import pytest

from hierarchicaldd_amd import domain
from tests.gpu_common import rel_err, same_bits

pytestmark = pytest.mark.gpu

SIGMA_TOL = 2e-4

CASES = {
    "box48_sub16": lambda: domain.small_box(48, 16, 3, seed=2),
    "sphere64_sub16": domain.config0,
}
PCG_CASES = {
    "box32_sub16": lambda: domain.small_box(32, 16, 3, seed=1),
    "sphere64_sub16": domain.config0,
}


def _solver(dom, cfg, mode):
    from hierarchicaldd_amd import hdd
    os.environ["HDD_DENSE_TOP"] = mode
    try:
        return hdd.DDSolver(dom, cfg)
    finally:
        os.environ.pop("HDD_DENSE_TOP", None)


@pytest.fixture(scope="module", params=list(CASES))
def case(request, oracle_lib):
    dom, cfg = CASES[request.param]()
    print(f"[case {request.param}] pages={dom.n_pages}", flush=True)
    o = oracle_lib.Oracle(dom, cfg)
    yield request.param, dom, cfg, o


@pytest.mark.parametrize("mode", ["256", ""])
def test_dense_top_sigma_smooth_vs_oracle(case, mode):
    name, dom, cfg, o = case
    g = _solver(dom, cfg, mode)
    try:
        rows = g.stats()["dense_top_rows"]
        print(f"[{name} HDD_DENSE_TOP={mode or 'default'}] dense top rows {rows:.0f}")
        if mode:
            assert rows > 0
        L, _ = g.sigma_info()
        rng = np.random.default_rng(11)
        for l in range(L):
            n = g.interface_dofs(l)
            b = rng.standard_normal(n).astype(np.float32)
            u = (0.1 * rng.standard_normal(n)).astype(np.float32)
            ug = g.sigma_smooth(l, b, u)
            uo = o.sigma_smooth(l, b.astype(np.float64), u.astype(np.float64))
            assert np.isfinite(ug).all()
            assert rel_err(ug, uo) < SIGMA_TOL, f"level {l}: {rel_err(ug, uo)}"
    finally:
        g.close()


@pytest.mark.parametrize("name", list(PCG_CASES))
def test_dense_top_pcg_vs_oracle(name, oracle_lib):
    from hierarchicaldd_amd import hdd
    dom, cfg = PCG_CASES[name]()
    o = oracle_lib.Oracle(dom, cfg)
    tol = 1e-6 * float(np.abs(dom.b).max())
    it_o, h_o, _ = o.solve(tol, 200)
    out = {}
    for mode in ("0", "256", "256"):
        g = _solver(dom, cfg, mode)
        try:
            it, hist, _ = g.solve(tol, 200)
            res = (it, hist, g.get(hdd.CH_X), g.stats()["dense_top_rows"])
        finally:
            g.close()
        if mode in out:
            assert res[0] == out[mode][0] and np.array_equal(res[1], out[mode][1]) and same_bits(res[2], out[mode][2])
        out[mode] = res
    (i0, h0, x0, r0), (i1, h1, x1, r1) = out["0"], out["256"]
    k = min(len(h1), len(h_o))
    rel = np.abs(h1[:k] - h_o[:k]) / np.maximum(h_o[:k], 1e-30)
    print(f"[{name}] oracle {it_o} iterations, per-height {i0}, dense tops ({r1:.0f} rows) {i1}; "
          f"history rel diff vs oracle {rel.max():.2e}; x rel diff vs per-height {rel_err(x1, x0):.2e}")
    assert r0 == 0 and r1 > 0
    assert abs(i1 - it_o) <= 1 and h1[-1] <= tol
    assert rel.max() < 1e-4
    assert rel_err(x1, x0) < 1e-5
