"""GPU sigma (interface Schur complement) multigrid against the oracle.

The oracle solves (A + 1e-5 I) x = b with a double-precision Cholesky; the GPU
applies float triangular-solve plans. The tolerance below bounds float
rounding amplified by the conditioning of the interface operators.
"""
import numpy as np
import pytest

from hierarchicaldd_amd import domain
from tests.gpu_common import rel_err

pytestmark = pytest.mark.gpu

SIGMA_TOL = 2e-4  # relative (max-norm), float factors vs the oracle's double factors

CASES = {
    "box32_sub16": lambda: domain.small_box(32, 16, 3, seed=1),
    "sphere64_sub16": domain.config0,
}


@pytest.fixture(scope="module", params=list(CASES))
def pair(request, oracle_lib):
    from hierarchicaldd_amd import hdd
    dom, cfg = CASES[request.param]()
    print(f"[case {request.param}] pages={dom.n_pages}", flush=True)
    o = oracle_lib.Oracle(dom, cfg)
    g = hdd.DDSolver(dom, cfg)
    yield request.param, dom, cfg, o, g
    g.close()


def test_sigma_smooth_each_level(pair):
    _, dom, cfg, o, g = pair
    L, _ = g.sigma_info()
    rng = np.random.default_rng(3)
    for l in range(L):
        n = g.interface_dofs(l)
        b = rng.standard_normal(n).astype(np.float32)
        u = (0.1 * rng.standard_normal(n)).astype(np.float32)
        ug = g.sigma_smooth(l, b, u)
        uo = o.sigma_smooth(l, b.astype(np.float64), u.astype(np.float64))
        assert np.isfinite(ug).all()
        assert rel_err(ug, uo) < SIGMA_TOL, f"level {l}: {rel_err(ug, uo)}"


def test_sigma_vcycle(pair):
    _, dom, cfg, o, g = pair
    n = g.interface_dofs(0)
    b = np.random.default_rng(4).standard_normal(n).astype(np.float32)
    ug = g.sigma_vcycle(b)
    uo = o.sigma_vcycle(b.astype(np.float64))
    assert rel_err(ug, uo) < SIGMA_TOL, rel_err(ug, uo)


def test_sigma_vcycle_linear(pair):
    _, dom, cfg, o, g = pair
    n = g.interface_dofs(0)
    rng = np.random.default_rng(5)
    b1 = rng.standard_normal(n).astype(np.float32)
    b2 = rng.standard_normal(n).astype(np.float32)
    u1, u2, u12 = g.sigma_vcycle(b1), g.sigma_vcycle(b2), g.sigma_vcycle(b1 + b2)
    assert rel_err(u12, u1 + u2) < 1e-4
    assert not np.any(g.sigma_vcycle(np.zeros(n, np.float32)))
