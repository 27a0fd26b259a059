"""GPU parity of the per-level setup and the matrix-free V-cycle against the oracle.

The V-cycle kernels reproduce the reference's float operation order, so the
GPU result must be bit-identical to the oracle's restatement of
LIB/SPGrid_V_Cycle_Helper.h:156-192 on the same inputs.
"""
import numpy as np
import pytest

from hierarchicaldd_amd import domain
from tests.gpu_common import random_rhs, same_bits

pytestmark = pytest.mark.gpu

CASES = {
    "box32_sub16": lambda: domain.small_box(32, 16, 3, seed=1),
    "box48_sub16_v2": lambda: domain.small_box(48, 16, 2, seed=2),
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


def test_page_order(pair):
    _, dom, cfg, o, g = pair
    assert np.array_equal(o.page_offsets(), g.page_offsets())


def test_level_flags(pair):
    _, dom, cfg, o, g = pair
    for l in range(cfg.vlevels):
        oo, of = o.level(l)
        go, gf = g.level(l)
        assert np.array_equal(oo, go), f"level {l} page sets differ"
        assert np.array_equal(of, gf), f"level {l} flags differ"


def test_smoothing_sweeps_bitwise(pair):
    _, dom, cfg, o, g = pair
    from oracle.pyoracle import CH_R, CH_Z
    flags = g.level(0)[1]
    b = random_rhs(flags, 5)
    u0 = random_rhs(flags, 6) * 0.1
    for which in (o, g):
        which.set(CH_R, b)
        which.set(CH_Z, u0)
    for mode in (1, 0, 2, 1):
        o.smooth0(mode)
        g.smooth(mode)
        assert same_bits(o.get(CH_Z), g.get(CH_Z)), f"sweep mode {mode} differs"


def test_vcycle_bitwise(pair):
    _, dom, cfg, o, g = pair
    flags = g.level(0)[1]
    for seed in (11, 12):
        b = random_rhs(flags, seed)
        zo = o.vcycle(b)
        zg = g.vcycle(b)
        assert np.isfinite(zg).all()
        assert same_bits(zo, zg), f"V-cycle differs (max abs diff {np.abs(zo - zg).max()})"


def test_vcycle_linear_and_zero(pair):
    """A V-cycle is a linear map: V(0) = 0 and V(2b) = 2 V(b) (powers of two scale exactly)."""
    _, dom, cfg, o, g = pair
    flags = g.level(0)[1]
    b = random_rhs(flags, 21)
    assert not np.any(g.vcycle(np.zeros_like(b)))
    z1 = g.vcycle(b)
    z2 = g.vcycle(2 * b)
    assert same_bits(2 * z1, z2)


LOCAL_BOTTOM_CASES = {
    # every subdomain's coarsest block inside one page
    "sphere64_sub16": domain.config0,
    # 32-cell subdomains, 4 levels: one cluster of two coupled pages at the coarsest level
    "sphere96_sub32": lambda: (domain.sphere_obstacle_box(96, [(48, 48, 48, 24)], delta=((24, 24, 24), 1.0)),
                               domain.DDConfig(sub_size=32, vlevels=4)),
}


@pytest.mark.parametrize("fp16", [False, True])
@pytest.mark.parametrize("name", list(LOCAL_BOTTOM_CASES))
def test_bottom_page_local_bitwise(name, fp16):
    """The coarsest level's 200 sweeps run in one launch, a workgroup per
    cluster of coupled pages, when no cluster has more than four pages. Results
    must equal those of the 200 sweep launches bit for bit (HDD_BOTTOM_LOCAL=0),
    in FP32 and in FP16 storage, for preconditioner applications and whole PCG
    solves."""
    import os
    from hierarchicaldd_amd import hdd
    dom, cfg = LOCAL_BOTTOM_CASES[name]()
    tol = 1e-6 * float(np.abs(dom.b).max())
    out = {}
    for mode in ("0", ""):
        os.environ["HDD_BOTTOM_LOCAL"] = mode
        try:
            g = hdd.DDSolver(dom, cfg)
        finally:
            os.environ.pop("HDD_BOTTOM_LOCAL", None)
        try:
            g.set_fp16(fp16)
            b = random_rhs(g.level(0)[1], 31)
            z = g.apply_preconditioner(b)
            it, hist, _ = g.solve(tol, 200)
            out[mode] = (g.stats()["bottom_page_local"], z, it, hist, g.get(hdd.CH_X))
        finally:
            g.close()
    (l0, z0, i0, h0, x0), (l1, z1, i1, h1, x1) = out["0"], out[""]
    print(f"[{name} fp16={fp16}] page-local bottom {l1:.0f}; iterations {i0} / {i1}")
    assert l0 == 0 and l1 == 1
    assert same_bits(z0, z1)
    assert i0 == i1 and np.array_equal(h0, h1) and same_bits(x0, x1)


def test_sweep_timer_without_setup():
    """hdd_profile_sweep_grid times the level-0 sweep of a page set without the solver's
    set-up: one launch moves 16 B per page slot, and its time is that of the same sweep
    inside a context to within a factor of two (both are back-to-back launches timed with
    HIP events; the page set alone lacks only the divider's interface flags)."""
    from hierarchicaldd_amd import hdd
    dom, cfg = domain.config0()
    ms_grid, by_grid = hdd.profile_sweep_grid(dom, cfg, 50)
    g = hdd.DDSolver(dom, cfg)
    try:
        ms_ctx, by_ctx = g.profile_sweep(50)
    finally:
        g.close()
    print(f"[sweep timer] page set alone {ms_grid * 1e3:.1f} us, in a context {ms_ctx * 1e3:.1f} us", flush=True)
    assert by_grid == by_ctx == 16.0 * 256 * dom.n_pages
    assert ms_grid > 0 and ms_ctx > 0
    assert 0.5 < ms_grid / ms_ctx < 2.0
