"""Host-side setup of the product (no GPU): level flags, page sets and every
interface operator must equal the oracle's exactly."""
import numpy as np
import pytest

from hierarchicaldd_amd import domain

CASES = {
    "box32_sub16": lambda: domain.small_box(32, 16, 3, seed=1),
    "box48_sub16_v2": lambda: domain.small_box(48, 16, 2, seed=2),
    "box24_sub8_sig2": lambda: (domain.small_box(24, 8, 2, seed=4)[0], domain.DDConfig(sub_size=8, vlevels=2, sigma_levels=2)),
    "sphere64_sub16": domain.config0,
    # the DD parameters of the 256^3 workload (sub_size 32: adaptive strides up to 8; vlevels 4)
    "box64_sub32_v4": lambda: domain.small_box(64, 32, 4, seed=3),
    "sphere96_sub32_v4": lambda: domain.sphere_case(96, 32, 4, delta=(32, 32, 32)),
}


@pytest.fixture(scope="module", params=list(CASES))
def pair(request, oracle_lib):
    from hierarchicaldd_amd import hdd
    dom, cfg = CASES[request.param]()
    return request.param, dom, cfg, oracle_lib.Oracle(dom, cfg), hdd.DDSolver(dom, cfg, host_only=True)


def test_level_pages_and_flags(pair):
    _, dom, cfg, o, g = pair
    assert np.array_equal(o.page_offsets(), g.page_offsets())
    for l in range(cfg.vlevels):
        oo, of = o.level(l)
        go, gf = g.level(l)
        assert np.array_equal(oo, go) and np.array_equal(of, gf), f"level {l}"


def test_interface_operators_exact(pair):
    _, dom, cfg, o, g = pair
    L, ns = g.sigma_info()
    assert (L, ns) == (o.sigma_levels, o.n_subdomains)
    for l in range(L):
        assert np.array_equal(g.interface_coords(l), o.interface_coords(l))
        mats = [("Arr", 0)] + [(w, s) for w in ("Aii", "Air") for s in range(ns)] + ([("Prr", 0)] if l < L - 1 else [])
        for w, s in mats:
            a, b = g.matrix(w, l, s), o.matrix(w, l, s)
            assert a[:2] == b[:2], (w, l, s)
            assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3]), (w, l, s)
            assert np.array_equal(a[4].astype(np.float64), b[4]), (w, l, s)


def test_host_only_context_rejects_device_calls(pair):
    _, dom, cfg, o, g = pair
    from hierarchicaldd_amd import hdd
    with pytest.raises(hdd.HddError):
        g.vcycle(np.zeros((g.n_pages, 256), np.float32))
