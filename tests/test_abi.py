"""The C-ABI library loads and exports every function include/hdd.h declares;
argument validation errors are returned as codes (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "hdd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hdd_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from hierarchicaldd_amd import hdd
    L = hdd.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(hdd.EXPORTS)


def test_invalid_arguments_return_codes():
    from hierarchicaldd_amd import hdd
    L = hdd.lib()
    h = C.c_void_p()
    assert L.hdd_create(None, C.byref(h)) == 1
    assert b"invalid" in L.hdd_last_error(None)
    d = hdd.ProblemDesc()
    d.vlevels = 0
    assert L.hdd_create_host_only(C.byref(d), C.byref(h)) == 1
    assert L.hdd_destroy(None) == 1
    assert L.hdd_pcg_solve(None, None, None, None, 0) == 1


def test_setup_failure_reports_error():
    """sigma_levels above the adaptive depth of a subdomain is rejected at setup."""
    from hierarchicaldd_amd import domain, hdd
    dom, _ = domain.small_box(16, 8, 2, seed=0)
    cfg = domain.DDConfig(sub_size=8, vlevels=2, sigma_levels=6)
    with pytest.raises(hdd.HddError, match="sigma_levels"):
        hdd.DDSolver(dom, cfg, host_only=True)


def test_host_threads_setting_keeps_setups_equal():
    """hdd_set_host_threads changes only how many host threads the calling thread's set-ups
    use: a host-only set-up on two threads gives the interface operators and V-cycle flags of
    one on all the host's threads; a negative count is an argument error."""
    from hierarchicaldd_amd import domain, hdd
    dom, cfg = domain.small_box(32, 16, 3, seed=5)
    a = hdd.DDSolver(dom, cfg, host_only=True)
    hdd.set_host_threads(2)
    try:
        b = hdd.DDSolver(dom, cfg, host_only=True)
    finally:
        hdd.set_host_threads(0)
    try:
        for lvl in range(cfg.vlevels):
            assert np.array_equal(a.level(lvl)[1], b.level(lvl)[1])
        assert a.stats()["interface_dofs"] == b.stats()["interface_dofs"] > 0
    finally:
        a.close()
        b.close()
    assert hdd.lib().hdd_set_host_threads(-1) == 1
