"""hdd_inner_product_noalloc against hdd_inner_product.

The entry point allocates nothing: its partial sums live in the library's
file-scope device array. It is called through the C ABI on device buffers the
test owns (hipMalloc'd here, as a caller would hold them), twice in a row, the
second time with a smaller input, and each value must carry the bits of
hdd_inner_product on a context holding the same pages. Arguments outside the
stated limits are rejected.
"""
import ctypes as C

import numpy as np
import pytest

from hierarchicaldd_amd import domain
from tests.gpu_common import HIP_D2H, HIP_H2D, hip_runtime

pytestmark = pytest.mark.gpu


def _context_value(hdd, dom, cfg, a, b):
    g = hdd.DDSolver(dom, cfg)
    try:
        g.set(hdd.CH_Z, a)
        g.set(hdd.CH_R, b)
        out = C.c_double()
        assert hdd.lib().hdd_inner_product(g.h, hdd.CH_Z, hdd.CH_R, C.byref(out)) == 0
        return out.value
    finally:
        g.close()


def _noalloc_value(hdd, H, a, b):
    n = a.size
    pa, pb, po = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert H.hipMalloc(C.byref(pa), n * 4) == 0
    assert H.hipMalloc(C.byref(pb), n * 4) == 0
    assert H.hipMalloc(C.byref(po), 8) == 0
    try:
        assert H.hipMemcpy(pa, a.ctypes.data_as(C.c_void_p), n * 4, HIP_H2D) == 0
        assert H.hipMemcpy(pb, b.ctypes.data_as(C.c_void_p), n * 4, HIP_H2D) == 0
        st = hdd.lib().hdd_inner_product_noalloc(pa, pb, n // 256, po, None)
        assert st == 0, hdd.lib().hdd_last_error(None)
        assert H.hipStreamSynchronize(None) == 0
        v = np.zeros(1, dtype=np.float64)
        assert H.hipMemcpy(v.ctypes.data_as(C.c_void_p), po, 8, HIP_D2H) == 0
        return float(v[0])
    finally:
        H.hipFree(pa)
        H.hipFree(pb)
        H.hipFree(po)


def test_noalloc_inner_product_matches_context():
    from hierarchicaldd_amd import hdd
    H = hip_runtime()
    rng = np.random.default_rng(3)
    for name, (dom, cfg) in (("box48_sub16", domain.small_box(48, 16, 3, seed=5)), ("box32_sub16", domain.small_box(32, 16, 3, seed=6))):
        a = rng.standard_normal((dom.n_pages, 256)).astype(np.float32)
        b = rng.standard_normal((dom.n_pages, 256)).astype(np.float32)
        ref = _context_value(hdd, dom, cfg, a, b)
        got = _noalloc_value(hdd, H, a, b)
        print(f"[{name}] pages={dom.n_pages} context {ref!r} no-allocation {got!r}", flush=True)
        assert np.float64(got).view(np.uint64) == np.float64(ref).view(np.uint64)


def test_noalloc_rejects_bad_arguments():
    from hierarchicaldd_amd import hdd
    L = hdd.lib()
    dummy = C.c_void_p(1)
    assert L.hdd_inner_product_noalloc(dummy, dummy, 0, dummy, None) == 1
    assert L.hdd_inner_product_noalloc(dummy, dummy, (1 << 24) + 1, dummy, None) == 1
    assert L.hdd_inner_product_noalloc(None, dummy, 4, dummy, None) == 1
    assert b"n_pages" in L.hdd_last_error(None)
