"""The Krylov plugin surface on device channels (hdd_multiply, hdd_inner_product,
hdd_convergence_norm, hdd_project, hdd_vector_*, hdd_apply_preconditioner_channels,
hdd_set_stream) against the library's own PCG.

The reference's loop, CONJUGATE_GRADIENT_DUMPING::Solve
(DDS/CONJUGATE_GRADIENT_DUMPING.cpp:146-212), is restated twice, driven only
through those entry points: in Python over ctypes below, and in C++ through the
CG_SYSTEM / CG_VECTOR-shaped classes of include/hdd_krylov.hpp
(tests/krylov_adapter_check.cpp). Both must reproduce hdd_pcg_solve's residual
history and solution bit for bit: the entry points do the same arithmetic in
the same order as the fused PCG, with the scalars (rho, beta, alpha) formed in
float as the reference forms them.
"""
import ctypes as C
import os

import numpy as np
import pytest

from hierarchicaldd_amd import domain, hdd
from tests.gpu_common import hip_runtime, random_rhs, same_bits

pytestmark = pytest.mark.gpu

Z, R, K, X, B, Q, S = range(7)
CASES = {
    "box32_sub16": lambda: domain.small_box(32, 16, 3, seed=1),
    "sphere64_sub16": domain.config0,
}


def _ok(solver, st):
    if st != 0:
        raise hdd.HddError(f"status {st}: {hdd.lib().hdd_last_error(solver.h).decode()}")


def restated_pcg(solver, tol, max_it):
    """CONJUGATE_GRADIENT_DUMPING::Solve through the channel entry points (Python)."""
    L, h = hdd.lib(), solver.h
    f32 = np.float32

    def inner(a, b):
        v = C.c_double()
        _ok(solver, L.hdd_inner_product(h, a, b, C.byref(v)))
        return v.value

    def norm(a):
        v = C.c_float()
        _ok(solver, L.hdd_convergence_norm(h, a, C.byref(v)))
        return f32(v.value)

    hist = []
    rho_old = f32(3.40282347e38)
    it = 0
    while True:
        restart = it == 0
        if restart:
            _ok(solver, L.hdd_vector_copy(h, R, 1.0, B))      # r = b
            _ok(solver, L.hdd_multiply(h, X, Q))              # q = A x
            _ok(solver, L.hdd_vector_minus_equal(h, R, Q))    # r -= q
            _ok(solver, L.hdd_project(h, R))
        conv = norm(R)
        hist.append(float(conv))
        if conv <= f32(tol):
            return it, np.array(hist)
        if it == max_it:
            return -it, np.array(hist)
        _ok(solver, L.hdd_apply_preconditioner_channels(h, Z, R, K))
        rho = f32(inner(Z, R))
        if restart:
            _ok(solver, L.hdd_vector_copy(h, S, 1.0, Z))
        else:
            _ok(solver, L.hdd_vector_copy2(h, S, float(rho / rho_old), S, Z))
        _ok(solver, L.hdd_multiply(h, S, Q))
        _ok(solver, L.hdd_project(h, Q))
        sdq = f32(inner(S, Q))
        alpha = rho / sdq if sdq != 0 else f32(3.40282347e38)
        _ok(solver, L.hdd_vector_copy2(h, X, float(alpha), S, X))
        _ok(solver, L.hdd_vector_copy2(h, R, float(-alpha), Q, R))
        rho_old = rho
        it += 1


def adapter_lib():
    path = os.path.join(os.path.dirname(hdd.LIB_PATH), "libhdd_krylov_check.so")
    A = C.CDLL(path)
    A.hdd_krylov_check_solve.argtypes = [C.c_void_p, C.c_float, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_char_p, C.c_int]
    A.hdd_krylov_check_solve.restype = C.c_int
    return A


def adapter_pcg(solver, tol, max_it):
    A = adapter_lib()
    hist = np.zeros(max_it + 1)
    its = C.c_int()
    err = C.create_string_buffer(256)
    ok = A.hdd_krylov_check_solve(solver.h, tol, max_it, hist.ctypes.data_as(C.c_void_p), max_it + 1, C.byref(its), err, 256)
    assert ok >= 0, err.value.decode()
    n = its.value + 1
    return (its.value if ok == 1 else -its.value), hist[:n]


@pytest.fixture(scope="module", params=list(CASES))
def case(request):
    dom, cfg = CASES[request.param]()
    print(f"[case {request.param}] pages={dom.n_pages}", flush=True)
    g = hdd.DDSolver(dom, cfg)
    tol = 1e-6 * float(np.abs(dom.b).max())
    it_ref, h_ref, _ = g.solve(tol, 100)
    x_ref = g.get(X).copy()
    yield request.param, g, tol, it_ref, h_ref, x_ref
    g.close()


def _zero_x(g):
    g.set(X, np.zeros((g.n_pages, 256), np.float32))


def test_restated_loop_reproduces_pcg_bitwise(case):
    name, g, tol, it_ref, h_ref, x_ref = case
    _zero_x(g)
    it, h = restated_pcg(g, tol, 100)
    print(f"[{name}] pcg {it_ref} iterations; restated loop {it}")
    assert it == it_ref
    assert np.array_equal(h, h_ref)
    assert same_bits(g.get(X), x_ref)


def test_cpp_adapter_reproduces_pcg_bitwise(case):
    name, g, tol, it_ref, h_ref, x_ref = case
    _zero_x(g)
    it, h = adapter_pcg(g, tol, 100)
    assert it == it_ref
    assert np.array_equal(h, h_ref)
    assert same_bits(g.get(X), x_ref)


def test_cpp_adapter_on_a_caller_stream(case):
    name, g, tol, it_ref, h_ref, x_ref = case
    H = hip_runtime()
    st = C.c_void_p()
    assert H.hipStreamCreate(C.byref(st)) == 0
    try:
        _ok(g, hdd.lib().hdd_set_stream(g.h, st))
        try:
            _zero_x(g)
            it, h = adapter_pcg(g, tol, 100)
        finally:
            _ok(g, hdd.lib().hdd_set_stream(g.h, None))
        assert H.hipStreamSynchronize(st) == 0
    finally:
        H.hipStreamDestroy(st)
    assert it == it_ref
    assert np.array_equal(h, h_ref)
    assert same_bits(g.get(X), x_ref)


def test_preconditioner_on_chosen_channels(case):
    """Apply_Preconditioner(result=q, rhs=s, tmp=x) equals the default channels'
    result bit for bit, and leaves every channel but result and tmp unchanged."""
    name, g, tol, it_ref, h_ref, x_ref = case
    flags = g.level(0)[1]
    r = random_rhs(flags, 21)
    z_default = g.apply_preconditioner(r)
    rng = np.random.default_rng(22)
    before = {c: rng.standard_normal((g.n_pages, 256)).astype(np.float32) for c in range(7)}
    before[S] = r
    for c, v in before.items():
        g.set(c, v)
    _ok(g, hdd.lib().hdd_apply_preconditioner_channels(g.h, Q, S, X))
    _ok(g, hdd.lib().hdd_synchronize(g.h))
    assert same_bits(g.get(Q), z_default)
    for c in (Z, R, K, B, S):
        assert same_bits(g.get(c), before[c]), f"channel {c} changed"
    # and again the default channels, after the exchange
    assert same_bits(g.apply_preconditioner(r), z_default)


def test_multiply_leaves_other_cells(case):
    """SPGrid_Laplace writes Active|Interface cells only."""
    name, g, tol, it_ref, h_ref, x_ref = case
    flags = g.level(0)[1]
    rng = np.random.default_rng(23)
    v = rng.standard_normal((g.n_pages, 256)).astype(np.float32)
    out0 = rng.standard_normal((g.n_pages, 256)).astype(np.float32)
    g.set(S, v)
    g.set(Q, out0)
    _ok(g, hdd.lib().hdd_multiply(g.h, S, Q))
    out = g.get(Q)
    act = (flags & (domain.ACTIVE | domain.INTERFACE)) != 0
    assert same_bits(out[~act], out0[~act])
    assert np.isfinite(out[act]).all() and np.abs(out[act]).max() > 0
