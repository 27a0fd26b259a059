"""Dense fronts of the interface solver's factorisation on the GPU (rocSOLVER
dpotrf, rocBLAS dtrsm / dsyrk, FP64) against the host's partial Cholesky.

With HDD_DEVICE_FRONT_MIN=64 every front of at least 64 rows of a 96^3 sub-32
problem is factored on the device: the FP64 factor solves must agree with
scipy's sparse direct solve of (A + 1e-5 I) to 1e-10, and the PCG solve must
take the iterations of a context whose fronts all stayed on the host
(HDD_DEVICE_FRONTS=0), with x within 1e-5 (the float plans built from the two
factors round differently)."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from hierarchicaldd_amd import domain

pytestmark = pytest.mark.gpu


def _with_env(env, fn):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return fn()
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def test_device_fronts_match_host_and_scipy():
    from hierarchicaldd_amd import hdd
    dom, cfg = domain.sphere_case(96, 32, 4, delta=(32, 32, 32))
    print(f"[device fronts] sphere96_sub32 pages={dom.n_pages}", flush=True)
    gd = _with_env({"HDD_DEVICE_FRONT_MIN": "64"}, lambda: hdd.DDSolver(dom, cfg))
    gh = _with_env({"HDD_DEVICE_FRONTS": "0"}, lambda: hdd.DDSolver(dom, cfg))
    try:
        rng = np.random.default_rng(40)
        levels, nsub = gd.sigma_info()
        for l in range(levels):
            for which, s in [("Arr", 0), ("Aii", 13)]:
                r, c, p, i, v = gd.matrix(which, l, s)
                if r == 0:
                    continue
                A = sp.csr_matrix((v.astype(np.float64), i, p), shape=(r, c))
                b = rng.standard_normal(r)
                x = gd.sigma_factor_solve(l, which, s, b)
                xs = spl.spsolve((A + 1e-5 * sp.identity(r)).tocsc(), b)
                err = np.abs(x - xs).max() / np.abs(xs).max()
                print(f"[device fronts] level {l} {which}: rows {r}, rel err vs scipy {err:.2e}")
                assert err <= 1e-10
        tol = 1e-6 * float(np.abs(dom.b).max())
        it_d, h_d, _ = gd.solve(tol, 200)
        x_d = gd.get(3)
        it_h, h_h, _ = gh.solve(tol, 200)
        x_h = gh.get(3)
        xr = float(np.abs(x_d - x_h).max() / np.abs(x_h).max())
        print(f"[device fronts] iterations device {it_d} host {it_h}, x rel diff {xr:.2e}")
        assert it_d == it_h and xr < 1e-5
    finally:
        gd.close()
        gh.close()
