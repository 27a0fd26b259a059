"""ORACLE — test infrastructure only.

ctypes wrapper around oracle/_build/libddoracle.so, the CPU restatement of the
reference dd_pcg path. Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libddoracle.so")
_lib = None


def build(force: bool = False) -> str:
    cmd = ["make", "-C", HERE] + (["-B"] if force else [])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P, I, D = C.c_void_p, C.c_int, C.c_double
        L.ora_create.restype = P
        L.ora_create.argtypes = [I, I, P, P, P, P, P, I, I, I, I, I, I, I, I, C.c_char_p, I]
        L.ora_destroy.argtypes = [P]
        L.ora_setup_seconds.restype = D
        L.ora_setup_seconds.argtypes = [P]
        for name in ["ora_num_pages", "ora_sigma_levels", "ora_num_subdomains"]:
            getattr(L, name).restype = I
            getattr(L, name).argtypes = [P]
        L.ora_page_offsets.argtypes = [P, P]
        L.ora_level_pages.restype = I
        L.ora_level_pages.argtypes = [P, I]
        L.ora_level_data.argtypes = [P, I, P, P]
        L.ora_get_channel.argtypes = [P, I, P]
        L.ora_set_channel.argtypes = [P, I, P]
        L.ora_vcycle.argtypes = [P]
        L.ora_smooth0.argtypes = [P, I]
        L.ora_apply_preconditioner.argtypes = [P]
        L.ora_interface_dofs.restype = I
        L.ora_interface_dofs.argtypes = [P, I]
        L.ora_interface_coords.argtypes = [P, I, P]
        L.ora_matrix_shape.restype = I
        L.ora_matrix_shape.argtypes = [P, I, I, I, P, P]
        L.ora_matrix_data.argtypes = [P, I, I, I, P, P, P]
        L.ora_sigma_vcycle.argtypes = [P, P, P]
        L.ora_sigma_smooth.argtypes = [P, I, P, P]
        L.ora_set_threads.argtypes = [I]
        L.ora_max_threads.restype = I
        L.ora_adaptive_stride.argtypes = [I, I, P, P]
        L.ora_level_get.argtypes = [P, I, I, P]
        L.ora_level_set.argtypes = [P, I, I, P]
        L.ora_level_kernel.argtypes = [P, I, I, I]
        L.ora_sigma_factor_solve.argtypes = [P, I, I, I, P, P]
        L.ora_solve.restype = I
        L.ora_solve.argtypes = [P, D, I, P, I, P]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


CH_Z, CH_R, CH_K, CH_X, CH_B, CH_Q, CH_S = range(7)


def set_threads(n: int) -> int:
    """OpenMP threads of later oracle calls; returns the previous setting."""
    prev = lib().ora_max_threads()
    lib().ora_set_threads(int(n))
    return prev


def adaptive_stride(lo: int, hi: int):
    """STRIDE_GENERATOR::Generate_Adaptive_Stride on one axis: (strides over [lo, hi], level)."""
    q = np.zeros(hi - lo + 1, dtype=np.int32)
    lev = C.c_int()
    lib().ora_adaptive_stride(lo, hi, _ptr(q), C.byref(lev))
    return q, lev.value


class Oracle:
    """Full reference setup + solver on one SPGrid domain (float or double).

    sigma_single: the interface solver's exact solves in single precision (factor
    rounded to float, float substitution), as the reference's single-precision
    PARDISO stores and computes them; default: exact (double) solves."""

    def __init__(self, dom, cfg, double: bool = False, sigma_single: bool = False):
        L = lib()
        self.dom, self.cfg = dom, cfg
        po = np.ascontiguousarray(dom.page_offsets, dtype=np.uint64)
        fl = np.ascontiguousarray(dom.flags, dtype=np.uint32)
        bb = np.ascontiguousarray(dom.b, dtype=np.float32)
        size = np.array(dom.size, dtype=np.int32)
        sub = np.array([cfg.sub_size] * 3, dtype=np.int32)
        err = C.create_string_buffer(512)
        self.h = L.ora_create(int(double), po.shape[0], _ptr(po), _ptr(fl), _ptr(bb), _ptr(size), _ptr(sub),
                              cfg.vlevels, cfg.sigma_levels, cfg.nVcycles, cfg.interior_smoothing,
                              cfg.boundary_smoothing, cfg.sigma_smoothing, cfg.sigma_mg, int(bool(sigma_single)), err, 512)
        if not self.h:
            raise RuntimeError("oracle setup failed: " + err.value.decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_destroy(self.h)
            self.h = None

    @property
    def setup_seconds(self) -> float:
        return lib().ora_setup_seconds(self.h)

    @property
    def n_pages(self) -> int:
        return lib().ora_num_pages(self.h)

    def page_offsets(self):
        out = np.zeros(self.n_pages, dtype=np.uint64)
        lib().ora_page_offsets(self.h, _ptr(out))
        return out

    def level(self, l):
        n = lib().ora_level_pages(self.h, l)
        off = np.zeros(n, dtype=np.uint64)
        fl = np.zeros((n, 256), dtype=np.uint32)
        lib().ora_level_data(self.h, l, _ptr(off), _ptr(fl))
        return off, fl

    def get(self, ch):
        out = np.zeros((self.n_pages, 256), dtype=np.float64)
        lib().ora_get_channel(self.h, ch, _ptr(out))
        return out

    def set(self, ch, v):
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(self.n_pages, 256)
        lib().ora_set_channel(self.h, ch, _ptr(v))

    def vcycle(self, b):
        self.set(CH_R, b)
        lib().ora_vcycle(self.h)
        return self.get(CH_Z)

    def smooth0(self, mode):
        lib().ora_smooth0(self.h, mode)

    def apply_preconditioner(self, r):
        self.set(CH_R, r)
        lib().ora_apply_preconditioner(self.h)
        return self.get(CH_Z)

    @property
    def sigma_levels(self):
        return lib().ora_sigma_levels(self.h)

    @property
    def n_subdomains(self):
        return lib().ora_num_subdomains(self.h)

    def interface_dofs(self, l):
        return lib().ora_interface_dofs(self.h, l)

    def interface_coords(self, l):
        n = self.interface_dofs(l)
        out = np.zeros((n, 3), dtype=np.int32)
        lib().ora_interface_coords(self.h, l, _ptr(out))
        return out

    def matrix(self, which: str, level: int, sub: int = 0):
        """Return (rows, cols, indptr, indices, data) of Arr/Aii/Air/Prr."""
        w = {"Arr": 0, "Aii": 1, "Air": 2, "Prr": 3}[which]
        r, c = C.c_int(), C.c_int()
        nnz = lib().ora_matrix_shape(self.h, w, level, sub, C.byref(r), C.byref(c))
        ptr = np.zeros(r.value + 1, dtype=np.int32)
        idx = np.zeros(nnz, dtype=np.int32)
        val = np.zeros(nnz, dtype=np.float64)
        lib().ora_matrix_data(self.h, w, level, sub, _ptr(ptr), _ptr(idx), _ptr(val))
        return r.value, c.value, ptr, idx, val

    def sigma_vcycle(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        u = np.zeros_like(b)
        lib().ora_sigma_vcycle(self.h, _ptr(b), _ptr(u))
        return u

    def sigma_smooth(self, level, b, u):
        b = np.ascontiguousarray(b, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64).copy()
        lib().ora_sigma_smooth(self.h, level, _ptr(b), _ptr(u))
        return u

    def level_get(self, l, c):
        """Channel c (0 u, 1 b, 2 r) of V-cycle level l, [pages, 256] in block order."""
        n = lib().ora_level_pages(self.h, l)
        out = np.zeros((n, 256), dtype=np.float64)
        lib().ora_level_get(self.h, l, c, _ptr(out))
        return out

    def level_set(self, l, c, v):
        n = lib().ora_level_pages(self.h, l)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(n, 256)
        lib().ora_level_set(self.h, l, c, _ptr(v))

    def level_kernel(self, l, kernel: str, mode: int = 2):
        """One V-cycle kernel on level l: residual, correction (mode 0 interior, 1
        boundary, 2 all), restriction (l -> l+1) or prolongation (l+1 -> l)."""
        k = {"residual": 0, "correction": 1, "restriction": 2, "prolongation": 3}[kernel]
        lib().ora_level_kernel(self.h, l, k, mode)

    def sigma_factor_solve(self, level, which: str, sub, b):
        """(A + 1e-5 I)^{-1} b with the oracle's factor of Arr or of subdomain sub's Aii."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        lib().ora_sigma_factor_solve(self.h, level, {"Arr": 0, "Aii": 1}[which], sub, _ptr(b), _ptr(x))
        return x

    def solve(self, tol_abs: float, max_it: int = 5000):
        hist = np.zeros(max_it + 1, dtype=np.float64)
        secs = C.c_double()
        it = lib().ora_solve(self.h, tol_abs, max_it, _ptr(hist), max_it + 1, C.byref(secs))
        n = abs(it) + 1
        return it, hist[:n], secs.value

    def x(self):
        return self.get(CH_X)
