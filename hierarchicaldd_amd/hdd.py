"""ctypes binding of the C ABI in include/hdd.h (library hierarchicaldd_amd/_build/libhdd.so).

The binding raises if the library has not been built; there is no CPU
fallback. Method names mirror the reference facade
(SPGrid_Domain_Decomposition_Solver: Read_SPGrid / Initialize_Preconditioner /
Solve, DDS/SPGrid_Domain_Decomposition_Solver.h:73-197).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libhdd.so")
_lib = None

CH_Z, CH_R, CH_K, CH_X, CH_B, CH_Q, CH_S = range(7)


class HddError(RuntimeError):
    pass


class ProblemDesc(C.Structure):
    _fields_ = [("size", C.c_int * 3), ("sub_size", C.c_int * 3), ("vlevels", C.c_int), ("sigma_levels", C.c_int),
                ("nVcycles", C.c_int), ("interior_smoothing", C.c_int), ("boundary_smoothing", C.c_int),
                ("sigma_smoothing", C.c_int), ("sigma_mg", C.c_int), ("n_pages", C.c_int),
                ("page_offsets", C.c_void_p), ("flags", C.c_void_p), ("b", C.c_void_p)]


class PcgParams(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("max_iterations", C.c_int)]


class PcgReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("seconds", C.c_double), ("final_norm", C.c_double), ("converged", C.c_int)]


class ServerParams(C.Structure):
    _fields_ = [("setup_threads", C.c_int), ("groups", C.c_int), ("max_batch", C.c_int), ("hbm_budget_bytes", C.c_double)]


class RequestReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("final_norm", C.c_double), ("latency_s", C.c_double), ("setup_s", C.c_double),
                ("queue_s", C.c_double), ("solve_s", C.c_double), ("batch_size", C.c_int), ("concurrent", C.c_int),
                ("cancelled", C.c_int)]


EXPORTS = ["hdd_create", "hdd_create_host_only", "hdd_destroy", "hdd_last_error", "hdd_num_pages", "hdd_page_offsets", "hdd_set_channel",
           "hdd_get_channel", "hdd_level_pages", "hdd_level_flags", "hdd_smooth", "hdd_vcycle",
           "hdd_apply_preconditioner", "hdd_sigma_info", "hdd_interface_dofs", "hdd_interface_coords",
           "hdd_interface_matrix", "hdd_sigma_vcycle", "hdd_sigma_smooth", "hdd_pcg_solve", "hdd_setup_seconds", "hdd_profile_sweep", "hdd_stats",
           "hdd_bottom_spectrum", "hdd_profile_copy", "hdd_sigma_plan_stats",
           "hdd_profile_solve", "hdd_create_batch", "hdd_batch_count", "hdd_batch_num_pages", "hdd_batch_page_offsets",
           "hdd_batch_pcg_solve", "hdd_batch_get_x", "hdd_create_batch_from", "hdd_server_create", "hdd_server_submit",
           "hdd_server_wait", "hdd_server_cancel", "hdd_server_destroy", "hdd_server_last_error", "hdd_server_budget", "hdd_set_graph", "hdd_set_vcycle_storage",
           "hdd_set_sigma_refinement", "hdd_sigma_factor_solve", "hdd_set_stream", "hdd_synchronize", "hdd_multiply",
           "hdd_inner_product", "hdd_convergence_norm", "hdd_project", "hdd_vector_copy", "hdd_vector_copy2",
           "hdd_vector_plus_equal", "hdd_vector_minus_equal", "hdd_vector_scale", "hdd_apply_preconditioner_channels",
           "hdd_set_pcg_precision", "hdd_get_x_f64", "hdd_set_dot_order", "hdd_inner_product_noalloc", "hdd_batches_pcg_solve",
           "hdd_set_host_threads", "hdd_profile_sweep_grid"]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HddError(f"{LIB_PATH} is missing: build it with `python -m hierarchicaldd_amd.build` "
                           "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        P, I = C.c_void_p, C.c_int
        L.hdd_create.argtypes = [C.POINTER(ProblemDesc), C.POINTER(C.c_void_p)]
        L.hdd_create_host_only.argtypes = [C.POINTER(ProblemDesc), C.POINTER(C.c_void_p)]
        L.hdd_destroy.argtypes = [P]
        L.hdd_last_error.argtypes = [P]
        L.hdd_last_error.restype = C.c_char_p
        L.hdd_num_pages.argtypes = [P, C.POINTER(I)]
        L.hdd_page_offsets.argtypes = [P, P]
        L.hdd_set_channel.argtypes = [P, I, P]
        L.hdd_get_channel.argtypes = [P, I, P]
        L.hdd_level_pages.argtypes = [P, I, C.POINTER(I)]
        L.hdd_level_flags.argtypes = [P, I, P, P]
        L.hdd_smooth.argtypes = [P, I]
        L.hdd_vcycle.argtypes = [P]
        L.hdd_apply_preconditioner.argtypes = [P]
        L.hdd_sigma_info.argtypes = [P, C.POINTER(I), C.POINTER(I)]
        L.hdd_interface_dofs.argtypes = [P, I, C.POINTER(I)]
        L.hdd_interface_coords.argtypes = [P, I, P]
        L.hdd_interface_matrix.argtypes = [P, I, I, I, C.POINTER(I), C.POINTER(I), C.POINTER(I), P, P, P]
        L.hdd_sigma_vcycle.argtypes = [P, P, P]
        L.hdd_sigma_smooth.argtypes = [P, I, P, P]
        L.hdd_pcg_solve.argtypes = [P, C.POINTER(PcgParams), C.POINTER(PcgReport), P, I]
        L.hdd_setup_seconds.argtypes = [P, C.POINTER(C.c_double)]
        L.hdd_profile_sweep.argtypes = [P, I, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.hdd_profile_copy.argtypes = [P, I, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.hdd_sigma_plan_stats.argtypes = [P, I, I, P, I, C.POINTER(I)]
        L.hdd_profile_solve.argtypes = [P, I, I, I, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.hdd_stats.argtypes = [P, P, I]
        L.hdd_bottom_spectrum.argtypes = [P, C.c_double, P, I, C.POINTER(I), C.POINTER(C.c_double)]
        L.hdd_create_batch.argtypes = [C.POINTER(ProblemDesc), I, C.POINTER(C.c_void_p)]
        L.hdd_batch_count.argtypes = [P, C.POINTER(I)]
        L.hdd_batch_num_pages.argtypes = [P, I, C.POINTER(I)]
        L.hdd_batch_page_offsets.argtypes = [P, I, P]
        L.hdd_batch_pcg_solve.argtypes = [P, C.POINTER(PcgParams), C.POINTER(PcgReport), P, I]
        L.hdd_batch_get_x.argtypes = [P, I, P]
        L.hdd_create_batch_from.argtypes = [C.POINTER(C.c_void_p), I, C.POINTER(C.c_void_p)]
        L.hdd_server_create.argtypes = [C.POINTER(ServerParams), C.POINTER(C.c_void_p)]
        L.hdd_server_submit.argtypes = [P, C.POINTER(ProblemDesc), C.POINTER(PcgParams), C.POINTER(C.c_int64)]
        L.hdd_server_wait.argtypes = [P, C.c_int64, P, C.POINTER(RequestReport)]
        L.hdd_server_cancel.argtypes = [P, C.c_int64]
        L.hdd_server_destroy.argtypes = [P]
        L.hdd_server_last_error.argtypes = [P]
        L.hdd_server_last_error.restype = C.c_char_p
        L.hdd_server_budget.argtypes = [P, C.POINTER(C.c_double)]
        L.hdd_set_graph.argtypes = [P, I]
        L.hdd_set_vcycle_storage.argtypes = [P, I]
        L.hdd_set_sigma_refinement.argtypes = [P, I]
        L.hdd_sigma_factor_solve.argtypes = [P, I, I, I, P, P]
        F = C.c_float
        L.hdd_set_stream.argtypes = [P, P]
        L.hdd_synchronize.argtypes = [P]
        L.hdd_multiply.argtypes = [P, I, I]
        L.hdd_inner_product.argtypes = [P, I, I, P]
        L.hdd_convergence_norm.argtypes = [P, I, P]
        L.hdd_project.argtypes = [P, I]
        L.hdd_vector_copy.argtypes = [P, I, F, I]
        L.hdd_vector_copy2.argtypes = [P, I, F, I, I]
        L.hdd_vector_plus_equal.argtypes = [P, I, I]
        L.hdd_vector_minus_equal.argtypes = [P, I, I]
        L.hdd_vector_scale.argtypes = [P, I, F]
        L.hdd_apply_preconditioner_channels.argtypes = [P, I, I, I]
        L.hdd_set_pcg_precision.argtypes = [P, I, I]
        L.hdd_get_x_f64.argtypes = [P, I, P]
        L.hdd_set_dot_order.argtypes = [P, I]
        L.hdd_batches_pcg_solve.argtypes = [C.POINTER(C.c_void_p), I, C.POINTER(PcgParams), C.POINTER(PcgReport)]
        L.hdd_inner_product_noalloc.argtypes = [P, P, C.c_int64, P, P]
        for name in EXPORTS:
            if name not in ("hdd_last_error", "hdd_server_last_error"):
                getattr(L, name).restype = I
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _desc(dom, cfg):
    keep = [np.ascontiguousarray(dom.page_offsets, dtype=np.uint64),
            np.ascontiguousarray(dom.flags, dtype=np.uint32),
            np.ascontiguousarray(dom.b, dtype=np.float32)]
    d = ProblemDesc()
    for a in range(3):
        d.size[a] = int(dom.size[a])
        d.sub_size[a] = int(cfg.sub_size)
    d.vlevels, d.sigma_levels, d.nVcycles = cfg.vlevels, cfg.sigma_levels, cfg.nVcycles
    d.interior_smoothing, d.boundary_smoothing = cfg.interior_smoothing, cfg.boundary_smoothing
    d.sigma_smoothing, d.sigma_mg = cfg.sigma_smoothing, cfg.sigma_mg
    d.n_pages = keep[0].shape[0]
    d.page_offsets, d.flags, d.b = (_ptr(a) for a in keep)
    return d, keep


class DDSolver:
    """Host facade over one hdd_ctx (Read_SPGrid + Initialize_Preconditioner in the constructor)."""

    def __init__(self, dom, cfg, host_only: bool = False):
        L = lib()
        d, self._keep = _desc(dom, cfg)
        h = C.c_void_p()
        st = (L.hdd_create_host_only if host_only else L.hdd_create)(C.byref(d), C.byref(h))
        if st != 0:
            raise HddError(f"hdd_create failed ({st}): {L.hdd_last_error(None).decode()}")
        self.h = h
        n = C.c_int()
        self._check(L.hdd_num_pages(self.h, C.byref(n)))
        self.n_pages = n.value

    def _check(self, st, ok=(0,)):
        if st not in ok:
            raise HddError(f"hdd call failed ({st}): {lib().hdd_last_error(self.h).decode()}")
        return st

    def close(self):
        if getattr(self, "h", None):
            lib().hdd_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def setup_seconds(self):
        s = C.c_double()
        self._check(lib().hdd_setup_seconds(self.h, C.byref(s)))
        return s.value

    def page_offsets(self):
        out = np.zeros(self.n_pages, dtype=np.uint64)
        self._check(lib().hdd_page_offsets(self.h, _ptr(out)))
        return out

    def set(self, ch, v):
        v = np.ascontiguousarray(v, dtype=np.float32).reshape(self.n_pages, 256)
        self._check(lib().hdd_set_channel(self.h, ch, _ptr(v)))

    def get(self, ch):
        out = np.zeros((self.n_pages, 256), dtype=np.float32)
        self._check(lib().hdd_get_channel(self.h, ch, _ptr(out)))
        return out

    def level(self, l):
        n = C.c_int()
        self._check(lib().hdd_level_pages(self.h, l, C.byref(n)))
        off = np.zeros(n.value, dtype=np.uint64)
        fl = np.zeros((n.value, 256), dtype=np.uint32)
        self._check(lib().hdd_level_flags(self.h, l, _ptr(off), _ptr(fl)))
        return off, fl

    def smooth(self, mode):
        self._check(lib().hdd_smooth(self.h, mode))

    def vcycle(self, b):
        self.set(CH_R, b)
        self._check(lib().hdd_vcycle(self.h))
        return self.get(CH_Z)

    def apply_preconditioner(self, r):
        self.set(CH_R, r)
        self._check(lib().hdd_apply_preconditioner(self.h))
        return self.get(CH_Z)

    def sigma_info(self):
        lv, ns = C.c_int(), C.c_int()
        self._check(lib().hdd_sigma_info(self.h, C.byref(lv), C.byref(ns)))
        return lv.value, ns.value

    def interface_dofs(self, l):
        n = C.c_int()
        self._check(lib().hdd_interface_dofs(self.h, l, C.byref(n)))
        return n.value

    def interface_coords(self, l):
        out = np.zeros((self.interface_dofs(l), 3), dtype=np.int32)
        self._check(lib().hdd_interface_coords(self.h, l, _ptr(out)))
        return out

    def matrix(self, which, level, sub=0):
        w = {"Arr": 0, "Aii": 1, "Air": 2, "Prr": 3}[which]
        r, c, nz = C.c_int(), C.c_int(), C.c_int()
        self._check(lib().hdd_interface_matrix(self.h, w, level, sub, C.byref(r), C.byref(c), C.byref(nz), None, None, None))
        ptr = np.zeros(r.value + 1, dtype=np.int32)
        idx = np.zeros(nz.value, dtype=np.int32)
        val = np.zeros(nz.value, dtype=np.float32)
        self._check(lib().hdd_interface_matrix(self.h, w, level, sub, C.byref(r), C.byref(c), C.byref(nz),
                                               _ptr(ptr), _ptr(idx), _ptr(val)))
        return r.value, c.value, ptr, idx, val

    def sigma_factor_solve(self, level, which, sub, b):
        """(A + 1e-5 I)^{-1} b in double with the host factor of Arr or of subdomain sub's Aii."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        self._check(lib().hdd_sigma_factor_solve(self.h, level, {"Arr": 0, "Aii": 1}[which], sub, _ptr(b), _ptr(x)))
        return x

    def sigma_vcycle(self, b):
        b = np.ascontiguousarray(b, dtype=np.float32)
        u = np.zeros_like(b)
        self._check(lib().hdd_sigma_vcycle(self.h, _ptr(b), _ptr(u)))
        return u

    def sigma_smooth(self, level, b, u):
        b = np.ascontiguousarray(b, dtype=np.float32)
        u = np.ascontiguousarray(u, dtype=np.float32).copy()
        self._check(lib().hdd_sigma_smooth(self.h, level, _ptr(b), _ptr(u)))
        return u

    def set_graph(self, enable: bool):
        """Replay the preconditioner apply from captured HIP graphs, or launch every kernel (default)."""
        self._check(lib().hdd_set_graph(self.h, int(bool(enable))))

    def set_fp16(self, enable: bool):
        """FP16 storage (FP32 arithmetic) for the V-cycle vectors of the preconditioner."""
        self._check(lib().hdd_set_vcycle_storage(self.h, int(bool(enable))))

    def set_outer_fp64(self, enable: bool, replace_every: int = 25):
        """FP64 PCG scalars and iterate with true-residual replacement (default: the reference's float scalars)."""
        self._check(lib().hdd_set_pcg_precision(self.h, int(bool(enable)), int(replace_every)))

    def x_f64(self, problem: int = 0):
        """The FP64 iterate of the last FP64-outer solve, [n_pages, 256] in page order."""
        out = np.zeros((self.n_pages, 256), dtype=np.float64)
        self._check(lib().hdd_get_x_f64(self.h, problem, _ptr(out)))
        return out

    def set_serial_dots(self, enable: bool):
        """Inner products in the reference's serial summation order (slow; for parity checks)."""
        self._check(lib().hdd_set_dot_order(self.h, int(bool(enable))))

    def set_sigma_refinement(self, steps: int):
        """Iterative-refinement steps after every exact solve of the interface solver (0: none, default)."""
        self._check(lib().hdd_set_sigma_refinement(self.h, int(steps)))

    def solve(self, tolerance: float, max_iterations: int = 5000):
        """PCG from x=0; returns (iterations, residual history, seconds)."""
        p = PcgParams(tolerance, max_iterations)
        rep = PcgReport()
        hist = np.zeros(max_iterations + 1, dtype=np.float64)
        self._check(lib().hdd_pcg_solve(self.h, C.byref(p), C.byref(rep), _ptr(hist), max_iterations + 1), ok=(0, 4))
        n = abs(rep.iterations) + 1
        return rep.iterations, hist[:n], rep.seconds

    def solve_status(self, tolerance: float, max_iterations: int = 5000):
        """PCG from x=0; returns (status code, report, residual history) without raising on
        HDD_ERR_NOT_CONVERGED (status 4)."""
        p = PcgParams(tolerance, max_iterations)
        rep = PcgReport()
        hist = np.zeros(max_iterations + 1, dtype=np.float64)
        st = lib().hdd_pcg_solve(self.h, C.byref(p), C.byref(rep), _ptr(hist), max_iterations + 1)
        self._check(st, ok=(0, 4))
        return st, rep, hist[:abs(rep.iterations) + 1]

    def profile_sweep(self, reps: int = 50):
        """Mean duration (ms) of one level-0 smoothing sweep and its algorithmic bytes."""
        ms, by = C.c_double(), C.c_double()
        self._check(lib().hdd_profile_sweep(self.h, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def profile_copy(self, reps: int = 20):
        """Mean duration (ms) of a plain copy of one level-0 channel and its bytes read plus written."""
        ms, by = C.c_double(), C.c_double()
        self._check(lib().hdd_profile_copy(self.h, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def stats(self):
        keys = ["setup_s", "factor_s", "sigma_device_bytes", "pages", "interface_dofs", "interior_dofs", "subdomains",
                "bottom_rho", "bottom_direct", "bottom_modes", "bottom_eig_s", "bottom_krylov_steps", "bottom_check",
                "tuned_launch_shapes", "dense_top_rows", "bottom_page_local", "fused_heights", "wave_heights"]
        out = np.zeros(len(keys), dtype=np.float64)
        self._check(lib().hdd_stats(self.h, _ptr(out), len(keys)))
        return dict(zip(keys, out.tolist()))

    def profile_solve(self, level: int, which: str, reps: int = 20):
        """Mean duration (ms) of one triangular-solve plan ("Aii", "Arr" or "bottom") and its algorithmic bytes."""
        ms, by = C.c_double(), C.c_double()
        self._check(lib().hdd_profile_solve(self.h, level, {"Aii": 0, "Arr": 1, "bottom": 2}[which], reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def plan_stats(self, level: int, which: str):
        """Per tree height of a triangular-solve plan ("Aii" batch or "Arr"): supernodes, rows,
        packed diagonal entries, L21 entries, largest supernode, most L21 rows of a supernode,
        most incidences of a row, all incidences."""
        out = np.zeros(8 * 64, dtype=np.float64)
        h = C.c_int()
        self._check(lib().hdd_sigma_plan_stats(self.h, level, {"Aii": 0, "Arr": 1}[which], _ptr(out), out.size, C.byref(h)))
        return out[:8 * min(h.value, 64)].reshape(-1, 8)

    def bottom_spectrum(self, lambda_cut: float, cap: int = 1024):
        """Eigenvalues >= lambda_cut of the bottom sigma smoother's iteration matrix (host, double)."""
        lam = np.zeros(cap, dtype=np.float64)
        k, lmax = C.c_int(), C.c_double()
        self._check(lib().hdd_bottom_spectrum(self.h, lambda_cut, _ptr(lam), cap, C.byref(k), C.byref(lmax)))
        return lam[:min(k.value, cap)].copy(), lmax.value

    # reference-facade spellings
    Solve = solve


class DDBatch(DDSolver):
    """Several independent problems in one context (hdd_create_batch): one launch per stage for all of them.

    ``problems`` is a list of (domain, config) pairs, or ``from_solvers`` a list of
    DDSolver contexts whose host setup is reused (hdd_create_batch_from)."""

    def __init__(self, problems=None, from_solvers=None):
        L = lib()
        h = C.c_void_p()
        self._keep = []
        if from_solvers is not None:
            arr = (C.c_void_p * len(from_solvers))(*[g.h.value for g in from_solvers])
            st = L.hdd_create_batch_from(arr, len(from_solvers), C.byref(h))
        else:
            descs = (ProblemDesc * len(problems))()
            for i, (dom, cfg) in enumerate(problems):
                d, keep = _desc(dom, cfg)
                descs[i] = d
                self._keep += keep
            st = L.hdd_create_batch(descs, len(problems), C.byref(h))
        if st != 0:
            raise HddError(f"hdd_create_batch failed ({st}): {L.hdd_last_error(None).decode()}")
        self.h = h
        n = C.c_int()
        self._check(L.hdd_num_pages(self.h, C.byref(n)))
        self.n_pages = n.value
        self._check(L.hdd_batch_count(self.h, C.byref(n)))
        self.count = n.value

    def problem_pages(self, p):
        n = C.c_int()
        self._check(lib().hdd_batch_num_pages(self.h, p, C.byref(n)))
        return n.value

    def problem_page_offsets(self, p):
        out = np.zeros(self.problem_pages(p), dtype=np.uint64)
        self._check(lib().hdd_batch_page_offsets(self.h, p, _ptr(out)))
        return out

    def solve_all(self, tolerances, max_iterations: int = 5000):
        """PCG from x=0 for every problem; returns a list of (iterations, residual history, seconds)."""
        n = self.count
        tol = np.broadcast_to(np.asarray(tolerances, dtype=np.float64), (n,))
        prm = (PcgParams * n)(*[PcgParams(float(t), int(max_iterations)) for t in tol])
        rep = (PcgReport * n)()
        hl = max_iterations + 1
        hist = np.zeros((n, hl), dtype=np.float64)
        self._check(lib().hdd_batch_pcg_solve(self.h, prm, rep, _ptr(hist), hl), ok=(0, 4))
        return [(rep[p].iterations, hist[p, :abs(rep[p].iterations) + 1].copy(), rep[p].seconds) for p in range(n)]

    def x(self, p):
        out = np.zeros((self.problem_pages(p), 256), dtype=np.float32)
        self._check(lib().hdd_batch_get_x(self.h, p, _ptr(out)))
        return out


def profile_sweep_grid(dom, cfg, reps: int = 50):
    """Mean duration (ms) of one FP32 level-0 sweep of a domain's page set on its own
    (hdd_profile_sweep_grid: no solver set-up) and its algorithmic bytes."""
    d, keep = _desc(dom, cfg)
    ms, by = C.c_double(), C.c_double()
    st = lib().hdd_profile_sweep_grid(C.byref(d), int(reps), C.byref(ms), C.byref(by))
    if st != 0:
        raise HddError(f"hdd_profile_sweep_grid failed ({st}): {lib().hdd_last_error(None).decode()}")
    del keep
    return ms.value, by.value


def set_host_threads(n: int):
    """Host threads of the set-ups the calling thread starts from now on (hdd_set_host_threads;
    0: all the host's)."""
    st = lib().hdd_set_host_threads(int(n))
    if st != 0:
        raise HddError(f"hdd_set_host_threads failed ({st}): {lib().hdd_last_error(None).decode()}")


def solve_concurrently(batches, tolerances, max_iterations: int = 5000):
    """PCG for every problem of several DDBatch / DDSolver contexts at once (hdd_batches_pcg_solve):
    each context iterates on its own stream, one iteration per context per round. tolerances: one
    per problem, contexts in order. Returns a list of (iterations, seconds, converged) per problem."""
    counts = [getattr(b, "count", 1) for b in batches]
    n = sum(counts)
    tol = np.broadcast_to(np.asarray(tolerances, dtype=np.float64), (n,))
    prm = (PcgParams * n)(*[PcgParams(float(t), int(max_iterations)) for t in tol])
    rep = (PcgReport * n)()
    arr = (C.c_void_p * len(batches))(*[b.h.value for b in batches])
    st = lib().hdd_batches_pcg_solve(arr, len(batches), prm, rep)
    if st not in (0, 4):
        raise HddError(f"hdd_batches_pcg_solve failed ({st}): {lib().hdd_last_error(None).decode()}")
    return [(rep[p].iterations, rep[p].seconds, bool(rep[p].converged)) for p in range(n)]


class Server:
    """Serving loop (hdd_server_*): submit problems as they arrive, wait for each result.
    Requests ready together form a unit (one batched context); up to `groups` units iterate
    on the device at once, and a unit admitted while others are mid-solve joins them at
    their next iteration (continuous batching)."""

    def __init__(self, setup_threads: int = 2, groups: int = 4, max_batch: int = 16, hbm_budget: float = 0.0):
        L = lib()
        prm = ServerParams(setup_threads, groups, max_batch, float(hbm_budget))
        h = C.c_void_p()
        st = L.hdd_server_create(C.byref(prm), C.byref(h))
        if st != 0:
            raise HddError(f"hdd_server_create failed ({st}): {L.hdd_server_last_error(None).decode()}")
        self.h = h
        self.n_pages = {}

    @property
    def budget(self):
        b = C.c_double()
        lib().hdd_server_budget(self.h, C.byref(b))
        return b.value

    def submit(self, dom, cfg, tolerance: float, max_iterations: int = 500) -> int:
        d, keep = _desc(dom, cfg)
        rid = C.c_int64()
        prm = PcgParams(float(tolerance), int(max_iterations))
        st = lib().hdd_server_submit(self.h, C.byref(d), C.byref(prm), C.byref(rid))
        if st != 0:
            raise HddError(f"hdd_server_submit failed ({st}): {lib().hdd_server_last_error(self.h).decode()}")
        self.n_pages[rid.value] = dom.n_pages
        return rid.value

    def wait(self, rid: int):
        """Blocks until request rid is done: (status, x [n_pages, 256] in the submitted page order, report dict)."""
        x = np.zeros((self.n_pages.pop(rid), 256), dtype=np.float32)
        rep = RequestReport()
        st = lib().hdd_server_wait(self.h, rid, _ptr(x), C.byref(rep))
        if st not in (0, 4, 5):
            raise HddError(f"request {rid} failed ({st}): {lib().hdd_server_last_error(self.h).decode()}")
        info = {k: getattr(rep, k) for k, _ in RequestReport._fields_}
        return st, x, info

    def cancel(self, rid: int) -> bool:
        """Asks request rid to stop (dropped if queued, stopped at its next residual test if
        running); False if it is unknown or already done. Its wait() then returns status 5."""
        return lib().hdd_server_cancel(self.h, int(rid)) == 0

    def close(self):
        if getattr(self, "h", None):
            lib().hdd_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
