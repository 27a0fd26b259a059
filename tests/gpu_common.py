"""Shared helpers of the GPU parity tests (problem construction, comparisons)."""
import numpy as np

from hierarchicaldd_amd import domain


def random_rhs(flags, seed, mask=domain.ACTIVE | domain.INTERFACE):
    """Random values on Active|Interface cells, zero elsewhere (a projected PCG residual)."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(flags.shape).astype(np.float32)
    v[(flags & mask) == 0] = 0
    return v


def same_bits(a, b):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def hip_runtime():
    """The HIP runtime the library itself links (libamdhip64 of /opt/rocm), for
    device buffers and streams the tests own as a caller would."""
    import ctypes as C
    import os
    path = "/opt/rocm/lib/libamdhip64.so"
    H = C.CDLL(path if os.path.exists(path) else "libamdhip64.so")
    P = C.c_void_p
    H.hipMalloc.argtypes = [C.POINTER(P), C.c_size_t]
    H.hipFree.argtypes = [P]
    H.hipMemcpy.argtypes = [P, P, C.c_size_t, C.c_int]
    H.hipStreamCreate.argtypes = [C.POINTER(P)]
    H.hipStreamDestroy.argtypes = [P]
    H.hipStreamSynchronize.argtypes = [P]
    for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamCreate", "hipStreamDestroy", "hipStreamSynchronize"):
        getattr(H, f).restype = C.c_int
    return H


HIP_H2D, HIP_D2H = 1, 2
