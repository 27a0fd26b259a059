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
