"""Independent numpy statements of the operators, used to pin the oracle.

Built from the problem definition only (cells, flags, faces), not from the
oracle's code: the global 7-point operator A (unit link weights, Dirichlet
neighbours contribute to the diagonal, exterior neighbours are Neumann), the
restriction / prolongation stencils, and cell lists in page order.
"""
import numpy as np
import scipy.sparse as sp

from hierarchicaldd_amd import domain

FACES = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]


def cell_coords(page_offsets):
    bx, by, bz = domain.morton_decode(np.asarray(page_offsets, dtype=np.uint64) >> np.uint64(12))
    c = np.arange(256)
    x = (bx[:, None] * 4 + (c >> 6)[None, :])
    y = (by[:, None] * 8 + ((c >> 3) & 7)[None, :])
    z = (bz[:, None] * 8 + (c & 7)[None, :])
    return x, y, z


def global_operator(page_offsets, flags, unknown_mask=domain.ACTIVE | domain.INTERFACE):
    """Sparse A over the cells selected by unknown_mask (page order), 7-point with
    neighbour classes taken from the flags themselves (not the face bits)."""
    x, y, z = cell_coords(page_offsets)
    f = flags.reshape(-1)
    x, y, z = x.reshape(-1), y.reshape(-1), z.reshape(-1)
    key = {(int(a), int(b), int(c)): i for i, (a, b, c) in enumerate(zip(x, y, z))}
    unk = np.nonzero(f & unknown_mask)[0]
    pos = -np.ones(f.shape[0], dtype=np.int64)
    pos[unk] = np.arange(unk.shape[0])
    rows, cols, vals = [], [], []
    for i in unk:
        d = 0.0
        for (dx, dy, dz) in FACES:
            j = key.get((int(x[i]) + dx, int(y[i]) + dy, int(z[i]) + dz))
            if j is None:
                continue
            fj = f[j]
            if fj & unknown_mask:
                rows.append(pos[i]); cols.append(pos[j]); vals.append(-1.0); d += 1.0
            elif fj & (domain.DIRICHLET | domain.ACTIVE | domain.INTERFACE):
                d += 1.0
        rows.append(pos[i]); cols.append(pos[i]); vals.append(d)
    n = unk.shape[0]
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n)), unk


def to_dense(page_offsets, values, pad=0):
    """Scatter per-page values [n_pages, 256] into a dense array indexed by cell
    coordinates (x, y, z) + pad; cells of missing pages are 0."""
    bx, by, bz = domain.morton_decode(np.asarray(page_offsets, dtype=np.uint64) >> np.uint64(12))
    X, Y, Z = int(bx.max() + 1) * 4, int(by.max() + 1) * 8, int(bz.max() + 1) * 8
    out = np.zeros((X + 2 * pad, Y + 2 * pad, Z + 2 * pad), dtype=np.asarray(values).dtype)
    blk = np.asarray(values).reshape(-1, 4, 8, 8)
    for i in range(len(bx)):
        x0, y0, z0 = int(bx[i]) * 4 + pad, int(by[i]) * 8 + pad, int(bz[i]) * 8 + pad
        out[x0:x0 + 4, y0:y0 + 8, z0:z0 + 8] = blk[i]
    return out


def from_dense(page_offsets, dense, pad=0):
    """Per-page values [n_pages, 256] gathered from a dense array (inverse of to_dense)."""
    bx, by, bz = domain.morton_decode(np.asarray(page_offsets, dtype=np.uint64) >> np.uint64(12))
    out = np.zeros((len(bx), 4, 8, 8), dtype=dense.dtype)
    for i in range(len(bx)):
        x0, y0, z0 = int(bx[i]) * 4 + pad, int(by[i]) * 8 + pad, int(bz[i]) * 8 + pad
        out[i] = dense[x0:x0 + 4, y0:y0 + 8, z0:z0 + 8]
    return out.reshape(-1, 256)


def face_residual(page_offsets, flags, x, b):
    """b - L x on Active|Interface cells (0 elsewhere), L x = sum over the cell's set
    face bits of (x_c - x_n): the reference's SPGrid_Laplace
    (DDS/SPGrid_Gemini_Arithmetic_Helpers.h:210-263), evaluated in float64 on a dense copy."""
    X = to_dense(page_offsets, np.asarray(x, dtype=np.float64), pad=1)
    out = np.asarray(b, dtype=np.float64).copy()
    for f, (dx, dy, dz) in enumerate(FACES):
        nb = from_dense(page_offsets, np.roll(X, shift=(-dx, -dy, -dz), axis=(0, 1, 2)), pad=1)
        has = ((flags >> np.uint32(20 + f)) & 1) != 0
        out -= np.where(has, x - nb, 0.0)
    unk = (flags & (domain.ACTIVE | domain.INTERFACE)) != 0
    return np.where(unk, out, 0.0)
