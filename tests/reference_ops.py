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
