"""Synthetic SPGrid Poisson domains in the reference's on-disk/in-memory format.

Rules follow the reference generator (demos/SPGrid_Create_Simple_Domain/
Create_3D_Domain/main.cpp:72-152, 217-306):

* generator size ``n`` -> cells ``[0, n-1]``; cells outside ``[1, n-2]`` are
  Dirichlet (``:91-92,118``);
* obstacle cells (inside a sphere, ``|i - c|^2 <= r^2`` on integer cell
  indices, ``:121``) carry flag 0, i.e. they are Neumann exterior;
* Dirichlet cells without an Active face neighbour are dropped (``:132-139``);
* a page is allocated iff it holds an Active or a kept Dirichlet cell;
* face bits are set by ``Flaging`` (DDS/SPGrid_Flag_Helper.h:42-74, called at
  ``:299``): bit ``f`` of an Active cell is set iff its neighbour across face
  ``f`` (order -x,+x,-y,+y,-z,+z) is Active, Interface or Dirichlet;
* the right-hand side is a delta (or any field) in the ``b`` channel.

The solver is then run with ``size = n + 1`` (README.md:124 pairs a domain
created with 256 with ``-size 257``).

Pages are 4(x) x 8(y) x 8(z) cells, z fastest inside a page, and pages are
listed in SPGrid Morton order (SPGrid/Core/SPGrid_Mask.h:20-140, 16-byte
struct, lob=1: page bit 0 = x, 1 = z, 2 = y).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

ACTIVE = 0x10000
DIRICHLET = 0x20000
BOUNDARY = 0x40000
INTERFACE = 0x80000
FACE_MX = 0x100000

BX, BY, BZ = 4, 8, 8


def morton_page(bx, by, bz):
    """SPGrid page index of block coordinates (x lowest, then z, then y)."""
    bx = np.asarray(bx, dtype=np.uint64)
    by = np.asarray(by, dtype=np.uint64)
    bz = np.asarray(bz, dtype=np.uint64)
    m = np.zeros(np.broadcast(bx, by, bz).shape, dtype=np.uint64)
    one = np.uint64(1)
    for k in range(21):
        kk = np.uint64(k)
        m |= ((bx >> kk) & one) << np.uint64(3 * k)
        m |= ((bz >> kk) & one) << np.uint64(3 * k + 1)
        m |= ((by >> kk) & one) << np.uint64(3 * k + 2)
    return m


def morton_decode(m):
    m = np.asarray(m, dtype=np.uint64)
    bx = np.zeros(m.shape, dtype=np.int64)
    by = np.zeros(m.shape, dtype=np.int64)
    bz = np.zeros(m.shape, dtype=np.int64)
    one = np.uint64(1)
    for k in range(21):
        bx |= ((m >> np.uint64(3 * k)) & one).astype(np.int64) << k
        bz |= ((m >> np.uint64(3 * k + 1)) & one).astype(np.int64) << k
        by |= ((m >> np.uint64(3 * k + 2)) & one).astype(np.int64) << k
    return bx, by, bz


@dataclass
class SPGridDomain:
    """A paged SPGrid problem: ``size`` is the solver size (n+1 per axis)."""

    size: tuple
    page_offsets: np.ndarray  # uint64 [nb], SPGrid linear offset of each page (Morton << 12)
    flags: np.ndarray         # uint32 [nb, 256]
    b: np.ndarray             # float32 [nb, 256]
    meta: dict = field(default_factory=dict)

    @property
    def n_pages(self) -> int:
        return int(self.page_offsets.shape[0])

    def dense(self, channel: np.ndarray, dtype=None) -> np.ndarray:
        """Scatter a per-page channel [nb,256] into a dense [X,Y,Z] array."""
        bx, by, bz = morton_decode(self.page_offsets >> np.uint64(12))
        X = int((bx.max() + 1) * BX)
        Y = int((by.max() + 1) * BY)
        Z = int((bz.max() + 1) * BZ)
        out = np.zeros((X, Y, Z), dtype=dtype or channel.dtype)
        blk = channel.reshape(-1, BX, BY, BZ)
        for i in range(self.n_pages):
            out[bx[i] * BX:(bx[i] + 1) * BX, by[i] * BY:(by[i] + 1) * BY, bz[i] * BZ:(bz[i] + 1) * BZ] = blk[i]
        return out


def _face_bits(flags: np.ndarray) -> np.ndarray:
    """Flaging<3>: OR face bits into Active|Interface cells of a dense array."""
    solid = (flags & (ACTIVE | INTERFACE | DIRICHLET)) != 0
    target = (flags & (ACTIVE | INTERFACE)) != 0
    out = flags.copy()
    pad = np.pad(solid, 1)
    X, Y, Z = flags.shape
    nbr = [pad[0:X, 1:Y + 1, 1:Z + 1], pad[2:X + 2, 1:Y + 1, 1:Z + 1],
           pad[1:X + 1, 0:Y, 1:Z + 1], pad[1:X + 1, 2:Y + 2, 1:Z + 1],
           pad[1:X + 1, 1:Y + 1, 0:Z], pad[1:X + 1, 1:Y + 1, 2:Z + 2]]
    for f, nb in enumerate(nbr):
        out[target & nb] |= np.uint32(FACE_MX << f)
    return out


def page_dense_grid(flags: np.ndarray, b: np.ndarray, size, meta=None) -> SPGridDomain:
    """Cut dense flags/b arrays (already padded to page multiples) into pages.

    A page is kept iff it holds a nonzero flag (Active or kept Dirichlet)."""
    X, Y, Z = flags.shape
    assert X % BX == 0 and Y % BY == 0 and Z % BZ == 0
    fb = flags.reshape(X // BX, BX, Y // BY, BY, Z // BZ, BZ).transpose(0, 2, 4, 1, 3, 5)
    bb = b.reshape(X // BX, BX, Y // BY, BY, Z // BZ, BZ).transpose(0, 2, 4, 1, 3, 5)
    keep = fb.reshape(X // BX, Y // BY, Z // BZ, -1).any(axis=3)
    bx, by, bz = np.nonzero(keep)
    key = morton_page(bx, by, bz)
    order = np.argsort(key, kind="stable")
    bx, by, bz, key = bx[order], by[order], bz[order], key[order]
    pf = fb[bx, by, bz].reshape(-1, 256).astype(np.uint32)
    pb = bb[bx, by, bz].reshape(-1, 256).astype(np.float32)
    return SPGridDomain(tuple(int(s) for s in size), key << np.uint64(12), np.ascontiguousarray(pf),
                        np.ascontiguousarray(pb), meta or {})


def _padded(n):
    return (-(-n // BX) * BX, -(-n // BY) * BY, -(-n // BZ) * BZ)


def sphere_obstacle_box(n: int, spheres, delta=None, rhs: np.ndarray | None = None, meta=None) -> SPGridDomain:
    """Box of n^3 cells with a Dirichlet ring and Neumann sphere obstacles.

    ``spheres`` is a list of (cx, cy, cz, r); ``delta`` is ((x,y,z), value)."""
    X, Y, Z = _padded(n)
    i = np.arange(X)[:, None, None]
    j = np.arange(Y)[None, :, None]
    k = np.arange(Z)[None, None, :]
    inside = (i < n) & (j < n) & (k < n)
    interior = (i >= 1) & (i <= n - 2) & (j >= 1) & (j <= n - 2) & (k >= 1) & (k <= n - 2)
    obstacle = np.zeros((X, Y, Z), dtype=bool)
    for (cx, cy, cz, r) in spheres:
        obstacle |= ((i - cx) ** 2 + (j - cy) ** 2 + (k - cz) ** 2) <= r * r
    flags = np.zeros((X, Y, Z), dtype=np.uint32)
    flags[inside & ~interior] = DIRICHLET
    flags[interior & ~obstacle] = ACTIVE
    # Dirichlet cells with no Active face neighbour are dropped
    act = np.pad(flags == ACTIVE, 1)
    has_act = (act[0:X, 1:Y + 1, 1:Z + 1] | act[2:X + 2, 1:Y + 1, 1:Z + 1] | act[1:X + 1, 0:Y, 1:Z + 1] |
               act[1:X + 1, 2:Y + 2, 1:Z + 1] | act[1:X + 1, 1:Y + 1, 0:Z] | act[1:X + 1, 1:Y + 1, 2:Z + 2])
    flags[(flags == DIRICHLET) & ~has_act] = 0
    flags = _face_bits(flags)
    b = np.zeros((X, Y, Z), dtype=np.float32)
    if rhs is not None:
        b[:n, :n, :n] = np.where((flags[:n, :n, :n] & ACTIVE) != 0, rhs, 0)
    if delta is not None:
        (dx, dy, dz), val = delta
        if not flags[dx, dy, dz] & ACTIVE:
            raise ValueError(f"delta position {(dx, dy, dz)} is not an Active cell")
        b[dx, dy, dz] = val
    return page_dense_grid(flags, b, (n + 1, n + 1, n + 1), meta)


@dataclass
class DDConfig:
    """Solver parameters of the reference's PARSE_ARGS flags (CPU/Test_Bench_3D/main.cpp:235-252)."""

    sub_size: int = 16
    vlevels: int = 3
    nVcycles: int = 1
    interior_smoothing: int = 1
    boundary_smoothing: int = 3
    sigma_levels: int = 3
    sigma_smoothing: int = 3
    sigma_mg: int = 1


def config0():
    """64^3 cells, sphere r=16 at (32,32,32), delta 1.0 at (16,16,16); sub 16, vlevels 3."""
    d = sphere_obstacle_box(64, [(32, 32, 32, 16)], delta=((16, 16, 16), 1.0),
                            meta={"workload": "sphere64_sub16"})
    return d, DDConfig(sub_size=16, vlevels=3)


def config1():
    """256^3 cells, sphere r=64 at (128,128,128), delta 1.0 at (64,64,64); sub 32, vlevels 4."""
    d = sphere_obstacle_box(256, [(128, 128, 128, 64)], delta=((64, 64, 64), 1.0),
                            meta={"workload": "sphere256_sub32"})
    return d, DDConfig(sub_size=32, vlevels=4)


def small_box(n: int = 32, sub: int = 16, vlevels: int = 3, seed: int = 0, sigma_levels: int = 3):
    """Small test problem: one sphere obstacle and a random RHS on Active cells."""
    rng = np.random.default_rng(seed)
    c = n // 2
    rhs = rng.standard_normal((n, n, n)).astype(np.float32)
    d = sphere_obstacle_box(n, [(c + 1, c - 2, c, max(2, n // 6))], rhs=rhs, meta={"workload": f"box{n}_sub{sub}"})
    return d, DDConfig(sub_size=sub, vlevels=vlevels, sigma_levels=sigma_levels)
