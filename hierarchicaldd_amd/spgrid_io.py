"""The reference's on-disk SPGrid directory format (SPGRID_READ_WRITE::Read_SPGrid /
Write_SPGrid, demos/Domain_Decomposition_CPU_Only/SPGrid_Domain_Decomposition_Solver/
SPGRID_READ_WRITE.h:31-106; the generators write it, the solver reads it).

A directory holds four little-endian binary files (PhysBAM Write_Binary with
float as the read/write type):

* ``size``          d int32: the padded grid size of the generator's allocator;
* ``block_offset``  int32 count, then an array view: int32 count and count
                    uint64 SPGrid linear offsets of the pages (Morton page
                    index << 12), in the set's (ascending) block order;
* ``flags``         per page an array view: int32 256, then 256 uint32 flags;
* ``data``          per page an array view: int32 256, then 256 float32 values
                    (the right-hand side channel).

Pages hold 256 cells: 4(x) x 8(y) x 8(z) in 3D, 16(x) x 16(y) in 2D, the last
axis fastest. This module reads and writes the format for d = 2 and 3, and
converts 3D directories to and from the solver's SPGridDomain.
"""
from __future__ import annotations

import os
import struct

import numpy as np

PAGE_CELLS = 256


def read_spgrid(path: str):
    """Returns (size tuple, page offsets uint64 [n], flags uint32 [n,256], data float32 [n,256])."""
    with open(os.path.join(path, "size"), "rb") as f:
        raw = f.read()
    if len(raw) not in (8, 12):
        raise ValueError(f"{path}/size: expected 2 or 3 int32, got {len(raw)} bytes")
    size = struct.unpack("<%di" % (len(raw) // 4), raw)
    with open(os.path.join(path, "block_offset"), "rb") as f:
        raw = f.read()
    n, m = struct.unpack("<2i", raw[:8])
    if n != m or len(raw) != 8 + 8 * n:
        raise ValueError(f"{path}/block_offset: inconsistent counts ({n}, {m}) for {len(raw)} bytes")
    offsets = np.frombuffer(raw, dtype="<u8", offset=8, count=n).astype(np.uint64)
    flags = _read_pages(os.path.join(path, "flags"), n, "<u4").astype(np.uint32)
    data = _read_pages(os.path.join(path, "data"), n, "<f4").astype(np.float32)
    return size, offsets, flags, data


def _read_pages(fn, n, dtype):
    raw = np.fromfile(fn, dtype=np.uint8)
    rec = 4 + 4 * PAGE_CELLS
    if raw.size != n * rec:
        raise ValueError(f"{fn}: expected {n} pages of {rec} bytes, got {raw.size} bytes")
    rows = raw.reshape(n, rec)
    counts = rows[:, :4].copy().view("<i4").ravel()
    if np.any(counts != PAGE_CELLS):
        raise ValueError(f"{fn}: page element counts must be {PAGE_CELLS}")
    return rows[:, 4:].copy().view(dtype).reshape(n, PAGE_CELLS)


def write_spgrid(path: str, size, offsets, flags, data) -> None:
    """Writes a directory the reference's Read_SPGrid accepts."""
    os.makedirs(path, exist_ok=True)
    offsets = np.ascontiguousarray(offsets, dtype="<u8")
    n = offsets.shape[0]
    flags = np.ascontiguousarray(flags, dtype="<u4").reshape(n, PAGE_CELLS)
    data = np.ascontiguousarray(data, dtype="<f4").reshape(n, PAGE_CELLS)
    with open(os.path.join(path, "size"), "wb") as f:
        f.write(struct.pack("<%di" % len(size), *[int(s) for s in size]))
    with open(os.path.join(path, "block_offset"), "wb") as f:
        f.write(struct.pack("<2i", n, n))
        f.write(offsets.tobytes())
    head = np.full((n, 1), PAGE_CELLS, dtype="<i4").view(np.uint8)
    for name, arr in (("flags", flags), ("data", data)):
        with open(os.path.join(path, name), "wb") as f:
            f.write(np.concatenate([head, arr.view(np.uint8)], axis=1).tobytes())


def _next_pow2(v: int) -> int:
    return 1 << max(0, int(v) - 1).bit_length()


def padded_size_3d(x: int, y: int, z: int, block=(4, 8, 8)):
    """Allocated size of a 3D SPGrid of x*y*z cells (SPGrid_Geometry<3>,
    external_libraries/Physbam/Common_Libraries/SPGrid/Core/SPGrid_Geometry.h:22-28)."""
    zp = _next_pow2(max(x, y, z, block[2]))
    yp = max(_next_pow2(max(y, x)), zp >> 1, block[1])
    xp = max(_next_pow2(x), zp >> 1, block[0])
    return xp, yp, zp


def write_domain(path: str, dom) -> None:
    """Writes a 3D SPGridDomain of generator size n = solver size - 1; the size
    file holds the generator's allocated size (the solver then runs with size
    n + 1, as the reference's README pairs a 256 domain with -size 257)."""
    n = [s - 1 for s in dom.size]
    write_spgrid(path, padded_size_3d(*n), dom.page_offsets, dom.flags, dom.b)


def read_domain(path: str, solver_size, meta=None):
    """Reads a 3D directory into an SPGridDomain for a solver of size ``solver_size``
    (the generator size + 1 per axis; the size file holds only the allocated size)."""
    from hierarchicaldd_amd.domain import SPGridDomain
    size, offsets, flags, data = read_spgrid(path)
    if len(size) != 3:
        raise ValueError("the solver is three-dimensional; this directory holds a 2D grid")
    order = np.argsort(offsets, kind="stable")
    return SPGridDomain(tuple(solver_size), offsets[order], flags[order], data[order], meta or {"source": path})
