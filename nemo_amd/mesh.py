"""Host-side uniform periodic hexahedral mesh (the upload format).

The reference builds its element/face connectivity from a p4est forest
(mesh/p4est/mesh_utils_mod.f90:160-441, types mesh/p4est/mesh_types_mod.f90:
30-100).  For the uniform periodic boxes of the benchmark configurations the
forest is a single tree refined uniformly, whose leaves p4est orders along the
Morton (Z-order) curve.  This module produces the same ordering for
power-of-two boxes (lexicographic order otherwise) and flattens the
connectivity into the table the device reads:

    nbr[e, f]   neighbour element across face f of element e, with faces
                0 = x- (NOR), 1 = x+ (SOU), 2 = y- (WES), 3 = y+ (EAS),
                4 = z- (FRO), 5 = z+ (BAC)   (utils/constants_mod.f90:34-40)
    delta[e, d] element edge length in direction d (quad%delta)
    lower[e, d] lower corner of element e
    ijk[e, d]   lattice coordinates of element e

All faces are conforming and aligned; periodic wrap in every direction, or
(``with_boundaries``) boundary faces on the sides of the box that are not
periodic:

    nbr[e, f] = -(b + 1)  face f of e is boundary face b
    bfaces[b] = (element, face, kind), kind 0 = transmissive, 1 = slip wall

(the BC hook of the reference's face pass, kernel/mcfv/3d/kernel_flux_mod.f90:
122-226, with the setup's boundary Riemann solvers).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def _part1by2(n: np.ndarray) -> np.ndarray:
    n = n.astype(np.uint64) & np.uint64(0x1FFFFF)
    n = (n | (n << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
    n = (n | (n << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
    n = (n | (n << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
    n = (n | (n << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
    n = (n | (n << np.uint64(2))) & np.uint64(0x1249249249249249)
    return n


def morton_key(i: np.ndarray, j: np.ndarray, k: np.ndarray) -> np.ndarray:
    return _part1by2(i) | (_part1by2(j) << np.uint64(1)) | (_part1by2(k) << np.uint64(2))


BC_TRANSMISSIVE, BC_SLIP_WALL = 0, 1


@dataclass
class Mesh:
    shape: tuple            # (nx, ny, nz) elements
    extent: tuple           # physical box lengths (Lx, Ly, Lz)
    nbr: np.ndarray         # int32 [nelem, 6]
    delta: np.ndarray       # float64 [nelem, 3]
    lower: np.ndarray       # float64 [nelem, 3]
    ijk: np.ndarray         # int64 [nelem, 3]
    morton: bool
    bfaces: np.ndarray = None   # int32 [nbound, 3]: element, face, kind (None: periodic)

    @property
    def nelem(self) -> int:
        return int(self.nbr.shape[0])

    def lex_id(self) -> np.ndarray:
        """Order-independent global element id (lexicographic, x fastest)."""
        nx, ny, _ = self.shape
        return self.ijk[:, 0] + nx * (self.ijk[:, 1] + ny * self.ijk[:, 2])

    def node_coords(self, xi: np.ndarray) -> np.ndarray:
        """Physical coordinates [nelem, 3, M^3] of the LGL nodes xi (on [-1,1])."""
        M = len(xi)
        ii, jj, kk = np.meshgrid(np.arange(M), np.arange(M), np.arange(M), indexing="ij")
        # node = i + M*j + M*M*k  -> flatten in Fortran order of (i,j,k)
        ref = np.stack([xi[ii].ravel(order="F"), xi[jj].ravel(order="F"), xi[kk].ravel(order="F")])
        return self.lower[:, :, None] + 0.5 * (ref[None, :, :] + 1.0) * self.delta[:, :, None]


def uniform_periodic(nx: int, ny: int = None, nz: int = None, extent=(1.0, 1.0, 1.0),
                     origin=(0.0, 0.0, 0.0), order: str = "auto") -> Mesh:
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ii, jj, kk = ii.ravel(order="F"), jj.ravel(order="F"), kk.ravel(order="F")
    pow2 = all(n & (n - 1) == 0 for n in (nx, ny, nz)) and nx == ny == nz
    use_morton = (order == "morton") or (order == "auto" and pow2)
    if use_morton:
        perm = np.argsort(morton_key(ii, jj, kk), kind="stable")
    else:
        perm = np.arange(nx * ny * nz)
    ii, jj, kk = ii[perm], jj[perm], kk[perm]
    nelem = nx * ny * nz
    lex = ii + nx * (jj + ny * kk)
    pos = np.empty(nelem, dtype=np.int64)
    pos[lex] = np.arange(nelem)

    def at(i, j, k):
        return pos[(i % nx) + nx * ((j % ny) + ny * (k % nz))]

    nbr = np.stack([at(ii - 1, jj, kk), at(ii + 1, jj, kk),
                    at(ii, jj - 1, kk), at(ii, jj + 1, kk),
                    at(ii, jj, kk - 1), at(ii, jj, kk + 1)], axis=1).astype(np.int32)
    d = np.array([extent[0] / nx, extent[1] / ny, extent[2] / nz])
    delta = np.tile(d, (nelem, 1))
    ijk = np.stack([ii, jj, kk], axis=1).astype(np.int64)
    lower = np.asarray(origin)[None, :] + ijk * d[None, :]
    return Mesh((nx, ny, nz), tuple(extent), np.ascontiguousarray(nbr), np.ascontiguousarray(delta),
                np.ascontiguousarray(lower), ijk, use_morton)


def with_boundaries(mesh: Mesh, kinds) -> Mesh:
    """The same box with boundary faces in place of the periodic wrap.

    kinds[f] for the six faces of the box (0 = x-, 1 = x+, ..., 5 = z+):
    None keeps that side periodic (then its opposite side must be None as
    well), BC_TRANSMISSIVE or BC_SLIP_WALL makes its element faces boundary
    faces of that kind."""
    kinds = list(kinds)
    if len(kinds) != 6:
        raise ValueError("one kind per box face")
    for d in range(3):
        if (kinds[2 * d] is None) != (kinds[2 * d + 1] is None):
            raise ValueError("a periodic direction needs both of its sides periodic")
    n = np.asarray(mesh.shape)
    nbr = mesh.nbr.copy()
    rows = []
    for f, kind in enumerate(kinds):
        if kind is None:
            continue
        d, plus = f // 2, f % 2
        on = np.nonzero(mesh.ijk[:, d] == (n[d] - 1 if plus else 0))[0]
        for e in on:
            rows.append((int(e), f, int(kind)))
    bfaces = np.array(rows, dtype=np.int32).reshape(-1, 3)
    # boundary faces in element order, then face order
    bfaces = bfaces[np.lexsort((bfaces[:, 1], bfaces[:, 0]))] if len(bfaces) else bfaces
    for b, (e, f, _) in enumerate(bfaces):
        nbr[e, f] = -(b + 1)
    return Mesh(mesh.shape, mesh.extent, np.ascontiguousarray(nbr), mesh.delta, mesh.lower, mesh.ijk, mesh.morton,
                np.ascontiguousarray(bfaces))
