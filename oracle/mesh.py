"""Mesh generators and geometric element data (oracle).

The reference builds meshes with scikit-fem (``mesh/toy_problem.py:10-37``)
which is absent here; these generators are this build's own numbering:

* nodes of an ``nx x ny x nz`` box grid are numbered x-fastest,
  ``n = i + (nx+1)*(j + (ny+1)*k)``, coordinates ``(i*hx, j*hy, k*hz)``;
* hex8 local nodes follow the corner order
  ``(0,0,0),(1,0,0),(1,1,0),(0,1,0),(0,0,1),(1,0,1),(1,1,1),(0,1,1)``;
* elements are numbered x-fastest, ``e = i + nx*(j + ny*k)``;
* a tet4 mesh splits every hex into the six tetrahedra around the
  0-6 diagonal (conforming on a structured grid), interior nodes optionally
  jittered by a seeded uniform perturbation.
"""
from __future__ import annotations

import numpy as np

HEX_CORNERS = np.array(
    [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
     [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.int64)

# Six tetrahedra sharing the 0-6 diagonal of a hex (local hex node ids).
HEX_TO_TETS = np.array(
    [[0, 1, 2, 6], [0, 2, 3, 6], [0, 3, 7, 6],
     [0, 7, 4, 6], [0, 4, 5, 6], [0, 5, 1, 6]], dtype=np.int64)


class Mesh:
    """Plain mesh container: ``p`` (3, n_nodes) float64, ``t`` (nen, n_el) int64."""

    def __init__(self, p: np.ndarray, t: np.ndarray, kind: str,
                 grid: tuple | None = None, h: tuple | None = None):
        self.p = np.ascontiguousarray(p, dtype=np.float64)
        self.t = np.ascontiguousarray(t, dtype=np.int64)
        self.kind = kind            # "hex8" or "tet4"
        self.grid = grid            # (nx, ny, nz) for structured meshes
        self.h = h                  # (hx, hy, hz) for structured meshes

    @property
    def nnodes(self) -> int:
        return self.p.shape[1]

    @property
    def nelements(self) -> int:
        return self.t.shape[1]


def grid_nodes(nx: int, ny: int, nz: int, hx=1.0, hy=1.0, hz=1.0) -> np.ndarray:
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1),
                          np.arange(nz + 1), indexing="ij")
    # x-fastest numbering
    i = i.transpose(2, 1, 0).ravel()
    j = j.transpose(2, 1, 0).ravel()
    k = k.transpose(2, 1, 0).ravel()
    return np.vstack([i * float(hx), j * float(hy), k * float(hz)])


def grid_hex_connectivity(nx: int, ny: int, nz: int) -> np.ndarray:
    ei, ej, ek = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz),
                             indexing="ij")
    ei = ei.transpose(2, 1, 0).ravel()
    ej = ej.transpose(2, 1, 0).ravel()
    ek = ek.transpose(2, 1, 0).ravel()
    t = np.empty((8, ei.size), dtype=np.int64)
    for a, (di, dj, dk) in enumerate(HEX_CORNERS):
        t[a] = (ei + di) + (nx + 1) * ((ej + dj) + (ny + 1) * (ek + dk))
    return t


def structured_hex(nx: int, ny: int, nz: int, hx=1.0, hy=1.0, hz=1.0) -> Mesh:
    return Mesh(grid_nodes(nx, ny, nz, hx, hy, hz),
                grid_hex_connectivity(nx, ny, nz), "hex8",
                grid=(nx, ny, nz), h=(float(hx), float(hy), float(hz)))


def jittered_tet(nx: int, ny: int, nz: int, jitter: float = 0.2,
                 seed: int = 0, h: float = 1.0) -> Mesh:
    """Hex grid split into 6 tets per hex; interior nodes moved by
    U(-jitter*h, +jitter*h) per coordinate from ``default_rng(seed)``."""
    p = grid_nodes(nx, ny, nz, h, h, h)
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter * h, jitter * h, size=p.shape)
        n = np.arange(p.shape[1])
        i = n % (nx + 1)
        j = (n // (nx + 1)) % (ny + 1)
        k = n // ((nx + 1) * (ny + 1))
        interior = (i > 0) & (i < nx) & (j > 0) & (j < ny) & (k > 0) & (k < nz)
        p = p + d * interior[None, :]
    th = grid_hex_connectivity(nx, ny, nz)
    t = np.empty((4, th.shape[1] * 6), dtype=np.int64)
    for s, loc in enumerate(HEX_TO_TETS):
        t[:, s::6] = th[loc]
    vol = tet_signed_volumes(p, t)
    neg = vol < 0
    if np.any(neg):
        tmp = t[1, neg].copy()
        t[1, neg] = t[2, neg]
        t[2, neg] = tmp
    return Mesh(p, t, "tet4", grid=(nx, ny, nz), h=(h, h, h))


def tet_signed_volumes(p: np.ndarray, t: np.ndarray) -> np.ndarray:
    """Signed tet volume ``((p1-p0) x (p2-p0)) . (p3-p0) / 6``
    (``fea/composer.py:164-180``)."""
    v1 = p[:, t[1]] - p[:, t[0]]
    v2 = p[:, t[2]] - p[:, t[0]]
    v3 = p[:, t[3]] - p[:, t[0]]
    return np.einsum("ij,ij->j", np.cross(v1.T, v2.T).T, v3) / 6.0


def elements_volume(mesh: Mesh) -> np.ndarray:
    """Element volumes (``fea/composer.py:314-322``).

    tet4: the signed volume, raising on a negative one
    (``composer.py:164-180``).  hex8: sum of six tetrahedra
    (``composer.py:191-248``) -- with this build's corner numbering the six
    tets around the 0-6 diagonal tile the hex exactly.
    """
    if mesh.kind == "tet4":
        vol = tet_signed_volumes(mesh.p, mesh.t)
        if np.any(vol < -1e-12):
            raise ValueError("tetrahedron with negative volume")
        return vol
    if mesh.kind == "hex8":
        vol = np.zeros(mesh.nelements)
        for loc in HEX_TO_TETS:
            vol += np.abs(tet_signed_volumes(mesh.p, mesh.t[loc]))
        return vol
    raise NotImplementedError(mesh.kind)


def element_centers(mesh: Mesh) -> np.ndarray:
    """Centres = mean of element node coordinates, shape (3, n_el)
    (``filters/spacial.py:13-16``)."""
    coords = mesh.p[:, mesh.t]
    return np.mean(coords, axis=1)


def elements_touching_nodes(mesh: Mesh, nodes: np.ndarray) -> np.ndarray:
    """Sorted unique elements containing any of ``nodes``
    (``mesh/utils.py:180-263``)."""
    flag = np.zeros(mesh.nnodes, dtype=bool)
    flag[np.asarray(nodes, dtype=np.int64)] = True
    hit = np.any(flag[mesh.t], axis=0)
    return np.nonzero(hit)[0]
