"""Structured hex8 grids and tet4 meshes (host numpy).

Node numbering of an ``nx x ny x nz`` grid is x-fastest,
``n = i + (nx+1)*(j + (ny+1)*k)``; hex8 local corners are
``(0,0,0),(1,0,0),(1,1,0),(0,1,0),(0,0,1),(1,0,1),(1,1,1),(0,1,1)``;
elements are x-fastest.  The matrix-free hex8 kernels rely on exactly this
numbering (connectivity is implicit).
"""
from __future__ import annotations

import numpy as np

CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                    [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.int64)
HEX_TETS = np.array([[0, 1, 2, 6], [0, 2, 3, 6], [0, 3, 7, 6],
                     [0, 7, 4, 6], [0, 4, 5, 6], [0, 5, 1, 6]], dtype=np.int64)


class _Mesh:
    kind = ""

    def __init__(self, p, t):
        self.p = np.ascontiguousarray(p, dtype=np.float64)
        self.t = np.ascontiguousarray(t, dtype=np.int64)

    @property
    def nvertices(self):
        return self.p.shape[1]

    @property
    def nelements(self):
        return self.t.shape[1]

    def element_centers(self) -> np.ndarray:
        """Mean of the element node coordinates, (3, n_el)
        (``filters/spacial.py:13-16``)."""
        return np.mean(self.p[:, self.t], axis=1)

    def elements_touching(self, nodes) -> np.ndarray:
        flag = np.zeros(self.nvertices, dtype=bool)
        flag[np.asarray(nodes, dtype=np.int64)] = True
        return np.nonzero(np.any(flag[self.t], axis=0))[0]


def _tet_volumes(p, t):
    v1 = p[:, t[1]] - p[:, t[0]]
    v2 = p[:, t[2]] - p[:, t[0]]
    v3 = p[:, t[3]] - p[:, t[0]]
    return np.einsum("ij,ij->j", np.cross(v1.T, v2.T).T, v3) / 6.0


class MeshHex(_Mesh):
    """Structured box grid of identical hx*hy*hz hexahedra."""
    kind = "hex8"

    def __init__(self, nx, ny, nz, hx=1.0, hy=1.0, hz=1.0):
        self.grid = (int(nx), int(ny), int(nz))
        self.h = (float(hx), float(hy), float(hz))
        i, j, k = self.node_ijk(nx, ny, nz)
        p = np.vstack([i * self.h[0], j * self.h[1], k * self.h[2]])
        e = np.arange(nx * ny * nz)
        ei, ej, ek = e % nx, (e // nx) % ny, e // (nx * ny)
        t = np.empty((8, e.size), dtype=np.int64)
        for a, (dx, dy, dz) in enumerate(CORNERS):
            t[a] = (ei + dx) + (nx + 1) * ((ej + dy) + (ny + 1) * (ek + dz))
        super().__init__(p, t)

    @staticmethod
    def node_ijk(nx, ny, nz):
        n = np.arange((nx + 1) * (ny + 1) * (nz + 1))
        return n % (nx + 1), (n // (nx + 1)) % (ny + 1), n // ((nx + 1) * (ny + 1))

    def elements_volume(self) -> np.ndarray:
        """Six-tetrahedron volume sum (``fea/composer.py:191-248``); exact
        for these box elements."""
        vol = np.zeros(self.nelements)
        for loc in HEX_TETS:
            vol += np.abs(_tet_volumes(self.p, self.t[loc]))
        return vol


class MeshTet(_Mesh):
    kind = "tet4"

    def elements_volume(self) -> np.ndarray:
        """Signed volumes; a negative one raises (``fea/composer.py:164-180``)."""
        vol = _tet_volumes(self.p, self.t)
        if np.any(vol < -1e-12):
            raise ValueError("tetrahedron with negative volume")
        return vol


def structured_hex(nx, ny, nz, hx=1.0, hy=1.0, hz=1.0) -> MeshHex:
    return MeshHex(nx, ny, nz, hx, hy, hz)


def jittered_tet(nx, ny, nz, jitter=0.2, seed=0, h=1.0) -> MeshTet:
    """Hex grid split into six tets per hex around the 0-6 diagonal, with
    interior nodes moved by U(-jitter*h, jitter*h) per coordinate drawn from
    ``numpy.random.default_rng(seed)``; tets are oriented positively."""
    i, j, k = MeshHex.node_ijk(nx, ny, nz)
    p = np.vstack([i * h, j * h, k * h]).astype(np.float64)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter * h, jitter * h, size=p.shape)
        inner = (i > 0) & (i < nx) & (j > 0) & (j < ny) & (k > 0) & (k < nz)
        p = p + d * inner[None, :]
    hexm = MeshHex(nx, ny, nz)
    t = np.empty((4, hexm.nelements * 6), dtype=np.int64)
    for s, loc in enumerate(HEX_TETS):
        t[:, s::6] = hexm.t[loc]
    vol = _tet_volumes(p, t)
    neg = vol < 0
    if np.any(neg):
        a = t[1, neg].copy()
        t[1, neg] = t[2, neg]
        t[2, neg] = a
    return MeshTet(p, t)
