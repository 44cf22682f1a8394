"""Structured hex8 grids and tet4 meshes (host numpy).

Node numbering of an ``nx x ny x nz`` grid is x-fastest,
``n = i + (nx+1)*(j + (ny+1)*k)``; hex8 local corners are
``(0,0,0),(1,0,0),(1,1,0),(0,1,0),(0,0,1),(1,0,1),(1,1,1),(0,1,1)``;
elements are x-fastest.  The matrix-free hex8 kernels rely on exactly this
numbering (connectivity is implicit).
"""
from __future__ import annotations

import copy

import numpy as np

CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                    [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.int64)
HEX_TETS = np.array([[0, 1, 2, 6], [0, 2, 3, 6], [0, 3, 7, 6],
                     [0, 7, 4, 6], [0, 4, 5, 6], [0, 5, 1, 6]], dtype=np.int64)
# element faces as cyclic corner lists (quads of a hex, triangles of a tet)
HEX_FACES = np.array([[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4],
                      [3, 7, 6, 2], [0, 4, 7, 3], [1, 2, 6, 5]], dtype=np.int64)
TET_FACES = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], dtype=np.int64)


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

    # --- facets (skfem's Mesh.facets / facets_satisfying / boundary_facets) --
    def _build_facets(self):
        faces = HEX_FACES if self.kind == "hex8" else TET_FACES
        allf = np.concatenate([self.t[f].T for f in faces])           # (nf_all, nen_f)
        key = np.sort(allf, axis=1)
        _, first, inv, cnt = np.unique(key, axis=0, return_index=True, return_inverse=True,
                                       return_counts=True)
        self._facets = np.ascontiguousarray(allf[first].T)           # cyclic node order
        self._facet_count = cnt

    @property
    def facets(self) -> np.ndarray:
        """(nodes per facet, n_facets): every facet once, numbered by its
        sorted node tuple; nodes in cyclic order."""
        if getattr(self, "_facets", None) is None:
            self._build_facets()
        return self._facets

    def boundary_facets(self) -> np.ndarray:
        self.facets
        return np.nonzero(self._facet_count == 1)[0]

    def facets_satisfying(self, test, boundaries_only=False) -> np.ndarray:
        """Facets whose midpoint (mean of its nodes) passes ``test``."""
        midp = np.sum(self.p[:, self.facets], axis=1) / self.facets.shape[0]
        f = np.nonzero(test(midp))[0]
        if boundaries_only:
            f = np.intersect1d(f, self.boundary_facets())
        return f

    def elements_satisfying(self, test) -> np.ndarray:
        """Elements whose midpoint (mean of its nodes) passes ``test``."""
        midp = np.sum(self.p[:, self.t], axis=1) / self.t.shape[0]
        return np.nonzero(test(midp))[0]

    def facet_areas(self, ids=None) -> np.ndarray:
        f = self.facets if ids is None else self.facets[:, np.asarray(ids, dtype=np.int64)]
        P = self.p[:, f]                                     # (3, nen_f, nf)
        if f.shape[0] == 3:
            c = np.cross((P[:, 1] - P[:, 0]).T, (P[:, 2] - P[:, 0]).T)
        else:                                                # planar quad: diagonals
            c = np.cross((P[:, 2] - P[:, 0]).T, (P[:, 3] - P[:, 1]).T)
        return 0.5 * np.linalg.norm(c, axis=1)

    def scaled(self, factor: float):
        """A copy with every node coordinate (and box size) multiplied by
        ``factor``; connectivity and facet numbering are unchanged."""
        out = copy.copy(self)
        out.p = self.p * factor
        if getattr(self, "h", None) is not None:
            out.h = tuple(float(h) * factor for h in self.h)
        return out

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
    """Structured box grid of identical hx*hy*hz hexahedra.  With
    ``extent=(x_len, y_len, z_len)`` the node coordinates are
    ``numpy.linspace(0, len, n + 1)`` per axis (the reference's
    ``create_box_hex``, ``mesh/toy_problem.py:10-37``) and h = len / n."""
    kind = "hex8"

    def __init__(self, nx, ny, nz, hx=1.0, hy=1.0, hz=1.0, extent=None):
        self.grid = (int(nx), int(ny), int(nz))
        i, j, k = self.node_ijk(nx, ny, nz)
        if extent is None:
            self.h = (float(hx), float(hy), float(hz))
            p = np.vstack([i * self.h[0], j * self.h[1], k * self.h[2]])
        else:
            L = [float(v) for v in extent]
            self.h = (L[0] / nx, L[1] / ny, L[2] / nz)
            xs = [np.linspace(0.0, L[a], n + 1) for a, n in enumerate(self.grid)]
            p = np.vstack([xs[0][i], xs[1][j], xs[2][k]])
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
    """Tetrahedral mesh.  ``cell_grid=(nx, ny, nz)`` records that the nodes
    are the (nx+1)(ny+1)(nz+1) corners of a grid of cells numbered x-fastest
    and that element 6 c + s is one of the six tets of cell c (the layout of
    ``jittered_tet``); the geometric multigrid uses that grid."""
    kind = "tet4"

    def __init__(self, p, t, cell_grid=None):
        super().__init__(p, t)
        self.cell_grid = None if cell_grid is None else tuple(int(v) for v in cell_grid)

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
    return MeshTet(p, t, cell_grid=(nx, ny, nz))
