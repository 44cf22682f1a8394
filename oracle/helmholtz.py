"""Nodal Helmholtz filter (oracle) -- TEST INFRASTRUCTURE ONLY.

Restates ``filters/helmholtz_filter_nodal.py`` (``HelmholtzFilterNodal``,
``element_to_node_density_averaging`` :31-59, ``node_to_element_density``
:27-28, ``solve_helmholtz`` :121-157, ``forward`` :183-201, ``gradient``
:203-233) with scipy in place of scikit-fem:

* the bilinear form ``u v + r^2 grad u . grad v`` assembled from the exact
  element mass and conduction matrices (the reference's default skfem
  quadrature integrates both exactly on affine hex8 / tet4 elements);
* the linear form ``int rho_field v`` with ``rho_field`` the interpolant of
  the nodal values, i.e. ``M rho_node``;
* ``skfem.condense`` + ``solve``: the free block solved with the fixed nodal
  values moved to the right-hand side.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import elements as el
from .mesh import HEX_CORNERS, Mesh


def hex8_mass_M0(h=(1.0, 1.0, 1.0), order: int = 2) -> np.ndarray:
    """8x8 unit-density mass matrix of a box element (Gauss, exact for
    order >= 2)."""
    x, w = el._gauss(order)
    detJ = h[0] * h[1] * h[2] / 8.0
    s = 2.0 * HEX_CORNERS - 1.0
    M = np.zeros((8, 8))
    for i, a in enumerate(x):
        for j, b in enumerate(x):
            for k, c in enumerate(x):
                N = (1 + s[:, 0] * a) * (1 + s[:, 1] * b) * (1 + s[:, 2] * c) / 8.0
                M += w[i] * w[j] * w[k] * detJ * np.outer(N, N)
    return M


def tet4_mass(p: np.ndarray, t: np.ndarray) -> np.ndarray:
    """(n_el, 4, 4) P1 mass matrices, V / 20 (1 + delta_ab)."""
    _, vol = el.tet4_grads(p, t)
    base = (np.ones((4, 4)) + np.eye(4)) / 20.0
    return vol[:, None, None] * base[None]


def _assemble(t: np.ndarray, Ke: np.ndarray, n: int) -> sp.csr_matrix:
    nen = t.shape[0]
    rows = np.repeat(t.T, nen, axis=1).ravel()
    cols = np.tile(t.T, (1, nen)).ravel()
    return sp.csr_matrix((Ke.reshape(Ke.shape[0], -1).ravel(), (rows, cols)), shape=(n, n))


def helmholtz_matrices(mesh: Mesh, radius: float):
    """(A = M + r^2 K, M) as CSR."""
    n = mesh.nnodes
    ne = mesh.nelements
    if mesh.kind == "hex8":
        M0 = hex8_mass_M0(mesh.h)
        K0 = el.hex8_conduction_K0(mesh.h)
        Me = np.broadcast_to(M0, (ne, 8, 8))
        Ke = np.broadcast_to(K0, (ne, 8, 8))
    else:
        Me = tet4_mass(mesh.p, mesh.t)
        Ke = el.tet4_conduction_K(mesh.p, mesh.t)
    M = _assemble(mesh.t, np.ascontiguousarray(Me), n)
    K = _assemble(mesh.t, np.ascontiguousarray(Ke), n)
    return (M + radius ** 2 * K).tocsr(), M


def element_to_node_density_averaging(mesh: Mesh, elements_volume, rho_elem, design_mask=None,
                                      weighted=True, fixed_value_for_design=1.0):
    """``helmholtz_filter_nodal.py:31-59``: element-major accumulation of
    ``w * value`` and ``w`` into the element's nodes, then the quotient."""
    t = mesh.t
    ne = t.shape[1]
    if design_mask is None:
        design_mask = np.ones(ne, dtype=bool)
    val = np.where(design_mask, rho_elem, fixed_value_for_design)
    w = np.asarray(elements_volume, dtype=np.float64) if weighted else np.ones(ne)
    idx = t.T.ravel()                      # element by element, as the loop
    rho_node = np.zeros(mesh.nnodes)
    wsum = np.zeros(mesh.nnodes)
    np.add.at(rho_node, idx, np.repeat(w * val, t.shape[0]))
    np.add.at(wsum, idx, np.repeat(w, t.shape[0]))
    wsum[wsum == 0.0] = 1.0
    return rho_node / wsum


def node_to_element_density(mesh: Mesh, rho_node):
    """``helmholtz_filter_nodal.py:27-28``."""
    return np.mean(rho_node[mesh.t], axis=0)


def solve_helmholtz(case: str, mesh: Mesh, rho_node, radius: float, design_mask=None, mats=None):
    """``helmholtz_filter_nodal.py:121-157``."""
    A, M = mats if mats is not None else helmholtz_matrices(mesh, radius)
    b = M @ rho_node
    if design_mask is None:
        fixed = np.array([], dtype=np.int64)
    else:
        fixed = np.unique(mesh.t[:, ~np.asarray(design_mask, dtype=bool)].ravel())
    if fixed.size == 0:
        return spla.spsolve(A.tocsc(), b)
    x = np.zeros(A.shape[0])
    x[fixed] = 1.0 if case == "forward" else 0.0
    free = np.setdiff1d(np.arange(A.shape[0]), fixed)
    A_ff = A[free][:, free]
    rhs = b[free] - A[free][:, fixed] @ x[fixed]
    x[free] = spla.spsolve(A_ff.tocsc(), rhs)
    return x


class HelmholtzFilterNodal:
    """``helmholtz_filter_nodal.py:162-233``."""

    def __init__(self, mesh: Mesh, elements_volume, radius: float, design_mask=None):
        self.mesh = mesh
        self.elements_volume = np.asarray(elements_volume, dtype=np.float64)
        self.design_mask = None if design_mask is None else np.asarray(design_mask, bool)
        self.radius = float(radius)
        self._mats = None
        self._mats_r = None

    def update_radius(self, radius, **_):
        self.radius = float(radius)

    def _matrices(self):
        if self._mats is None or self._mats_r != self.radius:
            self._mats = helmholtz_matrices(self.mesh, self.radius)
            self._mats_r = self.radius
        return self._mats

    def forward(self, rho_element):
        rho_node = element_to_node_density_averaging(
            self.mesh, self.elements_volume, rho_element, design_mask=self.design_mask,
            fixed_value_for_design=1.0)
        rn = solve_helmholtz("forward", self.mesh, rho_node, self.radius,
                             design_mask=self.design_mask, mats=self._matrices())
        return node_to_element_density(self.mesh, rn)

    def gradient(self, v_ele):
        v_node = element_to_node_density_averaging(
            self.mesh, self.elements_volume, v_ele, design_mask=self.design_mask,
            weighted=True, fixed_value_for_design=0.0)
        vn = solve_helmholtz("gradient", self.mesh, v_node, self.radius, design_mask=None,
                             mats=self._matrices())
        return np.minimum(node_to_element_density(self.mesh, vn), 0.0)
