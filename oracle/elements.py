"""Element matrices (oracle).

The reference assembles the bilinear forms ``lambda tr(eps u) tr(eps v) +
2 mu eps(u):eps(v)`` (``fea/composer.py:88-98``) and ``k grad u . grad v``
(``fea/composer.py:139-141``) with scikit-fem.  Here the same forms are
integrated element by element with Gauss-Legendre quadrature on this build's
own local node numbering (see ``oracle.mesh``); DOFs are node-major
``[u0x, u0y, u0z, u1x, ...]``.  For affine hex8 and tet4 elements any rule of
order >= 2 integrates the stiffness exactly.
"""
from __future__ import annotations

import numpy as np

from .mesh import HEX_CORNERS


def lame(E: float | np.ndarray, nu: float):
    """``lambda = nu E / ((1+nu)(1-2nu))``, ``mu = E / (2(1+nu))``
    (``fea/composer.py:82-83``)."""
    lam = (nu * E) / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    return lam, mu


def elasticity_D(nu: float, E: float = 1.0) -> np.ndarray:
    """Voigt matrix for strains ``[exx, eyy, ezz, gyz, gxz, gxy]``."""
    lam, mu = lame(E, nu)
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[[0, 1, 2], [0, 1, 2]] += 2.0 * mu
    D[[3, 4, 5], [3, 4, 5]] = mu
    return D


def _gauss(order: int):
    x, w = np.polynomial.legendre.leggauss(order)
    return x, w


def hex8_shape_grads(xi, eta, zeta, h=(1.0, 1.0, 1.0)) -> np.ndarray:
    """Physical gradients (8, 3) of the trilinear shape functions of an
    axis-aligned box element of size ``h`` at reference point (xi, eta, zeta)."""
    s = 2.0 * HEX_CORNERS - 1.0           # corner signs in {-1, +1}
    g = np.empty((8, 3))
    g[:, 0] = s[:, 0] * (1 + s[:, 1] * eta) * (1 + s[:, 2] * zeta) / 8.0
    g[:, 1] = s[:, 1] * (1 + s[:, 0] * xi) * (1 + s[:, 2] * zeta) / 8.0
    g[:, 2] = s[:, 2] * (1 + s[:, 0] * xi) * (1 + s[:, 1] * eta) / 8.0
    return g * (2.0 / np.asarray(h, dtype=np.float64))[None, :]


def strain_B(grads: np.ndarray) -> np.ndarray:
    """Voigt strain-displacement matrix (6, 3*nen) from shape gradients."""
    nen = grads.shape[0]
    B = np.zeros((6, 3 * nen))
    for a in range(nen):
        gx, gy, gz = grads[a]
        c = 3 * a
        B[0, c + 0] = gx
        B[1, c + 1] = gy
        B[2, c + 2] = gz
        B[3, c + 1] = gz
        B[3, c + 2] = gy
        B[4, c + 0] = gz
        B[4, c + 2] = gx
        B[5, c + 0] = gy
        B[5, c + 1] = gx
    return B


def hex8_elasticity_K0(nu: float, h=(1.0, 1.0, 1.0), order: int = 2) -> np.ndarray:
    """24x24 stiffness of a box element with E = 1."""
    x, w = _gauss(order)
    D = elasticity_D(nu)
    detJ = h[0] * h[1] * h[2] / 8.0
    K = np.zeros((24, 24))
    for i, a in enumerate(x):
        for j, b in enumerate(x):
            for k, c in enumerate(x):
                B = strain_B(hex8_shape_grads(a, b, c, h))
                K += w[i] * w[j] * w[k] * detJ * (B.T @ D @ B)
    return 0.5 * (K + K.T)


def hex8_conduction_K0(h=(1.0, 1.0, 1.0), order: int = 2) -> np.ndarray:
    """8x8 conduction matrix of a box element with k = 1."""
    x, w = _gauss(order)
    detJ = h[0] * h[1] * h[2] / 8.0
    K = np.zeros((8, 8))
    for i, a in enumerate(x):
        for j, b in enumerate(x):
            for k, c in enumerate(x):
                g = hex8_shape_grads(a, b, c, h)
                K += w[i] * w[j] * w[k] * detJ * (g @ g.T)
    return 0.5 * (K + K.T)


def hex8_load_weights(h=(1.0, 1.0, 1.0)) -> np.ndarray:
    """Consistent nodal weights of a unit volumetric source, ``int N_a``."""
    return np.full(8, h[0] * h[1] * h[2] / 8.0)


def tet4_grads(p: np.ndarray, t: np.ndarray):
    """Barycentric gradients (n_el, 4, 3) and volumes (n_el,) of tets."""
    x0 = p[:, t[0]].T
    J = np.stack([p[:, t[1]].T - x0, p[:, t[2]].T - x0, p[:, t[3]].T - x0],
                 axis=2)                      # (n, 3, 3), columns are edges
    vol = np.linalg.det(J) / 6.0
    Jinv = np.linalg.inv(J)                   # rows are grad(lambda_1..3)
    g = np.empty((t.shape[1], 4, 3))
    g[:, 1:, :] = Jinv
    g[:, 0, :] = -Jinv.sum(axis=1)
    return g, vol


def tet4_elasticity_K(p: np.ndarray, t: np.ndarray, nu: float) -> np.ndarray:
    """Per-element 12x12 stiffness with E = 1, shape (n_el, 12, 12)."""
    g, vol = tet4_grads(p, t)
    D = elasticity_D(nu)
    n = t.shape[1]
    B = np.zeros((n, 6, 12))
    for a in range(4):
        gx, gy, gz = g[:, a, 0], g[:, a, 1], g[:, a, 2]
        c = 3 * a
        B[:, 0, c + 0] = gx
        B[:, 1, c + 1] = gy
        B[:, 2, c + 2] = gz
        B[:, 3, c + 1] = gz
        B[:, 3, c + 2] = gy
        B[:, 4, c + 0] = gz
        B[:, 4, c + 2] = gx
        B[:, 5, c + 0] = gy
        B[:, 5, c + 1] = gx
    K = np.einsum("nij,jk,nkl->nil", B.transpose(0, 2, 1), D, B)
    return K * vol[:, None, None]


def tet4_conduction_K(p: np.ndarray, t: np.ndarray) -> np.ndarray:
    g, vol = tet4_grads(p, t)
    return np.einsum("nad,nbd->nab", g, g) * vol[:, None, None]
