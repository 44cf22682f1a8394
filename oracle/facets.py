"""Surface (facet) loads -- TEST INFRASTRUCTURE ONLY.

Restates ``assemble_surface_forces`` (``mesh/task_elastic.py:15-81``):
per load region ``A = int_facets 1 dS``, ``pressure = value / A`` and
``F_i = int_facets pressure N_i dS`` on component ``comp``, integrated with
Gauss quadrature on each facet's own parametrisation (bilinear quads on
[-1, 1]^2 with a 3x3 rule, linear triangles with a degree-2 rule), the way
skfem's ``FacetBasis`` integrates, rather than by any closed form.

Also a brute-force facet enumeration of a structured box grid (every
axis-aligned unit face listed by its lower corner) for checking the
product's unique-facet table.
"""
from __future__ import annotations

import numpy as np

_G3 = (np.array([-np.sqrt(0.6), 0.0, np.sqrt(0.6)]), np.array([5.0, 8.0, 5.0]) / 9.0)
# degree-2 rule on the reference triangle (area 1/2): edge midpoints
_T3 = (np.array([[0.5, 0.0], [0.5, 0.5], [0.0, 0.5]]), np.full(3, 1.0 / 6.0))


def _quad_integrals(P):
    """P (3, 4) cyclic corners -> (area, int N_i dS for i = 0..3)."""
    xs, ws = _G3
    sgn = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], dtype=float)
    area, Ni = 0.0, np.zeros(4)
    for a, wa in zip(xs, ws):
        for b, wb in zip(xs, ws):
            N = 0.25 * (1 + sgn[:, 0] * a) * (1 + sgn[:, 1] * b)
            dNa = 0.25 * sgn[:, 0] * (1 + sgn[:, 1] * b)
            dNb = 0.25 * sgn[:, 1] * (1 + sgn[:, 0] * a)
            J = np.linalg.norm(np.cross(P @ dNa, P @ dNb))
            area += wa * wb * J
            Ni += wa * wb * J * N
    return area, Ni


def _tri_integrals(P):
    pts, ws = _T3
    J = np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]))
    area, Ni = 0.0, np.zeros(3)
    for (r, s), w in zip(pts, ws):
        N = np.array([1 - r - s, r, s])
        area += w * J
        Ni += w * J * N
    return area, Ni


def surface_load(p, facets, comp, value, ncomp=3):
    """One load region: facets (nodes per facet, nf) in cyclic order."""
    integ = _quad_integrals if facets.shape[0] == 4 else _tri_integrals
    res = [integ(p[:, facets[:, f]]) for f in range(facets.shape[1])]
    A = sum(r[0] for r in res)
    pressure = float(value) / A
    F = np.zeros(p.shape[1] * ncomp)
    for f, (_, Ni) in enumerate(res):
        for a, node in enumerate(facets[:, f]):
            F[ncomp * node + comp] += pressure * Ni[a]
    return F


def box_grid_faces(nx, ny, nz):
    """Every unit face of an nx x ny x nz node-indexed grid as a sorted
    4-tuple of node ids (x-fastest numbering), enumerated by axis."""
    nid = lambda i, j, k: i + (nx + 1) * (j + (ny + 1) * k)  # noqa: E731
    out = set()
    for k in range(nz + 1):
        for j in range(ny):
            for i in range(nx):
                if True:
                    out.add(tuple(sorted((nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k),
                                          nid(i, j + 1, k)))))
    for k in range(nz):
        for j in range(ny + 1):
            for i in range(nx):
                out.add(tuple(sorted((nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j, k + 1),
                                      nid(i, j, k + 1)))))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx + 1):
                out.add(tuple(sorted((nid(i, j, k), nid(i, j + 1, k), nid(i, j + 1, k + 1),
                                      nid(i, j, k + 1)))))
    return out
