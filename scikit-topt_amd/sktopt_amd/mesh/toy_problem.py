"""Benchmark problem builders (host setup).

* ``cantilever``: x=0 face clamped, unit downward (-z) load split equally
  over the nodes of the bottom edge (x = nx, z = 0); hex8 grid or jittered
  tet4 mesh.
* ``multiload``: x=0 face clamped, ``n_loads`` unit point loads at distinct
  random nodes of the x = nx face with directions uniform on the sphere;
  ``default_rng(seed)`` draws the nodes first, then an (n_loads, 3) normal
  block that is normalised row-wise.
* ``heat_patch``: n^3 cube, uniform volumetric source, T = 0 on a
  ``patch x patch`` node patch centred on the z = 0 face.
"""
from __future__ import annotations

import numpy as np

from .meshes import MeshHex, jittered_tet, structured_hex
from .task import LinearElasticity, LinearHeatConduction


def cantilever(nx, ny, nz, kind="hex8", jitter=0.0, seed=0, E=1.0, nu=0.3):
    mesh = structured_hex(nx, ny, nz) if kind == "hex8" else jittered_tet(nx, ny, nz, jitter, seed)
    i, j, k = MeshHex.node_ijk(nx, ny, nz)
    clamp = np.nonzero(i == 0)[0]
    load = np.nonzero((i == nx) & (k == 0))[0]
    f = np.zeros(3 * mesh.nvertices)
    f[3 * load + 2] = -1.0 / load.size
    return LinearElasticity.from_nodes(mesh, clamp, "all", [f], load, None, E, nu)


def multiload(nx, ny, nz, n_loads=8, seed=0, E=1.0, nu=0.3):
    mesh = structured_hex(nx, ny, nz)
    i, j, k = MeshHex.node_ijk(nx, ny, nz)
    clamp = np.nonzero(i == 0)[0]
    face = np.nonzero(i == nx)[0]
    rng = np.random.default_rng(seed)
    nodes = rng.choice(face, size=n_loads, replace=False)
    g = rng.standard_normal((n_loads, 3))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    forces = []
    for a in range(n_loads):
        f = np.zeros(3 * mesh.nvertices)
        f[3 * nodes[a]:3 * nodes[a] + 3] = g[a]
        forces.append(f)
    return LinearElasticity.from_nodes(mesh, clamp, "all", forces, np.sort(nodes), None, E, nu)


def heat_patch(n, patch=16, source=1e-2, k=1.0):
    mesh = structured_hex(n, n, n)
    i, j, kk = MeshHex.node_ijk(n, n, n)
    lo = (n + 1 - patch) // 2
    sel = (kk == 0) & (i >= lo) & (i < lo + patch) & (j >= lo) & (j < lo + patch)
    nodes = np.nonzero(sel)[0]
    emit = np.zeros(mesh.nvertices)
    w = source * mesh.h[0] * mesh.h[1] * mesh.h[2] / 8.0
    for a in range(8):
        np.add.at(emit, mesh.t[a], w)
    return LinearHeatConduction(mesh, nodes, [0.0], emit, None, k)


# --- the reference's own toy problems (mesh/toy_problem.py:10-145) ----------

def get_points_in_range(x_rng, y_rng, z_rng):
    """``mesh/utils.py:8-15``: closed-box test on (3, n) points."""
    def in_range(x):
        return ((x[0] >= x_rng[0]) & (x[0] <= x_rng[1]) &
                (x[1] >= y_rng[0]) & (x[1] <= y_rng[1]) &
                (x[2] >= z_rng[0]) & (x[2] <= z_rng[1]))
    return in_range


def create_box_hex(x_len, y_len, z_len, mesh_size) -> MeshHex:
    """``mesh/toy_problem.py:10-37``: ceil(len / mesh_size) elements per
    axis, nodes at ``linspace(0, len, n + 1)``."""
    nx = int(np.ceil(x_len / mesh_size))
    ny = int(np.ceil(y_len / mesh_size))
    nz = int(np.ceil(z_len / mesh_size))
    return MeshHex(nx, ny, nz, extent=(x_len, y_len, z_len))


def toy_base(mesh_size: float, intorder: int = 2):
    """``mesh/toy_problem.py:51-91``: 8 x 6 x 4 box, x = 0 facets clamped,
    F = -100 in z spread over the facets whose midpoints lie in
    [6.8, 8.1] x [2.4, 3.6] x [2.8, 4] (interior facets included, as
    ``facets_satisfying`` selects them), E = 210e3, nu = 0.3."""
    x_len, y_len, z_len = 8.0, 6.0, 4.0
    eps = 1.2
    mesh = create_box_hex(x_len, y_len, z_len, mesh_size)
    dirichlet_facets = mesh.facets_satisfying(
        get_points_in_range((0.0, 0.03), (0.0, y_len), (0.0, z_len)))
    force_facets = mesh.facets_satisfying(get_points_in_range(
        (x_len - eps, x_len + 0.1), (y_len * 2 / 5, y_len * 3 / 5), (z_len - eps, z_len)))
    design_elements = mesh.elements_satisfying(
        get_points_in_range((0.0, x_len), (0.0, y_len), (0.0, z_len)))
    return LinearElasticity.from_facets(mesh, dirichlet_facets, "all", force_facets, "u^3",
                                        -100.0, design_elements, 210e3, 0.30)


def toy_test():
    return toy_base(1.0)


def toy1():
    return toy_base(0.3)


def toy1_fine():
    return toy_base(0.2)


def toy2():
    """``mesh/toy_problem.py:106-144``: 8 x 8 x 1 plate, x = 0 boundary
    clamped, two load cases of -1 / +1 in y on the x = 8 boundary strips
    y >= 8 - 0.3 and y <= 0.3 (boundary facets only, as ``with_boundaries``
    selects them)."""
    x_len, y_len, z_len = 8.0, 8.0, 1.0
    mesh_size = 0.3
    mesh = create_box_hex(x_len, y_len, z_len, mesh_size)
    eps = mesh_size
    sel = lambda rng: mesh.facets_satisfying(get_points_in_range(*rng),  # noqa: E731
                                             boundaries_only=True)
    dirichlet = sel(((0.0, 0.05), (0.0, y_len), (0.0, z_len)))
    n0 = sel(((x_len, x_len), (y_len - eps, y_len), (0, z_len)))
    n1 = sel(((x_len, x_len), (0, eps), (0, z_len)))
    return LinearElasticity.from_facets(mesh, dirichlet, "all", [n0, n1], ["u^2", "u^2"],
                                        [-1.0, 1.0], np.arange(mesh.nelements), 210e3, 0.30)
