"""Problem definitions for the benchmark workloads (oracle).

Design-set rule restated from ``mesh/task_common.py:104-270,398-401``:
design = all elements minus those touching a load node (Neumann), minus
those touching a Dirichlet node (``design_dirichlet=False``);
``dirichlet_neumann_elements`` (pinned to 1) is the concatenation of both
sets and ``fixed_elements`` = all minus the Neumann-excluded design set.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import mesh as m
from .fem import ElementOperator
from . import elements as el


@dataclass
class Task:
    mesh: m.Mesh
    physics: str                      # "elastic" or "heat"
    E0: float
    nu: float
    dirichlet_nodes: np.ndarray
    dirichlet_dofs: np.ndarray
    forces: list                      # list of (ndof,) arrays, one per load case
    neumann_nodes: np.ndarray | None
    dirichlet_values: list = field(default_factory=lambda: [0.0])
    quad_order: int = 2
    design_elements: np.ndarray = None
    design_elements_with_dirichlet: np.ndarray = None
    dirichlet_elements: np.ndarray = None
    neumann_elements: np.ndarray = None
    fixed_elements: np.ndarray = None
    dirichlet_neumann_elements: np.ndarray = None
    elements_volume: np.ndarray = None
    centers: np.ndarray = None

    def __post_init__(self):
        mesh = self.mesh
        all_el = np.arange(mesh.nelements)
        self.dirichlet_elements = m.elements_touching_nodes(mesh, self.dirichlet_nodes)
        if self.neumann_nodes is not None and len(self.neumann_nodes) > 0:
            self.neumann_elements = m.elements_touching_nodes(mesh, self.neumann_nodes)
        else:
            self.neumann_elements = np.array([], dtype=np.int64)
        design = all_el[~np.isin(all_el, self.neumann_elements)]
        self.fixed_elements = all_el[~np.isin(all_el, design)]
        # the task's design set (mesh/task_common.py:211-227) keeps the
        # elements touching Dirichlet nodes; DensityMethod drops them unless
        # design_dirichlet=True (common_density.py:546-547, task_common.py:398-401)
        self.design_elements_with_dirichlet = np.ascontiguousarray(design)
        design = design[~np.isin(design, self.dirichlet_elements)]
        self.design_elements = np.ascontiguousarray(design)
        valid = [a for a in (self.dirichlet_elements, self.neumann_elements) if len(a) > 0]
        self.dirichlet_neumann_elements = np.concatenate(valid) if valid \
            else np.array([], dtype=np.int64)
        self.elements_volume = m.elements_volume(mesh)
        self.centers = m.element_centers(mesh)

    @property
    def n_elements(self):
        return self.mesh.nelements

    @property
    def ncomp(self):
        return 3 if self.physics == "elastic" else 1

    @property
    def ndof(self):
        return self.mesh.nnodes * self.ncomp

    @property
    def n_tasks(self):
        return len(self.forces) if self.physics == "elastic" else len(self.dirichlet_values)

    @property
    def design_mask(self):
        return np.isin(np.arange(self.n_elements), self.design_elements)

    def element_operator(self):
        return ElementOperator(self.mesh, self.physics, self.nu, self.quad_order)


def _grid_index(nx, ny, nz):
    n = np.arange((nx + 1) * (ny + 1) * (nz + 1))
    return n % (nx + 1), (n // (nx + 1)) % (ny + 1), n // ((nx + 1) * (ny + 1))


def cantilever(nx, ny, nz, kind="hex8", jitter=0.0, seed=0, E0=1.0, nu=0.3):
    """x=0 face clamped; unit downward (-z) load split equally over the
    nodes of the bottom edge (x = nx, z = 0)."""
    if kind == "hex8":
        mesh = m.structured_hex(nx, ny, nz)
    else:
        mesh = m.jittered_tet(nx, ny, nz, jitter, seed)
    i, j, k = _grid_index(nx, ny, nz)
    clamp = np.nonzero(i == 0)[0]
    dofs = np.sort(np.concatenate([3 * clamp + c for c in range(3)]))
    load = np.nonzero((i == nx) & (k == 0))[0]
    f = np.zeros(3 * mesh.nnodes)
    f[3 * load + 2] = -1.0 / load.size
    return Task(mesh, "elastic", E0, nu, clamp, dofs, [f], load)


def multiload(nx, ny, nz, n_loads=8, seed=0, E0=1.0, nu=0.3):
    """x=0 face clamped; ``n_loads`` unit point loads at distinct random nodes
    of the x = nx face, directions uniform on the sphere (normalised
    Gaussian), all drawn from ``default_rng(seed)`` -- nodes first, then a
    (n_loads, 3) normal block."""
    mesh = m.structured_hex(nx, ny, nz)
    i, j, k = _grid_index(nx, ny, nz)
    clamp = np.nonzero(i == 0)[0]
    dofs = np.sort(np.concatenate([3 * clamp + c for c in range(3)]))
    face = np.nonzero(i == nx)[0]
    rng = np.random.default_rng(seed)
    nodes = rng.choice(face, size=n_loads, replace=False)
    g = rng.standard_normal((n_loads, 3))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    forces = []
    for a in range(n_loads):
        f = np.zeros(3 * mesh.nnodes)
        f[3 * nodes[a]:3 * nodes[a] + 3] = g[a]
        forces.append(f)
    return Task(mesh, "elastic", E0, nu, clamp, dofs, forces, np.sort(nodes))


def heat_patch(n, patch=16, source=1e-2, k0=1.0, quad_order=2, kind="hex8", jitter=0.0,
               seed=0):
    """Cube of n^3 unit hexes (or the six-tet split of each, interior nodes
    jittered), uniform volumetric source, T = 0 on a ``patch x patch`` node
    patch centred on the z = 0 face."""
    mesh = m.structured_hex(n, n, n) if kind == "hex8" else m.jittered_tet(n, n, n, jitter, seed)
    i, j, k = _grid_index(n, n, n)
    lo = (n + 1 - patch) // 2
    sel = (k == 0) & (i >= lo) & (i < lo + patch) & (j >= lo) & (j < lo + patch)
    nodes = np.nonzero(sel)[0]
    f = np.zeros(mesh.nnodes)
    if kind == "hex8":
        w = el.hex8_load_weights(mesh.h) * source
        for a in range(8):
            np.add.at(f, mesh.t[a], w[a])
    else:
        # int s N_a over a tet is s V / 4 for each of its nodes
        vol = m.tet_signed_volumes(mesh.p, mesh.t)
        for a in range(4):
            np.add.at(f, mesh.t[a], source * vol / 4.0)
    return Task(mesh, "heat", k0, 0.0, nodes, nodes.copy(), [f], None,
                dirichlet_values=[0.0], quad_order=quad_order)
