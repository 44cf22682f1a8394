"""FEMDomain-like task containers (``mesh/task_common.py:24-505``).

Attributes keep the reference's names (``design_elements``,
``dirichlet_elements``, ``neumann_elements``, ``fixed_elements``,
``dirichlet_neumann_elements``, ``elements_volume``, ``design_mask``,
``n_tasks``, ``neumann_linear``, ``dirichlet_dofs``...) so optimizer code reads
them the same way.
"""
from __future__ import annotations

import numpy as np


def _setdiff(a, b):
    """``mesh/task_common.py:18-21``: order-preserving ``a`` minus ``b``."""
    a = np.asarray(a)
    return np.ascontiguousarray(a[~np.isin(a, b)])


class FEMDomain:
    def __init__(self, mesh, ncomp, dirichlet_nodes, dirichlet_dofs, dirichlet_values,
                 neumann_nodes, design_elements):
        self.mesh = mesh
        self.ncomp = ncomp
        self.dirichlet_nodes = np.asarray(dirichlet_nodes, dtype=np.int64)
        self.dirichlet_dofs = np.asarray(dirichlet_dofs, dtype=np.int64)
        self.dirichlet_values = dirichlet_values
        self.neumann_nodes = None if neumann_nodes is None \
            else np.asarray(neumann_nodes, dtype=np.int64)
        self.dirichlet_elements = mesh.elements_touching(self.dirichlet_nodes)
        if self.neumann_nodes is not None and len(self.neumann_nodes) > 0:
            self.neumann_elements = mesh.elements_touching(self.neumann_nodes)
        else:
            self.neumann_elements = np.array([], dtype=np.int64)
        if design_elements is None:
            design_elements = np.arange(mesh.nelements)
        design = _setdiff(design_elements, self.neumann_elements)
        if len(design) == 0:
            raise ValueError("design_elements is empty")
        self.all_elements = np.arange(mesh.nelements)
        self.fixed_elements = _setdiff(self.all_elements, design)
        self.design_elements = design
        valid = [a for a in (self.dirichlet_elements, self.neumann_elements) if len(a) > 0]
        self.dirichlet_neumann_elements = np.concatenate(valid) if valid \
            else np.array([], dtype=np.int64)
        self.free_dofs = _setdiff(np.arange(self.ndof), self.dirichlet_dofs)
        self.elements_volume = mesh.elements_volume()

    def exlude_dirichlet_from_design(self):
        """``mesh/task_common.py:398-401`` (the reference's spelling)."""
        self.design_elements = _setdiff(self.design_elements, self.dirichlet_elements)

    exclude_dirichlet_from_design = exlude_dirichlet_from_design

    @property
    def ndof(self) -> int:
        return self.mesh.nvertices * self.ncomp

    @property
    def design_mask(self):
        return np.isin(self.all_elements, self.design_elements)

    def dof_fixed_mask(self) -> np.ndarray:
        m = np.zeros(self.ndof, dtype=np.uint8)
        m[self.dirichlet_dofs] = 1
        return m


def _dir_to_comp(d) -> int:
    """'u^1' -> 0 ... (``mesh/task_elastic.py:25-33``); ints pass through."""
    if isinstance(d, (int, np.integer)):
        return int(d)
    if not (isinstance(d, str) and d.startswith("u^") and d[2:].isdigit()):
        raise ValueError(f"force_dir_type must be like 'u^1','u^2','u^3', got: {d}")
    c = int(d[2:]) - 1
    if c < 0:
        raise ValueError(f"Invalid component index parsed from {d}")
    return c


def _as_list(x):
    return x if isinstance(x, list) else [x]


def assemble_surface_forces(mesh, force_facets_ids, force_dir_type, force_value,
                            ncomp=3) -> list:
    """``mesh/task_elastic.py:15-81``: per region, ``pressure = value / A``
    with A the total area of the region's facets, and the consistent load
    ``int_facets pressure v[comp] dS``.  On the planar facets of these
    meshes each facet node receives ``pressure * area / nodes_per_facet``
    (exact integral of a bilinear / linear shape function).  Returns one
    (ndof,) vector per region."""
    facets_list = _as_list(force_facets_ids)
    dirs_list = _as_list(force_dir_type)
    vals_list = _as_list(force_value)
    if not (len(facets_list) == len(dirs_list) == len(vals_list)):
        raise ValueError("Lengths of force_facets_ids, force_dir_type, and force_value "
                         "must match when lists.")
    out = []
    for ids, d, val in zip(facets_list, dirs_list, vals_list):
        comp = _dir_to_comp(d)
        ids = np.asarray(ids, dtype=np.int64)
        area = mesh.facet_areas(ids)
        pressure = float(val) / float(np.sum(area))
        nodes = mesh.facets[:, ids]                          # (nen_f, nf)
        f = np.zeros(mesh.nvertices * ncomp)
        share = pressure * area / nodes.shape[0]
        np.add.at(f, ncomp * nodes.T.ravel() + comp, np.repeat(share, nodes.shape[0]))
        out.append(f)
    return out


def _facet_nodes(mesh, ids):
    return np.unique(mesh.facets[:, np.asarray(ids, dtype=np.int64)].ravel())


class LinearElasticity(FEMDomain):
    """Linear elasticity task: ``E`` (``material_coef``), ``nu`` and one nodal
    load vector per load case (``neumann_linear``; ``n_tasks`` =
    number of load vectors, ``mesh/task_elastic.py:112-114``)."""

    def __init__(self, mesh, dirichlet_nodes, dirichlet_dofs, forces, neumann_nodes,
                 design_elements, E, nu):
        super().__init__(mesh, 3, dirichlet_nodes, dirichlet_dofs, 0.0, neumann_nodes,
                         design_elements)
        self.E = float(E)
        self.nu = float(nu)
        self.neumann_linear = [np.asarray(f, dtype=np.float64) for f in forces]

    @classmethod
    def from_nodes(cls, mesh, dirichlet_nodes, dirichlet_dir, forces, neumann_nodes,
                   design_elements, E, nu):
        """``dirichlet_dir``: "all", "u^1".."u^3" or a component index (or a
        list of them); with a list of node sets, one direction per set
        (``mesh/task_common.py:142-164``)."""
        if isinstance(dirichlet_nodes, list):
            dirs = dirichlet_dir if isinstance(dirichlet_dir, list) \
                else [dirichlet_dir] * len(dirichlet_nodes)
            sets = list(zip(dirichlet_nodes, dirs))
        else:
            sets = [(dirichlet_nodes, dirichlet_dir)]
        dofs, nodes = [], []
        for nd, d in sets:
            nd = np.asarray(nd, dtype=np.int64)
            comps = [0, 1, 2] if d == "all" else [_dir_to_comp(c) for c in np.atleast_1d(d)]
            dofs.append(np.concatenate([3 * nd + c for c in comps]))
            nodes.append(nd)
        dofs = np.unique(np.concatenate(dofs))
        nodes = np.unique(np.concatenate(nodes))
        return cls(mesh, nodes, dofs, forces, neumann_nodes, design_elements, E, nu)

    @classmethod
    def from_facets(cls, mesh, dirichlet_facets_ids, dirichlet_dir, force_facets_ids,
                    force_dir_type, force_value, design_elements, E, nu):
        """``mesh/task_elastic.py:150-250`` with ``FEMDomain.from_facets``
        (``mesh/task_common.py:273-347``): Dirichlet nodes are the nodes of
        the Dirichlet facets, load nodes those of all force facets, and each
        force region becomes one load case."""
        if isinstance(dirichlet_facets_ids, list):
            dn = [_facet_nodes(mesh, f) for f in dirichlet_facets_ids]
        else:
            dn = _facet_nodes(mesh, dirichlet_facets_ids)
        fids = np.concatenate(force_facets_ids) if isinstance(force_facets_ids, list) \
            else force_facets_ids
        nn = _facet_nodes(mesh, fids)
        forces = assemble_surface_forces(mesh, force_facets_ids, force_dir_type, force_value)
        return cls.from_nodes(mesh, dn, dirichlet_dir, forces, nn, design_elements, E, nu)

    @property
    def material_coef(self):
        return self.E

    @property
    def n_tasks(self):
        return len(self.neumann_linear)

    @property
    def force(self):
        return self.neumann_linear


class LinearHeatConduction(FEMDomain):
    """Steady heat conduction with ``objective='compliance'``: conductivity
    ``k``, Dirichlet temperature sets (one value per load case,
    ``mesh/task_heat.py:165-169``) and a nodal heat-source vector ``emit``
    (the reference injects sources through ``robin_linear``,
    ``fea/solver_heat.py:230-235``)."""

    def __init__(self, mesh, dirichlet_nodes, dirichlet_values, emit, design_elements, k,
                 objective="compliance"):
        if objective != "compliance":
            raise NotImplementedError("only objective='compliance' is built")
        vals = dirichlet_values if isinstance(dirichlet_values, list) else [dirichlet_values]
        super().__init__(mesh, 1, dirichlet_nodes, dirichlet_nodes, vals, None,
                         design_elements)
        self.k = float(k)
        self.objective = objective
        self.emit = np.asarray(emit, dtype=np.float64)
        self.robin_coefficient = None

    @property
    def material_coef(self):
        return self.k

    @property
    def n_tasks(self):
        return len(self.dirichlet_values)
