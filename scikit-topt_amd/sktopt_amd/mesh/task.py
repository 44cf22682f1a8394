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

    def scale(self, L_scale: float, F_scale: float):
        """``mesh/task_common.py:403-420``: node coordinates times
        ``L_scale`` (a new mesh), every load vector times ``F_scale`` in
        place.  Element volumes keep their construction-time values, as in
        the reference.  Only tasks with load vectors can be scaled."""
        forces = getattr(self, "neumann_linear", None)
        if forces is None:
            raise ValueError("should be ndarray or list of ndarray")
        self.mesh = self.mesh.scaled(L_scale)
        for f in forces:
            f *= F_scale

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


_QUAD_MASS = np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]],
                      dtype=np.float64) / 36.0


# packed index of entry (a, b) of a symmetric 4 x 4 element matrix, order
# 00 11 22 33 01 02 03 12 13 23 (the layout of skt_tet4_set_extra)
_SYM10 = np.array([[0, 4, 5, 6], [4, 1, 7, 8], [5, 7, 2, 9], [6, 8, 9, 3]], dtype=np.int64)


def _tet_facet_owners(mesh, ids):
    """(element (nf,), local corners (3, nf)) of boundary triangle facets:
    the tet each facet bounds and the corner slot of each facet node in it."""
    from .meshes import TET_FACES
    f = mesh.facets[:, ids]                                   # (3, nf)
    nn = np.int64(mesh.nvertices)
    code = lambda k: (k[0] * nn + k[1]) * nn + k[2]           # noqa: E731
    allf = np.concatenate([mesh.t[face] for face in TET_FACES], axis=1)   # (3, 4 n_el)
    ac = code(np.sort(allf, axis=0))
    order = np.argsort(ac, kind="stable")
    fc = code(np.sort(f, axis=0))
    pos = np.searchsorted(ac[order], fc)
    if np.any(pos >= ac.size) or np.any(ac[order][np.minimum(pos, ac.size - 1)] != fc):
        raise ValueError("a Robin facet is not a face of the mesh")
    elem = order[pos] % mesh.nelements
    loc = np.empty_like(f)
    for i in range(3):
        hit = mesh.t[:, elem] == f[i][None, :]                # (4, nf)
        loc[i] = np.argmax(hit, axis=0)
    return elem, loc


def assemble_surface_robin(mesh, robin_facets_ids, robin_coefficient, robin_bc_value):
    """``mesh/task_heat.py:60-139``: the facet forms ``int_G h u v`` and
    ``int_G h T_env v`` of every Robin region, summed, with a load vector.
    (The reference's density argument never reaches these forms, so they do
    not change with rho.)

    * Structured hex grids: a 27-point nodal stencil (27, n_nodes).  Facets
      are planar parallelograms: mass ``h A / 36 [[4 2 1 2] ...]`` in cyclic
      node order and load ``h T_env A / 4`` per node.
    * Tet4 meshes: element matrices (10, n_el) packed 00 11 22 33 01 02 03 12
      13 23, each triangle's mass ``h A / 12 (1 + d_ab)`` added to the tet it
      bounds; load ``h T_env A / 3`` per node."""
    kind = getattr(mesh, "kind", None)
    if not (kind == "tet4" or (kind == "hex8" and getattr(mesh, "grid", None) is not None)):
        raise NotImplementedError("Robin boundaries are built for structured hex8 grids and "
                                  "tet4 meshes")
    facets_list = _as_list(robin_facets_ids)
    hs = [float(robin_coefficient)] * len(facets_list) if not isinstance(
        robin_coefficient, list) else [float(v) for v in robin_coefficient]
    ts = [float(robin_bc_value)] * len(facets_list) if not isinstance(
        robin_bc_value, list) else [float(v) for v in robin_bc_value]
    if not (len(facets_list) == len(hs) == len(ts)):
        raise ValueError("Lengths of robin_facets_ids and robin_value must match when lists.")
    if kind == "tet4":
        R = np.zeros((10, mesh.nelements))
        load = np.zeros(mesh.nvertices)
        for ids, h, te in zip(facets_list, hs, ts):
            ids = np.asarray(ids, dtype=np.int64)
            f = mesh.facets[:, ids]
            area = mesh.facet_areas(ids)
            elem, loc = _tet_facet_owners(mesh, ids)
            for i in range(3):
                np.add.at(load, f[i], h * te * area / 3.0)
                for j in range(3):
                    w = (2.0 if i == j else 1.0) / 12.0
                    k = _SYM10[loc[i], loc[j]]
                    # each off-diagonal pair is visited twice (i, j) and (j, i)
                    np.add.at(R, (k, elem), h * area * w * (1.0 if i == j else 0.5))
        return R, load
    nx, ny, nz = mesh.grid
    nn = mesh.nvertices
    i, j, k = mesh.node_ijk(nx, ny, nz)
    st = np.zeros((27, nn))
    load = np.zeros(nn)
    for ids, h, te in zip(facets_list, hs, ts):
        ids = np.asarray(ids, dtype=np.int64)
        f = mesh.facets[:, ids]                               # (4, nf) cyclic
        area = mesh.facet_areas(ids)
        for a in range(4):
            np.add.at(load, f[a], h * te * area / 4.0)
            for b in range(4):
                na, nb = f[a], f[b]
                o = (i[nb] - i[na] + 1) + 3 * (j[nb] - j[na] + 1) + 9 * (k[nb] - k[na] + 1)
                np.add.at(st, (o, na), h * area * _QUAD_MASS[a, b])
    return st, load


class LinearHeatConduction(FEMDomain):
    """Steady heat conduction (``mesh/task_heat.py:141-250``): conductivity
    ``k``, Dirichlet temperature sets (one value per load case,
    ``:165-169``), a nodal heat-source vector ``emit`` (the reference
    injects sources through ``robin_linear``, ``fea/solver_heat.py:230-235``)
    and optional Robin facets with coefficient ``robin_coefficient`` and
    ambient value ``robin_bc_value``.  Objectives ``'compliance'``,
    ``'averaged_temp'`` and ``'heat_exchange'`` (with ``avg_temp_weight``)."""

    def __init__(self, mesh, dirichlet_nodes, dirichlet_values, emit, design_elements, k,
                 objective="compliance", robin_facets_ids=None, robin_coefficient=None,
                 robin_bc_value=None, design_robin_boundary=None, avg_temp_weight=0.0):
        if objective not in ("compliance", "heat_exchange", "averaged_temp"):
            raise ValueError("objective must be one of 'compliance', 'heat_exchange', or "
                             "'averaged_temp'")
        vals = dirichlet_values if isinstance(dirichlet_values, list) else [dirichlet_values]
        if robin_facets_ids is None:
            robin_coefficient = robin_bc_value = design_robin_boundary = None
        self.robin_facets_ids = robin_facets_ids
        self.robin_coefficient = robin_coefficient
        self.robin_bc_value = robin_bc_value
        self.design_robin_boundary = design_robin_boundary
        super().__init__(mesh, 1, dirichlet_nodes, dirichlet_nodes, vals, None,
                         design_elements)
        if robin_facets_ids is not None:
            fids = np.concatenate(robin_facets_ids) if isinstance(robin_facets_ids, list) \
                else np.asarray(robin_facets_ids)
            self.robin_nodes = _facet_nodes(mesh, fids)
            self.robin_elements = mesh.elements_touching(self.robin_nodes)
            if design_robin_boundary is False:
                d = _setdiff(self.design_elements, self.robin_elements)
                if len(d) == 0:
                    raise ValueError("design_elements is empty")
                self.design_elements = d
                self.fixed_elements = _setdiff(self.all_elements, d)
            if isinstance(robin_coefficient, list) or isinstance(robin_bc_value, list):
                raise NotImplementedError("one scalar Robin coefficient and ambient value "
                                          "(the virtual Robin terms take scalars)")
            mats, self.robin_load = assemble_surface_robin(
                mesh, robin_facets_ids, robin_coefficient, robin_bc_value)
            if mesh.kind == "tet4":
                self.robin_elem_mat = mats
            else:
                self.robin_stencil = mats
        else:
            self.robin_nodes = None
            self.robin_elements = None
        self.k = float(k)
        self.objective = objective
        self.avg_temp_weight = float(avg_temp_weight)
        self.emit = np.zeros(mesh.nvertices) if emit is None \
            else np.asarray(emit, dtype=np.float64)

    @classmethod
    def from_facets(cls, mesh, dirichlet_facets_ids, dirichlet_values, robin_facets_ids,
                    robin_coefficient, robin_bc_value, design_robin_boundary, design_elements,
                    k, objective="compliance", avg_temp_weight=0.0, emit=None):
        """``mesh/task_heat.py:170-250`` with nodes taken from facets
        (``mesh/task_common.py:273-347``).  One Dirichlet node set; a list of
        values makes one load case per value."""
        if isinstance(dirichlet_facets_ids, list):
            dn = np.unique(np.concatenate([_facet_nodes(mesh, f) for f in dirichlet_facets_ids]))
        else:
            dn = _facet_nodes(mesh, dirichlet_facets_ids)
        return cls(mesh, dn, dirichlet_values, emit, design_elements, k, objective,
                   robin_facets_ids, robin_coefficient, robin_bc_value, design_robin_boundary,
                   avg_temp_weight)

    @property
    def material_coef(self):
        return self.k

    @property
    def n_tasks(self):
        return len(self.dirichlet_values)
