"""Heat conduction with Robin boundaries (oracle) -- TEST INFRASTRUCTURE ONLY.

Restates, for structured box hex8 grids, ``fea/solver_heat.py``:

* ``get_robin_virtual`` (:552-573): the volume ("virtual") Robin forms
  ``int h a(rho) |grad rho| u v`` and ``int h a(rho) |grad rho| T_env v`` with
  ``a = rho^p (1 - rho)^q`` and ``rho`` the interpolant of nodal densities;
* the nodal densities of ``objectives_multi_load`` (:738-745) and
  ``_element_to_nodal_average`` (:600-613): ``np.add.at(rho_nodal,
  t.ravel(), np.repeat(rho, 8))`` -- note that ``t.ravel()`` runs corner by
  corner over all elements while ``np.repeat`` runs element by element, so
  the k-th term added is ``rho[k // 8]`` at node ``t.ravel()[k]``; the
  reference's pairing is reproduced as it is;
* ``_robin_compliance_explicit_grad_`` (:575-597) and
  ``_nodal_gradient_to_element_gradient`` (:616-624);
* ``objectives_multi_load`` (:705-895) for ``objective='compliance'``
  (``J = T^T K T`` with the unenforced total K, conduction + facet Robin +
  virtual Robin), ``'averaged_temp'`` (``J = sum(T)``, adjoint
  ``K lam = 1`` enforced with the state's Dirichlet values) and
  ``'heat_exchange'`` (``J = J_num / J_den``, adjoint load divided by
  ``J_den``, optional ``avg_temp_weight`` blend; functionals on a
  4-point-per-axis rule, see ``heat_exchange_terms``);
* ``assemble_surface_robin`` of ``mesh/task_heat.py:60-139``: the facet
  forms ``int_G h u v`` and ``int_G h T_env v`` (constant: the density
  field there is never set);

with Gauss-Legendre quadrature of ``npts`` points per axis for the volume
forms (the reference uses its basis' rule) and exact 2x2 Gauss on facets.

Tet4 meshes (P1 fields): the same forms with a tetrahedral rule of order
``npts`` (``tet_rule``: 1 the centroid, 2 the symmetric 4-point rule of
degree 2, which is the rule a P1 basis' default integration order 2 selects),
``int_e f = V sum_q w_q f(lambda_q)``; triangle facets with the
edge-midpoint rule (exact for the quadratic facet integrands).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import elements as el
from .fem import enforce, solve
from .mesh import HEX_CORNERS

_S = 2.0 * HEX_CORNERS - 1.0


def shape(xi):
    """N (8,) and reference gradients dN (8, 3) at xi (3,)."""
    f = 1.0 + _S * np.asarray(xi)[None, :]
    N = f[:, 0] * f[:, 1] * f[:, 2] / 8.0
    dN = np.empty((8, 3))
    dN[:, 0] = _S[:, 0] * f[:, 1] * f[:, 2] / 8.0
    dN[:, 1] = _S[:, 1] * f[:, 0] * f[:, 2] / 8.0
    dN[:, 2] = _S[:, 2] * f[:, 0] * f[:, 1] / 8.0
    return N, dN


def quadrature(npts):
    x, w = el._gauss(npts)
    pts = np.array([(a, b, c) for c in x for b in x for a in x])
    wts = np.array([wa * wb * wc for wc in w for wb in w for wa in w])
    return pts, wts


# symmetric 4-point rule of degree 2 on the tetrahedron, in barycentric
# coordinates: one coordinate (5 + 3 sqrt 5) / 20, the others (5 - sqrt 5) / 20
_TET_A = (5.0 + 3.0 * np.sqrt(5.0)) / 20.0
_TET_B = (5.0 - np.sqrt(5.0)) / 20.0


def tet_rule(order):
    """Barycentric points (nq, 4) and weights (nq,) summing to 1."""
    if order <= 1:
        return np.full((1, 4), 0.25), np.ones(1)
    if order == 2:
        lam = np.full((4, 4), _TET_B)
        np.fill_diagonal(lam, _TET_A)
        return lam, np.full(4, 0.25)
    raise ValueError("tetrahedral rules of order 1 and 2 only")


def _tet_fields(mesh, rho_nodal, order):
    """Tet4 counterpart of ``_element_fields``: barycentric values at the
    rule's points lam (nq, 4), barycentric gradients g (n_el, 4, 3), weights
    times volume wd (nq, n_el), rho_h (nq, n_el), grad rho_h (n_el, 3) and
    |grad rho_h| (n_el,)."""
    g, vol = el.tet4_grads(mesh.p, mesh.t)
    lam, w = tet_rule(order)
    rl = rho_nodal[mesh.t]                                  # (4, n_el)
    rh = lam @ rl
    grad = np.einsum("ead,ae->ed", g, rl)
    gn = np.sqrt(np.sum(grad ** 2, axis=1))
    return lam, g, w[:, None] * vol[None, :], rh, grad, gn


def nodal_average_reference(mesh, rho):
    """(rho_nodal, count) with the reference's add.at pairing."""
    t = mesh.t
    nen = t.shape[0]
    rn = np.zeros(mesh.nnodes)
    cnt = np.zeros(mesh.nnodes)
    np.add.at(rn, t.ravel(), np.repeat(rho, nen))
    np.add.at(cnt, t.ravel(), 1.0)
    rn /= np.maximum(cnt, 1.0)
    return rn, cnt


def _element_fields(mesh, rho_nodal, npts):
    """Per quadrature point: N (nq, 8), physical dN (nq, 8, 3), weights*detJ,
    rho_h (n_el, nq) and |grad rho_h| (n_el, nq)."""
    h = np.asarray(mesh.h, dtype=np.float64)
    detJ = h[0] * h[1] * h[2] / 8.0
    pts, wts = quadrature(npts)
    Ns, dNs = [], []
    for xi in pts:
        N, dN = shape(xi)
        Ns.append(N)
        dNs.append(dN * (2.0 / h)[None, :])
    Ns, dNs = np.array(Ns), np.array(dNs)
    rl = rho_nodal[mesh.t]                          # (8, n_el)
    rh = Ns @ rl                                    # (nq, n_el)
    g = np.einsum("qad,ae->qed", dNs, rl)           # (nq, n_el, 3)
    gn = np.sqrt(np.sum(g ** 2, axis=2))            # (nq, n_el)
    return Ns, dNs, wts * detJ, rh, g, gn


def _assemble(mesh, Ke):
    t = mesh.t
    nen = t.shape[0]
    rows = np.repeat(t.T, nen, axis=1).ravel()
    cols = np.tile(t.T, (1, nen)).ravel()
    n = mesh.nnodes
    return sp.csr_matrix((Ke.reshape(Ke.shape[0], -1).ravel(), (rows, cols)), shape=(n, n))


def virtual_robin(mesh, rho, p, q, h, T_env, npts=2):
    """(B_v, L_v) of get_robin_virtual with rho the element densities."""
    rn, _ = nodal_average_reference(mesh, rho)
    return virtual_robin_nodal(mesh, rn, p, q, h, T_env, npts)


def virtual_robin_nodal(mesh, rn, p, q, h, T_env, npts=2):
    """The same from nodal densities."""
    if mesh.kind == "tet4":
        lam, _, wd, rh, _, gn = _tet_fields(mesh, rn, npts)
        c = h * rh ** p * (1.0 - rh) ** q * gn[None, :]     # (nq, n_el)
        Ke = np.einsum("qe,qe,qa,qb->eab", wd, c, lam, lam)
        Le = np.einsum("qe,qe,qa->ea", wd, c * T_env, lam)
        L = np.zeros(mesh.nnodes)
        np.add.at(L, mesh.t.T.ravel(), Le.ravel())
        return _assemble(mesh, Ke), L
    Ns, dNs, wd, rh, g, gn = _element_fields(mesh, rn, npts)
    c = h * rh ** p * (1.0 - rh) ** q * gn          # (nq, n_el)
    Ke = np.einsum("q,qe,qa,qb->eab", wd, c, Ns, Ns)
    Le = np.einsum("q,qe,qa->ea", wd, c * T_env, Ns)
    L = np.zeros(mesh.nnodes)
    np.add.at(L, mesh.t.T.ravel(), Le.ravel())
    return _assemble(mesh, Ke), L


def facet_robin_tri(p_nodes, facets, h, T_env, n):
    """(B_f, L_f) of one region of triangle facets (3, nf), by the
    edge-midpoint rule: weight A / 3 at each edge midpoint, where the two
    edge nodes' shape functions are 1/2 and the third is 0."""
    rows, cols, vals = [], [], []
    L = np.zeros(n)
    mids = np.array([[0.5, 0.5, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5]])
    for f in range(facets.shape[1]):
        P = p_nodes[:, facets[:, f]]
        A = 0.5 * np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]))
        Me = sum((A / 3.0) * h * np.outer(N, N) for N in mids)
        Le = sum((A / 3.0) * h * T_env * N for N in mids)
        for i in range(3):
            L[facets[i, f]] += Le[i]
            for j in range(3):
                rows.append(facets[i, f])
                cols.append(facets[j, f])
                vals.append(Me[i, j])
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n)), L


def facet_robin(p_nodes, facets, h, T_env, n):
    """(B_f, L_f) of one facet region: facets (4, nf) cyclic quads, or
    (3, nf) triangles."""
    if facets.shape[0] == 3:
        return facet_robin_tri(p_nodes, facets, h, T_env, n)
    x, w = el._gauss(2)
    sg = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], dtype=float)
    rows, cols, vals = [], [], []
    L = np.zeros(n)
    for f in range(facets.shape[1]):
        P = p_nodes[:, facets[:, f]]
        Me = np.zeros((4, 4))
        Le = np.zeros(4)
        for a, wa in zip(x, w):
            for b, wb in zip(x, w):
                N = 0.25 * (1 + sg[:, 0] * a) * (1 + sg[:, 1] * b)
                dNa = 0.25 * sg[:, 0] * (1 + sg[:, 1] * b)
                dNb = 0.25 * sg[:, 1] * (1 + sg[:, 0] * a)
                J = np.linalg.norm(np.cross(P @ dNa, P @ dNb))
                Me += wa * wb * J * h * np.outer(N, N)
                Le += wa * wb * J * h * T_env * N
        for i in range(4):
            L[facets[i, f]] += Le[i]
            for j in range(4):
                rows.append(facets[i, f])
                cols.append(facets[j, f])
                vals.append(Me[i, j])
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n)), L


def robin_explicit_grad(mesh, rho, T, p, q, h, T_env, npts=2):
    """Element gradient (n_el,) of one load case: the nodal linear form of
    ``_robin_compliance_explicit_grad_`` mapped by
    ``_nodal_gradient_to_element_gradient``."""
    rn, cnt = nodal_average_reference(mesh, rho)
    nodal = robin_explicit_grad_nodal(mesh, rn, T, p, q, h, T_env, npts)
    out = np.zeros(mesh.nelements)
    for local in mesh.t:
        out += nodal[local] / np.maximum(cnt[local], 1.0)
    return out


def robin_functional(mesh, rn, T, p, q, h, T_env, npts=2):
    """int h a(rho) |grad rho| (2 T_env T - T^2): the functional whose
    derivative in the nodal densities (T fixed) the explicit gradient is."""
    if mesh.kind == "tet4":
        lam, _, wd, rh, _, gn = _tet_fields(mesh, rn, npts)
        Th = lam @ T[mesh.t]
        phi = 2.0 * T_env * Th - Th * Th
        return float(np.sum(wd * h * rh ** p * (1.0 - rh) ** q * gn[None, :] * phi))
    Ns, dNs, wd, rh, g, gn = _element_fields(mesh, rn, npts)
    Th = Ns @ T[mesh.t]
    phi = 2.0 * T_env * Th - Th * Th
    return float(np.einsum("q,qe->", wd, h * rh ** p * (1.0 - rh) ** q * gn * phi))


def robin_explicit_grad_nodal(mesh, rn, T, p, q, h, T_env, npts=2):
    """The nodal linear form of ``_robin_compliance_explicit_grad_``."""
    if mesh.kind == "tet4":
        lam, gb, wd, rh, grad, gn = _tet_fields(mesh, rn, npts)
        Th = lam @ T[mesh.t]                          # (nq, n_el)
        gs = np.maximum(gn, 1e-12)
        a = rh ** p * (1.0 - rh) ** q
        da = p * rh ** (p - 1.0) * (1.0 - rh) ** q - q * rh ** p * (1.0 - rh) ** (q - 1.0)
        phi = 2.0 * T_env * Th - Th * Th
        gv = np.einsum("ed,ead->ea", grad / gs[:, None], gb)     # constant on the tet
        Le = np.einsum("qe,qe,qa->ea", wd, h * da * gn[None, :] * phi, lam) + \
            np.einsum("qe,qe,ea->ea", wd, h * a * phi, gv)
        nodal = np.zeros(mesh.nnodes)
        np.add.at(nodal, mesh.t.T.ravel(), Le.ravel())
        return nodal
    Ns, dNs, wd, rh, g, gn = _element_fields(mesh, rn, npts)
    Tl = T[mesh.t]                                   # (8, n_el)
    Th = Ns @ Tl                                     # (nq, n_el)
    gs = np.maximum(gn, 1e-12)
    a = rh ** p * (1.0 - rh) ** q
    da = p * rh ** (p - 1.0) * (1.0 - rh) ** q - q * rh ** p * (1.0 - rh) ** (q - 1.0)
    phi = 2.0 * T_env * Th - Th * Th
    term1 = da * gn * phi                            # multiplies v
    gv = np.einsum("qed,qad->qea", g / gs[:, :, None], dNs)   # (grad rho / |.|) . grad N_a
    Le = np.einsum("q,qe,qa->ea", wd, h * term1, Ns) + \
        np.einsum("q,qe,qea->ea", wd, h * a * phi, gv)
    nodal = np.zeros(mesh.nnodes)
    np.add.at(nodal, mesh.t.T.ravel(), Le.ravel())
    return nodal


def heat_exchange_terms(mesh, rn, T, p, q, h, T_env, npts=4):
    """(J_den, J_num (k,), rhs (nn,)) of the heat-exchange objective
    (``solver_heat.py:306-322`` ``_hx_num_density_``, ``:380-390``
    ``_hx_den_density_``, ``:393-408`` ``_hx_adj_rhs_``, ``:411-436``
    ``heat_exchange_objective``): J_den = int |grad rho_h|, J_num_k =
    -T_env h int (T_k - T_env) |grad rho_h| with the scalar h, and the
    adjoint load int -T_env h_eff |grad rho_h| v with h_eff the interpolant
    of the nodal h rho^p (1 - rho)^q (``:824-829``) -- not divided by J_den.
    The reference evaluates them on ``skfem.CellBasis(mesh, elem)`` with its
    default rule; for ElementHex1 (maximum degree 3) skfem's default
    integration order is 2 * 3 = 6, a Gauss-Legendre rule of ceil(7 / 2) = 4
    points per axis, the default ``npts`` here.  For tet4 meshes (ElementTetP1,
    degree 1, default order 2) the integrands are linear in T and rho_h, so
    ``npts`` >= 1 gives the same values; order 2 is taken."""
    if mesh.kind == "tet4":
        lam, _, wd, rh, _, gn = _tet_fields(mesh, rn, 2)
        den = float(np.sum(wd * gn[None, :]))
        Th = np.einsum("qa,aek->qek", lam, T[mesh.t])
        num = -T_env * h * np.einsum("qe,qek,e->k", wd, Th - T_env, gn)
        heff = h * rn ** p * (1.0 - rn) ** q
        he = lam @ heff[mesh.t]
        Le = np.einsum("qe,qe,qa->ea", wd, -T_env * he * gn[None, :], lam)
        rhs = np.zeros(mesh.nnodes)
        np.add.at(rhs, mesh.t.T.ravel(), Le.ravel())
        return den, num, rhs
    Ns, dNs, wd, rh, g, gn = _element_fields(mesh, rn, npts)
    den = float(np.einsum("q,qe->", wd, gn))
    Th = np.einsum("qa,aek->qek", Ns, T[mesh.t])             # (nq, n_el, k)
    num = -T_env * h * np.einsum("q,qek,qe->k", wd, Th - T_env, gn)
    heff = h * rn ** p * (1.0 - rn) ** q
    he = Ns @ heff[mesh.t]                                    # (nq, n_el)
    Le = np.einsum("q,qe,qa->ea", wd, -T_env * he * gn, Ns)
    rhs = np.zeros(mesh.nnodes)
    np.add.at(rhs, mesh.t.T.ravel(), Le.ravel())
    return den, num, rhs


class HeatRobinFEM:
    """Oracle of ``FEM_SimpLinearHeatConduction`` with Robin data on a
    structured hex grid or a tet4 mesh.  ``robin_facets``: list of (4, nf)
    quad or (3, nf) triangle facet arrays; ``emit``: extra volumetric source
    vector added to the Robin loads.  On tets ``npts`` is the order of the
    tetrahedral rule of the virtual forms (``tet_rule``)."""

    def __init__(self, mesh, dirichlet_dofs, dirichlet_values, k, E_min_coeff, robin_facets,
                 h, T_env, emit=None, objective="compliance", q=4, npts=2, npts_cell=4,
                 avg_temp_weight=0.0):
        self.mesh = mesh
        self.D = np.asarray(dirichlet_dofs)
        self.values = list(dirichlet_values)
        self.k_max, self.k_min = k, k * E_min_coeff
        self.h, self.T_env, self.q, self.npts = h, T_env, q, npts
        self.objective = objective
        self.npts_cell = npts_cell
        self.avg_temp_weight = float(avg_temp_weight)
        # unit-conductivity element matrices, (n_el, nen, nen)
        if mesh.kind == "tet4":
            self.Ke0 = el.tet4_conduction_K(mesh.p, mesh.t)
        else:
            K0 = el.hex8_conduction_K0(mesh.h)
            self.Ke0 = np.broadcast_to(K0, (mesh.nelements,) + K0.shape)
        n = mesh.nnodes
        self.Bf = sp.csr_matrix((n, n))
        self.Lf = np.zeros(n)
        for fc in robin_facets:
            B, L = facet_robin(mesh.p, fc, h, T_env, n)
            self.Bf = self.Bf + B
            self.Lf = self.Lf + L
        self.emit0 = np.zeros(n) if emit is None else np.asarray(emit, float)
        self.lam = None

    def _K(self, rho, p):
        from . import transforms as tr
        E = tr.simp_interpolation(rho, self.k_max, self.k_min, p)
        Kc = _assemble(self.mesh, E[:, None, None] * self.Ke0)
        Bv, Lv = virtual_robin(self.mesh, rho, p, self.q, self.h, self.T_env, self.npts)
        return (Kc + self.Bf + Bv).tocsr(), self.emit0 + self.Lf + Lv

    def _solve(self, K, rhs, n_loads):
        cols = []
        for i in range(n_loads):
            Ke, fe = enforce(K, rhs, self.D, np.full(len(self.D), self.values[i]))
            cols.append(fe)
        return solve(Ke, np.column_stack(cols), "spsolve", 1e-12)

    def objectives_multi_load(self, rho, p, u_dofs, timer=None, force_scale=1.0):
        K, emit = self._K(rho, p)
        n_loads = u_dofs.shape[1]
        U = self._solve(K, emit, n_loads)
        u_dofs[:, :] = U
        if self.objective == "compliance":
            self.lam = -2.0 * U
            return np.array([float(U[:, i] @ (K @ U[:, i])) for i in range(n_loads)])
        if self.objective == "averaged_temp":
            self.lam = self._solve(K, np.ones_like(emit), n_loads)
            return np.array([float(np.sum(U[:, i])) for i in range(n_loads)])
        if self.objective == "heat_exchange":
            # solver_heat.py:791-869
            rn, _ = nodal_average_reference(self.mesh, rho)
            den, num, rhs = heat_exchange_terms(self.mesh, rn, U, p, self.q, self.h,
                                                self.T_env, self.npts_cell)
            if den <= 1e-16:
                self.lam = np.zeros_like(U)
                return np.zeros(n_loads)
            J = num / den
            self.lam = self._solve(K, rhs / den, n_loads)
            w = self.avg_temp_weight
            if w != 0.0:
                avg = U.sum(axis=0)
                lam_avg = self._solve(K, np.ones_like(emit), n_loads)
                if n_loads > 1:
                    hs = max(float(np.mean(np.abs(J))), 1e-12)
                    avs = max(float(np.mean(np.abs(avg))), 1e-12)
                    J = J / hs + w * (avg / avs)
                    self.lam = self.lam / hs + (w / avs) * lam_avg
                else:
                    J = J + w * avg
                    self.lam = self.lam + w * lam_avg
            return J
        raise ValueError(self.objective)

    def energy_multi_load(self, rho, p, u_dofs):
        from . import transforms as tr
        Ul = u_dofs[self.mesh.t]                                  # (8, n_el, k)
        if self.objective == "compliance":
            E = tr.simp_interpolation(rho, self.k_max, self.k_min, p)
            return 0.5 * E[:, None] * np.einsum("aek,eab,bek->ek", Ul, self.Ke0, Ul)
        Ll = self.lam[self.mesh.t]
        return np.einsum("aek,eab,bek->ek", Ul, self.Ke0, Ll)

    def compliance_sensitivity_multi_load(self, rho, p, u_dofs):
        from . import transforms as tr
        E = tr.simp_interpolation(rho, self.k_max, self.k_min, p)
        Ul = u_dofs[self.mesh.t]
        energy = 0.5 * E[:, None] * np.einsum("aek,eab,bek->ek", Ul, self.Ke0, Ul)
        out = np.column_stack([tr.dC_drho_simp(rho, energy[:, i], self.k_max, self.k_min, p)
                               for i in range(energy.shape[1])])
        if self.objective != "compliance":
            return out
        for i in range(out.shape[1]):
            out[:, i] += robin_explicit_grad(self.mesh, rho, u_dofs[:, i], p, self.q, self.h,
                                             self.T_env, self.npts)
        return out
