"""FEM state solve, compliance and element energies (oracle).

Restates:

* ``assemble_stiffness_matrix`` / ``assemble_conduction_matrix``
  (``fea/composer.py:53-145``): ``K = sum_e E(rho_e) K_e^0``;
* Dirichlet enforcement by zeroed rows, unit diagonal and ``b_D = x_D``
  (``skfem.enforce`` as called at ``fea/solver_elastic.py:211,361,366``,
  ``fea/solver_heat.py:48,107,118,160,168``);
* the solve with SuperLU -- ``spsolve`` for one load
  (``solver_elastic.py:106-108``), one ``splu`` factorisation for several
  (``solver_elastic.py:384-390``) -- or the ``cg_jacobi`` option
  (``solver_elastic.py:84-92``);
* elastic compliance ``C_k = F_e[:,k] . u_k`` (``solver_elastic.py:236,465``);
* heat compliance ``J_k = T_k^T K T_k`` with the unenforced K
  (``solver_heat.py:785-789``);
* element energy ``U_e = 1/2 E_e u_e^T K^0 u_e``
  (``solver_elastic.py:470-537``; heat ``solver_heat.py:256-303``).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import elements as el


def element_dofs(t: np.ndarray, ncomp: int) -> np.ndarray:
    """(nen*ncomp, n_el) global DOFs, node-major within the element."""
    nen, n_el = t.shape
    if ncomp == 1:
        return t.copy()
    d = np.empty((nen * ncomp, n_el), dtype=np.int64)
    for a in range(nen):
        for c in range(ncomp):
            d[ncomp * a + c] = ncomp * t[a] + c
    return d


class ElementOperator:
    """Unit-modulus element matrices for a mesh and a physics kind."""

    def __init__(self, mesh, physics: str, nu: float = 0.3, order: int = 2):
        self.mesh = mesh
        self.physics = physics
        self.ncomp = 3 if physics == "elastic" else 1
        self.nu = nu
        if mesh.kind == "hex8":
            if physics == "elastic":
                K0 = el.hex8_elasticity_K0(nu, mesh.h, order)
            else:
                K0 = el.hex8_conduction_K0(mesh.h, order)
            self.K0 = K0            # shared by all elements
            self.Ke = None
        else:
            self.K0 = None
            if physics == "elastic":
                self.Ke = el.tet4_elasticity_K(mesh.p, mesh.t, nu)
            else:
                self.Ke = el.tet4_conduction_K(mesh.p, mesh.t)
        self.edofs = element_dofs(mesh.t, self.ncomp)
        self.ndof = mesh.nnodes * self.ncomp

    def element_matrices(self) -> np.ndarray:
        if self.Ke is not None:
            return self.Ke
        return np.broadcast_to(self.K0, (self.mesh.nelements,) + self.K0.shape)

    def assemble(self, E: np.ndarray) -> sp.csr_matrix:
        Ke = self.element_matrices() * E[:, None, None]
        nd = self.edofs.shape[0]
        rows = np.repeat(self.edofs.T, nd, axis=1).ravel()
        cols = np.tile(self.edofs.T, (1, nd)).ravel()
        K = sp.coo_matrix((Ke.ravel(), (rows, cols)),
                          shape=(self.ndof, self.ndof)).tocsr()
        return K

    def element_quadratic(self, U: np.ndarray) -> np.ndarray:
        """``u_e^T K_e^0 u_e`` for every element and column, (n_el, k)."""
        Ue = U[self.edofs]                       # (nd, n_el, k)
        if self.Ke is not None:
            KU = np.einsum("nab,bnk->ank", self.Ke, Ue)
        else:
            KU = np.einsum("ab,bnk->ank", self.K0, Ue)
        return np.einsum("ank,ank->nk", Ue, KU)


def enforce(K: sp.csr_matrix, b: np.ndarray, D: np.ndarray, x_D=None):
    """Zero the rows in D, put 1 on their diagonal, set ``b_D = x_D``
    (default 0).  Returns new (K_e, b_e)."""
    K = K.tocsr()
    D = np.asarray(D, dtype=np.int64)
    keep = np.ones(K.shape[0])
    keep[D] = 0.0
    K = sp.diags(keep) @ K
    K = K + sp.diags(1.0 - keep)
    b = np.array(b, dtype=np.float64, copy=True)
    b[D] = 0.0 if x_D is None else x_D
    return K.tocsr(), b


def cg_jacobi(K, b, rtol=1e-8, maxiter=None, x0=None):
    """SciPy CG with a Jacobi preconditioner (``solver_elastic.py:84-92``)."""
    M_inv = 1.0 / K.diagonal()
    M = spla.LinearOperator(K.shape, matvec=lambda x: M_inv * x)
    n_it = [0]

    def cb(_):
        n_it[0] += 1
    u, info = spla.cg(K, b, M=M, rtol=rtol, maxiter=maxiter, x0=x0,
                      callback=cb)
    return u, info, n_it[0]


SOLVE_STATS: list = []


def solve(K_e: sp.csr_matrix, F: np.ndarray, solver: str = "spsolve",
          rtol: float = 1e-8, maxiter=None) -> np.ndarray:
    """Solve ``K_e U = F`` for (N,) or (N, k) right-hand sides."""
    F2 = F[:, None] if F.ndim == 1 else F
    if solver == "spsolve":
        if F2.shape[1] == 1:
            U = spla.spsolve(K_e.tocsc(), F2[:, 0])[:, None]
        else:
            U = spla.splu(K_e.tocsc()).solve(F2)
    elif solver == "cg_jacobi":
        cols = []
        for i in range(F2.shape[1]):
            u, info, n_it = cg_jacobi(K_e, F2[:, i], rtol, maxiter)
            SOLVE_STATS.append(n_it)
            cols.append(u)
        U = np.column_stack(cols)
    else:
        raise ValueError(f"Unknown solver: {solver}")
    return U if F.ndim == 2 else U[:, 0]


class ElasticFEM:
    """Oracle of ``FEM_SimpLinearElasticity`` (``solver_elastic.py:540-676``)."""

    def __init__(self, task, E_min_coeff: float, interpolation: str = "SIMP",
                 solver: str = "spsolve", rtol: float = 1e-12):
        from . import transforms as tr
        self.task = task
        self.rtol = rtol
        self.E_max = task.E0 * 1.0
        self.E_min = task.E0 * E_min_coeff
        self.interp = tr.simp_interpolation if interpolation == "SIMP" \
            else tr.ramp_interpolation
        self.solver = solver
        self.op = task.element_operator()
        self.last_K = None

    def objectives_multi_load(self, rho, p, u_dofs, timer=None, force_scale=1.0):
        F = np.column_stack([f * force_scale for f in self.task.forces])
        E = self.interp(rho, self.E_max, self.E_min, p)
        K = self.op.assemble(E)
        K_e, _ = enforce(K, F[:, 0], self.task.dirichlet_dofs)
        F_e = F.copy()
        F_e[self.task.dirichlet_dofs, :] = 0.0
        U = solve(K_e, F_e, self.solver, self.rtol, maxiter=200000)
        u_dofs[:, :] = U
        return np.einsum("ij,ij->j", F_e, U)

    def energy_multi_load(self, rho, p, u_dofs):
        E = self.interp(rho, self.E_max, self.E_min, p)
        return 0.5 * E[:, None] * self.op.element_quadratic(u_dofs)


class HeatFEM:
    """Oracle of ``FEM_SimpLinearHeatConduction`` for ``objective='compliance'``
    without Robin terms (``solver_heat.py:628-980``).  The volumetric source is
    supplied as the ``emit`` vector (``solver_heat.py:230-235``)."""

    def __init__(self, task, E_min_coeff: float, interpolation: str = "SIMP",
                 solver: str = "spsolve", rtol: float = 1e-12):
        from . import transforms as tr
        self.task = task
        self.rtol = rtol
        self.k_max = task.E0 * 1.0
        self.k_min = task.E0 * E_min_coeff
        self.interp = tr.simp_interpolation if interpolation == "SIMP" \
            else tr.ramp_interpolation
        self.solver = solver
        self.op = task.element_operator()

    def objectives_multi_load(self, rho, p, u_dofs, timer=None, force_scale=1.0):
        E = self.interp(rho, self.k_max, self.k_min, p)
        K = self.op.assemble(E)
        emit = np.sum(self.task.forces, axis=0) if len(self.task.forces) > 1 \
            else self.task.forces[0]
        n_loads = u_dofs.shape[1]
        cols = []
        for i in range(n_loads):
            val = self.task.dirichlet_values[i]
            K_e, f_e = enforce(K, emit, self.task.dirichlet_dofs,
                               np.full(len(self.task.dirichlet_dofs), val))
            cols.append(f_e)
        U = solve(K_e, np.column_stack(cols), self.solver, self.rtol,
                  maxiter=200000)
        u_dofs[:, :] = U
        return np.array([float(U[:, i] @ (K @ U[:, i])) for i in range(n_loads)])

    def energy_multi_load(self, rho, p, u_dofs):
        E = self.interp(rho, self.k_max, self.k_min, p)
        return 0.5 * E[:, None] * self.op.element_quadratic(u_dofs)

    def compliance_sensitivity_multi_load(self, rho, p, u_dofs):
        """Always SIMP-based (``solver_heat.py:928-950``)."""
        from . import transforms as tr
        energy = self.energy_multi_load(rho, p, u_dofs)
        return np.column_stack([
            tr.dC_drho_simp(rho, energy[:, i], self.k_max, self.k_min, p)
            for i in range(energy.shape[1])])
