"""Matrix-free GPU FEA with the reference's call surface.

``FEM_SimpLinearElasticity`` mirrors ``fea/solver_elastic.py:540-676`` and
``FEM_SimpLinearHeatConduction`` mirrors ``fea/solver_heat.py:628-980``
(``objective='compliance'``, no Robin terms):

* ``objectives_multi_load(rho, p, u_dofs, timer=None, force_scale=1.0)``
  solves ``K(E(rho)) u_k = f_k`` for every load case, writes ``u_dofs``
  (ndof, n_tasks) in place and returns the per-case objective;
* ``energy_multi_load(rho, p, u_dofs)`` returns element energies
  ``1/2 E_e u_e^T K_e^0 u_e`` (n_el, n_tasks);
* the heat class also has ``compliance_sensitivity_multi_load``.

Arguments may be float64 device tensors (the device-resident path) or numpy
arrays (copied to the GPU and back, ``u_dofs`` updated in place) so the
reference's own driver can call it unchanged.

The solve is a batched Jacobi-preconditioned CG (all load cases share each
operator apply): ``precision="mixed"`` runs float32 inner iterations inside
a float64 iterative refinement, ``precision="fp64"`` runs float64 CG.  The
reference's solver names are accepted; all of them select this solve.
"""
from __future__ import annotations

import ctypes as C
import logging
from dataclasses import dataclass

import numpy as np
import torch

from .. import _dev, _lib
from .composer import interpolation_kind, simp_interpolation

logger = logging.getLogger(__name__)

_SOLVER_NAMES = ("auto", "spsolve", "cg_jacobi", "cg_pyamg", "petsc", "petsc_spdirect", "pcg")


@dataclass(frozen=True)
class LinearSolverConfig:
    """``fea/solver_elastic.py:34-40`` plus the GPU solve's own knobs."""
    solver: str = "auto"
    rtol: float = 1e-8
    maxiter: int | None = None
    petsc_options: object = None
    allow_fallback_to_spsolve: bool = False
    precision: str = "mixed"          # "mixed" (fp32 PCG, fp64 reliable updates) or "fp64"
    inner_rtol: float = 0.0           # fp32 residual drop that triggers a reliable update
                                      # (<= 0: 1e-2 with multigrid, 1e-1 with Jacobi)
    warm_start: bool = True           # start from the previous u_dofs
    # the design loop's state solves start from U + theta (U - U_prev), U and
    # U_prev the last two solutions (0: the last solution as it is)
    warm_start_extrapolation: float = 0.5
    on_noconv: str = "raise"          # "raise" or "warn"
    preconditioner: str = "auto"      # "auto" (multigrid where possible), "mg", "jacobi"
    mg_smooth_degree: int = 2         # Chebyshev-Jacobi steps per smoothing
    mg_max_coarse: int = 2000         # DOFs of the coarsest (dense) multigrid level


def normalize_linear_solver_config(solver="auto", rtol=1e-8, maxiter=None,
                                   petsc_options=None, allow_fallback_to_spsolve=False,
                                   **kw) -> LinearSolverConfig:
    if isinstance(solver, LinearSolverConfig):
        return solver
    if solver not in _SOLVER_NAMES:
        raise ValueError(f"Unknown solver: {solver}")
    return LinearSolverConfig(solver=solver, rtol=rtol, maxiter=maxiter,
                              petsc_options=petsc_options,
                              allow_fallback_to_spsolve=allow_fallback_to_spsolve, **kw)


class _GpuFEM:
    physics = _lib.ELASTIC

    def __init__(self, task, E_min_coeff, density_interpolation=simp_interpolation,
                 solver_config=None, solver_option="spsolve", quad_order=2):
        self.task = task
        self.kind = interpolation_kind(density_interpolation)
        self.density_interpolation = density_interpolation
        cfg = solver_config if solver_config is not None else \
            normalize_linear_solver_config(solver_option)
        if cfg.solver not in _SOLVER_NAMES:
            raise ValueError(f"Unknown solver: {cfg.solver}")
        if cfg.precision not in ("mixed", "fp64"):
            raise ValueError("precision must be 'mixed' or 'fp64'")
        pcs = {"auto": _lib.PC_AUTO, "mg": _lib.PC_MG, "jacobi": _lib.PC_JACOBI}
        if cfg.preconditioner not in pcs:
            raise ValueError("preconditioner must be 'auto', 'mg' or 'jacobi'")
        self.solver_config = cfg
        self.solver_option = cfg.solver
        dev = _dev.device()
        self.device = dev
        mesh = task.mesh
        fixed = torch.as_tensor(task.dof_fixed_mask(), device=dev)
        h = C.c_void_p()
        nu = float(getattr(task, "nu", 0.0))
        if mesh.kind == "hex8":
            nx, ny, nz = mesh.grid
            hx, hy, hz = mesh.h
            _lib.call("skt_hex8_create", self.physics, nx, ny, nz, hx, hy, hz, nu,
                      int(quad_order), _dev.ptr(fixed), C.byref(h))
        else:
            coords = _dev.to_dev(mesh.p)
            conn = _dev.to_dev(mesh.t, torch.int32)
            _lib.call("skt_tet4_create", self.physics, mesh.nvertices, mesh.nelements,
                      _dev.ptr(coords), _dev.ptr(conn), nu, _dev.ptr(fixed), C.byref(h))
            grid = getattr(mesh, "cell_grid", None)
            if grid is not None:
                # a grid of cells, six tets each: the multigrid can coarsen it
                _lib.call("skt_tet4_set_grid", h, *grid)
        self._op = h
        ndof = C.c_int64()
        nel = C.c_int64()
        ncomp = C.c_int()
        _lib.call("skt_operator_info", h, C.byref(ndof), C.byref(nel), C.byref(ncomp))
        self.ndof, self.n_el, self.ncomp = ndof.value, nel.value, ncomp.value
        self.n_tasks = task.n_tasks
        s = C.c_void_p()
        _lib.call("skt_solver_create", h, self.n_tasks, C.byref(s))
        self._solver = s
        _lib.call("skt_solver_configure", s, pcs[cfg.preconditioner], int(cfg.mg_smooth_degree),
                  int(cfg.mg_max_coarse))
        self.E64 = torch.empty(self.n_el, dtype=torch.float64, device=dev)
        self.E32 = torch.empty(self.n_el, dtype=torch.float32, device=dev)
        self.F0 = self._build_rhs().contiguous()        # enforced RHS, (ndof, k)
        self.F = self.F0.clone()
        self._force_scale = 1.0
        self.last_solve = {}
        self.solve_history = []

    def __del__(self):
        try:
            if getattr(self, "_solver", None):
                _lib.load().skt_solver_destroy(self._solver)
            if getattr(self, "_op", None):
                _lib.load().skt_operator_destroy(self._op)
        except Exception:
            pass

    # --- helpers -------------------------------------------------------
    def _moduli(self, rho, p):
        emin, emax = self.E_min, self.E_max
        r = rho
        _lib.call("skt_interpolate", self.kind, self.n_el, _dev.ptr(r), emax, emin, float(p),
                  _dev.ptr(self.E64), _lib.F64, _dev.stream())
        if self.solver_config.precision == "mixed":
            _lib.call("skt_interpolate", self.kind, self.n_el, _dev.ptr(r), emax, emin,
                      float(p), _dev.ptr(self.E32), _lib.F32, _dev.stream())

    def _maxiter(self):
        m = self.solver_config.maxiter
        return int(m) if m is not None else 200000

    def solve(self, U: torch.Tensor):
        """Solve K(E64) U = F in place (U is the warm start)."""
        cfg = self.solver_config
        if not cfg.warm_start:
            U.zero_()
        mode = _lib.F32 if cfg.precision == "mixed" else _lib.F64
        iters = (C.c_int64 * 2)()
        relres = (C.c_double * self.n_tasks)()
        st = _lib.load().skt_solver_solve(
            self._solver, mode, self.n_tasks, _dev.ptr(self.E64),
            _dev.ptr(self.E32) if mode == _lib.F32 else None, _dev.ptr(self.F), _dev.ptr(U),
            float(cfg.rtol), self._maxiter(), float(cfg.inner_rtol), iters, relres,
            _dev.stream())
        pc, lev = C.c_int(), C.c_int()
        _lib.call("skt_solver_info", self._solver, C.byref(pc), C.byref(lev))
        self.last_solve = dict(iters=int(iters[0]), updates=int(iters[1]),
                               preconditioner=("mg" if pc.value == _lib.PC_MG else "jacobi"),
                               levels=int(lev.value),
                               relres=list(relres), converged=(st == _lib.SKT_OK))
        self.solve_history.append(self.last_solve)
        if st == _lib.SKT_ENOCONV and cfg.on_noconv == "warn":
            logger.warning("state solve did not converge: %s", _lib.last_error())
        else:
            _lib.check(st, "skt_solver_solve")

    def predict_warm_start(self, U: torch.Tensor, hist: dict):
        """Before a design loop's state solve: U holds the last solution;
        replace it by ``U + theta (U - U_prev)`` (theta =
        ``solver_config.warm_start_extrapolation``), U_prev the solution
        before it, kept in ``hist`` (the caller's, one per state).  The first
        two solves keep U as it is (the solution before the first one is only
        the initial guess).  The solve's tolerance fixes the result; a better
        start saves iterations."""
        cfg = self.solver_config
        theta = float(cfg.warm_start_extrapolation)
        if not cfg.warm_start or theta == 0.0:
            return
        if U.dtype != torch.float64:
            raise ValueError("predict_warm_start expects the float64 state U")
        S = _dev.stream()
        n = U.numel()
        prev = hist.get("prev")
        if prev is None or prev.shape != U.shape:
            hist["prev"] = U.clone()
            hist["spare"] = torch.empty_like(U)
            hist["count"] = 1
            return
        spare = hist["spare"]
        # spare <- U (the last solution), then U <- (1 + theta) U - theta U_prev
        _lib.call("skt_ewise", 0, n, 1.0, _dev.ptr(U), 0.0, None, _dev.ptr(spare), S)
        if hist["count"] >= 2:
            _lib.call("skt_ewise", 0, n, 1.0 + theta, _dev.ptr(spare), -theta, _dev.ptr(prev),
                      _dev.ptr(U), S)
        hist["prev"], hist["spare"] = spare, prev
        hist["count"] += 1

    def graph_stats(self) -> dict:
        """Replays and captures of the solver's PCG-chunk hipGraphs."""
        a, b = C.c_int64(), C.c_int64()
        _lib.call("skt_solver_graph_stats", self._solver, C.byref(a), C.byref(b))
        return dict(launches=int(a.value), captures=int(b.value))

    def _set_force_scale(self, force_scale):
        if force_scale != self._force_scale:
            torch.mul(self.F0, float(force_scale), out=self.F)
            self._force_scale = force_scale

    def energy(self, rho, p, U):
        """(n_el, k) float64 element energies on the device."""
        self._moduli(rho, p)
        out = torch.empty((self.n_el, self.n_tasks), dtype=torch.float64, device=self.device)
        _lib.call("skt_operator_energy", self._op, _lib.F64, self.n_tasks, _dev.ptr(self.E64),
                  _dev.ptr(U), _dev.ptr(out), _dev.stream())
        return out

    def col_reduce(self, op, X, Y=None):
        out = torch.empty(self.n_tasks, dtype=torch.float64, device=self.device)
        _lib.call("skt_reduce_cols", op, X.shape[0], self.n_tasks, _dev.ptr(X),
                  _dev.ptr(Y) if Y is not None else None, _dev.ptr(out), _dev.stream())
        return out

    def apply(self, X, E=None, dtype=torch.float64):
        """Y = K_e(E) X on the device (E defaults to the last moduli)."""
        Y = torch.empty_like(X)
        code = _lib.F64 if dtype == torch.float64 else _lib.F32
        if E is None:
            E = self.E64 if code == _lib.F64 else self.E32
        ncols = 1 if X.dim() == 1 else X.shape[1]
        _lib.call("skt_operator_apply", self._op, code, ncols, _dev.ptr(E), _dev.ptr(X),
                  _dev.ptr(Y), _dev.stream())
        return Y

    # --- host/device dispatch -----------------------------------------
    def _objectives_dev(self, rho, p, U, force_scale):
        raise NotImplementedError

    def objectives_multi_load(self, rho, p, u_dofs, timer=None, force_scale=1.0):
        if _dev.is_host(rho) or _dev.is_host(u_dofs):
            r = _dev.to_dev(rho)
            U = _dev.to_dev(u_dofs).reshape(self.ndof, self.n_tasks).contiguous()
            out = self._objectives_dev(r, p, U, force_scale)
            u_dofs[...] = U.cpu().numpy().reshape(np.shape(u_dofs))
            return out.cpu().numpy()
        return self._objectives_dev(rho, p, u_dofs, force_scale)

    def energy_multi_load(self, rho, p, u_dofs):
        if _dev.is_host(rho) or _dev.is_host(u_dofs):
            r = _dev.to_dev(rho)
            U = _dev.to_dev(u_dofs).reshape(self.ndof, self.n_tasks).contiguous()
            return self.energy(r, p, U).cpu().numpy()
        return self.energy(rho, p, u_dofs)


class FEM_SimpLinearElasticity(_GpuFEM):
    """GPU counterpart of ``fea/solver_elastic.py:540-676``."""
    physics = _lib.ELASTIC

    def __init__(self, task, E_min_coeff, density_interpolation=simp_interpolation,
                 solver_config=None, solver_option="spsolve", petsc_options=None,
                 quad_order=2):
        self.E_max = task.E * 1.0
        self.E_min = task.E * E_min_coeff
        super().__init__(task, E_min_coeff, density_interpolation, solver_config,
                         solver_option, quad_order)

    def _build_rhs(self):
        F = np.column_stack([np.asarray(f, dtype=np.float64) for f in self.task.neumann_linear])
        F[self.task.dirichlet_dofs, :] = 0.0          # enforce: b_D = 0
        return _dev.to_dev(F)

    def _objectives_dev(self, rho, p, U, force_scale=1.0):
        """compliance_k = F_e[:, k] . u_k (``solver_elastic.py:236,465``)."""
        self._set_force_scale(force_scale)
        self._moduli(rho, p)
        self.solve(U)
        return self.col_reduce(_lib.RED_DOT, self.F, U)


class FEM_SimpLinearHeatConduction(_GpuFEM):
    """GPU counterpart of ``fea/solver_heat.py:628-980``: objectives
    ``'compliance'`` and ``'averaged_temp'``, with or without Robin facets,
    and ``'heat_exchange'`` (Robin facets required).

    With Robin data (``task.robin_coefficient`` set) every solve uses
    ``K = K_cond(rho) + B_facets + B_virtual(rho)`` and the load
    ``emit + L_facets + L_virtual(rho)`` (``:724-760``, ``get_robin_virtual``
    :552-573).  On structured hex grids the Robin matrices are a 27-point
    stencil attached to the operator, and the volume integrals use ``npts``
    Gauss points per axis (``npts = 2`` integrates the basis products exactly
    on the reference's ``intorder`` 2 and 3); the heat-exchange functionals
    use ``npts_cell`` points per axis: the reference evaluates them on
    ``skfem.CellBasis`` with its default rule, 4 Gauss points per axis for
    trilinear hexahedra.  On tet4 meshes they are 4 x 4 element matrices
    attached to the operator, the volume forms use the tetrahedral rule of
    order ``tet_order`` (1: centroid, 2: the 4-point rule of degree 2, which a
    P1 basis of ``intorder`` 2 uses), and the heat-exchange functionals are
    exact."""
    physics = _lib.HEAT

    def __init__(self, task, E_min_coeff, density_interpolation=simp_interpolation,
                 solver_config=None, solver_option="spsolve", q=4, quad_order=2, npts=2,
                 npts_cell=4, tet_order=2):
        if getattr(task, "objective", "compliance") not in ("compliance", "averaged_temp",
                                                             "heat_exchange"):
            raise ValueError(f"Unknown objective: {task.objective}")
        self.npts_cell = int(npts_cell)
        self.k_max = task.k * 1.0
        self.k_min = task.k * E_min_coeff
        self.E_max, self.E_min = self.k_max, self.k_min
        self.q = q
        self.npts = int(npts)
        self.objective = getattr(task, "objective", "compliance")
        self.robin = getattr(task, "robin_coefficient", None) is not None
        self.tet = task.mesh.kind == "tet4"
        if tet_order not in (1, 2):
            raise ValueError("tet_order must be 1 or 2")
        self.tet_order = int(tet_order)
        if self.robin and not self.tet and getattr(task.mesh, "grid", None) is None:
            raise NotImplementedError("Robin boundaries are built for structured hex8 grids "
                                      "and tet4 meshes")
        super().__init__(task, E_min_coeff, density_interpolation, solver_config,
                         solver_option, quad_order)
        self.lam = None
        self._lam_warm = None
        self._D = torch.as_tensor(np.asarray(task.dirichlet_dofs, dtype=np.int64),
                                  device=self.device)
        if self.robin:
            dev, nn = self.device, task.mesh.nvertices
            z = lambda *sh: torch.zeros(sh, dtype=torch.float64, device=dev)  # noqa: E731
            self._L_facet = _dev.to_dev(task.robin_load)
            self._emit = _dev.to_dev(task.emit)
            self._rn, self._cnt = z(nn), z(nn)
            self._L_virt, self._emit_tot = z(nn), z(nn)
            if self.tet:
                ne = task.mesh.nelements
                self._st_facet = _dev.to_dev(task.robin_elem_mat)
                self._st_virt, self._st_tot = z(10, ne), z(10, ne)
                self._work = z(4 * ne + nn)
            else:
                self._st_facet = _dev.to_dev(task.robin_stencil)
                self._st_virt, self._st_tot = z(27, nn), z(27, nn)
                self._work = z(nn)

    def _build_rhs(self):
        cols = []
        for val in self.task.dirichlet_values:
            f = np.array(self.task.emit, dtype=np.float64)
            f[self.task.dirichlet_dofs] = val            # enforce: b_D = x_D
            cols.append(f)
        return _dev.to_dev(np.column_stack(cols))

    def _enforced(self, base):
        """(ndof, k): base in every column, b_D = the column's Dirichlet value."""
        F = base[:, None].repeat(1, self.n_tasks).contiguous()
        for c, val in enumerate(self.task.dirichlet_values):
            F[:, c].index_fill_(0, self._D, float(val))
        return F

    def _robin_update(self, rho, p):
        """Virtual Robin stencil and load for the current densities
        (``solver_heat.py:724-760``), attached to the operator; the enforced
        right-hand side is rebuilt from the total load."""
        t = self.task
        S = _dev.stream()
        self._nodal_pairing(rho)
        h, T_env = float(t.robin_coefficient), float(t.robin_bc_value)
        if self.tet:
            _lib.call("skt_tet4_robin_virtual", self._op, self.tet_order, _dev.ptr(self._rn),
                      float(p), float(self.q), h, T_env, _dev.ptr(self._st_virt),
                      _dev.ptr(self._L_virt), _dev.ptr(self._work), S)
        else:
            nx, ny, nz = t.mesh.grid
            _lib.call("skt_hex8_robin_virtual", nx, ny, nz, *t.mesh.h, self.npts,
                      _dev.ptr(self._rn), float(p), float(self.q), h, T_env,
                      _dev.ptr(self._st_virt), _dev.ptr(self._L_virt), S)
        _lib.call("skt_ewise", 0, self._st_tot.numel(), 1.0, _dev.ptr(self._st_facet), 1.0,
                  _dev.ptr(self._st_virt), _dev.ptr(self._st_tot), S)
        _lib.call("skt_tet4_set_extra" if self.tet else "skt_operator_set_extra", self._op,
                  _dev.ptr(self._st_tot), S)
        # emit + facet loads + virtual load, in the reference's list order
        _lib.call("skt_ewise", 0, self._emit.numel(), 1.0, _dev.ptr(self._emit), 1.0,
                  _dev.ptr(self._L_facet), _dev.ptr(self._emit_tot), S)
        _lib.call("skt_ewise", 0, self._emit.numel(), 1.0, _dev.ptr(self._emit_tot), 1.0,
                  _dev.ptr(self._L_virt), _dev.ptr(self._emit_tot), S)
        self.F = self._enforced(self._emit_tot)

    def _nodal_pairing(self, rho):
        """Nodal densities (and element counts) with the reference's add.at
        pairing (``solver_heat.py:738-745``) into ``self._rn``, ``self._cnt``."""
        S = _dev.stream()
        if self.tet:
            _lib.call("skt_tet4_nodal_average_pairing", self._op, _dev.ptr(rho),
                      _dev.ptr(self._rn), _dev.ptr(self._cnt), S)
        else:
            nx, ny, nz = self.task.mesh.grid
            _lib.call("skt_hex8_nodal_average_pairing", nx, ny, nz, _dev.ptr(rho),
                      _dev.ptr(self._rn), _dev.ptr(self._cnt), S)

    def _solve_rhs(self, F, U):
        saved = self.F
        self.F = F
        try:
            self.solve(U)
        finally:
            self.F = saved

    def _objectives_dev(self, rho, p, U, force_scale=1.0):
        """``compliance``: ``J_k = T_k^T K T_k`` with the unenforced total K
        (``solver_heat.py:785-789``): twice the conduction energies plus the
        Robin stencil's ``T^T S T``.  ``averaged_temp``: ``J_k = sum(T_k)``
        and the adjoint ``K lam = 1`` enforced with the state's Dirichlet
        values (``:869-882``)."""
        self._moduli(rho, p)
        if self.robin:
            self._robin_update(rho, p)
        self.solve(U)
        if self.objective == "averaged_temp":
            J = self.col_reduce(_lib.RED_SUM, U)
            self.lam = self._adjoint_avg()
            return J
        if self.objective == "heat_exchange":
            return self._heat_exchange(rho, p, U)
        en = torch.empty((self.n_el, self.n_tasks), dtype=torch.float64, device=self.device)
        _lib.call("skt_operator_energy", self._op, _lib.F64, self.n_tasks, _dev.ptr(self.E64),
                  _dev.ptr(U), _dev.ptr(en), _dev.stream())
        J = 2.0 * self.col_reduce(_lib.RED_SUM, en)
        self.lam = None
        if self.robin and self.tet:
            q = torch.empty((self.n_el, self.n_tasks), dtype=torch.float64, device=self.device)
            _lib.call("skt_tet4_extra_quadform", self._op, self.n_tasks, _dev.ptr(U),
                      _dev.ptr(q), _dev.stream())
            J = J + self.col_reduce(_lib.RED_SUM, q)
        elif self.robin:
            nx, ny, nz = self.task.mesh.grid
            SU = torch.empty_like(U)
            _lib.call("skt_stencil27_apply", nx, ny, nz, self.n_tasks, _dev.ptr(self._st_tot),
                      _dev.ptr(U), _dev.ptr(SU), _dev.stream())
            J = J + self.col_reduce(_lib.RED_DOT, U, SU)
        return J

    def _adjoint_avg(self):
        """The averaged-temperature adjoint ``K lam = 1``, enforced with the
        state's Dirichlet values (``solver_heat.py:869-882``)."""
        if self._lam_warm is None:
            self._lam_warm = torch.zeros((self.ndof, self.n_tasks), dtype=torch.float64,
                                         device=self.device)
        ones = torch.ones(self.ndof, dtype=torch.float64, device=self.device)
        self._solve_rhs(self._enforced(ones), self._lam_warm)
        return self._lam_warm

    def _heat_exchange(self, rho, p, U):
        """``objective='heat_exchange'`` (``solver_heat.py:791-869``):
        ``J_k = -T_env h int (T_k - T_env)|grad rho_h| / int |grad rho_h|``
        on the nodal densities of the reference's pairing, the adjoint
        ``K lam = rhs / J_den`` enforced with the state's Dirichlet values,
        and the ``avg_temp_weight`` blend with the averaged temperature."""
        t = self.task
        if not self.robin:
            raise RuntimeError("heat_exchange objective requires Robin boundary data.")
        S = _dev.stream()
        k = self.n_tasks
        dev = self.device
        self._nodal_pairing(rho)
        den = torch.empty(self.n_el, dtype=torch.float64, device=dev)
        num = torch.empty((self.n_el, k), dtype=torch.float64, device=dev)
        rhs = torch.empty(self.ndof, dtype=torch.float64, device=dev)
        h, T_env = float(t.robin_coefficient), float(t.robin_bc_value)
        Uc = U.contiguous()
        if self.tet:
            _lib.call("skt_tet4_heat_exchange", self._op, _dev.ptr(self._rn), k, _dev.ptr(Uc),
                      float(p), float(self.q), h, T_env, _dev.ptr(den), _dev.ptr(num),
                      _dev.ptr(rhs), _dev.ptr(self._work), S)
        else:
            nx, ny, nz = t.mesh.grid
            _lib.call("skt_hex8_heat_exchange", nx, ny, nz, *t.mesh.h, self.npts_cell,
                      _dev.ptr(self._rn), k, _dev.ptr(Uc), float(p), float(self.q), h, T_env,
                      _dev.ptr(den), _dev.ptr(num), _dev.ptr(rhs), S)
        s = torch.empty(1, dtype=torch.float64, device=dev)
        _lib.call("skt_reduce", _lib.RED_SUM, self.n_el, _dev.ptr(den), None, _dev.ptr(s), S)
        J_den = float(s.item())
        if J_den <= 1e-16:
            self.lam = torch.zeros_like(U)
            return torch.zeros(k, dtype=torch.float64, device=dev)
        J = (-T_env * h) * self.col_reduce(_lib.RED_SUM, num) / J_den
        lam = torch.zeros_like(U)
        self._solve_rhs(self._enforced(rhs / J_den), lam)
        w = t.avg_temp_weight
        if w != 0.0:
            avg = self.col_reduce(_lib.RED_SUM, U)
            lam_avg = self._adjoint_avg()
            if k > 1:
                hs = max(float(J.abs().mean()), 1e-12)
                avs = max(float(avg.abs().mean()), 1e-12)
                J = J / hs + w * (avg / avs)
                lam = lam / hs + (w / avs) * lam_avg
            else:
                J = J + w * avg
                lam = lam + w * lam_avg
        self.lam = lam
        return J

    def _cross_energy(self, A, B):
        """(n_el, k) ``int_e grad A . grad B`` with unit conductivity."""
        ones = torch.ones(self.n_el, dtype=torch.float64, device=self.device)
        def e(X):
            out = torch.empty((self.n_el, self.n_tasks), dtype=torch.float64, device=self.device)
            _lib.call("skt_operator_energy", self._op, _lib.F64, self.n_tasks, _dev.ptr(ones),
                      _dev.ptr(X.contiguous()), _dev.ptr(out), _dev.stream())
            return out
        return e(A + B) - e(A) - e(B)

    def energy(self, rho, p, U):
        """``energy_multi_load`` (``solver_heat.py:897-925``): conduction
        energies for ``compliance``; ``int grad T . grad lam`` for
        ``averaged_temp``."""
        if self.objective in ("averaged_temp", "heat_exchange"):
            if self.lam is None:
                raise RuntimeError("adjoint field λ_all is not computed.")
            return self._cross_energy(U, self.lam)
        return super().energy(rho, p, U)

    def sensitivity_energy(self, rho, p, U):
        """The energies the sensitivity uses: always the conduction ones
        (``solver_heat.py:928-950``)."""
        return _GpuFEM.energy(self, rho, p, U)

    def extra_sensitivity(self, rho, p, U):
        """(n_el, k) explicit Robin term of the compliance sensitivity
        (``solver_heat.py:955-980``), or None."""
        if not (self.robin and self.objective == "compliance"):
            return None
        t = self.task
        S = _dev.stream()
        self._nodal_pairing(rho)
        out = torch.empty((self.n_el, self.n_tasks), dtype=torch.float64, device=self.device)
        h, T_env = float(t.robin_coefficient), float(t.robin_bc_value)
        if self.tet:
            _lib.call("skt_tet4_robin_grad", self._op, self.tet_order, _dev.ptr(self._rn),
                      _dev.ptr(self._cnt), self.n_tasks, _dev.ptr(U.contiguous()), float(p),
                      float(self.q), h, T_env, _dev.ptr(self._work), _dev.ptr(out), S)
        else:
            nx, ny, nz = t.mesh.grid
            _lib.call("skt_hex8_robin_grad", nx, ny, nz, *t.mesh.h, self.npts,
                      _dev.ptr(self._rn), _dev.ptr(self._cnt), self.n_tasks,
                      _dev.ptr(U.contiguous()), float(p), float(self.q), h, T_env,
                      _dev.ptr(self._work), _dev.ptr(out), S)
        return out

    def compliance_sensitivity_multi_load(self, rho, p, u_dofs):
        """SIMP sensitivity of the conduction energies (``solver_heat.py:928-950``)
        plus, for the Robin compliance, the explicit Robin term (``:955-980``)."""
        host = _dev.is_host(rho) or _dev.is_host(u_dofs)
        r = _dev.to_dev(rho)
        U = _dev.to_dev(u_dofs).reshape(self.ndof, self.n_tasks).contiguous()
        en = self.sensitivity_energy(r, p, U)
        extra = self.extra_sensitivity(r, p, U)
        out = torch.empty_like(en)
        for k in range(self.n_tasks):
            col = en[:, k].contiguous()
            tmp = torch.empty_like(col)
            _lib.call("skt_dC_drho", _lib.SIMP, self.n_el, _dev.ptr(r), _dev.ptr(col), 1,
                      self.k_max, self.k_min, float(p), _dev.ptr(tmp), _dev.stream())
            out[:, k] = tmp
        if extra is not None:
            out += extra
        return out.cpu().numpy() if host else out
