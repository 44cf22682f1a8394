"""Density-method iteration driver (oracle).

Restates ``DensityMethod._optimize_impl`` (``core/optimizers/common_density.py:966-1243``),
``initialize_density`` (``:711-745``) and the partial-run bookkeeping of
``optimize_steps`` (``:849-859``, ``:997-1008``, ``:1240-1242``).  Output side
effects (VTU, images, npz, plots) are left out; the per-iteration records
the reference keeps (``:1136-1156``) are returned in ``history``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import fem as fem_mod
from . import filter as filt_mod
from . import helmholtz
from . import logmoc as logmoc_mod
from . import oc as oc_mod
from . import scheduler as sch
from . import transforms as tr


@dataclass
class Sched:
    kind: str = "Constant"
    init: float | None = None
    target: float | None = None
    num_steps: int | None = None
    curvature: float | None = None

    def value(self, it, iters_max):
        return sch.value(self.kind, it, iters_max, self.init, self.target,
                         self.num_steps, self.curvature)


@dataclass
class Config:
    """Field names and defaults of ``OC_Config`` / ``LogMOC_Config``
    (``common_density.py:207-389``, ``oc.py:98-108``, ``logmoc.py:13-33``)."""
    method: str = "OC"                      # "OC" or "LogMOC"
    interpolation: str = "SIMP"
    max_iters: int = 200
    beta_eta: float = 0.5
    eta: Sched = field(default_factory=lambda: Sched("Constant", 0.5, 0.5))
    p: Sched = field(default_factory=lambda: Sched("Step", 1.0, 3.0, 3))
    vol_frac: Sched = field(default_factory=lambda: Sched("Constant", 0.8, 0.8))
    beta: Sched = field(default_factory=lambda: Sched("StepAccelerating", 1.0, 2.0, 3, 2.0))
    neumann_scale: Sched = field(default_factory=lambda: Sched("ConstantOne", 1.0, 1.0))
    filter_type: str = "spacial"            # "spacial" or "helmholtz"
    filter_radius: Sched = field(default_factory=lambda: Sched("Constant", 0.01, 0.01))
    move_limit: Sched = field(default_factory=lambda: Sched("SawtoothDecay", 0.3, 0.1, 6))
    percentile: Sched = field(default_factory=lambda: Sched("None"))
    E_min_coeff: float = 1e-3
    rho_min: float = 1e-2
    rho_max: float = 1.0
    design_dirichlet: bool = False
    sensitivity_filter: bool = False
    lambda_lower: float = 1e-7
    lambda_upper: float = 1e7
    scaling_rate_min: float = 0.7
    scaling_rate_max: float = 1.3
    sensitivity_scale_floor: float = 1e-12
    # LogMOC
    mu_p: float = 5.0
    augmented_lagrangian_mu: float = 0.0
    lambda_v: float = 0.1
    lambda_decay: float = 0.90
    lagrangian_clip: float = 1.0
    lagrangian_percentile: float = 95.0
    lagrangian_scale_floor: float = 1e-8
    normalize_volume_chain: bool = False
    volume_chain_percentile: float = 95.0
    volume_chain_scale_floor: float = 1e-8
    volume_chain_gain: float = 1.0
    volume_chain_gain_under: float | None = None
    dual_update: str = "ema"
    filter_lagrangian: bool = False
    center_lagrangian: bool = False
    center_objective: bool = False
    solver: str = "spsolve"
    solver_rtol: float = 1e-12


class DensityLoop:
    def __init__(self, cfg: Config, task, fem=None):
        self.cfg = cfg
        self.task = task
        if fem is not None:
            self.fem = fem
        elif task.physics == "elastic":
            self.fem = fem_mod.ElasticFEM(task, cfg.E_min_coeff,
                                          cfg.interpolation, cfg.solver,
                                          cfg.solver_rtol)
        else:
            self.fem = fem_mod.HeatFEM(task, cfg.E_min_coeff,
                                       cfg.interpolation, cfg.solver,
                                       cfg.solver_rtol)
        self.design = task.design_elements_with_dirichlet if cfg.design_dirichlet \
            else task.design_elements
        design_mask = np.isin(np.arange(task.n_elements), self.design)
        vol = task.elements_volume[self.design]
        self.vol_design = vol
        self.vol_design_sum = np.sum(vol)
        self.dV_drho_design = vol / self.vol_design_sum
        r0 = cfg.filter_radius.init if isinstance(cfg.filter_radius.num_steps, (int, float)) \
            else cfg.filter_radius.target
        if cfg.filter_type == "helmholtz":
            self.filter = helmholtz.HelmholtzFilterNodal(task.mesh, task.elements_volume,
                                                         cfg.filter_radius.init,
                                                         design_mask)
        else:
            self.filter = filt_mod.SpacialFilter(task.centers, cfg.filter_radius.init,
                                                 design_mask)
        self.filter.update_radius(r0)
        self.update = oc_mod.OCUpdate(cfg) if cfg.method == "OC" \
            else logmoc_mod.LogMOCUpdate(cfg)
        self.rho = self.initialize_density()
        self.u_dofs = np.zeros((task.ndof, task.n_tasks))
        self.iter_next = 1
        self.iter_end = cfg.max_iters + 1
        self.history = []

    def initialize_density(self):
        cfg, task = self.cfg, self.task
        v = cfg.vol_frac.init if cfg.vol_frac.init is not None else cfg.vol_frac.target
        rho = np.zeros(task.n_elements) + v
        np.clip(rho, cfg.rho_min, cfg.rho_max, out=rho)
        if cfg.design_dirichlet:
            rho[task.neumann_elements] = 1.0
        else:
            rho[task.dirichlet_neumann_elements] = 1.0
        rho[task.fixed_elements] = 1.0
        return rho

    def optimize_steps(self, n: int | None = None):
        cfg, task = self.cfg, self.task
        start = self.iter_next
        limit = self.iter_end if n is None else min(start + n, self.iter_end)
        for it in range(start, limit):
            self.step(it)
        self.iter_next = max(start, limit)

    def optimize(self):
        self.optimize_steps(None)

    def step(self, it: int):
        cfg, task = self.cfg, self.task
        M = cfg.max_iters
        neumann_scale = cfg.neumann_scale.value(it, M)
        p = cfg.p.value(it, M)
        vol_frac = cfg.vol_frac.value(it, M)
        beta = cfg.beta.value(it, M)
        move_limit = cfg.move_limit.value(it, M)
        eta = cfg.eta.value(it, M)
        percentile = cfg.percentile.value(it, M)
        radius = cfg.filter_radius.value(it, M)
        if radius != self.filter.radius:
            self.filter.update_radius(radius)
        rho = self.rho
        rho_prev = rho.copy()
        rho_filtered = self.filter.forward(rho)
        rho_projected = tr.heaviside_projection(rho_filtered, beta, cfg.beta_eta)
        comp = self.fem.objectives_multi_load(rho_projected, p, self.u_dofs,
                                              force_scale=neumann_scale)
        compliance_avg = comp.mean()
        energy = self.fem.energy_multi_load(rho_projected, p, self.u_dofs)
        custom = None
        if hasattr(self.fem, "compliance_sensitivity_multi_load"):
            custom = self.fem.compliance_sensitivity_multi_load(
                rho_projected, p, self.u_dofs)
        dC_full = np.zeros(task.n_elements)
        dCfn = tr.dC_drho_simp if cfg.interpolation == "SIMP" else tr.dC_drho_ramp
        u_max = []
        dH = tr.heaviside_projection_derivative(rho_filtered, beta, cfg.beta_eta)
        for k in range(task.n_tasks):
            u_max.append(np.abs(self.u_dofs[:, k]).max())
            if custom is not None:
                dC = custom[:, k]
            else:
                dC = dCfn(rho_projected, energy[:, k], task.E0,
                          task.E0 * cfg.E_min_coeff, p)
            dC_full += self.filter.gradient(dC * dH)
        dC_full /= task.n_tasks
        if cfg.sensitivity_filter:
            dC_full = self.filter.forward(dC_full)
        dC_design = dC_full[self.design].copy()
        rho_design = rho[self.design].copy()
        if cfg.method == "OC":
            self.update(it, rho_design, dC_design, move_limit, eta, beta,
                        percentile, self.vol_design, self.vol_design_sum,
                        vol_frac, rho, self.design, self.filter,
                        self.dV_drho_design)
        else:
            self.update(it, rho_design, dC_design, rho_projected, rho_filtered,
                        move_limit, eta, beta, percentile, self.vol_design,
                        self.vol_design_sum, vol_frac, self.design, self.filter)
        rho[self.design] = rho_design
        if cfg.design_dirichlet:
            rho[task.neumann_elements] = 1.0
        else:
            rho[task.dirichlet_neumann_elements] = 1.0
        rec = dict(iter=it, compliance=float(compliance_avg),
                   compliance_each=comp.copy(),
                   rho_change_max=float(np.max(np.abs(rho[self.design] - rho_prev[self.design]))),
                   u_max=np.array(u_max), energy_mean=energy.mean(axis=1),
                   kkt_residual=self.update.kkt_residual, p=p, beta=beta,
                   vol_frac=vol_frac, move_limit=move_limit, eta=eta,
                   rho_projected=rho_projected, dC_full=dC_full.copy())
        rec.update(self.update.last)
        self.history.append(rec)
        return rec
