"""Device-resident density-method driver.

Restates ``DensityMethod`` (``core/optimizers/common_density.py:468-1265``)
with every per-element and per-DOF array kept in HBM as float64 torch
tensors and every numeric step done by the HIP library:

  filter forward -> Heaviside projection -> FEA solve (objectives) ->
  element energies -> fused dC/drho * H'(rho_filtered) over load cases ->
  filter adjoint -> subclass ``rho_update`` -> scatter into rho.

Config field names and defaults are the reference's (``:62-389``); the GPU
solve adds ``precision`` / ``solver_rtol`` / ``inner_rtol`` knobs.  Output
side effects of the reference (VTU, images, plots) are not reproduced;
the per-iteration records (``:1136-1156``) are kept in ``self.recorder`` and
``data/{iter:06d}-rho.npz`` checkpoints are written when
``write_checkpoints`` is set (``:1186-1233``).
"""
from __future__ import annotations

import inspect
import json
import logging
import os
from dataclasses import asdict, dataclass, field
from typing import Literal

import numpy as np
import torch

from ... import _dev, _lib
from ... import fea, filters
from ...fea import composer
from ...mesh import LinearElasticity, LinearHeatConduction
from ...tools import SchedulerConfig, Schedulers
from ...tools.history import HistoryCollection  # noqa: F401
from .. import visualization

logger = logging.getLogger(__name__)


@dataclass
class DensityMethodConfig:
    dst_path: str = "./result/pytests"
    interpolation: Literal["SIMP", "RAMP"] = "SIMP"
    record_times: int = 20
    max_iters: int = 200
    beta_eta: float = 0.50
    eta: object = 0.6
    p: SchedulerConfig = field(default_factory=lambda: SchedulerConfig.step(
        init_value=1.0, target_value=3.0, num_steps=3))
    vol_frac: SchedulerConfig = field(default_factory=lambda: SchedulerConfig.constant(
        target_value=0.8))
    beta: SchedulerConfig = field(default_factory=lambda: SchedulerConfig.step_accelerating(
        init_value=1.0, target_value=2.0, num_steps=3, curvature=2.0))
    neumann_scale: SchedulerConfig = field(default_factory=lambda: SchedulerConfig.constant_one(
        name="neumann_scale"))
    filter_type: Literal["spacial", "helmholtz"] = "helmholtz"
    filter_radius: SchedulerConfig = field(default_factory=lambda: SchedulerConfig.constant(
        target_value=0.01))
    E_min_coeff: float = 1e-3
    rho_min: float = 1e-2
    rho_max: float = 1.0
    restart: bool = False
    restart_from: int = -1
    export_img: bool = False
    export_img_opaque: bool = False
    design_dirichlet: bool = False
    sensitivity_filter: bool = False
    solver_option: str = "spsolve"
    petsc_options: object = None
    scaling: bool = False
    check_convergence: bool = False
    tol_rho_change: float = 2e-1
    tol_kkt_residual: float = 5e-3
    # GPU state solve
    precision: Literal["mixed", "fp64"] = "mixed"
    solver_rtol: float = 1e-8
    inner_rtol: float = 0.0
    preconditioner: str = "auto"
    solver_maxiter: int | None = None
    write_checkpoints: bool = False

    @classmethod
    def from_defaults(cls, **args):
        keys = inspect.signature(cls).parameters.keys()
        return cls(**{k: v for k, v in args.items() if k in keys})

    def __post_init__(self):
        self.solver_config = fea.LinearSolverConfig(
            solver=self.solver_option, rtol=self.solver_rtol, maxiter=self.solver_maxiter,
            precision=self.precision, inner_rtol=self.inner_rtol,
            preconditioner=self.preconditioner)
        if self.solver_option not in fea.solver._SOLVER_NAMES:
            raise ValueError(f"Unknown solver: {self.solver_option}")

    def export(self, path: str):
        """``common_density.py:322-324``: ``cfg.json`` (schedules as dicts)."""
        d = {k: v for k, v in asdict(self).items()}
        os.makedirs(path, exist_ok=True)
        with open(f"{path}/cfg.json", "w") as f:
            json.dump(d, f, indent=2, default=str)

    @classmethod
    def import_from(cls, path: str):
        """``common_density.py:312-319``; schedule dicts become
        ``SchedulerConfig`` again."""
        with open(f"{path}/cfg.json", "r") as f:
            data = json.load(f)
        data.pop("record_timing", None)
        keys = inspect.signature(cls).parameters.keys()
        out = {}
        for k, v in data.items():
            if k not in keys:
                continue
            if isinstance(v, dict) and "scheduler_type" in v:
                v = SchedulerConfig(**v)
            out[k] = v
        return cls(**out)

    def vtu_path(self, iter_num: int):
        """``common_density.py:338-339``."""
        return f"{self.dst_path}/mesh_rho/info_mesh-{iter_num:08d}.vtu"


@dataclass
class DensityMethod_OC_Config(DensityMethodConfig):
    lambda_lower: float = 1e-7
    lambda_upper: float = 1e7
    percentile: SchedulerConfig = field(default_factory=SchedulerConfig.none)
    move_limit: SchedulerConfig = field(default_factory=lambda: SchedulerConfig.sawtooth_decay(
        "move_limit", 0.3, 0.1, 3))


def interpolation_funcs(cfg):
    if cfg.interpolation == "SIMP":
        return composer.simp_interpolation, _lib.SIMP
    if cfg.interpolation == "RAMP":
        return composer.ramp_interpolation, _lib.RAMP
    raise ValueError("Interpolation method must be SIMP or RAMP.")


class DensityMethod:
    def __init__(self, cfg: DensityMethodConfig, tsk):
        self.cfg = cfg
        self.tsk = tsk
        if cfg.scaling:
            self.scale()
        os.makedirs(cfg.dst_path, exist_ok=True)
        if cfg.design_dirichlet is False:
            tsk.exlude_dirichlet_from_design()
        design = np.asarray(tsk.design_elements)
        if design.size > 1 and not np.all(np.diff(design) > 0):
            raise ValueError("design_elements must be sorted and unique")
        self.interp, self.kind = interpolation_funcs(cfg)
        if isinstance(tsk, LinearElasticity):
            self.fem = fea.FEM_SimpLinearElasticity(
                tsk, cfg.E_min_coeff, density_interpolation=self.interp,
                solver_config=cfg.solver_config)
        elif isinstance(tsk, LinearHeatConduction):
            self.fem = fea.FEM_SimpLinearHeatConduction(
                tsk, cfg.E_min_coeff, density_interpolation=self.interp,
                solver_config=cfg.solver_config)
        else:
            raise NotImplementedError("")
        self.schedulers = Schedulers(cfg.dst_path)
        self._schedulers_initialized = False
        cfg.export(cfg.dst_path)
        self.recorder = self.add_recorder(tsk)
        self.dev = _dev.device()
        vol = tsk.elements_volume[tsk.design_elements]
        self._vol_design_sum = float(np.sum(vol))
        self._vol_design = _dev.to_dev(vol)
        self._dV_drho_design = _dev.to_dev(vol / self._vol_design_sum)
        self._design_ids = torch.as_tensor(np.asarray(tsk.design_elements, dtype=np.int32),
                                           device=self.dev)
        self.kkt_residual = None
        self._rec_host = None      # pinned host copy of an iteration's records
        self._rec_pending = None   # (copy, event, context) not yet fed to the recorder
        self._state = None
        self._iter_next = None
        self._iter_end = None
        self._completed = False
        self.filter = None
        self.timings = []
        # device phases of an iteration captured as CUDA graphs (OC with the
        # sparse-matrix filter); SKT_STEP_GRAPH=0 runs them eagerly
        self._step_graphs = os.environ.get("SKT_STEP_GRAPH", "1") != "0"
        self._phases = {}
        self._cap_stream = None

    # --- setup ------------------------------------------------------------
    def add_recorder(self, tsk) -> HistoryCollection:
        """History names and types of ``common_density.py:591-611``."""
        r = HistoryCollection(self.cfg.dst_path)
        r.add("rho_projected", plot_type="min-max-mean-std")
        r.add("energy", plot_type="min-max-mean-std")
        r.add("vol_error")
        r.add("u_max", plot_type="min-max-mean-std" if tsk.n_tasks > 1 else "value")
        r.add(self._objective_history_name(tsk), ylog=self._objective_history_ylog(tsk))
        r.add("scaling_rate", plot_type="min-max-mean-std")
        r.add("neumann_scale")
        r.add("rho_change_max")
        r.add("kkt_residual")
        return r

    @staticmethod
    def _objective_history_name(tsk) -> str:
        """``common_density.py:613-621``."""
        if isinstance(tsk, LinearHeatConduction):
            return str(getattr(tsk, "objective", "compliance"))
        return "compliance"

    @staticmethod
    def _objective_history_ylog(tsk) -> bool:
        if isinstance(tsk, LinearHeatConduction):
            return getattr(tsk, "objective", "compliance") == "compliance"
        return True

    def scale(self):
        """``common_density.py:626-636``: lengths divided by the largest
        bounding-box side, loads by 1e5, before the FEA and filter objects
        are built from the task."""
        bbox = np.ptp(self.tsk.mesh.p, axis=1)
        self.L_scale = float(np.max(bbox))
        self.F_scale = 10 ** 5
        self.tsk.scale(1.0 / self.L_scale, 1.0 / self.F_scale)

    def unscale(self):
        """``common_density.py:638-641``."""
        self.tsk.scale(self.L_scale, self.F_scale)

    def params_latest(self):
        return self.recorder.as_object_latest()

    def init_schedulers(self, export: bool = True):
        cfg = self.cfg
        s = self.schedulers
        s.add_object_from_config(cfg.p, "p")
        s.add_object_from_config(cfg.vol_frac, "vol_frac")
        s.add_object_from_config(cfg.move_limit, "move_limit")
        s.add_object_from_config(cfg.beta, "beta")
        s.add_object_from_config(cfg.percentile, "percentile")
        s.add_object_from_config(cfg.filter_radius, "filter_radius")
        allowed = {"ConstantOne", "StepToOne", "StepAcceleratingToOne", "StepDeceleratingToOne"}
        if cfg.neumann_scale.scheduler_type not in allowed:
            raise ValueError(f"neumann_scale must use one of {allowed}")
        s.add_object_from_config(cfg.neumann_scale, "neumann_scale")
        if isinstance(cfg.eta, SchedulerConfig):
            s.add_object_from_config(cfg.eta, "eta")
        else:
            s.add("eta", cfg.eta, cfg.eta, -1, cfg.max_iters)
        s.set_iters_max(cfg.max_iters)
        self._schedulers_initialized = True

    def parameterize(self):
        cfg = self.cfg
        if cfg.filter_type == "spacial":
            self.filter = filters.SpacialFilter.from_defaults(
                self.tsk.mesh, self.tsk.elements_volume, cfg.filter_radius.init_value,
                design_mask=self.tsk.design_mask)
        elif cfg.filter_type == "helmholtz":
            self.filter = filters.HelmholtzFilterNodal.from_defaults(
                self.tsk.mesh, self.tsk.elements_volume, cfg.filter_radius.init_value,
                design_mask=self.tsk.design_mask)
        else:
            raise ValueError("should be spacial or helmholtz")

    def initialize_density(self):
        cfg, tsk = self.cfg, self.tsk
        v0 = cfg.vol_frac.init_value if cfg.vol_frac.init_value is not None \
            else cfg.vol_frac.target_value
        rho = np.zeros(tsk.mesh.nelements, dtype=np.float64)
        iter_begin = 1
        if cfg.restart:
            data_dir = f"{cfg.dst_path}/data"
            if cfg.restart_from > 0:
                it = cfg.restart_from
            else:
                files = sorted(f for f in os.listdir(data_dir) if f.endswith("-rho.npz"))
                if not files:
                    raise FileNotFoundError(f"no checkpoint in {data_dir}")
                it = int(files[-1].split("-")[0])
            with np.load(f"{data_dir}/{it:06d}-rho.npz") as d:
                rho[tsk.design_elements] = d["rho_design_elements"]
            iter_begin = it + 1
        else:
            rho += v0
            np.clip(rho, cfg.rho_min, cfg.rho_max, out=rho)
        if cfg.design_dirichlet:
            rho[tsk.neumann_elements] = 1.0
        else:
            rho[tsk.dirichlet_neumann_elements] = 1.0
        rho[tsk.fixed_elements] = 1.0
        return rho, iter_begin, cfg.max_iters + 1

    def _ensure_state_initialized(self):
        if self._completed:
            return False
        if not self._schedulers_initialized:
            self.init_schedulers()
        if self.filter is None:
            self.parameterize()
        if self._state is None:
            rho, it0, it1 = self.initialize_density()
            n_el = rho.size
            nd = len(self.tsk.design_elements)
            k = self.tsk.n_tasks
            z = lambda *s: torch.zeros(s, dtype=torch.float64, device=self.dev)  # noqa: E731
            fr = self.cfg.filter_radius
            r0 = fr.init_value if isinstance(fr.num_steps, (int, float)) else fr.target_value
            self.filter.update_radius(r0)
            self._state = dict(
                rho=_dev.to_dev(rho), rho_filtered=z(n_el), rho_projected=z(n_el),
                g=z(n_el), dC_full=z(n_el), dC_design=z(nd), rho_design=z(nd),
                rho_design_prev=z(nd), u_dofs=z(self.fem.ndof, k), energy=None,
                scal=z(16))
            self._iter_next = it0
            self._iter_end = it1
        return True

    # --- run --------------------------------------------------------------
    def optimize(self):
        self._optimize_impl()

    def optimize_steps(self, num_steps: int):
        if num_steps <= 0:
            return
        self._optimize_impl(num_steps)

    def _optimize_impl(self, max_steps=None):
        if not self._ensure_state_initialized():
            return
        start = self._iter_next
        end = self._iter_end
        if start >= end:
            self._completed = True
            return
        limit = end if max_steps is None else min(start + max_steps, end)
        self._iter_limit = limit
        converged = False
        for it in range(start, limit):
            converged = self.step(it)
            self._iter_next = it + 1
            if converged:
                break
        self._flush_records()
        if converged or self._iter_next >= end:
            self._finalize()
            self._completed = True

    def step(self, iter_num: int) -> bool:
        cfg, tsk, st = self.cfg, self.tsk, self._state
        (neumann_scale, p, vol_frac, beta, move_limit, eta, percentile,
         filter_radius) = self.schedulers.values_as_list(
            iter_num, ["neumann_scale", "p", "vol_frac", "beta", "move_limit", "eta",
                       "percentile", "filter_radius"])
        if filter_radius != self.filter.radius:
            self.filter.update_radius(filter_radius)
            self._phases.clear()
        k = tsk.n_tasks
        # filter -> projection (a replayed graph once captured)
        self._phase("pre", (float(beta),), lambda: self._step_pre(beta))
        predict = getattr(self.fem, "predict_warm_start", None)
        if predict is not None:
            predict(st["u_dofs"], st.setdefault("u_hist", {}))
        comp = self.fem.objectives_multi_load(st["rho_projected"], p, st["u_dofs"],
                                              force_scale=neumann_scale)
        if "comp" not in st:
            st["comp"] = torch.empty_like(comp)
        st["comp"].copy_(comp)
        # energies -> sensitivities -> filter adjoint -> update -> records
        # (a replayed graph once captured; the key holds every host scalar
        # its kernels take)
        key = (float(p), float(beta), float(move_limit), float(eta), float(vol_frac),
               percentile if percentile is None else float(percentile), iter_num <= 1)
        out = self._phase("post", key, lambda: self._step_post(iter_num, p, beta, move_limit,
                                                                eta, percentile, vol_frac))
        st["energy"] = out["energy"]
        names, rec = out["names"], out["rec"]
        # the previous iteration's records, while the GPU runs this one's
        # update
        self._flush_records()
        ctx = dict(names=names, k=k, neumann_scale=neumann_scale,
                   solve_iters=self.fem.last_solve.get("iters"),
                   obj=self._objective_history_name(tsk), host=self._host_records())
        if cfg.check_convergence:
            # the stop test needs this iteration's values now
            converged = self._process_records(rec.cpu().numpy(), ctx)
        else:
            # copied asynchronously and fed to the recorder during the next
            # iteration (or when the run returns), so the host never waits here
            buf = self._rec_host
            if buf is None or buf.numel() != rec.numel():
                buf = self._rec_host = torch.empty(rec.numel(), dtype=rec.dtype,
                                                   pin_memory=True)
            buf.copy_(rec, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record()
            self._rec_pending = (buf, ev, ctx)
            converged = False
        rt = max(cfg.max_iters // max(cfg.record_times, 1), 1)
        # save points of common_density.py:1186-1193 (the last iteration of
        # this call, as the reference does for partial runs)
        last = getattr(self, "_iter_limit", self._iter_end) - 1
        if cfg.write_checkpoints and (iter_num % rt == 0 or iter_num == 1 or converged or
                                      iter_num == last):
            os.makedirs(f"{cfg.dst_path}/data", exist_ok=True)
            visualization.export_mesh_with_info(
                tsk.mesh, cell_data_names=["rho_projected", "energy"],
                cell_data_values=[st["rho_projected"].cpu().numpy(),
                                  out["energy_mean"].cpu().numpy()],
                filepath=cfg.vtu_path(iter_num))
            np.savez_compressed(f"{cfg.dst_path}/data/{iter_num:06d}-rho.npz",
                                rho_design_elements=self.rho_host()[tsk.design_elements])
        return converged

    # --- the iteration's device phases --------------------------------------
    def _step_pre(self, beta):
        """rho_prev for the change record, rho_filtered = F(rho), projection."""
        cfg, st = self.cfg, self._state
        rho, rf, rp = st["rho"], st["rho_filtered"], st["rho_projected"]
        self.gather_design(rho, st["rho_design_prev"])
        self.filter.forward_into(rho, rf)
        _lib.call("skt_heaviside_project", rho.numel(), _dev.ptr(rf), float(beta),
                  float(cfg.beta_eta), _dev.ptr(rp), _dev.stream())

    def _step_post(self, iter_num, p, beta, move_limit, eta, percentile, vol_frac) -> dict:
        """Energies, the sensitivity chain, the filter adjoint, the design
        update and the record values of one iteration (device work only;
        ``comp`` is the state's persistent copy of the objectives)."""
        cfg, tsk, st = self.cfg, self.tsk, self._state
        S = _dev.stream()
        filt = self.filter
        rho, rf, rp = st["rho"], st["rho_filtered"], st["rho_projected"]
        n_el = rho.numel()
        nd = st["rho_design"].numel()
        k = tsk.n_tasks
        energy = self.fem.energy_multi_load(rp, p, st["u_dofs"])
        u_max = self.fem.col_reduce(_lib.RED_MAXABS, st["u_dofs"])
        # sensitivity: the heat FEM always uses the SIMP form (solver_heat.py:945-947)
        kind = _lib.SIMP if isinstance(tsk, LinearHeatConduction) else self.kind
        E0 = tsk.material_coef
        # the heat FEM's custom sensitivity (common_density.py:1067-1081): the
        # conduction energies, plus an explicit Robin term per load case
        sens_e = self.fem.sensitivity_energy(rp, p, st["u_dofs"]) \
            if hasattr(self.fem, "sensitivity_energy") else energy
        extra = self.fem.extra_sensitivity(rp, p, st["u_dofs"]) \
            if hasattr(self.fem, "extra_sensitivity") else None
        args = (float(beta), float(cfg.beta_eta), float(E0), float(E0 * cfg.E_min_coeff),
                float(p))
        if filt.linear_adjoint or k == 1:
            # sum over load cases first, then one adjoint application
            _lib.call("skt_sensitivity_chain_extra", kind, n_el, k, _dev.ptr(rp),
                      _dev.ptr(sens_e), _dev.ptr(extra), _dev.ptr(rf), *args, 1.0 / k, 0,
                      _dev.ptr(st["g"]), S)
            filt.gradient_into(st["g"], st["dC_full"])
        else:
            # one adjoint per load case (batched), averaged
            G = st.get("g_cols")
            if G is None:
                G = st["g_cols"] = torch.empty((n_el, k), dtype=torch.float64, device=self.dev)
            _lib.call("skt_sensitivity_chain_extra", kind, n_el, k, _dev.ptr(rp),
                      _dev.ptr(sens_e), _dev.ptr(extra), _dev.ptr(rf), *args, 1.0, 1,
                      _dev.ptr(G), S)
            filt.gradient_mean_into(G, st["dC_full"])
        if cfg.sensitivity_filter:
            tmp = st.get("dC_tmp")
            if tmp is None:
                tmp = st["dC_tmp"] = torch.empty_like(st["dC_full"])
            filt.forward_into(st["dC_full"], tmp)
            st["dC_full"].copy_(tmp)
        self.gather_design(st["dC_full"], st["dC_design"])
        self.gather_design(rho, st["rho_design"])
        self.rho_update(iter_num, st["rho_design"], rp, st["dC_design"], st["u_dofs"], None,
                        None, move_limit, eta, beta, None, None, percentile, self._vol_design,
                        self._vol_design_sum, vol_frac)
        self.scatter_design(st["rho_design"], rho)
        # record values, reduced on the device
        sc = st["scal"]
        _lib.call("skt_reduce", _lib.RED_MAXABSDIFF, nd, _dev.ptr(st["rho_design"]),
                  _dev.ptr(st["rho_design_prev"]), _dev.ptr(sc[0:1]), S)
        rpd = st.get("rp_design")
        if rpd is None:
            rpd = st["rp_design"] = torch.empty_like(st["rho_design"])
        self.gather_design(rp, rpd)
        e_mean = torch.empty(n_el, dtype=torch.float64, device=self.dev)
        _lib.call("skt_row_mean", n_el, k, _dev.ptr(energy), _dev.ptr(e_mean), S)
        arrays = {"rho_projected": rpd, "energy": e_mean}
        arrays.update(self._record_arrays(st))
        names = list(arrays)
        # [min, mean, max, std] (population std, as numpy's) per array
        sbuf = st.get("stats4")
        if sbuf is None or sbuf.numel() != 4 * len(names):
            sbuf = st["stats4"] = torch.empty(4 * len(names), dtype=torch.float64,
                                              device=self.dev)
        for i, n in enumerate(names):
            x = arrays[n].reshape(-1).to(torch.float64).contiguous()
            _lib.call("skt_stats4", x.numel(), _dev.ptr(x), _dev.ptr(sbuf[4 * i:4 * i + 4]), S)
        rec = torch.cat([sc[:8], st["comp"].reshape(-1), u_max.reshape(-1), sbuf])
        return dict(energy=energy, energy_mean=e_mean, names=names, rec=rec)

    def _graphable(self) -> bool:
        """The device phases can be captured when every decision inside them
        is made on the device: the OC update with the sparse-matrix filter
        on the elastic FEM (the Helmholtz filter's bisection keeps its
        bracket on the host; LogMOC reads scalars back)."""
        from .oc import OC_Optimizer
        return (self._step_graphs and isinstance(self, OC_Optimizer) and
                hasattr(self.filter, "handle") and
                isinstance(self.fem, fea.FEM_SimpLinearElasticity))

    def _phase(self, name, key, fn):
        """Run a device phase: eagerly the first time a key is seen (and
        capture it as a CUDA graph for the next iterations), then by graph
        replay while the key stays the same.  The graph's kernels take the
        same pointers and host scalars as the eager run, so both give the
        same bits."""
        if not self._graphable():
            return fn()
        ph = self._phases.get(name)
        if ph is not None and ph["key"] == key:
            ph["graph"].replay()
            return ph["out"]
        out = fn()
        self._capture(name, key, fn)
        return out

    def _capture(self, name, key, fn):
        s = self._cap_stream
        if s is None:
            s = self._cap_stream = torch.cuda.Stream()
            # the reductions keep per-stream partials, allocated on a
            # stream's first use: create them before any capture on s
            one = torch.ones(1, dtype=torch.float64, device=self.dev)
            res = torch.empty(1, dtype=torch.float64, device=self.dev)
            s.wait_stream(torch.cuda.current_stream())
            _lib.call("skt_reduce", _lib.RED_SUM, 1, _dev.ptr(one), None, _dev.ptr(res),
                      s.cuda_stream)
            torch.cuda.current_stream().wait_stream(s)
            torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            out = fn()
        self._phases[name] = dict(key=key, graph=g, out=out)

    def _flush_records(self):
        """Feed the recorder with a pending iteration's records (waits for
        their copy only)."""
        pending = self._rec_pending
        if pending is None:
            return
        buf, ev, ctx = pending
        self._rec_pending = None
        # queried in a loop: a blocking wait's wake-up costs tens of
        # microseconds of idle GPU between iterations
        while not ev.query():
            pass
        self._process_records(buf.numpy().copy(), ctx)

    def _process_records(self, rec: np.ndarray, ctx: dict) -> bool:
        """Records of one iteration from its host copy; returns the
        convergence flag (common_density.py:1140-1185)."""
        cfg = self.cfg
        k = ctx["k"]
        upd = self._update_records(rec[:8], ctx["host"])
        rho_change_max = float(rec[0])
        comp_h = rec[8:8 + k]
        umax_h = rec[8 + k:8 + 2 * k]
        stats_h = rec[8 + 2 * k:].reshape(-1, 4)
        compliance_avg = float(comp_h.mean())
        r = self.recorder
        r.feed_data("rho_change_max", rho_change_max)
        for n, sv in zip(ctx["names"], stats_h):
            r.feed_stats(n, sv)
        r.feed_data(ctx["obj"], compliance_avg)
        r.feed_data("u_max", float(umax_h[0]) if k == 1 else umax_h.copy())
        r.feed_data("neumann_scale", ctx["neumann_scale"])
        for name, v in upd.items():
            r.feed_data(name, v)
        r.feed_data("compliance_each", comp_h.copy())
        r.feed_data("solve_iters", ctx["solve_iters"])
        if not cfg.check_convergence:
            return False
        kkt = self.kkt_residual
        if kkt is None:
            raise ValueError("kkt_residual is not computed in rho_update")
        conv_rho = rho_change_max < cfg.tol_rho_change
        conv_kkt = abs(kkt) < cfg.tol_kkt_residual if np.isfinite(kkt) else True
        return conv_rho and conv_kkt

    def _record_arrays(self, st) -> dict:
        """Subclass hook: device arrays recorded as [min, mean, max, std]."""
        return {}

    def _finalize(self):
        """``common_density.py:944-964``: ``histories.npz`` always; the
        ``cubic_top.vtu`` submesh (rho > 0.5) when checkpoints are written."""
        cfg, tsk = self.cfg, self.tsk
        if cfg.scaling:
            self.unscale()
        if cfg.write_checkpoints and self._state is not None:
            visualization.export_submesh(tsk, self.rho_host(), 0.5,
                                         f"{cfg.dst_path}/cubic_top.vtu")
        self.recorder.export_histories(fname="histories.npz")

    def _host_records(self) -> dict:
        """Subclass hook: host scalars of this iteration's update, taken
        when the iteration runs (its records are processed later)."""
        return {}

    def _update_records(self, scal: np.ndarray, host: dict) -> dict:
        """Subclass hook: turn the device scalar slots (copied to the host
        once per iteration) and the host snapshot of the same iteration
        into named records."""
        return {}

    def gather_design(self, full, design):
        """``design = full[design_elements]`` on the device."""
        _lib.call("skt_design_gather", design.numel(), _dev.ptr(self._design_ids),
                  _dev.ptr(full), _dev.ptr(design), _dev.stream())

    def scatter_design(self, design, full):
        """``full[design_elements] = design`` on the device."""
        _lib.call("skt_design_scatter", design.numel(), _dev.ptr(self._design_ids),
                  _dev.ptr(design), _dev.ptr(full), _dev.stream())

    # --- accessors ----------------------------------------------------------
    def rho_host(self) -> np.ndarray:
        return self._state["rho"].cpu().numpy()

    @property
    def state(self):
        return self._state

    def rho_update(self, *args, **kw):
        raise NotImplementedError("")
