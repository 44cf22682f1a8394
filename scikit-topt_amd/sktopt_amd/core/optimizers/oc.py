"""OC update with the lambda bisection on the physical volume, on the GPU.

``OC_Config`` / ``OC_Optimizer`` mirror ``core/optimizers/oc.py:98-250``.
With the sparse-matrix (spacial) filter the update runs without a host
round trip: the sensitivity scale (max or percentile of |dC|), its running
average, the whole bisection (``bisection_with_physical_volume``,
``oc.py:16-95``: candidate update, filter gather, projection, float64
volume sum, bracket update, exit test) and the KKT residual are device
kernels writing device scalars.  With the nodal Helmholtz filter every
bisection evaluation is a filter solve, so the host keeps the bracket and
reads one volume scalar per evaluation.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Literal

import numpy as np
import torch

from ... import _dev, _lib
from ...tools import SchedulerConfig
from . import common_density


@dataclass
class OC_Config(common_density.DensityMethod_OC_Config):
    interpolation: Literal["SIMP"] = "SIMP"
    eta: SchedulerConfig = field(default_factory=lambda: SchedulerConfig.constant(
        target_value=0.5))
    scaling_rate_min: float = 0.7
    scaling_rate_max: float = 1.3
    sensitivity_scale_floor: float = 1e-12


# device scalar slots (DensityMethod._state["scal"])
S_CHANGE, S_LMID, S_VOLERR, S_NEVAL, S_WIDTH, S_KKT, S_RSCALE, S_SCALE = range(8)


class OC_Optimizer(common_density.DensityMethod):
    def __init__(self, cfg: OC_Config, tsk):
        assert cfg.lambda_lower < cfg.lambda_upper
        super().__init__(cfg, tsk)
        self.running_scale = 0
        self._dC_raw = None
        self._rho_e = None

    def rho_update(self, iter_num, rho_design_eles, rho_projected, dC_drho_design_eles,
                   u_dofs, strain_energy_mean, scaling_rate, move_limit, eta, beta,
                   rho_clip_lower, rho_clip_upper, percentile, elements_volume_design,
                   elements_volume_design_sum, vol_frac):
        cfg = self.cfg
        st = self._state
        if st is None:
            raise RuntimeError("Optimizer state is not initialized.")
        self._last_eta = float(eta)
        S = _dev.stream()
        n = rho_design_eles.numel()
        sc = st["scal"]
        if self._dC_raw is None:
            self._dC_raw = torch.empty_like(dC_drho_design_eles)
            self._rho_e = torch.empty_like(rho_design_eles)
        self._dC_raw.copy_(dC_drho_design_eles)
        self._rho_e.copy_(rho_design_eles)
        scale = sc[S_SCALE:S_SCALE + 1]
        if isinstance(percentile, float):
            _lib.call("skt_percentile_abs", n, _dev.ptr(dC_drho_design_eles), float(percentile),
                      _dev.ptr(scale), S)
        else:
            _lib.call("skt_reduce", _lib.RED_MAXABS, n, _dev.ptr(dC_drho_design_eles), None,
                      _dev.ptr(scale), S)
        rs = sc[S_RSCALE:S_RSCALE + 1]
        _lib.call("skt_oc_normalize", n, _dev.ptr(dC_drho_design_eles), _dev.ptr(scale),
                  float(cfg.sensitivity_scale_floor), 0.6, 1 if iter_num <= 1 else 0,
                  _dev.ptr(rs), _dev.ptr(dC_drho_design_eles), S)
        res = sc[S_LMID:S_LMID + 4]
        if hasattr(self.filter, "handle"):
            # sparse-matrix filter: the whole bisection runs on the device
            _lib.call("skt_oc_bisect", self.filter.handle, _dev.ptr(st["rho"]),
                      _dev.ptr(self._rho_e), _dev.ptr(dC_drho_design_eles),
                      _dev.ptr(elements_volume_design), float(elements_volume_design_sum),
                      float(vol_frac), float(move_limit), float(eta), float(cfg.rho_min),
                      float(cfg.rho_max), float(cfg.scaling_rate_min),
                      float(cfg.scaling_rate_max), float(beta), float(cfg.beta_eta),
                      float(cfg.lambda_lower), float(cfg.lambda_upper), 1e-5, 1e-4, 1000,
                      _dev.ptr(rho_design_eles), _dev.ptr(res), S)
        else:
            self._bisect_host(self._rho_e, dC_drho_design_eles, rho_design_eles,
                              elements_volume_design, float(elements_volume_design_sum),
                              float(vol_frac), float(move_limit), float(eta), float(beta), res)
        _lib.call("skt_oc_kkt", n, _dev.ptr(self._dC_raw), _dev.ptr(rho_design_eles),
                  _dev.ptr(self._dV_drho_design), _dev.ptr(res), _dev.ptr(rs),
                  float(cfg.rho_min), float(cfg.rho_max), _dev.ptr(sc[S_KKT:S_KKT + 1]), S)

    def _bisect_host(self, rho_e, dC, rho_out, vol_design, vol_sum, vol_frac, move_limit, eta,
                     beta, res, tolerance=1e-5, vol_tol=1e-4, max_iter=1000):
        """``bisection_with_physical_volume`` (``oc.py:16-95``) for filters
        that are not a sparse matrix (the nodal Helmholtz solve): the host
        keeps the bracket, every evaluation runs the candidate update, the
        filter, the projection and the float64 volume sum on the device and
        reads back one scalar."""
        cfg, st = self.cfg, self._state
        S = _dev.stream()
        nd = rho_e.numel()
        n_el = st["rho"].numel()
        b = getattr(self, "_bis", None)
        if b is None:
            z = lambda m: torch.empty(m, dtype=torch.float64, device=self.dev)  # noqa: E731
            b = self._bis = dict(cand=z(n_el), rf=z(n_el), rp=z(n_el), rpd=z(nd), vol=z(1))
        b["cand"].copy_(st["rho"])      # non-design entries are carried over
        l1, l2 = float(cfg.lambda_lower), float(cfg.lambda_upper)
        lmid = 0.5 * (l1 + l2)
        vol_error = 0.0
        it = 0
        n_evals = 0
        while True:
            _lib.call("skt_oc_candidate", nd, _dev.ptr(self._design_ids), _dev.ptr(rho_e),
                      _dev.ptr(dC), lmid, eta, float(cfg.rho_min), float(cfg.rho_max),
                      float(cfg.scaling_rate_min), float(cfg.scaling_rate_max), move_limit,
                      _dev.ptr(b["cand"]), _dev.ptr(rho_out), S)
            self.filter.forward_into(b["cand"], b["rf"])
            _lib.call("skt_heaviside_project", n_el, _dev.ptr(b["rf"]), beta,
                      float(cfg.beta_eta), _dev.ptr(b["rp"]), S)
            self.gather_design(b["rp"], b["rpd"])
            _lib.call("skt_reduce", _lib.RED_DOT, nd, _dev.ptr(b["rpd"]), _dev.ptr(vol_design),
                      _dev.ptr(b["vol"]), S)
            n_evals += 1
            vol_error = float(b["vol"].item()) / vol_sum - vol_frac
            if abs(vol_error) < vol_tol or it >= max_iter or abs(l2 - l1) <= tolerance:
                break
            if vol_error > 0:
                l1 = lmid
            else:
                l2 = lmid
            it += 1
            lmid = 0.5 * (l1 + l2)
        res.copy_(torch.tensor([lmid, vol_error, float(n_evals), l2 - l1], dtype=torch.float64))

    def add_recorder(self, tsk):
        """``oc.py:118-122``."""
        r = super().add_recorder(tsk)
        r.add("-dC", plot_type="min-max-mean-std", ylog=True)
        r.add("lmid", ylog=True)
        return r

    def _record_arrays(self, st) -> dict:
        """``oc.py:247-250`` and the driver's ``scaling_rate`` record: the
        normalised sensitivities and the rates at the final multiplier."""
        cfg = self.cfg
        dC = st["dC_design"]
        lmid = st["scal"][S_LMID]
        rate = torch.clamp(torch.pow(-dC / (lmid + 1e-12), self._last_eta),
                           cfg.scaling_rate_min, cfg.scaling_rate_max)
        return {"-dC": -dC, "scaling_rate": rate}

    def _update_records(self, scal: np.ndarray, host: dict) -> dict:
        self.kkt_residual = float(scal[S_KKT])
        self.running_scale = float(scal[S_RSCALE])
        return {"lmid": float(scal[S_LMID]), "vol_error": float(scal[S_VOLERR]),
                "bisection_evals": int(scal[S_NEVAL]), "kkt_residual": self.kkt_residual,
                "running_scale": self.running_scale}
