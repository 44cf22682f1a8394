"""OC update with the lambda bisection on the physical volume, on the GPU.

``OC_Config`` / ``OC_Optimizer`` mirror ``core/optimizers/oc.py:98-250``.
The update runs without a host round trip: the sensitivity scale (max or
percentile of |dC|), its running average, the whole bisection
(``bisection_with_physical_volume``, ``oc.py:16-95``: candidate update,
filter gather, projection, float64 volume sum, bracket update, exit test)
and the KKT residual are device kernels writing device scalars.
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
        _lib.call("skt_oc_bisect", self.filter.handle, _dev.ptr(st["rho"]),
                  _dev.ptr(self._rho_e), _dev.ptr(dC_drho_design_eles),
                  _dev.ptr(elements_volume_design), float(elements_volume_design_sum),
                  float(vol_frac), float(move_limit), float(eta), float(cfg.rho_min),
                  float(cfg.rho_max), float(cfg.scaling_rate_min), float(cfg.scaling_rate_max),
                  float(beta), float(cfg.beta_eta), float(cfg.lambda_lower),
                  float(cfg.lambda_upper), 1e-5, 1e-4, 1000, _dev.ptr(rho_design_eles),
                  _dev.ptr(res), S)
        _lib.call("skt_oc_kkt", n, _dev.ptr(self._dC_raw), _dev.ptr(rho_design_eles),
                  _dev.ptr(self._dV_drho_design), _dev.ptr(res), _dev.ptr(rs),
                  float(cfg.rho_min), float(cfg.rho_max), _dev.ptr(sc[S_KKT:S_KKT + 1]), S)

    def _update_records(self, scal: np.ndarray) -> dict:
        self.kkt_residual = float(scal[S_KKT])
        self.running_scale = float(scal[S_RSCALE])
        return {"lmid": float(scal[S_LMID]), "vol_error": float(scal[S_VOLERR]),
                "bisection_evals": int(scal[S_NEVAL]), "kkt_residual": self.kkt_residual,
                "running_scale": self.running_scale}
