"""Log-space modified OC on the GPU (``core/optimizers/logmoc.py:13-258``).

Vectors stay on the device; the handful of scalars the update needs (volume
error, percentile scales, dual variable) are reduced on the device and read
back once each, so this update has a few host round trips per iteration.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Literal

import numpy as np
import torch

from ... import _dev, _lib
from . import common_density


@dataclass
class LogMOC_Config(common_density.DensityMethod_OC_Config):
    interpolation: Literal["SIMP"] = "SIMP"
    mu_p: float = 5.0
    augmented_lagrangian_mu: float = 0.0
    lambda_v: float = 0.1
    lambda_decay: float = 0.90
    lambda_lower: float = -1e7
    lambda_upper: float = 1e7
    lagrangian_clip: float = 1.0
    lagrangian_percentile: float = 95.0
    lagrangian_scale_floor: float = 1e-8
    normalize_volume_chain: bool = False
    volume_chain_percentile: float = 95.0
    volume_chain_scale_floor: float = 1e-8
    volume_chain_gain: float = 1.0
    volume_chain_gain_under: float | None = None
    dual_update: Literal["ema", "augmented"] = "ema"
    filter_lagrangian: bool = False
    center_lagrangian: bool = False
    center_objective: bool = False


class LogMOC_Optimizer(common_density.DensityMethod):
    def __init__(self, cfg: LogMOC_Config, tsk):
        super().__init__(cfg, tsk)
        self.lambda_v = cfg.lambda_v
        self.running_scale = None
        self._buf = None
        self._last = {}

    def _scalar(self, op, n, x, y=None) -> float:
        out = self._buf["s"]
        _lib.call("skt_reduce", op, n, _dev.ptr(x), _dev.ptr(y) if y is not None else None,
                  _dev.ptr(out), _dev.stream())
        return float(out.item())

    def _percentile(self, x, q) -> float:
        out = self._buf["s"]
        _lib.call("skt_percentile_abs", x.numel(), _dev.ptr(x), float(q), _dev.ptr(out),
                  _dev.stream())
        return float(out.item())

    def _ewise(self, op, a, x, b, y, out):
        _lib.call("skt_ewise", op, x.numel(), float(a), _dev.ptr(x), float(b),
                  _dev.ptr(y) if y is not None else None, _dev.ptr(out), _dev.stream())

    def rho_update(self, iter_num, rho_design_eles, rho_projected, dC_drho_design_eles,
                   u_dofs, strain_energy_mean, scaling_rate, move_limit, eta, beta,
                   rho_clip_lower, rho_clip_upper, percentile, elements_volume_design,
                   elements_volume_design_sum, vol_frac):
        cfg, tsk, st = self.cfg, self.tsk, self._state
        S = _dev.stream()
        filt = self.filter
        nd = rho_design_eles.numel()
        n_el = rho_projected.numel()
        vsum = float(elements_volume_design_sum)
        if self._buf is None:
            z = lambda m: torch.zeros(m, dtype=torch.float64, device=self.dev)  # noqa: E731
            dVp = z(n_el)
            self.scatter_design(self._dV_drho_design, dVp)
            self._buf = dict(dC_raw=z(nd), dL=z(nd), chain=z(nd), dVp=dVp, dVf=z(n_el),
                             back=z(n_el), full=z(n_el), full2=z(n_el), tmp=z(nd),
                             s=z(1), zero1=z(1), one1=torch.ones(1, dtype=torch.float64,
                                                                 device=self.dev))
        b = self._buf
        b["dC_raw"].copy_(dC_drho_design_eles)
        if cfg.center_objective:
            mean = self._scalar(_lib.RED_DOT, nd, b["dC_raw"], elements_volume_design) / vsum
            b["tmp"].fill_(mean)
            self._ewise(0, 1.0, b["dC_raw"], -1.0, b["tmp"], b["dC_raw"])
        # volume chain dV = F^T (H'(rho_filtered) * v / sum v)
        _lib.call("skt_heaviside_derivative", n_el, _dev.ptr(st["rho_filtered"]), float(beta),
                  float(cfg.beta_eta), _dev.ptr(b["dVf"]), S)
        self._ewise(1, 1.0, b["dVf"], 0.0, b["dVp"], b["dVf"])
        filt.gradient_into(b["dVf"], b["back"])
        back = b["back"]
        # np.allclose(back, 0) and any(dVf > 0): forward fallback (logmoc.py:145-153)
        if self._scalar(_lib.RED_MAXABS, n_el, back) <= 1e-8 and \
                self._scalar(_lib.RED_MAXABS, n_el, b["dVf"]) > 0.0:
            filt.forward_into(b["dVf"], b["full"])
            back = b["full"]
        self.gather_design(back, b["chain"])
        self.gather_design(rho_projected, b["tmp"])
        volume = self._scalar(_lib.RED_DOT, nd, b["tmp"], elements_volume_design) / vsum
        vol_error = volume - vol_frac
        chain_scale = 1.0
        if cfg.normalize_volume_chain:
            obj = max(self._percentile(b["dC_raw"], cfg.lagrangian_percentile),
                      cfg.lagrangian_scale_floor)
            vs = max(self._percentile(b["chain"], cfg.volume_chain_percentile),
                     cfg.volume_chain_scale_floor)
            gain = cfg.volume_chain_gain
            if cfg.volume_chain_gain_under is not None and vol_error < 0.0:
                gain = cfg.volume_chain_gain_under
            chain_scale = gain * (obj / vs)
            self._ewise(0, chain_scale, b["chain"], 0.0, None, b["chain"])
        penalty = cfg.mu_p * vol_error
        if cfg.dual_update == "augmented":
            self.lambda_v = self.lambda_v + penalty if iter_num > 1 else penalty
        else:
            self.lambda_v = cfg.lambda_decay * self.lambda_v + \
                (1.0 - cfg.lambda_decay) * penalty if iter_num > 1 else penalty
        self.lambda_v = float(np.clip(self.lambda_v, cfg.lambda_lower, cfg.lambda_upper))
        coeff = self.lambda_v + cfg.augmented_lagrangian_mu * vol_error
        self._ewise(0, 1.0, b["dC_raw"], coeff, b["chain"], b["dL"])
        if cfg.filter_lagrangian:
            b["full"].zero_()
            self.scatter_design(b["dL"], b["full"])
            filt.forward_into(b["full"], b["full2"])
            self.gather_design(b["full2"], b["dL"])
        if cfg.center_lagrangian:
            mean = self._scalar(_lib.RED_DOT, nd, b["dL"], elements_volume_design) / vsum
            b["tmp"].fill_(mean)
            self._ewise(0, 1.0, b["dL"], -1.0, b["tmp"], b["dL"])
        q = cfg.lagrangian_percentile
        if isinstance(percentile, float):
            q = percentile
        scale = max(self._percentile(b["dL"], q), cfg.lagrangian_scale_floor)
        self.running_scale = 0.2 * self.running_scale + (1.0 - 0.2) * scale \
            if iter_num > 1 else scale
        self._ewise(0, 1.0 / self.running_scale, b["dL"], 0.0, None, b["dL"])
        # KKT-like residual: max |dL| over interior densities (before the update)
        kkt = st["scal"][5:6]
        _lib.call("skt_oc_kkt", nd, _dev.ptr(b["dL"]), _dev.ptr(rho_design_eles),
                  _dev.ptr(b["dL"]), _dev.ptr(b["zero1"]), _dev.ptr(b["one1"]),
                  float(cfg.rho_min), float(cfg.rho_max), _dev.ptr(kkt), S)
        _lib.call("skt_logmoc_update", nd, _dev.ptr(rho_design_eles), _dev.ptr(b["dL"]),
                  float(eta), float(move_limit), float(cfg.rho_min), float(cfg.rho_max),
                  float(cfg.lagrangian_clip), S)
        self._last = dict(lambda_v=self.lambda_v, constraint_coeff=coeff, vol_error=vol_error,
                          volume_chain_scale=chain_scale, running_scale=self.running_scale)

    def add_recorder(self, tsk):
        """``logmoc.py:75-83``."""
        r = super().add_recorder(tsk)
        r.add("dL", plot_type="min-max-mean-std", ylog=False)
        r.add("-dC", plot_type="min-max-mean-std", ylog=True)
        r.add("lambda_v", ylog=False)
        r.add("constraint_coeff", ylog=False)
        r.add("dV_chain", plot_type="min-max-mean-std", ylog=True)
        r.add("volume_chain_scale", ylog=True)
        return r

    def _record_arrays(self, st) -> dict:
        """``logmoc.py:238-245``."""
        b = self._buf
        return {"-dC": -b["dC_raw"], "dV_chain": b["chain"].abs(), "dL": b["dL"]}

    def _host_records(self) -> dict:
        # the dual variable, constraint coefficient, volume error and scales
        # of the iteration that just ran (logmoc.py:238-245 records them in
        # the same iteration's update)
        return dict(self._last)

    def _update_records(self, scal: np.ndarray, host: dict) -> dict:
        self.kkt_residual = float(scal[5])
        d = dict(host)
        d["kkt_residual"] = self.kkt_residual
        return d
