"""Log-space modified OC update (oracle).

Restates ``core/optimizers/logmoc.py:36-66`` (``lagrangian_log_update``) and
``:90-258`` (``LogMOC_Optimizer.rho_update``).
"""
from __future__ import annotations

import numpy as np

from . import transforms as tr


def lagrangian_log_update(rho, dL, eta, move_limit, rho_min, rho_max,
                          lagrangian_clip):
    """In-place on ``rho``; returns it (``logmoc.py:36-66``)."""
    step = np.clip(dL, -lagrangian_clip, lagrangian_clip)
    np.clip(rho, rho_min, rho_max, out=rho)
    log_r = np.log(rho)
    width = np.log(move_limit / np.exp(log_r) + 1.0)
    lower = log_r - width
    upper = lower + 2.0 * width
    x = np.log(rho) - eta * step
    x = np.clip(x, lower, upper)
    rho[:] = np.clip(np.exp(x), rho_min, rho_max)
    return rho


class LogMOCUpdate:
    def __init__(self, cfg):
        self.cfg = cfg
        self.lambda_v = cfg.lambda_v
        self.running_scale = None
        self.kkt_residual = None
        self.last = {}

    def __call__(self, iter_num, rho_design, dC_design, rho_projected,
                 rho_filtered, move_limit, eta, beta, percentile, vol_design,
                 vol_design_sum, vol_frac, design_elements, filt):
        cfg = self.cfg
        dC_raw = dC_design.copy()
        if cfg.center_objective:
            dC_raw -= np.sum(dC_raw * vol_design) / vol_design_sum
        n_el = rho_projected.shape[0]
        dV_proj = np.zeros(n_el)
        dV_proj[design_elements] = vol_design / vol_design_sum
        dV_filt = tr.heaviside_projection_derivative(rho_filtered, beta,
                                                     cfg.beta_eta)
        dV_filt *= dV_proj
        back = filt.gradient(dV_filt)
        if np.allclose(back, 0.0) and np.any(dV_filt > 0.0):
            back = filt.forward(dV_filt)
        dV_chain = back[design_elements].copy()
        volume = np.sum(rho_projected[design_elements] * vol_design) \
            / vol_design_sum
        vol_error = volume - vol_frac
        chain_scale = 1.0
        if cfg.normalize_volume_chain:
            obj_scale = max(np.percentile(np.abs(dC_raw),
                                          cfg.lagrangian_percentile),
                            cfg.lagrangian_scale_floor)
            vol_scale = max(np.percentile(np.abs(dV_chain),
                                          cfg.volume_chain_percentile),
                            cfg.volume_chain_scale_floor)
            gain = cfg.volume_chain_gain
            if cfg.volume_chain_gain_under is not None and vol_error < 0.0:
                gain = cfg.volume_chain_gain_under
            chain_scale = gain * (obj_scale / vol_scale)
            dV_chain *= chain_scale
        penalty = cfg.mu_p * vol_error
        if cfg.dual_update == "augmented":
            self.lambda_v = self.lambda_v + penalty if iter_num > 1 else penalty
        else:
            self.lambda_v = cfg.lambda_decay * self.lambda_v + \
                (1.0 - cfg.lambda_decay) * penalty if iter_num > 1 else penalty
        self.lambda_v = float(np.clip(self.lambda_v, cfg.lambda_lower,
                                      cfg.lambda_upper))
        coeff = self.lambda_v + cfg.augmented_lagrangian_mu * vol_error
        dL = dC_raw + coeff * dV_chain
        if cfg.filter_lagrangian:
            full = np.zeros(n_el)
            full[design_elements] = dL
            dL = filt.forward(full)[design_elements]
        if cfg.center_lagrangian:
            dL = dL - np.sum(dL * vol_design) / vol_design_sum
        q = cfg.lagrangian_percentile
        if isinstance(percentile, float):
            q = percentile
        scale = max(np.percentile(np.abs(dL), q), cfg.lagrangian_scale_floor)
        self.running_scale = 0.2 * self.running_scale + (1.0 - 0.2) * scale \
            if iter_num > 1 else scale
        dL = dL / self.running_scale
        mask = (rho_design > cfg.rho_min + 1e-6) & (rho_design < cfg.rho_max - 1e-6)
        self.kkt_residual = float(np.max(np.abs(dL[mask]))) if np.any(mask) else 0.0
        self.last = dict(lambda_v=self.lambda_v, vol_error=vol_error,
                         running_scale=self.running_scale, coeff=coeff)
        lagrangian_log_update(rho_design, dL, eta, move_limit, cfg.rho_min,
                              cfg.rho_max, cfg.lagrangian_clip)
        return rho_design
