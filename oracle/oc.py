"""OC update with the lambda bisection on the physical volume (oracle).

Restates ``core/optimizers/oc.py:16-95`` (bisection) and ``:149-250``
(``OC_Optimizer.rho_update``).
"""
from __future__ import annotations

import numpy as np

from . import transforms as tr


def bisection_with_physical_volume(dC, rho_e, rho_full, design_elements,
                                   filt, rho_min, rho_max, move_limit, eta,
                                   eps, vol_frac, beta, beta_eta,
                                   elements_volume, elements_volume_sum,
                                   scaling_rate_min, scaling_rate_max,
                                   max_iter=100, tolerance=1e-4,
                                   vol_tol=1e-4, l1=1e-7, l2=1e7):
    """Returns (lmid, vol_error, rho_design_candidate, n_evals)."""
    it = 0
    lmid = 0.5 * (l1 + l2)
    cand_full = np.copy(rho_full)
    lower = np.maximum(rho_e - move_limit, rho_min)
    upper = np.minimum(rho_e + move_limit, rho_max)
    n_evals = 0
    while True:
        with np.errstate(invalid="ignore"):
            rate = np.power(-dC / (lmid + eps), eta)
        rate = np.clip(rate, scaling_rate_min, scaling_rate_max)
        cand = np.clip(rho_e * rate, lower, upper)
        cand_full[design_elements] = cand
        proj = tr.heaviside_projection(filt.forward(cand_full), beta, beta_eta)
        vol_error = np.sum(proj[design_elements] * elements_volume) \
            / elements_volume_sum - vol_frac
        n_evals += 1
        if abs(vol_error) < vol_tol:
            break
        if it >= max_iter:
            break
        if abs(l2 - l1) <= tolerance:
            break
        if vol_error > 0:
            l1 = lmid
        else:
            l2 = lmid
        it += 1
        lmid = 0.5 * (l1 + l2)
    return lmid, vol_error, cand, n_evals


def kkt_residual(dC_raw, rho_new, lmid, running_scale, dV, rho_min, rho_max):
    """``oc.py:230-240``: max |dC_raw + lmid * running_scale * dV| over the
    design elements whose updated density is strictly inside
    (rho_min + 1e-6, rho_max - 1e-6); 0 when there is none."""
    mask = (rho_new > rho_min + 1e-6) & (rho_new < rho_max - 1e-6)
    if not np.any(mask):
        return 0.0
    dL = dC_raw[mask] + (lmid * running_scale) * dV[mask]
    return float(np.linalg.norm(dL, ord=np.inf))


class OCUpdate:
    """State of ``OC_Optimizer.rho_update`` (``oc.py:149-250``)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.running_scale = 0.0
        self.kkt_residual = None
        self.last = {}

    def __call__(self, iter_num, rho_design, dC_design, move_limit, eta, beta,
                 percentile, vol_design, vol_design_sum, vol_frac,
                 rho_full, design_elements, filt, dV_drho_design):
        cfg = self.cfg
        dC_raw = dC_design.copy()
        rho_e = rho_design.copy()
        eps = 1e-12
        if isinstance(percentile, float):
            scale = np.percentile(np.abs(dC_design), percentile)
        else:
            scale = np.max(np.abs(dC_design))
        scale = max(scale, cfg.sensitivity_scale_floor)
        self.running_scale = 0.6 * self.running_scale + (1 - 0.6) * scale \
            if iter_num > 1 else scale
        dC = dC_design / self.running_scale
        lmid, vol_error, cand, n_evals = bisection_with_physical_volume(
            dC, rho_e, rho_full, design_elements, filt, cfg.rho_min,
            cfg.rho_max, move_limit, eta, eps, vol_frac, beta, cfg.beta_eta,
            vol_design, vol_design_sum, cfg.scaling_rate_min,
            cfg.scaling_rate_max, max_iter=1000, tolerance=1e-5,
            l1=cfg.lambda_lower, l2=cfg.lambda_upper)
        self.kkt_residual = kkt_residual(dC_raw, cand, lmid, self.running_scale,
                                         dV_drho_design, cfg.rho_min, cfg.rho_max)
        self.last = dict(lmid=lmid, vol_error=vol_error, n_evals=n_evals,
                         running_scale=self.running_scale)
        rho_design[:] = cand
        return rho_design
