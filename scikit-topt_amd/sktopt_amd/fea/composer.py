"""SIMP / RAMP material interpolation (``fea/composer.py:19-39``).

The callables keep the reference's signature ``f(rho, E0, Emin, p)`` and are
passed to the FEM classes as ``density_interpolation``; their ``kind``
attribute selects the device kernel.
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _dev, _lib


def _interp(kind, rho, E0, Emin, p, out_dtype=torch.float64):
    host = _dev.is_host(rho)
    r = _dev.to_dev(rho)
    out = torch.empty(r.shape, dtype=out_dtype, device=r.device)
    code = _lib.F64 if out_dtype == torch.float64 else _lib.F32
    _lib.call("skt_interpolate", kind, r.numel(), _dev.ptr(r), float(E0), float(Emin),
              float(p), _dev.ptr(out), code, _dev.stream())
    return out.cpu().numpy() if host else out


def simp_interpolation(rho, E0, Emin, p):
    """``Emin + (E0 - Emin) rho^p`` (no clamp)."""
    return _interp(_lib.SIMP, rho, E0, Emin, p)


def ramp_interpolation(rho, E0, Emin, p):
    """``Emin + (E0 - Emin) rho / (1 + p (1 - rho))``."""
    return _interp(_lib.RAMP, rho, E0, Emin, p)


simp_interpolation.kind = _lib.SIMP
ramp_interpolation.kind = _lib.RAMP


def interpolation_kind(func) -> int:
    k = getattr(func, "kind", None)
    if k is None:
        name = getattr(func, "__name__", "")
        if "simp" in name:
            return _lib.SIMP
        if "ramp" in name:
            return _lib.RAMP
        raise ValueError("Interpolation method must be SIMP or RAMP.")
    return k
