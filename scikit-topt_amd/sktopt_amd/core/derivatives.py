"""Compliance sensitivities on the GPU (``core/derivatives.py:42-68``)."""
from __future__ import annotations

import numpy as np
import torch

from .. import _dev, _lib


def _dC(kind, rho, strain_energy, E0, Emin, p):
    host = _dev.is_host(rho)
    r = _dev.to_dev(rho)
    U = _dev.to_dev(strain_energy)
    out = torch.empty_like(r)
    _lib.call("skt_dC_drho", kind, r.numel(), _dev.ptr(r), _dev.ptr(U), 1, float(E0),
              float(Emin), float(p), _dev.ptr(out), _dev.stream())
    return out.cpu().numpy() if host else out


def dC_drho_simp(rho, strain_energy, E0, Emin, p):
    """``-2 U dE/drho / max(E, 1e-12)`` with rho clamped at 1e-6
    (``core/derivatives.py:42-52``)."""
    return _dC(_lib.SIMP, rho, strain_energy, E0, Emin, p)


def dC_drho_ramp(rho, strain_energy, E0, Emin, p):
    """RAMP counterpart (``core/derivatives.py:65-68``)."""
    return _dC(_lib.RAMP, rho, strain_energy, E0, Emin, p)
