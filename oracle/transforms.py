"""Design-variable transforms (oracle): projection, interpolation, dC/drho."""
from __future__ import annotations

import numpy as np


def heaviside_projection(rho, beta, eta=0.5, out=None):
    """Smooth Heaviside with the reference's ``+1e-12`` in the denominator
    (``core/projection.py:27-77``)::

        (tanh(beta eta) + tanh(beta (rho - eta)))
        / (tanh(beta eta) + tanh(beta (1 - eta)) + 1e-12)
    """
    if out is None:
        out = np.empty_like(rho)
    t = np.tanh(beta * (np.asarray(rho) - eta))
    out[...] = np.tanh(beta * eta) + t
    out /= np.tanh(beta * eta) + np.tanh(beta * (1.0 - eta)) + 1e-12
    return out


def heaviside_projection_derivative(rho, beta, eta=0.5, out=None):
    """``beta / cosh^2(beta (rho - eta)) / (tanh(beta eta) + tanh(beta (1 - eta))
    + 1e-12)`` (``core/projection.py:80-118``); cosh overflow gives 0."""
    if out is None:
        out = np.empty_like(rho)
    with np.errstate(over="ignore"):
        c = np.cosh(beta * (np.asarray(rho) - eta))
        out[...] = 1.0 / (c * c)
    out *= beta
    out /= np.tanh(beta * eta) + np.tanh(beta * (1.0 - eta)) + 1e-12
    return out


def simp_interpolation(rho, E0, Emin, p):
    """``Emin + (E0 - Emin) rho^p`` with no clamp (``fea/composer.py:19-22``)."""
    return Emin + (E0 - Emin) * (rho ** p)


def ramp_interpolation(rho, E0, Emin, p):
    """``Emin + (E0 - Emin) rho / (1 + p (1 - rho))`` (``fea/composer.py:25-39``)."""
    return Emin + (E0 - Emin) * (rho / (1.0 + p * (1.0 - rho)))


def dC_drho_simp(rho, strain_energy, E0, Emin, p):
    """``-2 U dE/drho / max(E, 1e-12)`` with ``rho`` clamped at 1e-6 inside
    both E and dE (``core/derivatives.py:25-52``)."""
    r = np.maximum(rho, 1e-6)
    dE = p * (E0 - Emin) * r ** (p - 1)
    E = Emin + (E0 - Emin) * r ** p
    return -2.0 * strain_energy * dE / np.maximum(E, 1e-12)


def dC_drho_ramp(rho, strain_energy, E0, Emin, p):
    """RAMP counterpart, no clamp on rho (``core/derivatives.py:55-68``)."""
    denom = 1.0 + p * (1.0 - rho)
    dE = (E0 - Emin) * (denom - p * rho) / (denom ** 2)
    E = Emin + (E0 - Emin) * (rho / denom)
    return -2.0 * strain_energy * dE / np.maximum(E, 1e-12)


def dC_drho_ramp_exact(rho, strain_energy, E0, Emin, p):
    """The derivative of the compliance through ``ramp_interpolation``:
    dE/drho = (E0 - Emin) (1 + p) / (1 + p (1 - rho))^2.  Not a reference
    function (the reference's expression is ``dC_drho_ramp``); it is what
    the differentiable compliance must reproduce."""
    denom = 1.0 + p * (1.0 - rho)
    dE = (E0 - Emin) * (1.0 + p) / (denom ** 2)
    E = Emin + (E0 - Emin) * (rho / denom)
    return -2.0 * strain_energy * dE / np.maximum(E, 1e-12)
