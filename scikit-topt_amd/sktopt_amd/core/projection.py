"""Smooth Heaviside projection on the GPU.

Same call surface as ``core/projection.py:27-118`` of the reference:
``heaviside_projection_inplace(rho, beta, eta=0.5, out=None)`` and
``heaviside_projection_derivative_inplace(rho, beta, eta=0.5, out=None)``.
``rho``/``out`` may be float64 device tensors (device-resident path) or numpy
arrays (host buffers are copied over PCIe and back).
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _dev, _lib


def _run(name, rho, beta, eta, out):
    host = _dev.is_host(rho)
    x = _dev.to_dev(rho)
    if out is None:
        y = torch.empty_like(x)
    elif _dev.is_host(out):
        y = torch.empty_like(x)
    else:
        y = out
    _lib.call(name, x.numel(), _dev.ptr(x), float(beta), float(eta), _dev.ptr(y), _dev.stream())
    if out is not None and _dev.is_host(out):
        out[...] = y.cpu().numpy().reshape(np.shape(out))
        return out
    if out is None and host:
        return y.cpu().numpy().reshape(np.shape(rho))
    return y


def heaviside_projection_inplace(rho, beta, eta=0.5, out=None):
    """``(tanh(beta eta) + tanh(beta (rho - eta))) / (tanh(beta eta) +
    tanh(beta (1 - eta)) + 1e-12)`` (``core/projection.py:27-77``)."""
    return _run("skt_heaviside_project", rho, beta, eta, out)


def heaviside_projection_derivative_inplace(rho, beta, eta=0.5, out=None):
    """``beta sech^2(beta (rho - eta)) / (tanh(beta eta) + tanh(beta (1 - eta))
    + 1e-12)`` (``core/projection.py:80-118``)."""
    return _run("skt_heaviside_derivative", rho, beta, eta, out)


def heaviside_projection(rho, beta, eta=0.5):
    return heaviside_projection_inplace(rho, beta, eta)


def heaviside_projection_derivative(rho, beta, eta=0.5):
    return heaviside_projection_derivative_inplace(rho, beta, eta)
