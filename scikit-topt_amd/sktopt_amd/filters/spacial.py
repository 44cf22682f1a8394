"""Gaussian density filter as a precomputed neighbour-list gather.

Call surface of ``filters/spacial.py:106-166``: ``from_defaults(mesh,
elements_volume, radius, design_mask)``, ``forward(v)``, ``gradient(v)``,
``update_radius(r)`` and ``radius``.  The reference rebuilds its CSR matrix on
every ``forward`` (``spacial.py:143-149``); the matrix depends only on the
radius and the design mask, so here it is built once per radius on the GPU
(neighbour search, Gaussian weights, row normalisation, transposed values)
and reused.  ``forward`` / ``gradient`` accept float64 device tensors or
numpy arrays (host copies go over PCIe and back).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _dev, _lib


def make_kernel(kind: str, r_min: float, sigma: float | None = None):
    """Kernel support and sigma exactly as ``spacial.py:19-48`` computes
    them (only the Gaussian kind is on the GPU path)."""
    if kind != "gaussian":
        raise ValueError("unknown kernel kind" if kind not in ("linear", "quadratic")
                         else f"{kind} kernel is not built on the GPU path")
    s = (r_min / 3.0) if sigma is None else float(sigma)
    return s, 3.0 * s


class SpacialFilter:
    def __init__(self, mesh, elements_volume, radius, design_mask=None, element_centers=None):
        self.mesh = mesh
        self.elements_volume = elements_volume
        self.radius = float(radius)
        n = mesh.nelements if element_centers is None else element_centers.shape[1]
        self.design_mask = np.ones(n, dtype=bool) if design_mask is None \
            else np.asarray(design_mask, dtype=bool)
        c = mesh.element_centers() if element_centers is None else element_centers
        self.element_centers = np.ascontiguousarray(c, dtype=np.float64)
        self.n_el = self.element_centers.shape[1]
        self._dev_centers = None
        self._dev_mask = None
        self._h = None
        self._built_radius = None
        self.n_design = 0
        self.nnz = 0
        # structured hex grids: apply W as a verified stencil (no CSR reads)
        self.use_stencil = True
        self.stencil_active = False

    @classmethod
    def from_defaults(cls, mesh, elements_volume, radius=0.3, design_mask=None):
        return cls(mesh, elements_volume, radius, design_mask)

    def update_radius(self, radius, **args):
        self.radius = float(radius)

    # --- device matrix --------------------------------------------------
    def _ensure(self):
        if self._h is not None and self._built_radius == self.radius:
            return
        if self._h is not None:
            _lib.load().skt_filter_destroy(self._h)
            self._h = None
        if self._dev_centers is None:
            self._dev_centers = _dev.to_dev(self.element_centers)
            self._dev_mask = torch.as_tensor(self.design_mask.astype(np.uint8),
                                             device=_dev.device())
        sigma, support = make_kernel("gaussian", self.radius)
        h = C.c_void_p()
        _lib.call("skt_filter_create", self.n_el, _dev.ptr(self._dev_centers),
                  _dev.ptr(self._dev_mask), float(sigma), float(support), C.byref(h))
        self._h = h
        self._built_radius = self.radius
        self.stencil_active = False
        if self.use_stencil and getattr(self.mesh, "kind", None) == "hex8" and \
                getattr(self.mesh, "grid", None) is not None:
            act = C.c_int32(0)
            _lib.call("skt_filter_set_grid", h, *self.mesh.grid, *self.mesh.h, C.byref(act))
            self.stencil_active = bool(act.value)
        nd, nnz = C.c_int64(), C.c_int64()
        _lib.call("skt_filter_info", h, C.byref(nd), C.byref(nnz))
        self.n_design, self.nnz = nd.value, nnz.value

    @property
    def handle(self):
        self._ensure()
        return self._h

    def __del__(self):
        try:
            if self._h is not None:
                _lib.load().skt_filter_destroy(self._h)
        except Exception:
            pass

    def _run(self, name, v):
        self._ensure()
        host = _dev.is_host(v)
        x = _dev.to_dev(v)
        y = torch.empty_like(x)
        _lib.call(name, self._h, _dev.ptr(x), _dev.ptr(y), _dev.stream())
        return y.cpu().numpy() if host else y

    def forward(self, rho_element):
        """``out = rho; out[design] = W rho[design]`` (``spacial.py:135-157``)."""
        return self._run("skt_filter_forward", rho_element)

    def gradient(self, v):
        """``out = 0; out[design] = W^T v[design]`` (``spacial.py:159-166``)."""
        return self._run("skt_filter_gradient", v)

    # --- device interface of the density loop ------------------------------
    linear_adjoint = True     # a fixed sparse matrix: the load cases can be summed first

    def forward_into(self, x, y):
        _lib.call("skt_filter_forward", self.handle, _dev.ptr(x), _dev.ptr(y), _dev.stream())

    def gradient_into(self, v, g):
        _lib.call("skt_filter_gradient", self.handle, _dev.ptr(v), _dev.ptr(g), _dev.stream())

    def csr(self):
        """(indptr, indices, data, design_ids) as numpy, for inspection."""
        self._ensure()
        dev = _dev.device()
        indptr = torch.empty(self.n_design + 1, dtype=torch.int64, device=dev)
        indices = torch.empty(self.nnz, dtype=torch.int32, device=dev)
        data = torch.empty(self.nnz, dtype=torch.float64, device=dev)
        ids = torch.empty(self.n_design, dtype=torch.int32, device=dev)
        _lib.call("skt_filter_export", self._h, _dev.ptr(indptr), _dev.ptr(indices),
                  _dev.ptr(data), _dev.ptr(ids), _dev.stream())
        return (indptr.cpu().numpy(), indices.cpu().numpy(), data.cpu().numpy(),
                ids.cpu().numpy())
