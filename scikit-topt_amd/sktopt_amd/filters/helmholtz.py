"""Nodal Helmholtz density filter on the GPU.

Call surface of ``filters/helmholtz_filter_nodal.py:162-233``
(``HelmholtzFilterNodal``): ``from_defaults(mesh, elements_volume, radius,
design_mask)``, ``forward(v)``, ``gradient(v)``, ``update_radius(r)`` and
``radius``.

forward (``:183-201``): element -> node volume-weighted averaging with
non-design elements counted as 1, the nodal solve ``(M + r^2 K) x = M rho``
with ``x = 1`` on the nodes of non-design elements, node -> element mean.
gradient (``:203-233``): the same averaging with 0 for non-design elements,
the solve without Dirichlet nodes, the mean, then ``min(v, 0)``.  As in the
reference the gradient is not the exact adjoint of the forward map.

On structured hex8 grids ``M + r^2 K`` and ``M`` are matrix-free scalar
operators (diagonal in the Walsh-Hadamard corner basis, the same modal
kernel as the state operator) and the solves use the batched CG with the
geometric-multigrid preconditioner, warm-started from the previous call.
On tet4 meshes (``ElementTetP1``) they are the matrix-free P1 operator with
the exact element mass ``V/20 (1 + delta_ab)`` and conduction
``V grad N_a . grad N_b``, gathered node by node over the node's elements;
``M + r^2 K`` is well conditioned (its spectrum is the mass's, shifted by
r^2 times the Laplacian's), so the Jacobi-preconditioned CG solves it in a
few dozen iterations.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _dev, _lib


class HelmholtzFilterNodal:
    def __init__(self, mesh, elements_volume, radius, design_mask=None, rtol=1e-12,
                 precision="fp64"):
        kind = getattr(mesh, "kind", None)
        if not ((kind == "hex8" and hasattr(mesh, "grid")) or kind == "tet4"):
            raise NotImplementedError("the nodal Helmholtz filter is built for structured "
                                      "hex8 grids and tet4 meshes")
        self.kind = kind
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.mesh = mesh
        self.elements_volume = elements_volume
        self.radius = float(radius)
        self.design_mask = None if design_mask is None else np.asarray(design_mask, dtype=bool)
        self.rtol = float(rtol)
        self.precision = precision
        dev = _dev.device()
        if kind == "hex8":
            self.grid = tuple(int(v) for v in mesh.grid)
            self.h = tuple(float(v) for v in mesh.h)
            nx, ny, nz = self.grid
            self.n_el = nx * ny * nz
            self.n_nodes = (nx + 1) * (ny + 1) * (nz + 1)
        else:
            self.n_el = int(mesh.nelements)
            self.n_nodes = int(mesh.nvertices)
            self._coords = _dev.to_dev(mesh.p)
            self._conn = _dev.to_dev(mesh.t, torch.int32)
        self._vol = _dev.to_dev(np.asarray(elements_volume, dtype=np.float64))
        self._design = None if self.design_mask is None else \
            torch.as_tensor(self.design_mask.astype(np.uint8), device=dev)
        # nodes of non-design elements carry x = 1 in the forward solve
        fixed = np.zeros(self.n_nodes, dtype=np.uint8)
        if self.design_mask is not None and (~self.design_mask).any():
            fixed[np.unique(mesh.t[:, ~self.design_mask].ravel())] = 1
        self._fixed_fwd = torch.as_tensor(fixed, device=dev)
        self._fixed_ids = torch.as_tensor(np.nonzero(fixed)[0], dtype=torch.int64, device=dev)
        self._n_fixed = int(fixed.sum())
        self._fixed_none = torch.zeros(self.n_nodes, dtype=torch.uint8, device=dev)
        self._ones64 = torch.ones(self.n_el, dtype=torch.float64, device=dev)
        self._ones32 = torch.ones(self.n_el, dtype=torch.float32, device=dev)
        self._mass = self._scalar_op(1.0, 0.0, self._fixed_none)
        self._ops = {}       # case -> (operator, solver, max_cols) for the current radius
        self._built_radius = None
        self._warm = {}      # case -> previous nodal solution (warm start)
        self.last_solve = {}

    @classmethod
    def from_defaults(cls, mesh, elements_volume, radius=0.3, design_mask=None):
        return cls(mesh, elements_volume, radius, design_mask)

    def update_radius(self, radius, **args):
        self.radius = float(radius)

    # --- device objects -----------------------------------------------
    def _scalar_op(self, a, b, fixed):
        h = C.c_void_p()
        if self.kind == "hex8":
            nx, ny, nz = self.grid
            _lib.call("skt_hex8_create_scalar", nx, ny, nz, *self.h, float(a), float(b),
                      _dev.ptr(fixed), C.byref(h))
        else:
            _lib.call("skt_tet4_create_scalar", self.n_nodes, self.n_el, _dev.ptr(self._coords),
                      _dev.ptr(self._conn), float(a), float(b), _dev.ptr(fixed), C.byref(h))
        return h

    def _to_nodes(self, k, fixed_value, x, rn, S):
        design = _dev.ptr(self._design) if self._design is not None else None
        if self.kind == "hex8":
            nx, ny, nz = self.grid
            _lib.call("skt_hex8_elem_to_node", nx, ny, nz, k, _dev.ptr(self._vol), design,
                      fixed_value, _dev.ptr(x), _dev.ptr(rn), S)
        else:
            _lib.call("skt_tet4_elem_to_node", self._mass, k, _dev.ptr(self._vol), design,
                      fixed_value, _dev.ptr(x), _dev.ptr(rn), S)

    def _to_elements(self, k, U, clamp, mean_cols, out, S):
        if self.kind == "hex8":
            nx, ny, nz = self.grid
            _lib.call("skt_hex8_node_to_elem", nx, ny, nz, k, _dev.ptr(U), clamp, mean_cols,
                      _dev.ptr(out), S)
        else:
            _lib.call("skt_tet4_node_to_elem", self._mass, k, _dev.ptr(U), clamp, mean_cols,
                      _dev.ptr(out), S)

    def _ensure(self, case, ncols):
        if self._built_radius != self.radius:
            self._destroy_ops()
            self._built_radius = self.radius
        have = self._ops.get(case)
        if have is not None and have[2] >= ncols:
            return have
        if have is not None:
            self._destroy_one(case)
        fixed = self._fixed_fwd if case == "forward" else self._fixed_none
        op = self._scalar_op(1.0, self.radius ** 2, fixed)
        s = C.c_void_p()
        _lib.call("skt_solver_create", op, int(ncols), C.byref(s))
        self._ops[case] = (op, s, int(ncols))
        return self._ops[case]

    def _destroy_one(self, case):
        op, s, _ = self._ops.pop(case)
        lib = _lib.load()
        lib.skt_solver_destroy(s)
        lib.skt_operator_destroy(op)

    def _destroy_ops(self):
        for case in list(self._ops):
            self._destroy_one(case)

    def __del__(self):
        try:
            self._destroy_ops()
            if getattr(self, "_mass", None):
                _lib.load().skt_operator_destroy(self._mass)
        except Exception:
            pass

    # --- the filter ----------------------------------------------------
    def _apply(self, case, x, out, mean_cols=False):
        """x: (n_el,) or (n_el, k) float64 device tensor; out: (n_el,) when
        x is one column or mean_cols, else (n_el, k)."""
        k = 1 if x.dim() == 1 else int(x.shape[1])
        op, solver, _ = self._ensure(case, k)
        S = _dev.stream()
        dev = x.device
        rn = torch.empty((self.n_nodes, k), dtype=torch.float64, device=dev)
        self._to_nodes(k, 1.0 if case == "forward" else 0.0, x, rn, S)
        b = torch.empty((self.n_nodes, k), dtype=torch.float64, device=dev)
        _lib.call("skt_operator_apply", self._mass, _lib.F64, k, _dev.ptr(self._ones64),
                  _dev.ptr(rn), _dev.ptr(b), S)
        if case == "forward" and self._n_fixed:
            b.index_fill_(0, self._fixed_ids, 1.0)     # prescribed x_D = 1 (enforced rhs)
        U = self._warm.get(case)
        if U is None or U.shape[1] != k:
            U = torch.zeros((self.n_nodes, k), dtype=torch.float64, device=dev)
            self._warm[case] = U
        mode = _lib.F32 if self.precision == "mixed" else _lib.F64
        iters = (C.c_int64 * 2)()
        relres = (C.c_double * k)()      # one final residual per column
        st = _lib.load().skt_solver_solve(
            solver, mode, k, _dev.ptr(self._ones64),
            _dev.ptr(self._ones32) if mode == _lib.F32 else None, _dev.ptr(b), _dev.ptr(U),
            self.rtol, 100000, 0.0, iters, relres, S)
        self.last_solve = dict(case=case, iters=int(iters[0]), relres=max(relres))
        _lib.check(st, "skt_solver_solve")
        self._to_elements(k, U, 1 if case == "gradient" else 0,
                          1 if (mean_cols or k == 1) else 0, out, S)
        return out

    def _run(self, case, v):
        host = _dev.is_host(v)
        x = _dev.to_dev(v).contiguous()
        out = torch.empty(x.shape, dtype=torch.float64, device=x.device)
        self._apply(case, x, out)
        return out.cpu().numpy() if host else out

    def forward(self, rho_element):
        """``helmholtz_filter_nodal.py:183-201``."""
        return self._run("forward", rho_element)

    def gradient(self, v_ele):
        """``helmholtz_filter_nodal.py:203-233`` (clamped to <= 0)."""
        return self._run("gradient", v_ele)

    # --- device interface of the density loop ------------------------------
    linear_adjoint = False    # the gradient clamps: applied per load case

    def forward_into(self, x, y):
        self._apply("forward", x, y)

    def gradient_into(self, v, g):
        self._apply("gradient", v, g)

    def gradient_mean_into(self, G, g):
        """g = mean over the columns of G (n_el, k) of gradient(G[:, j]),
        the per-load-case accumulation of common_density.py:1075-1101, with
        the k solves batched in one multi-column solve."""
        self._apply("gradient", G, g, mean_cols=True)
