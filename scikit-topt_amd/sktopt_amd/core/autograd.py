"""Differentiable compliance: filter -> Heaviside projection -> SIMP/RAMP
interpolation -> state solve -> compliance, as a ``torch.autograd.Function``.

Forward is the chain the density loop runs each iteration
(``core/optimizers/common_density.py:1038-1064``):

    rho_f = W rho                     (``filters/spacial.py:135-157``)
    rho_p = H_beta,eta(rho_f)         (``core/projection.py:27-77``)
    E     = interp(rho_p)             (``fea/composer.py:19-39``)
    K(E) u_k = f_k,  C = mean_k f_k . u_k

and backward is its gradient (``common_density.py:1073-1108``):

    dC/drho = W^T ( (1/k) sum_k dC_k/drho_p * H'(rho_f) ),
    dC_k/drho_p,e = -u_{k,e}^T dK_e/drho_p u_{k,e} = -2 U_{k,e} E'_e / E_e.

Compliance is self-adjoint, so the backward reuses the displacement of the
forward solve instead of solving an adjoint system, and it recomputes the
element strain energies U_{k,e} from that displacement (one matrix-free
energy pass) instead of storing them.  All of it runs in the library's HIP
kernels on the device.  The RAMP branch uses the exact derivative
(1 + p) / (1 + p (1 - rho))^2, not the expression of
``core/derivatives.py:56-61`` that the optimizer reproduces, so that the
gradient is the derivative of the forward map.
"""
from __future__ import annotations

import torch

from .. import _dev, _lib


class ComplianceProblem:
    """Everything fixed while rho varies: the FEA object (its operator,
    solver and load block), the filter, the projection parameters and the
    penalisation.  ``U`` holds the last displacement block (ndof, k) and is
    the warm start of the next solve."""

    def __init__(self, fem, filt, beta: float, eta: float = 0.5, p: float = 3.0):
        # the backward below is the self-adjoint compliance gradient; other
        # objectives (heat averaged_temp / heat_exchange) and the rho-dependent
        # Robin terms need an adjoint solve and extra terms it does not have
        if fem.physics == _lib.HEAT and (getattr(fem, "objective", "compliance") != "compliance"
                                         or getattr(fem, "robin", False)):
            raise NotImplementedError(
                "ComplianceProblem supports elastic compliance and heat compliance without "
                "Robin terms")
        self.fem = fem
        self.filter = filt
        self.beta = float(beta)
        self.eta = float(eta)
        self.p = float(p)
        self.U = torch.zeros((fem.ndof, fem.n_tasks), dtype=torch.float64, device=fem.device)

    @property
    def sensitivity_kind(self) -> int:
        # the derivative of the interpolation the forward pass used (the
        # optimizer instead reproduces the reference's SIMP-only heat
        # sensitivity, fea/solver_heat.py:945-947, and its RAMP expression)
        return _lib.SIMP if self.fem.kind == _lib.SIMP else _lib.RAMP_EXACT

    def moduli_range(self):
        return self.fem.E_max, self.fem.E_min

    def compliance(self, rho: torch.Tensor) -> torch.Tensor:
        """C(rho) as a differentiable scalar (float64 device tensor)."""
        return Compliance.apply(rho, self)


class Compliance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, prob: ComplianceProblem):
        if rho.dtype != torch.float64 or not rho.is_cuda:
            raise TypeError("rho must be a float64 device tensor")
        rho = rho.contiguous()
        n = rho.numel()
        S = _dev.stream()
        rf = prob.filter.forward(rho)
        rp = torch.empty_like(rf)
        _lib.call("skt_heaviside_project", n, _dev.ptr(rf), prob.beta, prob.eta, _dev.ptr(rp), S)
        comp = prob.fem.objectives_multi_load(rp, prob.p, prob.U)   # (k,) device
        ctx.prob = prob
        # this forward's displacement: prob.U is the next solve's warm start
        # and is overwritten by the next forward
        ctx.save_for_backward(rf, rp, prob.U.clone())
        return comp.mean()

    @staticmethod
    def backward(ctx, grad_out):
        prob = ctx.prob
        rf, rp, U = ctx.saved_tensors
        fem = prob.fem
        n = rf.numel()
        k = fem.n_tasks
        S = _dev.stream()
        # element energies recomputed from the forward displacement
        energy = fem.energy(rp, prob.p, U)
        e_max, e_min = prob.moduli_range()
        g = torch.empty_like(rf)
        _lib.call("skt_sensitivity_chain", prob.sensitivity_kind, n, k, _dev.ptr(rp),
                  _dev.ptr(energy), _dev.ptr(rf), prob.beta, prob.eta, float(e_max),
                  float(e_min), prob.p, 1.0 / k, _dev.ptr(g), S)
        grad = prob.filter.gradient(g)
        return grad * grad_out, None
