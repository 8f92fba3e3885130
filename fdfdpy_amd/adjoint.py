"""Differentiable eps_r -> fields solve with an adjoint backward pass.

No reference counterpart (fdfdpy has no gradients); dA/deps follows the Ez
assembly of linalg.py:53-66, where eps enters only the diagonal
omega^2 eps0 L0 eps. For a real objective J and the forward field E with
A E = b, torch hands backward G = dJ/dconj(E) * 2, and

    dJ/deps_p = -Re(lam_p * omega^2 eps0 L0 * E_p),   A^T lam = conj(G).

The transpose solve reuses the forward solve's resident factor through
A^T = D A D^-1 (D the stretch symmetrizer), and the product is formed by the
fdfd_grad_eps kernel.
"""

from __future__ import annotations

import torch

from .device import DeviceOperator


class _EpsToField(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps, b, op, precision, tol):
        op.set_eps(eps.detach())
        x, rep = op.solve(b, precision=precision, tol=tol)
        ctx.op = op
        ctx.version = op.version
        ctx.precision = precision
        ctx.tol = tol
        ctx.eps = eps.detach()
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, grad_x):
        (x,) = ctx.saved_tensors
        op = ctx.op
        if op.version != ctx.version:
            op.set_eps(ctx.eps)      # operator changed since forward: rebuild
        g = torch.conj_physical(grad_x.to(torch.complex128)).contiguous()
        if not bool(torch.any(g != 0)):
            return torch.zeros_like(ctx.eps, dtype=torch.float64), None, None, None, None
        lam, _ = op.solve(g, precision=ctx.precision, tol=ctx.tol, transpose=True)
        grad = op.grad_eps(lam, x)
        return grad.to(ctx.eps.dtype), None, None, None, None


class FieldSolver:
    """eps_r (real tensor, requires_grad) -> Ez for a fixed source (pol='Ez')."""

    def __init__(self, shape, omega, dl, NPML, L0=1e-6, device=None, precision='fp64', tol=1e-12):
        self.op = DeviceOperator(shape, omega, dl, NPML, 'Ez', L0, device)
        self.omega = float(omega)
        self.precision = precision
        self.tol = tol

    def __call__(self, eps, src):
        """Ez = A(eps)^-1 (1j omega src), differentiable with respect to eps."""
        src = torch.as_tensor(src, device=self.op.device).to(torch.complex128)
        b = src * (1j * self.omega)
        return _EpsToField.apply(eps, b, self.op, self.precision, self.tol)
