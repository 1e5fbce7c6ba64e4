"""torch.autograd front end of the gis solve and its hand-written adjoint.

    x = solve(gis, diag, lower, upper, b)        # differentiable in all four

Forward: one device solve of A x = b (CG for symmetric A, BiCG otherwise).
Backward: gis_adjoint -- A^T lambda = dL/dx on the device-built transpose,
dL/db = lambda, dL/dA_ij = -lambda_i x_j on the pattern folded to the LDU
coefficients (diag, lower, upper).  The matrix pattern (face addressing) must
already be set on the handle; this module only moves coefficients and vectors.
A PyTorch-ROCm harness for tests and gradient checks: every array is a float64
tensor on the handle's device.
"""
from __future__ import annotations

import torch

from .gis import Gis


class GisSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, diag, lower, upper, b, gis: Gis, method: str, error: float, max_steps: int):
        gis.set_values(diag.detach().contiguous(), lower.detach().contiguous(),
                       upper.detach().contiguous())
        x = torch.zeros_like(b)
        x, _, _ = gis.solve(method, b.detach().contiguous(), x, error=error,
                            max_steps=max_steps, hist_cap=1)
        ctx.gis, ctx.method, ctx.error, ctx.max_steps = gis, method, error, max_steps
        ctx.save_for_backward(diag, lower, upper, x)
        return x

    @staticmethod
    def backward(ctx, grad_x):
        diag, lower, upper, x = ctx.saved_tensors
        if not torch.any(grad_x):
            # A^T lambda = 0 has lambda = 0; BiCG itself would form 0/0 on a zero
            # right-hand side (as the reference's loop does), so skip the solve
            return (torch.zeros_like(diag), torch.zeros_like(lower), torch.zeros_like(upper),
                    torch.zeros_like(x), None, None, None, None)
        g = ctx.gis
        # the handle may have been re-valued since forward: restore this system
        g.set_values(diag.detach().contiguous(), lower.detach().contiguous(),
                     upper.detach().contiguous())
        lam, dd, dlo, dup, _ = g.adjoint(grad_x.detach().contiguous(), x, method=ctx.method,
                                         error=ctx.error, max_steps=ctx.max_steps)
        return dd[0], dlo[0], dup[0], lam[0], None, None, None, None


def solve(gis: Gis, diag, lower, upper, b, method: str | None = None, error: float = 0.0,
          max_steps: int = 1000):
    """x = A(diag, lower, upper)^-1 b, differentiable w.r.t. all four tensors.

    A fixed iteration count (error = 0) makes x a smooth function of the inputs,
    which finite-difference checks need; error > 0 stops on ||r||_2 <= error."""
    if max_steps <= 0 and not error > 0:
        raise ValueError("need max_steps > 0 or error > 0 (the solve would not stop)")
    if method is None:
        method = "CG" if torch.equal(lower, upper) else "BiCG"
    return GisSolve.apply(diag, lower, upper, b, gis, method, error, max_steps)
