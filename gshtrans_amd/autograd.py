"""Differentiable synthesis: torch.autograd.Function whose backward pass is
the hand-written adjoint kernel (gsh_adjoint_inverse), for least-squares
problems on the sphere such as min_c ||S_n(c) - d||^2.

The backward pass recomputes the Wigner values with the same on-the-fly FP64
recurrence as the forward pass; nothing is saved between the two except the
call's parameters.

Gradient convention (torch's for complex tensors): for a real loss the
gradient of a complex coefficient is dL/dRe(c) + i dL/dIm(c), i.e. for the
linear synthesis y = S c the backward pass returns S^H(grad_y).  For real
fields S is the real-linear c2r synthesis and S^H its adjoint for the inner
product Re <a, b>.
"""
import torch

__all__ = ["Synthesis", "synthesize"]


class Synthesis(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coeffs, grid, lmax, n, real, precision):
        ctx.grid, ctx.lmax, ctx.n, ctx.real, ctx.precision = grid, lmax, n, real, precision
        return grid.inverse(lmax, n, coeffs.contiguous(), real=real, precision=precision)

    @staticmethod
    def backward(ctx, grad_field):
        g = ctx.grid.adjoint(ctx.lmax, ctx.n, grad_field.contiguous(), real=ctx.real,
                             precision=ctx.precision)
        return g, None, None, None, None, None


def synthesize(grid, lmax, n, coeffs, real=False, precision="fp64"):
    """field = S_n(coeffs) with gradients through the adjoint kernel."""
    return Synthesis.apply(coeffs, grid, lmax, n, real, precision)
