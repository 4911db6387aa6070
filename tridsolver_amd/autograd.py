"""Differentiable batched tridiagonal solve: a torch.autograd.Function whose
forward is tridXmtsvStridedBatch and whose backward is the library's adjoint
kernel (tridXmtsvStridedBatchAdjoint).

    x = TridSolve.apply(a, b, c, d, solvedim)

a, b, c, d are dense tensors of one shape (x fastest = last torch dimension,
so torch shape [nz, ny, nx] is the C API's dims [nx, ny, nz]); gradients are
returned for every input that requires one. The forward leaves d untouched
(it solves a copy); only the gradients the autograd graph needs are formed.
"""
from __future__ import annotations

import torch

from . import api


class TridSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a, b, c, d, solvedim):
        tensors = [t.contiguous() for t in (a, b, c, d)]
        a, b, c, d = tensors
        dims = list(d.shape[::-1])
        # torch dimension k is C dimension ndim-1-k
        x = d.clone()
        api.trid_strided_batch(a, b, c, x, solvedim=solvedim, dims=dims)
        ctx.solvedim = solvedim
        ctx.save_for_backward(a, b, c, x)
        return x

    @staticmethod
    def backward(ctx, gx):
        a, b, c, x = ctx.saved_tensors
        need = ctx.needs_input_grad
        lam, ga, gb, gc = api.adjoint_strided_batch(
            a, b, c, x, gx.contiguous(), solvedim=ctx.solvedim, dims=list(x.shape[::-1]),
            grad_a=True if need[0] else None, grad_b=True if need[1] else None,
            grad_c=True if need[2] else None)
        return ga, gb, gc, (lam if need[3] else None), None


def trid_solve(a, b, c, d, solvedim):
    """x = A^-1 d along `solvedim` (C API numbering: 0 = x = last torch dim)."""
    return TridSolve.apply(a, b, c, d, solvedim)
