"""Reverse-mode differentiation of the KernelMatrix product, u = H(x, y) b.

`kernel_matvec` is a torch.autograd.Function: the role ChainRulesCore's
`rrule` plays for a Julia caller of the reference. Its pullback, for an output
cotangent r, returns

  grad_b = H^T r                          (transpose product)
  grad_x = r .* d/dx_i (H b)_i            (hm_matvec_grad_x)
  grad_y = b .* d/dy_j (H^T r)_j          (hm_matvec_t_grad, same pass as grad_b)

These are the exact derivatives of the computed product: the block tree (which
depends on x and y only through comparisons) is held fixed, the near field is
regenerated with df/dx, and the far field is differentiated through its
Chebyshev-form basis. The forward plan is matrix-free, so nothing is stored
between forward and backward beyond the points and b.
"""
import torch

from .kernelmatrix import KernelMatrix


class _KernelMatVec(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y, b, kernel, intervals, blocksize, rank):
        H = KernelMatrix(kernel, x.detach().cpu().numpy(), y.detach().cpu().numpy(), *intervals,
                         blocksize=blocksize, rank=rank, matrix_free=True)
        b = b.detach().contiguous()
        u = H.matvec(b)
        ctx.H = H
        ctx.save_for_backward(b)
        return u

    @staticmethod
    def backward(ctx, r):
        (b,) = ctx.saved_tensors
        H = ctx.H
        r = r.contiguous()
        gx = H.grad_x(b, r) if ctx.needs_input_grad[0] else None
        gb, gy = H.tgrad(r, b)
        return (gx, gy if ctx.needs_input_grad[1] else None,
                gb if ctx.needs_input_grad[2] else None, None, None, None, None)


def kernel_matvec(kernel, x, y, b, intervals, blocksize, rank):
    """u = H(x, y) b with H = KernelMatrix(kernel, x, y, *intervals) (x, y
    non-increasing CUDA tensors, float64 or float32: the element type of x
    selects a Float64 or a Float32 plan); differentiable in x, y and b."""
    return _KernelMatVec.apply(x, y, b, kernel, tuple(intervals), blocksize, rank)
