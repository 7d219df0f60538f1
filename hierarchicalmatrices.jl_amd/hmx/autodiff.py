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

Plans are reused from step to step: a forward takes a plan of the same shape
(points per side, kernel, intervals, leaf threshold, rank, element type) from
a small pool and gives it the new points on the device
(hm_plan_update_points), which checks there that they keep the tree; only
when they do not (or on the first step) is a plan built from host copies of
the points. A plan stays with its autograd context until that context is
freed (the graph and the output that holds it are gone), so two live graphs
never share one and a retained graph can be differentiated again.
"""
import weakref

import torch

from .kernelmatrix import KernelMatrix


class _PlanPool:
    def __init__(self, per_key=4):
        self.free = {}
        self.per_key = per_key

    def take(self, key):
        lst = self.free.get(key)
        return lst.pop() if lst else None

    def give(self, key, H):
        lst = self.free.setdefault(key, [])
        if len(lst) < self.per_key:
            lst.append(H)
        else:
            H.close()

    def clear(self):
        for lst in self.free.values():
            for H in lst:
                H.close()
        self.free.clear()


_POOL = _PlanPool()


def _forward_plan(x, y, b, kernel, intervals, blocksize, rank):
    key = (str(kernel), x.numel(), y.numel(), tuple(float(v) for v in intervals), blocksize,
           rank, x.dtype)
    H = _POOL.take(key)
    u = None
    if H is not None:
        H.update_points(x, y, check=False)
        u = H.matvec(b)
        if not H.tree_status():  # read after the product is queued
            H.close()
            H = u = None
    if H is None:
        H = KernelMatrix(kernel, x.cpu().numpy(), y.cpu().numpy(), *intervals,
                         blocksize=blocksize, rank=rank, matrix_free=True)
        u = H.matvec(b)
    return key, H, u


class _KernelMatVec(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y, b, kernel, intervals, blocksize, rank):
        x, y, b = (v.detach().contiguous() for v in (x, y, b))
        key, H, u = _forward_plan(x, y, b, kernel, intervals, blocksize, rank)
        ctx.H = H
        # back to the pool once the graph is gone (stream-ordered: products
        # that reuse it run after this graph's work)
        weakref.finalize(ctx, _POOL.give, key, H)
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
