"""Reverse-mode differentiation of the KernelMatrix product, u = H(x, y) b.

`kernel_matvec` is a torch.autograd.Function: the role ChainRulesCore's
`rrule` plays for a Julia caller of the reference. Its pullback, for an output
cotangent r, returns

  grad_b = H^T r                          (transpose product)
  grad_x = r .* d/dx_i (H b)_i            (the sums d/dx_i (H b)_i come out of
                                           the forward pass, hm_matvec_grad)
  grad_y = b .* d/dy_j (H^T r)_j          (hm_matvec_t_grad, same pass as grad_b)

These are the exact derivatives of the computed product: the block tree (which
depends on x and y only through comparisons) is held fixed, the near field is
regenerated with df/dx, and the far field is differentiated through its
Chebyshev-form basis. The forward plan is matrix-free, so nothing is stored
between forward and backward beyond the points, b and (when x needs a
gradient) the derivative sums: the forward then runs hm_matvec_grad, one pass
for H b and d/dx_i (H b)_i, and the backward's x-gradient is r .* those sums
with no pass of its own.

Plans are reused from step to step: a forward takes a plan of the same shape
(points per side, kernel, intervals, leaf threshold, rank, element type) from
a small pool and gives it the new points on the device
(hm_plan_update_points), which checks there that they keep the tree; only
when they do not (or on the first step) is a plan built from host copies of
the points. A plan stays with its autograd context until that context is
freed (the graph and the output that holds it are gone), so two live graphs
never share one and a retained graph can be differentiated again.

kernel_matvec reads the device's tree check back on every forward (one
synchronisation). KernelProduct holds one plan and leaves the check on the
device until asked (tree_ok), so that a whole step can be captured into a
CUDA graph and replayed.
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


def _product(H, b, dx):
    """u = H b, and with dx the sums d/dx_i (H b)_i from the same pass."""
    return H.matvec_grad(b) if dx else (H.matvec(b), None)


def _forward_plan(x, y, b, kernel, intervals, blocksize, rank, dx):
    key = (str(kernel), x.numel(), y.numel(), tuple(float(v) for v in intervals), blocksize,
           rank, x.dtype)
    H = _POOL.take(key)
    u = d = None
    if H is not None:
        H.update_points(x, y, check=False)
        u, d = _product(H, b, dx)
        if not H.tree_status():  # read after the product is queued
            H.close()
            H = u = d = None
    if H is None:
        H = KernelMatrix(kernel, x.cpu().numpy(), y.cpu().numpy(), *intervals,
                         blocksize=blocksize, rank=rank, matrix_free=True)
        u, d = _product(H, b, dx)
    return key, H, u, d


class _KernelMatVec(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y, b, kernel, intervals, blocksize, rank):
        x, y, b = (v.detach().contiguous() for v in (x, y, b))
        key, H, u, d = _forward_plan(x, y, b, kernel, intervals, blocksize, rank,
                                     ctx.needs_input_grad[0])
        ctx.H = H
        # back to the pool once the graph is gone (stream-ordered: products
        # that reuse it run after this graph's work)
        weakref.finalize(ctx, _POOL.give, key, H)
        ctx.save_for_backward(b, d)
        return u

    @staticmethod
    def backward(ctx, r):
        b, d = ctx.saved_tensors
        H = ctx.H
        r = r.contiguous()
        gx = r * d if ctx.needs_input_grad[0] else None
        gb, gy = H.tgrad(r, b)
        return (gx, gy if ctx.needs_input_grad[1] else None,
                gb if ctx.needs_input_grad[2] else None, None, None, None, None)


class _KernelProductFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y, b, kp):
        x, y, b = (v.detach().contiguous() for v in (x, y, b))
        H = kp.H
        H.update_points(x, y, check=False)
        torch.bitwise_or(kp._bad, H._status, out=kp._bad)
        u, d = _product(H, b, ctx.needs_input_grad[0])
        ctx.H = H
        ctx.save_for_backward(b, d)
        return u

    @staticmethod
    def backward(ctx, r):
        b, d = ctx.saved_tensors
        H = ctx.H
        r = r.contiguous()
        gx = r * d if ctx.needs_input_grad[0] else None
        gb, gy = H.tgrad(r, b)
        return (gx, gy if ctx.needs_input_grad[1] else None,
                gb if ctx.needs_input_grad[2] else None, None)


class KernelProduct:
    """u = H(x, y) b, differentiable in x, y and b, with one matrix-free plan
    held by the object (built from the first points on the host).

    Every call gives the plan the new points on the device
    (hm_plan_update_points) and records there whether they keep the plan's
    tree; nothing is read back and nothing is allocated outside torch, so a
    training step that calls it (forward and backward) can be captured into a
    CUDA graph (torch.cuda.graph) and replayed. tree_ok() reads, with one
    synchronisation, whether every call since the previous tree_ok() kept the
    tree: when one did not, the products of those calls followed the old tree
    (not the reference's for the new points), and rebuild(x, y) gives the
    object a plan built for the new points (a graph captured before must be
    captured again). kernel_matvec makes that check on every call instead."""

    def __init__(self, kernel, x, y, intervals, blocksize, rank):
        self.kernel, self.intervals = kernel, tuple(float(v) for v in intervals)
        self.blocksize, self.rank = blocksize, rank
        self.H = None
        self.rebuild(x, y)

    def rebuild(self, x, y):
        if self.H is not None:
            self.H.close()
        self.H = KernelMatrix(self.kernel, x.detach().cpu().numpy(), y.detach().cpu().numpy(),
                              *self.intervals, blocksize=self.blocksize, rank=self.rank,
                              matrix_free=True)
        self.H._status = torch.zeros(1, dtype=torch.int32, device=x.device)
        self._bad = torch.zeros(1, dtype=torch.int32, device=x.device)

    def __call__(self, x, y, b):
        return _KernelProductFn.apply(x, y, b, self)

    def tree_ok(self):
        ok = int(self._bad.item()) == 0
        self._bad.zero_()
        return ok

    def close(self):
        if self.H is not None:
            self.H.close()
            self.H = None


def kernel_matvec(kernel, x, y, b, intervals, blocksize, rank):
    """u = H(x, y) b with H = KernelMatrix(kernel, x, y, *intervals) (x, y
    non-increasing CUDA tensors, float64 or float32: the element type of x
    selects a Float64 or a Float32 plan); differentiable in x, y and b."""
    return _KernelMatVec.apply(x, y, b, kernel, tuple(intervals), blocksize, rank)
