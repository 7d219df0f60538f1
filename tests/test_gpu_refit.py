"""hm_plan_update_points: new point values for a plan's tree, checked and
refitted on the device.

- Points moved within their clusters (the gradient workload's perturbation
  never crosses a split): status "same tree", and every product (forward,
  transpose, 16 right-hand sides, gradients) equals that of a plan built from
  the new points on the host, bit for bit (the same kernels on the same
  factors: dense blocks and bases are formed again, F depends on the
  intervals only). Stored, hybrid and matrix-free; Float64 and Float32.
- A point moved across a split, or a pair out of order: status "tree
  changed".
- The autograd function reuses its plans this way: gradients of consecutive
  steps with moving points equal those of freshly built plans."""
import numpy as np
import pytest
import torch

import hmx
from hmx import autodiff
from hmx.autodiff import kernel_matvec

pytestmark = pytest.mark.gpu


def same(a, b):
    """Bit for bit, NaN where the other has NaN: a Float32 point can land on
    a Chebyshev node exactly, and the basis then holds lam/0 as the
    reference's does (in both plans alike)."""
    return a.shape == b.shape and bool(((a == b) | (torch.isnan(a) & torch.isnan(b))).all())


def moved(w, seed, n, dtype):
    rng = np.random.default_rng(seed)
    x = w["x"] + (rng.random(n) - 0.5) * 0.2 / n
    y = w["y"] + (rng.random(n) - 0.5) * 0.2 / n
    return x.astype(dtype), y.astype(dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("mode", ["stored", "hybrid", "free"])
def test_refit_equals_rebuild(cuda, dtype, mode):
    n = 4096
    w = hmx.workloads.config_plain(n=n, blocksize=64, rank=16)
    kw = dict(blocksize=64, rank=16, matrix_free=mode == "free", hybrid=mode == "hybrid")
    x0, y0 = w["x"].astype(dtype), w["y"].astype(dtype)
    H = hmx.KernelMatrix("cauchykernel", x0, y0, *w["intervals"], **kw)
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    v = torch.from_numpy(w["b"]).to(cuda).to(tdt)
    V16 = torch.from_numpy(np.random.default_rng(2).standard_normal((n, 16))).to(cuda).to(tdt)
    for step in range(3):
        x1, y1 = moved(w, step, n, dtype)
        ok = H.update_points(torch.from_numpy(x1).to(cuda), torch.from_numpy(y1).to(cuda))
        assert ok
        R = hmx.KernelMatrix("cauchykernel", x1, y1, *w["intervals"], **kw)
        assert same(H @ v, R @ v)
        assert same(H.matvec(v, transpose=True), R.matvec(v, transpose=True))
        if dtype == np.float64:
            assert same(H @ V16, R @ V16)
        s = v.flip(0).contiguous()
        assert same(H.grad_x(v, s), R.grad_x(v, s))
        R.close()


def test_refit_detects_a_changed_tree(cuda):
    n = 4096
    w = hmx.workloads.config_plain(n=n, blocksize=64, rank=16)
    H = hmx.KernelMatrix("cauchykernel", w["x"], w["y"], *w["intervals"], blocksize=64, rank=16,
                         matrix_free=True)
    x = torch.from_numpy(w["x"]).to(cuda)
    y = torch.from_numpy(w["y"]).to(cuda)
    assert H.update_points(x, y)
    # x_i = (i - 1/4)/n: the point just above 1/2 moved below it crosses the
    # root split
    i = int(np.argmin(np.abs(w["x"] - 0.5)))
    xs = x.clone()
    xs[i] = 0.5 - 0.1 / n if w["x"][i] > 0.5 else 0.5 + 0.1 / n
    assert not H.update_points(xs, y)
    # two points out of order
    ys = y.clone()
    ys[10], ys[11] = y[11], y[10]
    assert not H.update_points(x, ys)
    # and back: the same tree again
    assert H.update_points(x, y)
    # a non-finite point
    xs = x.clone()
    xs[5] = float("nan")
    assert not H.update_points(xs, y)


def test_autograd_reuses_plans(cuda):
    n = 4096
    w = hmx.workloads.config_gradient(n=n, blocksize=64, rank=16)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    b, t = T(w["b"]), T(w["t"])

    def grads(x, y):
        xg, yg, bg = (v.clone().requires_grad_(True) for v in (x, y, b))
        u = kernel_matvec("cauchykernel", xg, yg, bg, w["intervals"], 64, 16)
        (0.5 * torch.sum((u - t) ** 2)).backward()
        return u.detach(), xg.grad, yg.grad, bg.grad

    pts = [moved(w, s, n, np.float64) for s in range(3)]
    autodiff._POOL.clear()
    fresh = []
    for x1, y1 in pts:
        autodiff._POOL.clear()  # every step builds its plan on the host
        fresh.append(grads(T(x1), T(y1)))
    autodiff._POOL.clear()
    for (x1, y1), want in zip(pts, fresh):  # steps after the first refit a pooled plan
        got = grads(T(x1), T(y1))
        for a, c in zip(got, want):
            assert torch.equal(a, c)
    assert sum(len(v) for v in autodiff._POOL.free.values()) == 1
