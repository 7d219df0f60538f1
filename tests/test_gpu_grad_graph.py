"""KernelProduct: the differentiable product with a held plan and the tree
check left on the device, eager and captured into a CUDA graph.

- Eager steps with moving points (the gradient workload's perturbation, which
  keeps every split) give exactly the forward values and gradients of
  kernel_matvec (the same kernels on the same plan data), and tree_ok() is
  True.
- A step captured with torch.cuda.graph and replayed after the points, b and
  t are overwritten in place gives, on every replay, exactly the eager step's
  values and gradients for those inputs.
- A point moved across a split: tree_ok() turns False (and clears); after
  rebuild the product equals kernel_matvec's, which builds the new tree.
Float64 and Float32."""
import numpy as np
import pytest
import torch

import hmx
from hmx.autodiff import KernelProduct, kernel_matvec

pytestmark = pytest.mark.gpu


def moved(w, seed, n):
    rng = np.random.default_rng(seed)
    x = w["x"] + (rng.random(n) - 0.5) * 0.2 / n
    y = w["y"] + (rng.random(n) - 0.5) * 0.2 / n
    return x, y


def grads(fn, x, y, b, t):
    xg, yg, bg = (v.clone().requires_grad_(True) for v in (x, y, b))
    u = fn(xg, yg, bg)
    (0.5 * torch.sum((u - t) ** 2)).backward()
    return u.detach(), xg.grad, yg.grad, bg.grad


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_eager_equals_kernel_matvec(cuda, dtype):
    n = 4096
    w = hmx.workloads.config_gradient(n=n, blocksize=64, rank=16)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda).to(dtype)
    b, t = T(w["b"]), T(w["t"])
    kp = KernelProduct("cauchykernel", T(w["x"]), T(w["y"]), w["intervals"], 64, 16)
    for s in range(3):
        x1, y1 = (T(v) for v in moved(w, s, n))
        got = grads(kp, x1, y1, b, t)
        want = grads(lambda a, c, d: kernel_matvec("cauchykernel", a, c, d, w["intervals"], 64, 16),
                     x1, y1, b, t)
        for g, h in zip(got, want):
            assert torch.equal(g, h)
    assert kp.tree_ok()
    kp.close()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_graph_replays_equal_eager(cuda, dtype):
    n = 4096
    w = hmx.workloads.config_gradient(n=n, blocksize=64, rank=16)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda).to(dtype)
    kp = KernelProduct("cauchykernel", T(w["x"]), T(w["y"]), w["intervals"], 64, 16)
    # static inputs of the captured step
    x, y, b, t = T(w["x"]), T(w["y"]), T(w["b"]), T(w["t"])
    for v in (x, y, b):
        v.requires_grad_(True)

    def step():
        x.grad = y.grad = b.grad = None
        u = kp(x, y, b)
        (0.5 * torch.sum((u - t) ** 2)).backward()
        return u

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(3):
            step()
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        u_static = step()
    gx, gy, gb = x.grad, y.grad, b.grad
    rng = np.random.default_rng(7)
    for s in range(4):
        x1, y1 = (T(v) for v in moved(w, 10 + s, n))
        b1 = T(rng.standard_normal(n))
        with torch.no_grad():
            x.copy_(x1)
            y.copy_(y1)
            b.copy_(b1)
        g.replay()
        torch.cuda.synchronize()
        want = grads(kp, x1, y1, b1, t)
        for got, h in zip((u_static, gx, gy, gb), want):
            assert torch.equal(got, h), s
        # eager work between replays
        _ = kp(x1, y1, b1)
        torch.cuda.synchronize()
    assert kp.tree_ok()
    del g
    kp.close()


def test_tree_change_is_reported_and_rebuilt(cuda):
    n = 4096
    w = hmx.workloads.config_gradient(n=n, blocksize=64, rank=16)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    b = T(w["b"])
    x, y = T(w["x"]), T(w["y"])
    kp = KernelProduct("cauchykernel", x, y, w["intervals"], 64, 16)
    kp(x, y, b)
    assert kp.tree_ok()
    # the point nearest 1/2 moved across the root split
    i = int(np.argmin(np.abs(w["x"] - 0.5)))
    xs = x.clone()
    xs[i] = 0.5 - 0.1 / n if w["x"][i] > 0.5 else 0.5 + 0.1 / n
    kp(xs, y, b)
    assert not kp.tree_ok()
    assert kp.tree_ok()  # cleared by the read
    kp.rebuild(xs, y)
    u = kp(xs, y, b)
    assert kp.tree_ok()
    want = kernel_matvec("cauchykernel", xs, y, b, w["intervals"], 64, 16)
    assert torch.equal(u, want)
    kp.close()
