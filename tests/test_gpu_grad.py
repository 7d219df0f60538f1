"""Gradients of L(x, y, b) = 1/2 ||H(x, y) b - t||^2 through hmx.autodiff
(the pullback of the KernelMatrix product). No reference function exists for
these; they are pinned by the dense FP64 gradient (within the approximation
error, gate 1e-6 relative) and by central finite differences of the computed
product itself (gate 1e-6 relative).

Finite differences use h = 1e-6 in units of the point spacing 1/n: the closest
x-y pairs are ~0.3/n apart, so an absolute h = 1e-6 at n = 2^16 would sit at a
fifth of that distance and the O((h/d)^2) truncation term would dominate.
"""
import numpy as np
import pytest
import torch

import hmx
import oracle as O
from hmx.autodiff import kernel_matvec

pytestmark = pytest.mark.gpu

DERIV = {  # df/dx of the kernels as functions of d = x - y (df/dy = -df/dx)
    "cauchykernel": (lambda d: 1.0 / d, lambda d: -1.0 / (d * d)),
    "coulombkernel": (lambda d: 1.0 / (d * d), lambda d: -2.0 / (d * d * d)),
    "coulombprimekernel": (lambda d: 1.0 / (d * d * d), lambda d: -3.0 / (d * d * d * d)),
    "logkernel": (lambda d: torch.log(torch.abs(d)), lambda d: 1.0 / d),
}
KID = {"cauchykernel": O.CAUCHY, "coulombkernel": O.COULOMB,
       "coulombprimekernel": O.COULOMB3, "logkernel": O.LOG}


def dense(kernel, x, y, b, r, chunk=4096):
    """Dense FP64 products on the GPU, row tiles: K b, K^T r, (K_x b), (K_y^T r)."""
    f, fx = DERIV[kernel]
    n, m = x.numel(), y.numel()
    Kb = torch.zeros(n, dtype=torch.float64, device=x.device)
    Kxb = torch.zeros_like(Kb)
    KTr = torch.zeros(m, dtype=torch.float64, device=x.device)
    KyTr = torch.zeros_like(KTr)
    for i0 in range(0, n, chunk):
        d = x[i0:i0 + chunk, None] - y[None, :]
        K = f(d)
        Kx = fx(d)
        Kb[i0:i0 + chunk] = K @ b
        Kxb[i0:i0 + chunk] = Kx @ b
        KTr += K.T @ r[i0:i0 + chunk]
        KyTr -= Kx.T @ r[i0:i0 + chunk]
    return Kb, KTr, Kxb, KyTr


def rel(a, b):
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def setup(n, kernel="cauchykernel", bs=64, r=16):
    w = hmx.workloads.config_gradient(n=n, blocksize=bs, rank=r)
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    return w, T(w["x"]), T(w["y"]), T(w["b"]), T(w["t"])


def grads(kernel, w, x, y, b, t):
    xg, yg, bg = (v.clone().requires_grad_(True) for v in (x, y, b))
    u = kernel_matvec(kernel, xg, yg, bg, w["intervals"], w["blocksize"], w["rank"])
    L = 0.5 * torch.sum((u - t) ** 2)
    L.backward()
    return u.detach(), xg.grad, yg.grad, bg.grad


def test_dense_helper_matches_oracle(cuda):
    w, x, y, b, t = setup(2048)
    Kb, _, _, _ = dense("cauchykernel", x, y, b, t)
    rows = np.arange(0, 2048, 97)
    ref = O.dense_matvec(O.CAUCHY, w["x"], w["y"], w["b"], rows=rows)
    assert np.max(np.abs(Kb.cpu().numpy()[rows] - ref) / np.abs(ref)) < 1e-12


@pytest.mark.parametrize("kernel", list(DERIV))
def test_grad_vs_dense_small(cuda, kernel):
    w, x, y, b, t = setup(4096, kernel)
    print(f"case: grad dense n=4096 kernel={kernel}", flush=True)
    u, gx, gy, gb = grads(kernel, w, x, y, b, t)
    Kb, _, _, _ = dense(kernel, x, y, b, t)
    r = Kb - t
    _, KTr, Kxb, KyTr = dense(kernel, x, y, b, r)
    assert rel(u, Kb) < 1e-10
    assert rel(gb, KTr) < 1e-6
    assert rel(gx, r * Kxb) < 1e-6
    assert rel(gy, b * KyTr) < 1e-6


def test_grad_config_n65536_dense_and_fd(cuda):
    """n = 2^16, leaf 64, rank 16, perturbed grid: gradients against the dense
    FP64 gradient and central differences of the computed L."""
    n = 1 << 16
    w, x, y, b, t = setup(n)
    print("case: grad n=2^16", flush=True)
    u, gx, gy, gb = grads("cauchykernel", w, x, y, b, t)
    Kb, _, _, _ = dense("cauchykernel", x, y, b, t)
    r = Kb - t
    _, KTr, Kxb, KyTr = dense("cauchykernel", x, y, b, r)
    errs = dict(u=rel(u, Kb), b=rel(gb, KTr), x=rel(gx, r * Kxb), y=rel(gy, b * KyTr))
    print("dense relative errors", errs, flush=True)
    assert all(v < 1e-6 for v in errs.values())

    iv, bs, rk = w["intervals"], w["blocksize"], w["rank"]

    def prod(xx, yy, bb):
        H = hmx.KernelMatrix("cauchykernel", xx.cpu().numpy(), yy.cpu().numpy(), *iv,
                             blocksize=bs, rank=rk, matrix_free=True)
        return H @ bb

    def dL(up, um):  # L(+) - L(-) without forming the large L itself
        return float(torch.sum(0.5 * (up - um) * (up + um - 2 * t)))

    h = 1e-6 / n
    rng = np.random.default_rng(7)
    for i in rng.integers(0, n, 6):
        e = torch.zeros(n, dtype=torch.float64, device=x.device)
        e[i] = h
        fd = dL(prod(x + e, y, b), prod(x - e, y, b)) / (2 * h)
        assert abs(fd - float(gx[i])) <= 1e-6 * abs(float(gx[i])), ("x", i, fd, float(gx[i]))
        fd = dL(prod(x, y + e, b), prod(x, y - e, b)) / (2 * h)
        assert abs(fd - float(gy[i])) <= 1e-6 * abs(float(gy[i])), ("y", i, fd, float(gy[i]))
        hb = 1e-3
        eb = torch.zeros_like(b)
        eb[i] = hb
        fd = dL(prod(x, y, b + eb), prod(x, y, b - eb)) / (2 * hb)
        assert abs(fd - float(gb[i])) <= 1e-6 * abs(float(gb[i])), ("b", i, fd, float(gb[i]))
    # one random direction in x and y together
    dx = torch.from_numpy(rng.standard_normal(n)).to(x.device) * h
    dy = torch.from_numpy(rng.standard_normal(n)).to(x.device) * h
    fd = dL(prod(x + dx, y + dy, b), prod(x - dx, y - dy, b)) / 2
    an = float(torch.dot(gx, dx) + torch.dot(gy, dy))
    assert abs(fd - an) <= 1e-6 * (float(torch.dot(gx.abs(), dx.abs()) + torch.dot(gy.abs(), dy.abs())))


def test_gradcheck_small(cuda):
    """torch.autograd.gradcheck on a tiny instance (all three inputs)."""
    w, x, y, b, t = setup(200, bs=16, r=8)
    h = 1e-6 / 200

    def f(xx, yy, bb):
        return kernel_matvec("cauchykernel", xx, yy, bb, w["intervals"], 16, 8)

    inputs = (x.clone().requires_grad_(True), y.clone().requires_grad_(True),
              b.clone().requires_grad_(True))
    assert torch.autograd.gradcheck(f, inputs, eps=h, atol=1e-3, rtol=1e-5)


@pytest.mark.parametrize("kernel", ["cauchykernel", "logkernel"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_matvec_grad_one_pass(cuda, kernel, dtype):
    """hm_matvec_grad: H b and s .* d/dx_i (H b)_i from one pass. The sums are
    the ones hm_matvec_grad_x forms (bit for bit: the same arithmetic), the
    product agrees with hm_matvec's (<= 1e-13 FP64, 1e-5 FP32: the near field
    through the kernel-and-derivative routine) and with the dense product;
    beta = 1 accumulates into both outputs; s NULL means ones."""
    import ctypes
    from hmx._lib import lib
    n = 4096
    w, x, y, b, _ = setup(n, kernel)
    print(f"case: matvec_grad kernel={kernel} dtype={dtype}", flush=True)
    np_dt = np.float32 if dtype == torch.float32 else np.float64
    H = hmx.KernelMatrix(kernel, w["x"].astype(np_dt), w["y"].astype(np_dt), *w["intervals"],
                         blocksize=64, rank=16, matrix_free=True)
    bd = b.to(dtype)
    s = torch.from_numpy(np.random.default_rng(5).standard_normal(n)).to(cuda).to(dtype)
    u, g = H.matvec_grad(bd, s)
    assert torch.equal(g, H.grad_x(bd, s))
    u1, g1 = H.matvec_grad(bd)
    assert torch.equal(u1, u)
    assert torch.equal(s * g1, H.grad_x(bd, s))
    tol = 1e-13 if dtype == torch.float64 else 1e-5
    assert rel(u.double(), H.matvec(bd).double()) <= tol
    # the dense product of the plan's own points (Float32-rounded for Float32)
    xp, yp, bp = (v.to(dtype).double() for v in (x, y, b))
    Kb = dense(kernel, xp, yp, bp, torch.zeros_like(xp))[0]
    assert rel(u.double(), Kb) <= (1e-6 if dtype == torch.float64 else 3e-5)
    # beta = 1: y += H b, g += s .* d/dx (H b)
    u0 = torch.ones_like(u)
    g0 = torch.full_like(u, 2.0)
    fn = lib.hm_matvec_grad_f32 if dtype == torch.float32 else lib.hm_matvec_grad
    assert fn(H._h, ctypes.c_void_p(u0.data_ptr()), ctypes.c_void_p(g0.data_ptr()),
              ctypes.c_void_p(bd.data_ptr()), ctypes.c_void_p(s.data_ptr()), 1, 1, 1, 1, 1,
              None) == 0
    torch.cuda.synchronize()
    # (the kernel may fuse the add into a multiply-add: equal to rounding)
    assert rel(u0.double(), (1 + u).double()) <= 1e-7 and rel(g0.double(), (2 + g).double()) <= 1e-7
    H.close()
