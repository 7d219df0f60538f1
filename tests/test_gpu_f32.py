"""Float32 plans (KernelMatrix{Float32}) on the GPU against the Float32 oracle.

The reference's path is generic in T, so for Float32 points it splits at
Float32 midpoints, forms every factor in Float32 and accumulates mul! in
Float32. Bars:
  compression: bit-exact against the Float32 oracle (hm_entries_f32);
  products: <= 1e-5 relative 2-norm against the oracle's Float32 traversal
  (the two sum in different orders, each with Float32 rounding, ~1e-7 per
  term), and <= 3e-5 against the dense FP64 product of the same points;
  gradients: <= 1e-4 relative against the dense FP64 gradient.
"""
import numpy as np
import pytest
import torch

import hmx
import oracle as O
from hmx.autodiff import kernel_matvec

pytestmark = pytest.mark.gpu
f32 = np.float32
KIDS = {"cauchykernel": O.CAUCHY, "coulombkernel": O.COULOMB, "logkernel": O.LOG}


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def grid32(n):
    x, y, iv = hmx.workloads.shifted_grid(n)
    return x.astype(f32), y.astype(f32), iv


def test_f32_compression_bit_exact(cuda):
    x, y, iv = grid32(4096)
    print("case: f32 compression n=4096 r=16", flush=True)
    H = hmx.KernelMatrix("cauchykernel", x, y, *iv, blocksize=64, rank=16)
    assert H.info["float32"] == 1 and H.info["nonfinite"] == 0
    t = O.Tree(O.CAUCHY, x, y, *iv, 64, 16, f32=True)
    rng = np.random.default_rng(3)
    ii = rng.integers(0, 4096, 3000)
    jj = rng.integers(0, 4096, 3000)
    got = H[ii, jj]
    assert got.dtype == f32
    want = np.array([t.getindex(int(i), int(j)) for i, j in zip(ii, jj)], dtype=f32)
    assert np.array_equal(got, want)
    # the diagonal band (dense blocks) and far entries both occur
    ints, _ = t.blocks()
    assert (ints[:, 0] == O.CODE_DENSE).any() and (ints[:, 0] == O.CODE_BARY).any()


@pytest.mark.parametrize("kernel", list(KIDS))
@pytest.mark.parametrize("mode", ["stored", "hybrid", "free"])
def test_f32_products_vs_oracle(cuda, kernel, mode):
    x, y, iv = grid32(4096)
    print(f"case: f32 product kernel={kernel} mode={mode}", flush=True)
    H = hmx.KernelMatrix(kernel, x, y, *iv, blocksize=64, rank=16,
                         matrix_free=mode == "free", hybrid=mode == "hybrid")
    b = hmx.workloads.normal(0, 4096).astype(f32)
    u = H.matvec(torch.from_numpy(b).to(cuda)).cpu().numpy()
    assert u.dtype == f32
    ref = O.Tree(KIDS[kernel], x, y, *iv, 64, 16, f32=True).matvec(b)
    assert rel(u, ref) <= 1e-5
    dense = O.dense_matvec(KIDS[kernel], x.astype(np.float64), y.astype(np.float64),
                           b.astype(np.float64))
    assert rel(u, dense) <= 3e-5


@pytest.mark.parametrize("mode", ["hybrid", "free"])
def test_f32_side_stream_split(cuda, mode):
    """Float32 hybrid and matrix-free plans with at least 2^24 near-field
    entries: one right-hand side runs the near field on a side stream and the
    far field beside it into a plan buffer, added onto y after the join.
    Against the one-stream product of the same plan (pass profiling keeps one
    stream; <= 1e-6, the two add near and far field in different orders) for
    forward, transpose, accumulate and repeated products (identical), and
    against the dense FP64 product on sampled rows (<= 3e-5; on this grid the
    Float32 nodes coincide with points, so the oracle's stored Float32
    traversal holds Inf factors, as the reference's Float32 update! does)."""
    from hmx._lib import lib
    n = 1 << 16
    x, y, iv = grid32(n)
    print(f"case: f32 side-stream split n=2^16 leaf 256 mode={mode}", flush=True)
    H = hmx.KernelMatrix("cauchykernel", x, y, *iv, blocksize=256, rank=16,
                         matrix_free=mode == "free", hybrid=mode == "hybrid")
    assert H.info["dense_entries"] >= 1 << 24
    b = hmx.workloads.normal(0, n).astype(f32)
    bd = torch.from_numpy(b).to(cuda)
    u0 = torch.from_numpy(hmx.workloads.normal(9, n).astype(f32)).to(cuda)
    s = torch.cuda.Stream()

    def products():
        with torch.cuda.stream(s):
            out = [H.matvec(bd, stream=s), H.matvec(bd, transpose=True, stream=s),
                   H.mul(u0.clone(), bd, stream=s)]
            out += [H.matvec(bd, stream=s) for _ in range(2)]
        s.synchronize()
        return [o.cpu().numpy() for o in out]

    split = products()
    assert lib.hm_plan_profile(H._h, 1) == 0
    one = products()
    assert lib.hm_plan_profile(H._h, 0) == 0
    H.close()
    for a, e in zip(split, one):
        assert np.all(np.isfinite(a))
        assert rel(a, e) <= 1e-6
    for q in (3, 4):
        assert np.array_equal(split[q], split[0])
    rows = np.arange(0, n, 61)
    ref = O.dense_matvec(O.CAUCHY, x.astype(np.float64), y.astype(np.float64),
                         b.astype(np.float64), rows=rows)
    assert rel(split[0][rows], ref) <= 3e-5


def test_f32_default_rank(cuda):
    """BLOCKRANK(Float32) = 10, BLOCKSIZE(Float32) = 40 at n = 20000 (shifted
    grid: random Float32 points with y = x + 0.5/n would coincide with other
    x points and give Inf, as in the reference)."""
    x, y, iv = grid32(20000)
    H = hmx.KernelMatrix("cauchykernel", x, y, *iv)
    assert (H.rank, H.blocksize_) == (10, 40)
    assert H.info["nonfinite"] == 0
    b = hmx.workloads.normal(1, 20000).astype(f32)
    u = H.matvec(torch.from_numpy(b).to(cuda)).cpu().numpy()
    ref = O.Tree(O.CAUCHY, x, y, *iv, 40, 10, f32=True).matvec(b)
    assert rel(u, ref) <= 1e-5


def test_f32_multi_rhs_strides_and_accumulate(cuda):
    x, y, iv = grid32(4096)
    H = hmx.KernelMatrix("cauchykernel", x, y, *iv, blocksize=64, rank=16, hybrid=True)
    B = torch.from_numpy(hmx.workloads.normal(2, 4096 * 7).reshape(4096, 7).astype(f32)).to(cuda)
    Y = H.matvec(B)
    for q in range(7):
        yq = H.matvec(B[:, q].contiguous())
        assert rel(Y[:, q].cpu().numpy(), yq.cpu().numpy()) <= 1e-5
    # mul!(u, H, v, istart, jstart) with strides: u[istart + 2(i-1)] += (H v)_i
    u = torch.ones(2 * 4096 + 5, dtype=torch.float32, device=cuda)
    v = torch.zeros(3 * 4096 + 2, dtype=torch.float32, device=cuda)
    v[2::3] = B[:, 0]
    hmx.mul_(u, H, v, 4, 3, 3, 2)
    want = 1 + Y[:, 0]
    assert rel(u[3:3 + 2 * 4096:2].cpu().numpy(), want.cpu().numpy()) <= 1e-6
    assert torch.all(u[:3] == 1)


def test_f32_transpose(cuda):
    """H^T(x, y) = -H(y, x) for the Cauchy kernel on equal intervals, and the
    adjoint identity <H^T r, b> = <r, H b>."""
    x, y, iv = grid32(4096)
    H = hmx.KernelMatrix("cauchykernel", x, y, *iv, blocksize=64, rank=16)
    Hyx = hmx.KernelMatrix("cauchykernel", y, x, iv[2], iv[3], iv[0], iv[1], blocksize=64, rank=16)
    r = torch.from_numpy(hmx.workloads.normal(5, 4096).astype(f32)).to(cuda)
    b = torch.from_numpy(hmx.workloads.normal(6, 4096).astype(f32)).to(cuda)
    ht = H.matvec(r, transpose=True)
    assert rel(ht.cpu().numpy(), (-(Hyx @ r)).cpu().numpy()) <= 1e-5
    lhs = float(torch.dot(ht.double(), b.double()))
    rhs = float(torch.dot(r.double(), (H @ b).double()))
    assert abs(lhs - rhs) <= 1e-5 * float(torch.linalg.norm(r.double()) * torch.linalg.norm((H @ b).double()))


def test_f32_graph_replays(cuda):
    x, y, iv = grid32(4096)
    H = hmx.KernelMatrix("cauchykernel", x, y, *iv, blocksize=64, rank=16, matrix_free=True)
    B = torch.from_numpy(hmx.workloads.normal(7, 4096 * 4).reshape(4096, 4).astype(f32)).to(cuda)
    eager = H.matvec(B)
    Y = torch.zeros_like(eager)
    g = H.graph(Y, B)
    for _ in range(3):
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(Y, eager)
        Y.zero_()
        _ = H.matvec(B[:, 0].contiguous())
        torch.cuda.synchronize()
    g.close()


def test_f32_gradients_vs_dense_fp64(cuda):
    """L = 1/2 ||H(x, y) b - t||^2 with Float32 points, b and t: the pullback
    through Float32 plans against the dense FP64 gradient."""
    n = 4096
    w = hmx.workloads.config_gradient(n=n)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    x, y, b, t = (T(w[k]).float() for k in ("x", "y", "b", "t"))
    xg, yg, bg = (v.clone().requires_grad_(True) for v in (x, y, b))
    u = kernel_matvec("cauchykernel", xg, yg, bg, w["intervals"], 64, 16)
    assert u.dtype == torch.float32
    L = 0.5 * torch.sum((u.double() - t.double()) ** 2)
    L.backward()
    xd, yd, bd, td = (v.double() for v in (x, y, b, t))
    d = xd[:, None] - yd[None, :]
    K, Kx = 1.0 / d, -1.0 / (d * d)
    r = K @ bd - td
    errs = dict(u=rel(u.detach().cpu(), (K @ bd).cpu()),
                b=rel(bg.grad.cpu(), (K.T @ r).cpu()),
                x=rel(xg.grad.cpu(), (r * (Kx @ bd)).cpu()),
                y=rel(yg.grad.cpu(), (bd * (-(Kx.T @ r))).cpu()))
    print("f32 gradient relative errors", errs, flush=True)
    assert all(v <= 1e-4 for v in errs.values()), errs
