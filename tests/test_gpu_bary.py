"""BarycentricPoly2D / evaluate, BarycentricMatrix2D and KernelMatrix with an
arbitrary host f on the device, against the oracle.

- evaluate (src/BarycentricMatrix.jl:180-205): bit for bit against the
  oracle's restatement (same operation order, IEEE division, no contraction),
  Float64 and Float32.
- BarycentricMatrix2D: U and V (update!) bit for bit against the oracle
  tree's far blocks, F bit for bit for the Cauchy f; getindex bit for bit
  against the oracle's getindex; the leaf mul! (src/algebra.jl:243-277) with
  offsets within 1e-14 of the oracle's loop.
- KernelMatrix(f, ...) with a host f: a plan whose tree and bases are built on
  the device and whose dense leaves and F come from f on the host. For
  f = 1/(x - y) as a plain function it equals the named Cauchy plan bit for
  bit; for a Gaussian f it matches the oracle's traversal mul! over leaves
  built from the oracle's blocks (its U, V, which do not depend on f, and f at
  the points and the mapped nodes) within 1e-12, also when those leaves are
  handed to hm_plan_from_blocks as barycentric leaves (kind 4)."""
import numpy as np
import pytest
import torch

import hmx
import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def gauss(x, y):
    return np.exp(-4.0 * (x - y) ** 2)


def cauchy_fn(x, y):
    return 1.0 / (x - y)


@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("r", [5, 16, 20])
def test_evaluate_bit_exact(cuda, T, r):
    B = hmx.BarycentricPoly2D(gauss, 1.0, 0.5, 0.25, -0.75, rank=r, dtype=T)
    rng = np.random.default_rng(r)
    px = (0.5 + 0.5 * rng.random(300)).astype(T)
    py = (-0.75 + rng.random(300)).astype(T)
    got = B(px, py)
    want = O.bary_evaluate(B.x, B.y, B.lamx, B.lamy, B.F, px, py, f32=T == np.float32)
    assert got.dtype == T and np.array_equal(got, want)
    if T == np.float64 and r == 20:  # rank 20 interpolates the Gaussian to round-off
        assert np.max(np.abs(got - gauss(px, py))) <= 5e-13


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_barycentric_matrix_2d_vs_oracle_tree(cuda, T):
    n, bs, r = 4096, 64, 16
    w = hmx.workloads.config_plain(n=n, blocksize=bs, rank=r)
    x, y = w["x"].astype(T), w["y"].astype(T)
    t = O.Tree(O.CAUCHY, x, y, *w["intervals"], bs, r, f32=T == np.float32)
    ints, dbls = t.blocks()
    far = [k for k in range(len(ints)) if ints[k][0] == 2]
    rng = np.random.default_rng(5)
    for k in [far[0], far[len(far) // 2], far[-1]]:
        code, _, i0, j0, m, nn = ints[k]
        a, b, c, d = dbls[k]
        L = hmx.BarycentricMatrix2D(cauchy_fn, a, b, c, d, x, y, range(i0 + 1, i0 + m + 1),
                                    range(j0 + 1, j0 + nn + 1), rank=r, dtype=T)
        ref = t.block_data(k)
        assert np.array_equal(L.U, ref["U"]) and np.array_equal(L.V, ref["V"])
        assert np.array_equal(L.F, ref["F"])
        for i, j in zip(rng.integers(0, m, 5), rng.integers(0, nn, 5)):
            assert L[int(i), int(j)] == t.getindex(int(i0 + i), int(j0 + j))
        if T == np.float64:
            u0 = rng.standard_normal(n + 9)
            v = rng.standard_normal(n + 9)
            for istart, jstart in ((1, 1), (4, 7)):
                u = torch.from_numpy(u0.copy()).to(cuda)
                L.mul_(u, torch.from_numpy(v).to(cuda), istart, jstart)
                want = O.bary_mul(u0.copy(), L.U, L.F, L.V, v, istart, jstart)
                assert rel(u.cpu().numpy(), want) <= 1e-14


@pytest.mark.parametrize("hybrid", [False, True])
def test_plain_function_equals_named_kernel(cuda, hybrid):
    """f = 1/(x - y) given as a plain function (kernel HM_USER, dense leaves
    and F from the host) computes the named Cauchy plan's product bit for bit
    (same factors in the same layout)."""
    w = hmx.workloads.config_plain(n=4096, blocksize=64, rank=16)
    Hn = hmx.KernelMatrix("cauchykernel", w["x"], w["y"], *w["intervals"], blocksize=64,
                          rank=16, hybrid=hybrid)
    Hu = hmx.KernelMatrix(cauchy_fn, w["x"], w["y"], *w["intervals"], blocksize=64, rank=16,
                          hybrid=hybrid)
    assert Hu.info["kernel"] == 4 and Hu.info["nonfinite"] == 0
    v = torch.from_numpy(w["b"]).to(cuda)
    assert torch.equal(Hu.matvec(v), Hn.matvec(v))
    assert torch.equal(Hu.matvec(v, transpose=True), Hn.matvec(v, transpose=True))
    V16 = torch.from_numpy(np.random.default_rng(3).standard_normal((4096, 16))).to(cuda)
    assert torch.equal(Hu.matvec(V16), Hn.matvec(V16))
    ii = np.arange(0, 4096, 37)
    jj = (ii * 7919) % 4096
    assert np.array_equal(Hu[ii, jj], Hn[ii, jj])
    with pytest.raises(hmx.HMError):
        Hu.matvec(v, matrix_free=True)
    with pytest.raises(TypeError):
        hmx.KernelMatrix(cauchy_fn, w["x"], w["y"], *w["intervals"], matrix_free=True)


def oracle_leaves(f, x, y, iv, bs, r):
    """The leaves of KernelMatrix(f, x, y, ...) from the oracle's tree: its
    block list, its U and V (f-independent), and f at the points and at the
    mapped nodes t = (a+b)/2 + (b-a)/2 x_hat (checked against the oracle's
    Cauchy F of the same block)."""
    t = O.Tree(O.CAUCHY, x, y, *iv, bs, r)
    ints, dbls = t.blocks()
    xh = O.chebyshevpoints(r)
    out = []
    for k, ((code, _, i0, j0, m, n), (a, b, c, d)) in enumerate(zip(ints, dbls)):
        if m == 0 or n == 0:
            continue
        if code == 3:
            out.append(("dense", i0, j0, f(x[i0:i0 + m, None], y[None, j0:j0 + n])))
        else:
            tx = 0.5 * (a + b) + 0.5 * (b - a) * xh
            ty = 0.5 * (c + d) + 0.5 * (d - c) * xh
            dat = t.block_data(k)
            assert np.array_equal(1.0 / (tx[:, None] - ty[None, :]), dat["F"])
            out.append(("far", i0, j0, (dat["U"], f(tx[:, None], ty[None, :]), dat["V"])))
    return out


def oracle_mul(leaves, n, v):
    """The traversal mul! (src/KernelMatrix.jl:17-45) over the leaves: dense
    leaves by src/algebra.jl:37-48, far leaves by :243-277."""
    u = np.zeros(n)
    for kind, i0, j0, dat in leaves:
        if kind == "dense":
            O.dense_mul(u, dat, v, i0 + 1, j0 + 1)
        else:
            U, F, V = dat
            O.bary_mul(u, U, F, V, v, i0 + 1, j0 + 1)
    return u


@pytest.mark.parametrize("points", ["grid", "random"])
def test_arbitrary_f_vs_oracle(cuda, points):
    n, bs, r = 4096, 64, 16
    if points == "grid":
        x, y, iv = hmx.workloads.shifted_grid(n)
    else:
        x, y, iv = hmx.workloads.sorted_uniform_points(n, 4)
    leaves = oracle_leaves(gauss, x, y, iv, bs, r)
    v = hmx.workloads.normal(8, n)
    want = oracle_mul(leaves, n, v)
    vd = torch.from_numpy(v).to(cuda)
    for hybrid in (False, True):
        H = hmx.KernelMatrix(gauss, x, y, *iv, blocksize=bs, rank=r, hybrid=hybrid)
        got = H.matvec(vd).cpu().numpy()
        assert rel(got, want) <= 1e-12, (hybrid, rel(got, want))
    # the same leaves through hm_plan_from_blocks: dense (kind 3) and
    # barycentric U F V' (kind 4) leaves
    flat = []
    for kind, i0, j0, dat in leaves:
        flat.append((np.asfortranarray(dat) if kind == "dense" else hmx.BarycentricFactors(*dat),
                     i0, j0))
    P = hmx.blocks_plan(n, n, flat)
    got = P.matvec(vd).cpu().numpy()
    assert rel(got, want) <= 1e-12
    # its transpose against the dense transpose of the same leaves
    D = np.zeros((n, n))
    for kind, i0, j0, dat in leaves:
        blk = dat if kind == "dense" else dat[0] @ dat[1] @ dat[2].T
        D[i0:i0 + blk.shape[0], j0:j0 + blk.shape[1]] = blk
    gt = P.matvec(vd, transpose=True).cpu().numpy()
    assert rel(gt, D.T @ v) <= 1e-12
