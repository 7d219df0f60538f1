"""EvenBarycentricMatrix on the GPU against the oracle's restatement of
src/BarycentricMatrix.jl:5-89 and src/algebra.jl:168-239 (pinned in
test_oracle_even.py).

Bars: W and w bit-exact (the device forms them in the reference's operation
order with IEEE division); products <= 1e-14 relative 2-norm against the
oracle's mul! (the same factors, summed in another order)."""
import ctypes

import numpy as np
import pytest
import torch

import hmx
import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb else float(np.linalg.norm(a - b))


def f_smooth(T, t, j):
    return 1.0 / (t + j + 100.0)


def f_osc(T, t, j):
    return np.cos(0.01 * t * j) / (1.0 + 0.001 * (t - j) ** 2)


@pytest.mark.parametrize("a,b,n", [(1, 40, 20), (3, 1000, 20), (-7, 120, 16), (100, 163, 12),
                                   (0, 1, 8), (1, 5000, 64), (2, 300, 24), (1, 77, 48)])
def test_weights_bit_exact(cuda, a, b, n):
    B = hmx.EvenBarycentricMatrix(f_smooth, a, b, 1, 10, rank=n)
    w, W = O.even_weights(a, b, n)
    assert np.array_equal(B.W, W), np.max(np.abs(B.W - W))
    assert np.array_equal(B.w, w)
    L = hmx.barycentricmatrix(f_smooth, a, b, 1, 10, rank=n)
    assert np.array_equal(L.U, O.barycentricmatrix_weights(a, b, n))
    assert np.array_equal(L.S, np.ones(n)) and np.array_equal(L.V, B.F)


def test_single_point_grid_is_nan_like_the_reference(cuda):
    B = hmx.EvenBarycentricMatrix(f_smooth, 5, 5, 1, 4, rank=20)
    assert np.all(np.isnan(B.W)) and np.isnan(B.w[0])


def test_F_is_f_at_the_nodes(cuda):
    B = hmx.EvenBarycentricMatrix(f_osc, 3, 90, -4, 60, rank=20)
    t = O.even_nodes(3, 90, 20)
    j = np.arange(-4, 61)
    assert np.array_equal(B.F, f_osc(np.float64, t[None, :], j[:, None]))


@pytest.mark.parametrize("istart,jstart", [(1, 1), (2, 1), (1, 2), (4, 6), (3, 8)])
@pytest.mark.parametrize("a,b,c,d", [(1, 40, 1, 33), (3, 400, 7, 900), (10, 30, 2, 2)])
def test_leaf_mul_vs_oracle(cuda, istart, jstart, a, b, c, d):
    B = hmx.EvenBarycentricMatrix(f_osc, a, b, c, d, rank=20)
    m, p = B.shape
    rng = np.random.default_rng(istart * 31 + jstart)
    v = rng.standard_normal(p + jstart + 3)
    u0 = rng.standard_normal(m + istart + 3)
    want = O.even_mul(u0.copy(), B.W, B.F, v, istart, jstart)
    u = torch.from_numpy(u0.copy()).to(cuda)
    B.mul_(u, torch.from_numpy(v).to(cuda), istart, jstart)
    got = u.cpu().numpy()
    assert rel(got, want) <= 1e-14, rel(got, want)
    # rows outside the block untouched
    assert np.array_equal(got[:istart - 1], u0[:istart - 1])
    assert np.array_equal(got[istart - 1 + m:], u0[istart - 1 + m:])
    # numpy in, numpy out (copied back)
    un = u0.copy()
    B.mul_(un, v, istart, jstart)
    assert np.array_equal(un, got)


@pytest.mark.parametrize("m,p", [(120, 100), (120, 99), (2, 7), (33, 1)])
def test_matmul_is_the_getindex_product(cuda, m, p):
    """B * v falls back to the generic product over getindex, whose parity
    uses size(B,1) + size(B,2) (src/BarycentricMatrix.jl:49-58)."""
    B = hmx.EvenBarycentricMatrix(f_osc, 1, m, 1, p, rank=16)
    G = np.array([[O.even_getindex(B.W, B.F, i, j) for j in range(1, p + 1)]
                  for i in range(1, m + 1)])
    v = np.random.default_rng(m + p).standard_normal(p)
    assert rel(B @ v, G @ v) <= 1e-13
    V = np.random.default_rng(1).standard_normal((p, 3))
    assert rel(B @ V, G @ V) <= 1e-13
    i, j = m - 1, p - 1
    assert B[i, j] == pytest.approx(G[i, j], rel=1e-15, abs=0)


def test_plus_minus(cuda):
    B1 = hmx.EvenBarycentricMatrix(f_osc, 1, 200, 1, 150, rank=20)
    B2 = hmx.EvenBarycentricMatrix(f_smooth, 1, 200, 1, 150, rank=20)
    v = torch.from_numpy(np.random.default_rng(2).standard_normal(150)).to(cuda)
    u1, u2, us, ud, un = (torch.zeros(200, dtype=torch.float64, device=cuda) for _ in range(5))
    B1.mul_(u1, v)
    B2.mul_(u2, v)
    (B1 + B2).mul_(us, v)
    (B1 - B2).mul_(ud, v)
    (-B1).mul_(un, v)
    assert rel(us.cpu().numpy(), (u1 + u2).cpu().numpy()) <= 1e-14
    assert rel(ud.cpu().numpy(), (u1 - u2).cpu().numpy()) <= 1e-14
    assert torch.equal(un, -u1)
    with pytest.raises(AssertionError):
        _ = B1 + hmx.EvenBarycentricMatrix(f_osc, 1, 201, 1, 150, rank=20)


def hodlr_leaves(N, levels, n, f):
    """Weakly admissible off-diagonal blocks of an N x N matrix: at level l,
    the 2^l blocks (2s, 2s+1) and (2s+1, 2s) of size N / 2^l (grids are the
    global 1-based indices of each block)."""
    out = []
    for l in range(1, levels + 1):
        s = N >> l
        for q in range(0, 1 << l, 2):
            for (bi, bj) in ((q, q + 1), (q + 1, q)):
                r0, c0 = bi * s, bj * s
                B = hmx.EvenBarycentricMatrix(f, r0 + 1, r0 + s, c0 + 1, c0 + s, rank=n)
                out.append((r0, c0, B))
    return out


def oracle_leaves(u, leaves, v, ioff=0, joff=0):
    for r0, c0, B in leaves:
        O.even_mul(u, B.W, B.F, v, ioff + r0 + 1, joff + c0 + 1)
    return u


@pytest.mark.parametrize("n", [8, 16, 20, 32, 64])
def test_leaves_vs_oracle(cuda, n):
    N, levels = 4096, 5
    leaves = hodlr_leaves(N, levels, n, f_osc)
    H = hmx.EvenBarycentricLeaves(N, N, leaves)
    info = H.info
    assert info["nleaves"] == 2 * (2 ** levels - 1) and info["rank"] == n
    v = np.random.default_rng(n).standard_normal(N)
    want = oracle_leaves(np.zeros(N), leaves, v)
    got = H.matvec(torch.from_numpy(v).to(cuda)).cpu().numpy()
    assert rel(got, want) <= 1e-14, rel(got, want)
    # one stored W per leaf size (the values depend on the rows only through
    # i - a and b - a), bit-identical to each leaf's own W formed apart
    assert info["w_blocks"] == levels
    assert info["w_entries"] == sum((N >> l) * n for l in range(1, levels + 1))
    assert H.leaf_w_entries() == sum(B.shape[0] * n for _, _, B in leaves)
    for t in range(len(leaves)):
        W, F = H.leaf_data(t)
        assert np.array_equal(W, leaves[t][2].W) and np.array_equal(F, leaves[t][2].F), t


def test_offsets_strides_and_many_rhs(cuda):
    """u[q*ldu + (ioff+I)*incu] += ..., k = 7 right-hand sides (groups of the
    launch width), row-major (n, k) tensors (inc = k, ld = 1), odd offsets
    flip the parity pairing."""
    N = 2048
    leaves = hodlr_leaves(N, 4, 20, f_osc)
    H = hmx.EvenBarycentricLeaves(N, N, leaves)
    k = 7
    rng = np.random.default_rng(3)
    V = rng.standard_normal((N + 5, k))
    U0 = rng.standard_normal((N + 4, k))
    U = torch.from_numpy(U0.copy()).to(cuda)
    H.mul_(U, torch.from_numpy(V).to(cuda), istart=4, jstart=6)
    got = U.cpu().numpy()
    for q in range(k):
        want = oracle_leaves(U0[:, q].copy(), leaves, np.ascontiguousarray(V[:, q]), 3, 5)
        assert rel(got[:, q], want) <= 1e-14, (q, rel(got[:, q], want))


def test_overwrite_zeroes_uncovered_rows(cuda):
    """beta = 0 overwrites every row; rows no leaf covers become 0."""
    B = hmx.EvenBarycentricMatrix(f_osc, 1, 300, 1, 200, rank=20)
    H = hmx.EvenBarycentricLeaves(500, 400, [(100, 50, B)])
    v = np.random.default_rng(4).standard_normal(400)
    u = H.matvec(v)
    want = O.even_mul(np.zeros(500), B.W, B.F, v, 101, 51)
    assert np.all(u[:100] == 0) and np.all(u[400:] == 0)
    assert rel(u, want) <= 1e-14


def test_wide_leaf_goes_through_the_reduce(cuda):
    """p = 30000 columns: several upward tasks (<= 4096 columns each), summed
    by the reduce kernel."""
    B = hmx.EvenBarycentricMatrix(f_osc, 1, 3000, 1, 30000, rank=24)
    H = hmx.EvenBarycentricLeaves(3000, 30000, [(0, 0, B)])
    assert H.info["n_up"] == 8 and H.info["n_reduce"] == 1
    v = np.random.default_rng(5).standard_normal(30000)
    want = O.even_mul(np.zeros(3000), B.W, B.F, v, 1, 1)
    assert rel(H.matvec(v), want) <= 1e-14


def test_interpolates_f_on_the_grid(cuda):
    """End to end: the device product of one leaf against the dense masked
    matrix of f itself. f = exp(-((t - j)/512)^2) is entire and varies on the
    scale of the interval, so rank 20 interpolates it to round-off."""
    a, b, c, d = 1, 512, 1, 256
    f = hmx.workloads.gaussian_even_kernel(512)
    B = hmx.EvenBarycentricMatrix(f, a, b, c, d, rank=20)
    i = np.arange(a, b + 1)[:, None]
    j = np.arange(c, d + 1)[None, :]
    M = np.where((i + j) % 2 == 0, np.exp(-((i - j) / 512.0) ** 2), 0.0)
    v = np.random.default_rng(6).standard_normal(d - c + 1)
    u = np.zeros(b - a + 1)
    B.mul_(u, v)
    assert rel(u, M @ v) <= 1e-12


# ---------------------------------------------------------------- Float32
# EvenBarycentricMatrix{Float32}: rank BLOCKRANK(Float32) = 10 by default;
# nodes, W and products in Float32 through the _f32 entry points. Bars: W
# and w bit-exact against the oracle's Float32 restatement (the same IEEE
# Float32 operations in the same order); products <= 2e-6 relative 2-norm
# against the Float64 product of the same Float32 factors, and <= 1e-5
# against the oracle's Float32 mul!, whose sequential sums over p columns
# carry errors of order eps(Float32) * sqrt(p) themselves.

F32_REL = 2e-6
F32_REL_ORACLE = 1e-5


def f64_product(leaves, v, N, ioff=0, joff=0):
    """The same Float32 factors multiplied in Float64 (the exact product up
    to Float64 round-off)."""
    u = np.zeros(N)
    for r0, c0, B in leaves:
        O.even_mul(u, B.W.astype(np.float64), B.F.astype(np.float64), v.astype(np.float64),
                   ioff + r0 + 1, joff + c0 + 1)
    return u


@pytest.mark.parametrize("a,b,n", [(1, 40, 10), (3, 1000, 10), (-7, 120, 16), (0, 1, 8),
                                   (1, 5000, 64), (2, 300, 24), (1, 200000, 10)])
def test_weights_f32_bit_exact(cuda, a, b, n):
    B = hmx.EvenBarycentricMatrix(f_smooth, a, b, 1, 10, rank=n, dtype=np.float32)
    w, W = O.even_weights(a, b, n, f32=True)
    assert B.W.dtype == np.float32 and B.F.dtype == np.float32
    assert np.array_equal(B.W, W), np.max(np.abs(B.W - W))
    assert np.array_equal(B.w, w)
    assert np.array_equal(B.x, O.chebyshevpoints(n, f32=True))
    assert np.array_equal(B.lam, O.chebyshevbarycentricweights(n, f32=True))
    L = hmx.barycentricmatrix(f_smooth, a, b, 1, 10, rank=n, dtype=np.float32)
    assert L.U.dtype == np.float32
    assert np.array_equal(L.U, O.barycentricmatrix_weights(a, b, n, f32=True))


def test_f32_default_rank_and_F(cuda):
    B = hmx.EvenBarycentricMatrix(f_osc, 3, 90, -4, 60, dtype=np.float32)
    assert B.rank() == hmx.BLOCKRANK(np.float32) == 10
    t = O.even_nodes(3, 90, 10, f32=True)
    j = np.arange(-4, 61)
    assert np.array_equal(B.F, f_osc(np.float32, t[None, :], j[:, None]).astype(np.float32))


@pytest.mark.parametrize("istart,jstart", [(1, 1), (2, 1), (1, 2), (4, 6)])
@pytest.mark.parametrize("a,b,c,d", [(1, 40, 1, 33), (3, 400, 7, 900), (10, 30, 2, 2)])
def test_leaf_mul_f32_vs_oracle(cuda, istart, jstart, a, b, c, d):
    B = hmx.EvenBarycentricMatrix(f_osc, a, b, c, d, dtype=np.float32)
    m, p = B.shape
    rng = np.random.default_rng(istart * 31 + jstart)
    v = rng.standard_normal(p + jstart + 3).astype(np.float32)
    u0 = rng.standard_normal(m + istart + 3).astype(np.float32)
    want = O.even_mul(u0.copy(), B.W, B.F, v, istart, jstart)
    exact = u0.astype(np.float64)
    O.even_mul(exact, B.W.astype(np.float64), B.F.astype(np.float64), v.astype(np.float64),
               istart, jstart)
    u = torch.from_numpy(u0.copy()).to(cuda)
    B.mul_(u, torch.from_numpy(v).to(cuda), istart, jstart)
    got = u.cpu().numpy()
    assert got.dtype == np.float32
    assert rel(got, want) <= F32_REL_ORACLE, rel(got, want)
    assert rel(got, exact) <= F32_REL, rel(got, exact)
    assert np.array_equal(got[:istart - 1], u0[:istart - 1])
    assert np.array_equal(got[istart - 1 + m:], u0[istart - 1 + m:])


@pytest.mark.parametrize("n", [10, 16, 32, 64])
def test_leaves_f32_vs_oracle(cuda, n):
    N, levels = 4096, 5
    out = []
    for l in range(1, levels + 1):
        s = N >> l
        for q in range(0, 1 << l, 2):
            for (bi, bj) in ((q, q + 1), (q + 1, q)):
                r0, c0 = bi * s, bj * s
                out.append((r0, c0, hmx.EvenBarycentricMatrix(f_osc, r0 + 1, r0 + s, c0 + 1,
                                                               c0 + s, rank=n,
                                                               dtype=np.float32)))
    H = hmx.EvenBarycentricLeaves(N, N, out)
    assert H.dtype == np.float32 and H.info["f32"] == 1
    assert H.bytes_per_product() == 4 * (H.info["w_entries"] + H.info["f_entries"] + 3 * N)
    v = np.random.default_rng(n).standard_normal(N).astype(np.float32)
    want = np.zeros(N, dtype=np.float32)
    for r0, c0, B in out:
        O.even_mul(want, B.W, B.F, v, r0 + 1, c0 + 1)
    got = H.matvec(torch.from_numpy(v).to(cuda)).cpu().numpy()
    assert rel(got, want) <= F32_REL_ORACLE, rel(got, want)
    assert rel(got, f64_product(out, v, N)) <= F32_REL
    # many right-hand sides with offsets
    k = 5
    V = np.random.default_rng(1).standard_normal((N + 3, k)).astype(np.float32)
    U = torch.zeros((N + 2, k), dtype=torch.float32, device=cuda)
    H.mul_(U, torch.from_numpy(V).to(cuda), istart=3, jstart=4)
    G = U.cpu().numpy()
    for q in range(k):
        e = np.zeros(N + 2)
        for r0, c0, B in out:
            O.even_mul(e, B.W.astype(np.float64), B.F.astype(np.float64),
                       V[:, q].astype(np.float64), r0 + 3, c0 + 4)
        assert rel(G[:, q], e) <= F32_REL, (q, rel(G[:, q], e))
    W, F = H.leaf_data(5)
    assert W.dtype == np.float32
    assert np.array_equal(W, out[5][2].W) and np.array_equal(F, out[5][2].F)


def test_f32_wide_leaf_and_overwrite(cuda):
    B = hmx.EvenBarycentricMatrix(f_osc, 1, 3000, 1, 30000, dtype=np.float32)
    H = hmx.EvenBarycentricLeaves(3100, 30000, [(50, 0, B)])
    assert H.info["n_reduce"] == 1
    v = np.random.default_rng(5).standard_normal(30000).astype(np.float32)
    got = H.matvec(v)
    assert got.dtype == np.float32 and np.all(got[:50] == 0)
    assert rel(got, f64_product([(50, 0, B)], v, 3100)) <= F32_REL


def test_f32_matmul_and_plus(cuda):
    B1 = hmx.EvenBarycentricMatrix(f_osc, 1, 120, 1, 99, dtype=np.float32)
    B2 = hmx.EvenBarycentricMatrix(f_smooth, 1, 120, 1, 99, dtype=np.float32)
    G = np.array([[O.even_getindex(B1.W, B1.F, i, j) for j in range(1, 100)]
                  for i in range(1, 121)])
    v = np.random.default_rng(7).standard_normal(99).astype(np.float32)
    assert rel(B1 @ v, G.astype(np.float64) @ v) <= F32_REL
    assert isinstance(B1[3, 4], np.float32)
    u1, u2, us = (torch.zeros(120, dtype=torch.float32, device=cuda) for _ in range(3))
    vt = torch.from_numpy(v).to(cuda)
    B1.mul_(u1, vt)
    B2.mul_(u2, vt)
    (B1 + B2).mul_(us, vt)
    assert rel(us.cpu().numpy(), (u1 + u2).cpu().numpy()) <= F32_REL
    with pytest.raises(TypeError):
        _ = B1 + hmx.EvenBarycentricMatrix(f_osc, 1, 120, 1, 99)


def test_element_types_do_not_mix(cuda):
    B = hmx.EvenBarycentricMatrix(f_osc, 1, 64, 1, 64, dtype=np.float32)
    H = B.plan()
    with pytest.raises(TypeError):
        H.matvec(torch.zeros(64, dtype=torch.float64, device=cuda))
    u = torch.zeros(64, dtype=torch.float64, device=cuda)
    rc = hmx._lib.lib.hm_even_matvec(H._h, ctypes.c_void_p(u.data_ptr()),
                                      ctypes.c_void_p(u.data_ptr()), 0, 0, 1, 1, 1, 0, 0, 0, None)
    assert rc == hmx._lib.HM_ERR_INVALID
    assert b"Float32" in hmx._lib.lib.hm_last_error()
    with pytest.raises(TypeError):
        hmx.EvenBarycentricLeaves(64, 64, [(0, 0, B)], dtype=np.float64)
