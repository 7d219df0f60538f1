"""The reference's own known-answer tests of the leaf mul! (test/runtests.jl:
15-52) run through the device, plus the transposed and adjoint leaf products
with offsets, strides and accumulation (src/algebra.jl:50-82,133-159) and the
transposed KernelMatrix product with offsets and strides.

Each case builds a plan of hm_plan_from_blocks (or its Float32 twin) holding
the one leaf, and calls hm_matvec / hm_matvec_t (beta = 1, the reference's
accumulating mul!) with the reference's istart, jstart, INCX, INCY.

Bars:
- random leaves, as the reference: ||y - ref|| <= 4 eps(T) ||ref|| against the
  oracle's restatement of the reference's loop. The reference's own bar is
  eps(T) for its loop against BLAS; the device sums each entry in a fused
  multiply-add chain, which rounds up to n/2 ulp apart from a loop that
  rounds products and sums separately (n = 5 or 10 terms here).
- exact twins (the reference's BigFloat cases, :35-52): small-integer leaves
  and vectors, whose products and sums are exact in both types, must match
  the exact result bit for bit.
- the transposed KernelMatrix product: <= 1e-12 relative 2-norm against
  -H(y, x) v from the oracle's traversal (for Cauchy on equal intervals
  H(x, y)^T equals -H(y, x) block for block, F' = -F^T exactly)."""
import numpy as np
import pytest
import torch

import hmx
import oracle as O
from hmx import Block, HierarchicalMatrix, LowRankMatrix

pytestmark = pytest.mark.gpu

TYPES = [np.float32, np.float64]
# (transposed, istart, jstart, INCX, INCY) of test/runtests.jl:20-33
CASES = [(False, 1, 1, 1, 1), (False, 5, 5, 2, 2), (True, 1, 5, 2, 1), (True, 6, 3, 1, 3)]


def one_leaf(leaf):
    H = HierarchicalMatrix(1, 1)
    H[Block(1), Block(1)] = leaf
    return H.plan()


def run(P, y0, x, tr, istart, jstart, incx, incy, cuda):
    y = torch.from_numpy(y0.copy()).to(cuda)
    xd = torch.from_numpy(x).to(cuda)
    (P.tmul if tr else P.mul)(y, xd, istart, jstart, incx, incy)
    return y.cpu().numpy()


def oracle(A, y0, x, tr, istart, jstart, incx, incy):
    y = y0.copy()
    (O.dense_tmul if tr else O.dense_mul)(y, A, x, istart, jstart, incx, incy)
    return y


def touched(y, A, tr, istart, incy):
    m = A.shape[1] if tr else A.shape[0]
    return y[istart - 1:istart - 1 + incy * (m - 1) + 1:incy]


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("case", CASES)
def test_reference_leaf_kats(cuda, T, case):
    tr, istart, jstart, incx, incy = case
    print(f"case: KAT {np.dtype(T).name} {case}", flush=True)
    rng = np.random.default_rng(0)
    A = rng.random((10, 5)).astype(T)
    x = rng.random(40).astype(T)
    P = one_leaf(np.asfortranarray(A))
    y0 = np.zeros(40, dtype=T)
    got = run(P, y0, x, tr, istart, jstart, incx, incy, cuda)
    want = oracle(A, y0, x, tr, istart, jstart, incx, incy)
    gt, wt = touched(got, A, tr, istart, incy), touched(want, A, tr, istart, incy)
    assert np.linalg.norm(gt - wt) <= 4 * np.finfo(T).eps * np.linalg.norm(wt)
    # the entries outside the strided output are untouched
    mask = np.ones(40, dtype=bool)
    mask[istart - 1:istart - 1 + incy * ((5 if tr else 10) - 1) + 1:incy] = False
    assert np.all(got[mask] == 0)
    # the reference's formula itself: y[istart:INCY:..] == A x[jstart:INCX:..]
    xs = x[jstart - 1:jstart - 1 + incx * ((10 if tr else 5) - 1) + 1:incx].astype(np.float64)
    ref = (A.T.astype(np.float64) if tr else A.astype(np.float64)) @ xs
    assert np.linalg.norm(gt - ref) <= 4 * np.finfo(T).eps * np.linalg.norm(ref)


@pytest.mark.parametrize("T", TYPES)
@pytest.mark.parametrize("case", CASES)
def test_reference_leaf_kats_exact(cuda, T, case):
    """The BigFloat twins (test/runtests.jl:35-52): exact arithmetic, so the
    device result equals A x bit for bit."""
    tr, istart, jstart, incx, incy = case
    rng = np.random.default_rng(1)
    A = rng.integers(-8, 9, (10, 5)).astype(T)
    x = rng.integers(-8, 9, 40).astype(T)
    P = one_leaf(np.asfortranarray(A))
    got = run(P, np.zeros(40, dtype=T), x, tr, istart, jstart, incx, incy, cuda)
    xs = x[jstart - 1:jstart - 1 + incx * ((10 if tr else 5) - 1) + 1:incx].astype(np.int64)
    exact = (A.T.astype(np.int64) if tr else A.astype(np.int64)) @ xs
    assert np.array_equal(touched(got, A, tr, istart, incy), exact.astype(T))


@pytest.mark.parametrize("T", TYPES)
def test_accumulates_onto_y(cuda, T):
    """mul! adds onto y (src/algebra.jl:37-48): y0 + A x on a nonzero y0."""
    rng = np.random.default_rng(2)
    A = rng.integers(-8, 9, (10, 5)).astype(T)
    x = rng.integers(-8, 9, 40).astype(T)
    y0 = rng.integers(-8, 9, 40).astype(T)
    P = one_leaf(np.asfortranarray(A))
    for tr, istart, jstart, incx, incy in CASES:
        got = run(P, y0, x, tr, istart, jstart, incx, incy, cuda)
        want = oracle(A, y0, x, tr, istart, jstart, incx, incy)
        assert np.array_equal(got, want), (tr, istart, jstart, incx, incy)


@pytest.mark.parametrize("istart,jstart,incx,incy", [(1, 1, 1, 1), (3, 2, 2, 3), (7, 11, 3, 2)])
def test_lowrank_adjoint_leaf(cuda, istart, jstart, incx, incy):
    """mul!(y, L', x, istart, jstart, INCX, INCY) (src/algebra.jl:133-159):
    y += V S U' x on the device (hm_matvec_t of a LowRank leaf) against the
    oracle's loop, and the forward leaf (:110-131) likewise."""
    rng = np.random.default_rng(istart + 10 * jstart)
    m, n, r = 37, 23, 6
    U, S, V = rng.standard_normal((m, r)), rng.standard_normal(r), rng.standard_normal((n, r))
    P = one_leaf(LowRankMatrix(U, S, V))
    x = rng.standard_normal(200)
    y0 = rng.standard_normal(200)
    got = run(P, y0, x, True, istart, jstart, incx, incy, cuda)
    want = O.lowrank_mul(y0.copy(), U, S, V, x, istart, jstart, incx, incy, adjoint=True)
    assert np.linalg.norm(got - want) <= 1e-14 * np.linalg.norm(want)
    got = run(P, y0, x, False, istart, jstart, incx, incy, cuda)
    want = O.lowrank_mul(y0.copy(), U, S, V, x, istart, jstart, incx, incy)
    assert np.linalg.norm(got - want) <= 1e-14 * np.linalg.norm(want)


@pytest.mark.parametrize("mode", ["stored", "hybrid", "free"])
def test_kernelmatrix_transpose_offsets_strides(cuda, mode):
    """u[istart + (i-1) INCY] += (H^T v[jstart:INCX:...])_i for a Cauchy
    KernelMatrix on the shifted grid against -H(y, x) v by the oracle's
    traversal mul!."""
    n = 4096
    w = hmx.workloads.config_plain(n=n, blocksize=64, rank=16)
    x, y = w["x"], w["y"]
    H = hmx.KernelMatrix("cauchykernel", x, y, *w["intervals"], blocksize=64, rank=16,
                         matrix_free=mode == "free", hybrid=mode == "hybrid")
    istart, jstart, incx, incy = 3, 5, 2, 3
    rng = np.random.default_rng(9)
    v = rng.standard_normal(jstart - 1 + incx * (n - 1) + 1 + 7)
    u0 = rng.standard_normal(istart - 1 + incy * (n - 1) + 1 + 5)
    u = torch.from_numpy(u0.copy()).to(cuda)
    H.tmul(u, torch.from_numpy(v).to(cuda), istart, jstart, incx, incy)
    got = u.cpu().numpy()
    t = O.Tree(O.CAUCHY, y, x, *w["intervals"][2:], *w["intervals"][:2], 64, 16)
    vs = np.ascontiguousarray(v[jstart - 1::incx][:n])
    want = u0.copy()
    want[istart - 1::incy][:n] -= t.matvec(vs)
    sl = slice(istart - 1, istart - 1 + incy * (n - 1) + 1, incy)
    assert np.linalg.norm(got[sl] - want[sl]) <= 1e-12 * np.linalg.norm(want[sl])
    mask = np.ones(len(u0), dtype=bool)
    mask[sl] = False
    assert np.array_equal(got[mask], u0[mask])
