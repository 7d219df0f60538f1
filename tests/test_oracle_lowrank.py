"""Pins the numpy restatement of the LowRankMatrix algebra (oracle/lowrank.py,
src/LowRankMatrix.jl:64-171) with checks that do not reuse its steps: getrank
on hand-worked singular values, and for every operation the dense result
(formed directly from the operands), its singular values from a plain SVD of
that dense matrix, and the orthonormality of the returned factors."""
import numpy as np
import pytest

import lowrank as LR


def rand_lr(rng, m, n, r, scale=1.0):
    U, _ = np.linalg.qr(rng.standard_normal((m, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)))
    S = np.sort(rng.random(r) + 0.1)[::-1] * scale
    return U, S, V


def test_getrank_known_answers():
    eps1 = np.spacing(1.0)  # 2^-52
    assert LR.getrank(np.array([])) == 0
    assert LR.getrank(np.array([0.0, 0.0])) == 0
    assert LR.getrank(np.array([1.0, 1e-20])) == 1
    # tol = 2 * eps(1.0) = 4.44e-16
    assert LR.getrank(np.array([1.0, 2 * eps1])) == 1        # equal to tol: dropped
    assert LR.getrank(np.array([1.0, 2.0001 * eps1])) == 2
    # eps of the first value, not of 1: s1 = 3 has eps 2^-51
    assert LR.getrank(np.array([3.0, 4 * eps1])) == 1
    assert LR.getrank(np.array([3.0, 4.001 * eps1])) == 2
    # only trailing values are dropped
    assert LR.getrank(np.array([1.0, 1e-30, 0.5])) == 3


def check(out, want_dense, rng_rank=None):
    U, S, V = out
    assert np.allclose(U.T @ U, np.eye(len(S)), atol=1e-13)
    assert np.allclose(V.T @ V, np.eye(len(S)), atol=1e-13)
    assert np.all(np.diff(S) <= 0)
    D = LR.dense(out)
    assert np.linalg.norm(D - want_dense) <= 1e-13 * np.linalg.norm(want_dense)
    sv = np.linalg.svd(want_dense, compute_uv=False)
    assert np.allclose(S, sv[:len(S)], rtol=0, atol=1e-13 * sv[0])
    if rng_rank is not None:
        assert len(S) == rng_rank


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_add_sub(sign):
    rng = np.random.default_rng(1)
    L1, L2 = rand_lr(rng, 200, 150, 12), rand_lr(rng, 200, 150, 9)
    check(LR.add(L1, L2, sign), LR.dense(L1) + sign * LR.dense(L2), 21)


def test_add_shared_columns_truncates():
    """L + L has the rank of L: the doubled columns leave singular values at
    round-off, below getrank's tolerance."""
    rng = np.random.default_rng(2)
    L = rand_lr(rng, 120, 100, 10)
    out = LR.add(L, L)
    check(out, 2 * LR.dense(L), 10)
    assert np.allclose(out[1], 2 * L[1], rtol=1e-13)


def test_mul_adjmul():
    rng = np.random.default_rng(3)
    L1, L2 = rand_lr(rng, 90, 70, 8), rand_lr(rng, 70, 60, 11)
    check(LR.mul(L1, L2), LR.dense(L1) @ LR.dense(L2), 8)
    L3 = rand_lr(rng, 90, 40, 6)
    check(LR.adjmul(L1, L3), LR.dense(L1).T @ LR.dense(L3), 6)


def test_dense_times_lowrank():
    rng = np.random.default_rng(4)
    L = rand_lr(rng, 80, 50, 7)
    A = rng.standard_normal((60, 80))
    check(LR.dmul(A, L), A @ LR.dense(L), 7)
    B = rng.standard_normal((50, 45))
    check(LR.muld(L, B), LR.dense(L) @ B, 7)
    d = rng.random(80) + 0.5
    check(LR.dmul(np.diag(d), L), d[:, None] * LR.dense(L), 7)


def test_svdtrunc_of_a_low_rank_matrix():
    rng = np.random.default_rng(5)
    L = rand_lr(rng, 64, 48, 5)
    out = LR.svdtrunc(LR.dense(L))
    check(out, LR.dense(L), 5)
    assert np.allclose(out[1], L[1], rtol=1e-13)
