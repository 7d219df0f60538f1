"""Pins the numpy restatement of the hierarchical Cholesky factorization
(oracle/cholesky.py, src/cholesky.jl:12-228) by identities that do not reuse
its recursion: R is block upper triangular with R' R = A (the dense matrix of
the upper-stored A), R equals the unique Cholesky factor of that dense
matrix, and solvetransposed solves R' x = b, for all eight block-kind cases
of the reference."""
import numpy as np
import pytest

import cholesky as OC
import hmx


def to_oracle(H):
    """hmx.HierarchicalMatrix (host data) -> OTree."""
    if isinstance(H, hmx.HierarchicalMatrix):
        T = OC.OTree(*H.assigned.shape)
        M, N = H.assigned.shape
        for m in range(M):
            for n in range(N):
                if H.assigned[m, n]:
                    T.set(m, n, to_oracle(H.blocks[m][n]))
        return T
    if isinstance(H, hmx.LowRankMatrix):
        return (H.U, H.S, H.V)
    return np.array(H)


def case_matrix(n, leaf, dense_offdiag=(), dense_diag_at=(), seed=0):
    x = hmx.workloads.spd_points(n, seed)
    K = hmx.workloads.spd_kernel(x)
    return K, hmx.workloads.spd_hierarchical(K, leaf, dense_offdiag=dense_offdiag,
                                             dense_diag_at=dense_diag_at)


# (A11, A12, A22) kinds at the root: 1 tree, 2 LowRank, 3 dense
CASES = {
    (1, 2, 1): dict(),
    (1, 3, 1): dict(dense_offdiag=(0,)),
    (3, 2, 3): dict(dense_diag_at=(1,)),
    (3, 3, 3): dict(dense_offdiag=(0,), dense_diag_at=(1,)),
}


def build_case(kinds, n=256, leaf=48):
    if kinds == (1, 2, 3) or kinds == (1, 3, 3) or kinds == (3, 2, 1) or kinds == (3, 3, 1):
        # a dense half beside a tree: assemble the root by hand
        x = hmx.workloads.spd_points(n, 3)
        K = hmx.workloads.spd_kernel(x)
        s = n // 2
        first = (hmx.workloads.spd_hierarchical(K[:s, :s], leaf) if kinds[0] == 1
                 else np.asfortranarray(K[:s, :s]))
        second = (hmx.workloads.spd_hierarchical(K[s:, s:], leaf) if kinds[2] == 1
                  else np.asfortranarray(K[s:, s:]))
        B = K[:s, s:]
        if kinds[1] == 2:
            U, S, Vt = np.linalg.svd(B, full_matrices=False)
            r = int(np.sum(S > 1e-13 * S[0]))
            off = hmx.LowRankMatrix(U[:, :r], S[:r], Vt[:r].T)
        else:
            off = np.asfortranarray(B)
        H = hmx.HierarchicalMatrix(2, 2)
        H[hmx.Block(1), hmx.Block(1)] = first
        H[hmx.Block(1), hmx.Block(2)] = off
        H[hmx.Block(2), hmx.Block(2)] = second
        return K, H
    return case_matrix(n, leaf, **CASES[kinds])


ALL = [(a, b, c) for a in (1, 3) for b in (2, 3) for c in (1, 3)]


@pytest.mark.parametrize("kinds", ALL)
def test_factor_identities(kinds):
    K, H = build_case(kinds)
    A = to_oracle(H)
    assert tuple(int(v) for v in (A.codes[0, 0], A.codes[0, 1], A.codes[1, 1])) == kinds
    Ad = OC.symmetric_upper(OC.densify(A))
    R = OC.chol(A)
    Rd = OC.densify(R)
    assert np.array_equal(np.tril(Rd, -1), np.zeros_like(Rd))
    assert np.linalg.norm(Rd.T @ Rd - Ad) <= 1e-13 * np.linalg.norm(Ad)
    Rc = np.linalg.cholesky(Ad).T
    assert np.linalg.norm(Rd - Rc) <= 1e-11 * np.linalg.norm(Rc)
    # block kinds of R follow A's: R12 like A12; R22 a tree only if A22 is a
    # tree and A12 is low rank (dense R12' R12 makes the difference dense)
    assert int(R.codes[0, 1]) == kinds[1]
    assert int(R.codes[1, 1]) == (1 if kinds[2] == 1 and kinds[1] == 2 else 3)
    assert int(R.codes[1, 0]) == 0


@pytest.mark.parametrize("kinds", [(1, 2, 1), (3, 3, 3), (1, 3, 1)])
def test_solvetransposed(kinds):
    K, H = build_case(kinds)
    R = OC.chol(to_oracle(H))
    Rd = OC.densify(R)
    rng = np.random.default_rng(1)
    b = rng.standard_normal(K.shape[0])
    x = OC.solvet(R, b)
    assert np.linalg.norm(Rd.T @ x - b) <= 1e-12 * np.linalg.norm(b) * np.linalg.cond(Rd)
    Bm = rng.standard_normal((K.shape[0], 3))
    X = OC.solvet(R, Bm)
    assert np.allclose(Rd.T @ X, Bm, atol=1e-10)
    L = (rng.standard_normal((K.shape[0], 4)), np.array([3.0, 2.0, 1.0, 0.5]),
         rng.standard_normal((7, 4)))
    Lx = OC.solvet(R, L)
    assert np.allclose(Rd.T @ Lx[0], L[0], atol=1e-10) and Lx[1] is L[1] and Lx[2] is L[2]


def test_deep_tree_with_low_rank_schur_updates():
    """n = 1024, leaf 64: four levels; the Schur complements A22 - R12'R12
    go down the tree leaf by leaf (LowRank - LowRank recompressed)."""
    K, H = case_matrix(1024, 64)
    A = to_oracle(H)
    R = OC.chol(A)
    Rd = OC.densify(R)
    Ad = OC.symmetric_upper(OC.densify(A))
    assert np.linalg.norm(Rd.T @ Rd - Ad) <= 1e-12 * np.linalg.norm(Ad)


def test_not_positive_definite_raises():
    with pytest.raises(np.linalg.LinAlgError):
        OC.chol(np.array([[1.0, 2.0], [2.0, 1.0]]))
