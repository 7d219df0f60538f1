"""Hierarchical Cholesky factorization of HierarchicalMatrices.jl, restated
in numpy (src/cholesky.jl:12-228).

TEST INFRASTRUCTURE ONLY: tests/ and the cpu_baseline leg of bench.py may
import this module; the product library never calls it.

A tree is OTree (2 x 2 blocks, `codes` as the reference's `assigned`:
1 tree, 2 LowRankMatrix (U, S, V), 3 dense ndarray, 0 unassigned). Dense
Cholesky is LAPACK potrf through numpy (the reference: cholesky(Symmetric(A)),
which reads the upper triangle), triangular solves are LAPACK trtrs through
scipy (the reference: UpperTriangular(R)' \\ b), and the LowRankMatrix
products and differences are oracle/lowrank.py.
"""
import numpy as np
import scipy.linalg as sla

import lowrank as LR


class OTree:
    def __init__(self, M=2, N=2):
        self.blocks = [[None] * N for _ in range(M)]
        self.codes = np.zeros((M, N), dtype=np.int64)

    def set(self, m, n, x):
        self.blocks[m][n] = x
        self.codes[m, n] = 1 if isinstance(x, OTree) else (2 if isinstance(x, tuple) else 3)

    def blocksize(self, m, n, k):
        c = self.blocks[m][n]
        if c is None:
            return 0
        return shape(c)[k - 1]

    def shape(self):
        M, N = self.codes.shape
        return (sum(self.blocksize(m, N - 1, 1) for m in range(M)),
                sum(self.blocksize(0, n, 2) for n in range(N)))


def shape(x):
    if isinstance(x, OTree):
        return x.shape()
    if isinstance(x, tuple):
        return (x[0].shape[0], x[2].shape[0])
    return x.shape


def densify(x):
    """Dense matrix; unassigned blocks are zero (getindex of the reference)."""
    if isinstance(x, tuple):
        return LR.dense(x)
    if not isinstance(x, OTree):
        return np.array(x, dtype=np.float64)
    m, n = x.shape()
    out = np.zeros((m, n))
    M, N = x.codes.shape
    p = 0
    for a in range(M):
        q = 0
        for b in range(N):
            c = x.blocks[a][b]
            if c is not None:
                r, s = shape(c)
                out[p:p + r, q:q + s] = densify(c)
            q += x.blocksize(0, b, 2)
        p += x.blocksize(a, N - 1, 1)
    return out


def symmetric_upper(A):
    """Symmetric(A) of Julia (upper triangle)."""
    A = np.asarray(A)
    return np.triu(A) + np.triu(A, 1).T


def _check(A):
    c = (int(A.codes[0, 0]), int(A.codes[0, 1]), int(A.codes[1, 1]))
    if c[0] not in (1, 3) or c[1] not in (2, 3) or c[2] not in (1, 3):
        raise TypeError(f"no method for the block kinds {c}")


def chol(A):
    """hierarchicalcholesky, src/cholesky.jl:12-95."""
    if not isinstance(A, OTree):
        return np.linalg.cholesky(symmetric_upper(A)).T.copy()
    _check(A)
    R = OTree()
    R11 = chol(A.blocks[0][0])
    R.set(0, 0, R11)
    R12 = solvet(R11, A.blocks[0][1])
    R.set(0, 1, R12)
    R.set(1, 1, chol(schur(A.blocks[1][1], R12)))
    return R


def solvet(R, B):
    """solvetransposed, src/cholesky.jl:97-228: R^{-T} B (B vector, matrix or
    LowRankMatrix)."""
    if isinstance(B, tuple):
        return (solvet(R, B[0]), B[1], B[2])
    if not isinstance(R, OTree):
        return sla.solve_triangular(R, B, trans="T", lower=False)
    _check(R)
    s = shape(R.blocks[0][0])[1]
    x1 = solvet(R.blocks[0][0], B[:s])
    R12 = R.blocks[0][1]
    if isinstance(R12, tuple):
        U, S, V = R12
        t = U.T @ x1
        t = (S[:, None] * t) if t.ndim == 2 else S * t
        rhs = B[s:] - V @ t
    else:
        rhs = B[s:] - R12.T @ x1
    x2 = solvet(R.blocks[1][1], rhs)
    return np.concatenate([x1, x2])


def sub_lowrank(A, L):
    """A - L for a tree or dense A and a LowRankMatrix L: src/algebra.jl:460-491
    leaf by leaf (L[pr, qr] = (U[pr], S, V[qr]), src/LowRankMatrix.jl:60-62);
    LowRank - LowRank recompresses; Matrix - LowRank is dense."""
    U, S, V = L
    if isinstance(A, OTree):
        M, N = A.codes.shape
        G = OTree(M, N)
        p = 0
        for a in range(M):
            q = 0
            for b in range(N):
                c = A.blocks[a][b]
                if c is not None:
                    r, s = shape(c)
                    G.set(a, b, sub_lowrank(c, (U[p:p + r], S, V[q:q + s])))
                q += A.blocksize(0, b, 2)
            p += A.blocksize(a, N - 1, 1)
        return G
    if isinstance(A, tuple):
        return LR.add(A, L, -1.0)
    return np.asarray(A) - LR.dense(L)


def schur(A22, R12):
    """A22 - R12' R12."""
    if isinstance(R12, tuple):
        return sub_lowrank(A22, LR.adjmul(R12, R12))
    return densify(A22) - R12.T @ R12
