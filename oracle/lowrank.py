"""LowRankMatrix algebra of HierarchicalMatrices.jl, restated in numpy.

TEST INFRASTRUCTURE ONLY: tests/ and the cpu_baseline leg of bench.py may
import this module; the product library never calls it.

The reference recompresses with LAPACK through Julia's LinearAlgebra
(`qr!` -> geqrt, `svd!` -> gesdd, thin factors); numpy.linalg.qr (geqrf +
orgqr, mode "reduced") and numpy.linalg.svd (gesdd, full_matrices=False) are
the same LAPACK algorithms. A LowRankMatrix is (U, S, V) with L = U diag(S) V^T
(src/LowRankMatrix.jl:28-38).
"""
import numpy as np


def getrank(s):
    """src/LowRankMatrix.jl:70-80: tol = length(s) * eps(s[1]); the largest r
    with s[r] > tol (s non-increasing)."""
    r = len(s)
    if r == 0:
        return 0
    tol = r * np.spacing(np.float64(s[0]))
    while r >= 1:
        if s[r - 1] > tol:
            return r
        r -= 1
    return r


def _svd(M):
    U, S, Vt = np.linalg.svd(M, full_matrices=False)
    return U, S, Vt.T


def svdtrunc(A):
    """src/LowRankMatrix.jl:82-86 (convert(LowRankMatrix, A))."""
    U, S, V = _svd(np.asarray(A, dtype=np.float64))
    r = getrank(S)
    return U[:, :r], S[:r], V[:, :r]


def add(L1, L2, sign=1.0):
    """L1 + L2 (sign = 1) and L1 - L2 (sign = -1), src/LowRankMatrix.jl:95-111."""
    (U1, S1, V1), (U2, S2, V2) = L1, L2
    QU, RU = np.linalg.qr(np.hstack([U1, U2]))
    QV, RV = np.linalg.qr(np.hstack([V1, V2]))
    SU, S, SV = _svd(RU @ np.diag(np.concatenate([S1, sign * np.asarray(S2)])) @ RV.T)
    r = getrank(S)
    return (QU @ SU)[:, :r], S[:r], (QV @ SV)[:, :r]


def mul(L1, L2):
    """L1 * L2, src/LowRankMatrix.jl:113-118."""
    (U1, S1, V1), (U2, S2, V2) = L1, L2
    SU, S, SV = _svd(np.diag(S1) @ V1.T @ U2 @ np.diag(S2))
    r = getrank(S)
    return (U1 @ SU)[:, :r], S[:r], (V2 @ SV)[:, :r]


def adjmul(L1, L2):
    """L1' * L2, src/LowRankMatrix.jl:120-126."""
    (U1, S1, V1), (U2, S2, V2) = L1, L2
    SU, S, SV = _svd(np.diag(S1) @ U1.T @ U2 @ np.diag(S2))
    r = getrank(S)
    return (V1 @ SU)[:, :r], S[:r], (V2 @ SV)[:, :r]


def dmul(A, L):
    """A * L for a dense (or Diagonal) A, src/LowRankMatrix.jl:128-134,144-150."""
    U, S, V = L
    Q, R = np.linalg.qr(A @ U)
    SU, Sn, SV = _svd(R @ np.diag(S))
    r = getrank(Sn)
    return (Q @ SU)[:, :r], Sn[:r], (V @ SV)[:, :r]


def muld(L, A):
    """L * A for a dense (or Diagonal) A, src/LowRankMatrix.jl:136-142,152-158."""
    U, S, V = L
    Q, R = np.linalg.qr(A.T @ V)
    SU, Sn, SV = _svd(np.diag(S) @ R.T)
    r = getrank(Sn)
    return (U @ SU)[:, :r], Sn[:r], (Q @ SV)[:, :r]


def dense(L):
    U, S, V = L
    return (U * S[None, :]) @ V.T
