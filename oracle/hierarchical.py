"""scale! and add_col! of HierarchicalMatrices.jl's HierarchicalMatrix,
restated in numpy (src/HierarchicalMatrix.jl:24-146 and the leaf forms of
src/algebra.jl:282-380), with the strided traversal mul! they are checked
through.

TEST INFRASTRUCTURE ONLY: tests/ may import this module; the product library
never calls it.

A tree is an OTree of oracle/cholesky.py (codes as the reference's
`assigned`: 1 tree, 2 LowRankMatrix (U, S, V), 3 dense ndarray, 0
unassigned); the functions return updated trees and leave their input as it
was (the reference updates in place; the restatement copies so that a test
keeps the tree before and after).
"""
import copy

import numpy as np

import oracle as O
from cholesky import OTree


def blocksize(T, m, n, k):
    """blocksize(H, m, n, k) with 0-based m, n (src/hierarchical.jl:33-47)."""
    return T.blocksize(m, n, k)


def mul(T, y, x, istart, jstart, incx=1, incy=1):
    """mul!(y, H, x, istart, jstart, INCX, INCY), src/HierarchicalMatrix.jl:24-52:
    y += H x with the leaf loops of src/algebra.jl:37-48,110-131."""
    M, N = T.codes.shape
    p = 0
    for m in range(M):
        q = 0
        for n in range(N):
            code, c = T.codes[m, n], T.blocks[m][n]
            if code == 1:
                mul(c, y, x, istart + p, jstart + q, incx, incy)
            elif code == 2:
                O.lowrank_mul(y, c[0], c[1], c[2], x, istart + p, jstart + q, incx, incy)
            elif code == 3:
                O.dense_mul(y, c, x, istart + p, jstart + q, incx, incy)
            q += incx * blocksize(T, 0, n, 2)
        p += incy * blocksize(T, m, N - 1, 1)
    return y


def _scale_right_inplace(T, b, jstart):
    M, N = T.codes.shape
    q = 0
    for n in range(N):          # columns outer, as src/HierarchicalMatrix.jl:57-77
        for m in range(M):
            code, c = T.codes[m, n], T.blocks[m][n]
            if code == 1:
                _scale_right_inplace(c, b, jstart + q)
            elif code == 2:     # scale!(C.V, b, A.V, jstart): V[i, k] = V[i, k] b[jshift + i]
                U, S, V = c
                js = jstart + q - 1
                T.blocks[m][n] = (U, S, V * b[js:js + V.shape[0], None])
            elif code == 3:     # C[i, j] = A[i, j] b[jshift + j], src/algebra.jl:284-295
                js = jstart + q - 1
                T.blocks[m][n] = c * b[None, js:js + c.shape[1]]
        q += blocksize(T, 0, n, 2)


def scale_right(T, b, jstart=1):
    """scale!(H, b, jstart): H := H Diagonal(b[jstart:...])."""
    out = copy.deepcopy(T)
    _scale_right_inplace(out, np.asarray(b, dtype=np.float64), jstart)
    return out


def _scale_left_inplace(b, T, istart):
    M, N = T.codes.shape
    p = 0
    for m in range(M):          # rows outer, as src/HierarchicalMatrix.jl:80-102
        for n in range(N):
            code, c = T.codes[m, n], T.blocks[m][n]
            if code == 1:
                _scale_left_inplace(b, c, istart + p)
            elif code == 2:     # scale!(C.U, b, A.U, istart)
                U, S, V = c
                ii = istart + p - 1
                T.blocks[m][n] = (U * b[ii:ii + U.shape[0], None], S, V)
            elif code == 3:     # C[i, j] = A[i, j] b[ishift + i], src/algebra.jl:306-315
                ii = istart + p - 1
                T.blocks[m][n] = c * b[ii:ii + c.shape[0], None]
        p += blocksize(T, m, N - 1, 1)


def scale_left(b, T, istart=1):
    """scale!(b, H, istart): H := Diagonal(b[istart:...]) H."""
    out = copy.deepcopy(T)
    _scale_left_inplace(np.asarray(b, dtype=np.float64), out, istart)
    return out


def _add_col_lowrank(L, u, istart, j):
    """add_col(L, u, istart, j), the generic method of src/algebra.jl:352-360:
    U1 = [U u[istart:...]], Sigma1 = [Sigma; 1], V1 = [V 0] with V1[j, r+1] = 1
    (the new term last; the Float64 method of :362-380 builds the same matrix
    with the term first, but its Vector{T}(r+1) call no longer exists in Julia)."""
    U, S, V = L
    m, n = U.shape[0], V.shape[0]
    U1 = np.concatenate([U, u[istart - 1:istart - 1 + m, None]], axis=1)
    S1 = np.concatenate([S, [1.0]])
    V1 = np.concatenate([V, np.zeros((n, 1))], axis=1)
    V1[j - 1, -1] = 1.0
    return (U1, S1, V1)


def _add_col_inplace(T, u, istart, j):
    M, N = T.codes.shape
    n = 0
    q = blocksize(T, 0, n, 2)
    while q < j:                # src/HierarchicalMatrix.jl:111-116
        n += 1
        q += blocksize(T, 0, n, 2)
    q -= blocksize(T, 0, n, 2)
    p = 0
    for m in range(M):
        code, c = T.codes[m, n], T.blocks[m][n]
        if code == 1:
            _add_col_inplace(c, u, istart + p, j - q)
        elif code == 2:         # immutable leaf: replaced by add_col
            T.blocks[m][n] = _add_col_lowrank(c, u, istart + p, j - q)
        elif code == 3:         # add_col!: A[i + (j-1) m] += u[i + ishift]
            ii = istart + p - 1
            c = np.array(c, copy=True)
            c[:, j - q - 1] += u[ii:ii + c.shape[0]]
            T.blocks[m][n] = c
        p += blocksize(T, m, N - 1, 1)


def add_col(T, u, istart, j):
    """add_col!(H, u, istart, j): column j (1-based) of H gains u[istart:...]."""
    out = copy.deepcopy(T)
    _add_col_inplace(out, np.asarray(u, dtype=np.float64), istart, j)
    return out


def from_tree(H):
    """An OTree with copies of the leaves of a tree that has `assigned` and
    `blocks` (codes 1 tree, 2 (U, S, V), 3 dense)."""
    M, N = H.assigned.shape
    T = OTree(M, N)
    for m in range(M):
        for n in range(N):
            code, c = H.assigned[m, n], H.blocks[m][n]
            if code == 1:
                T.set(m, n, from_tree(c))
            elif code == 2:
                T.set(m, n, (np.array(c.U, dtype=np.float64), np.array(c.S, dtype=np.float64),
                             np.array(c.V, dtype=np.float64)))
            elif code == 3:
                T.set(m, n, np.array(c, dtype=np.float64))
    return T
