"""Hierarchical Cholesky factorization and transposed triangular solve.

Mirrors src/cholesky.jl:12-228: `hierarchicalcholesky(A)` for a 2 x 2
HierarchicalMatrix A stored by its upper blocks (A[1,1] and A[2,2]
HierarchicalMatrix or Matrix, A[1,2] LowRankMatrix or Matrix; the eight
cases of the reference), or for a dense SPD block; `solvetransposed(R, b)`
for R from it and b a vector, Matrix or LowRankMatrix. The recursion runs on
the host as the reference's does; every numerical step runs on the device:

  dense leaf Cholesky            hm_dense_potrf
  R^{-T} on dense right sides    hm_dense_trsm_t
  R12' R12, A22 - (...), R12' x  hm_dense_gemm (LowRankMatrix factors through
                                 its diagonal argument) and hm_lowrank_op
                                 (L' * L, L - L with recompression)

The factors live in device memory during the factorization (torch float64
tensors, column-major); the host API converts at its boundary, and the
`*_device` functions work on device trees directly.
"""
import ctypes

import numpy as np

from ._lib import HM_ERR_INVALID, HMError, check, lib
from .hierarchical import HierarchicalMatrix, LowRankMatrix
from .kernelmatrix import Block, _stream_ptr, _torch
from . import lowrank as _lr


# ------------------------------------------------------------ device data

class DLowRank:
    """A LowRankMatrix on the device: U (m x r), S (r), V (n x r)."""

    def __init__(self, U, S, V):
        self.U, self.S, self.V = U, S, V

    @property
    def shape(self):
        return (self.U.shape[0], self.V.shape[0])


class DTree:
    """An M x N HierarchicalMatrix on the device: children DTree, DLowRank or
    dense column-major tensors; codes as the reference's `assigned`
    (1 HierarchicalMatrix, 2 LowRankMatrix, 3 Matrix, 0 unassigned)."""

    def __init__(self, M, N):
        self.blocks = [[None] * N for _ in range(M)]
        self.codes = np.zeros((M, N), dtype=np.int64)

    def set(self, m, n, x):
        self.blocks[m][n] = x
        self.codes[m, n] = 1 if isinstance(x, DTree) else (2 if isinstance(x, DLowRank) else 3)

    def blocksize(self, m, n, k):
        c = self.blocks[m][n]
        if c is None:
            return 0
        return c.size(k) if isinstance(c, DTree) else c.shape[k - 1]

    def size(self, k=None):
        M, N = self.codes.shape
        rows = sum(self.blocksize(m, N - 1, 1) for m in range(M))
        cols = sum(self.blocksize(0, n, 2) for n in range(N))
        return (rows, cols) if k is None else (rows, cols)[k - 1]

    @property
    def shape(self):
        return self.size()


def _colmajor(t):
    """The same values as a column-major (Fortran-order) tensor."""
    if t.dim() == 1:
        t = t.reshape(-1, 1)
    if t.stride(0) == 1 and (t.shape[1] == 1 or t.stride(1) >= max(1, t.shape[0])):
        return t
    return t.t().contiguous().t()


def _new(m, n):
    torch = _torch()
    return torch.zeros(n, m, dtype=torch.float64, device="cuda").t()  # column-major m x n


def _empty(m, n):
    """An uninitialized column-major m x n tensor (for outputs written whole)."""
    torch = _torch()
    return torch.empty(n, m, dtype=torch.float64, device="cuda").t()


def _copy(t):
    """A new column-major tensor with the values of t (a matrix or vector)."""
    t2 = t.reshape(t.shape[0], -1) if t.dim() == 1 else t
    out = _empty(*t2.shape)
    out.copy_(t2)
    return out


def _same(a, b):
    return (a.data_ptr() == b.data_ptr() and a.shape == b.shape and a.stride() == b.stride())


def _ld(t):
    """Leading dimension of a column-major matrix (a single column's stride
    along the size-1 dimension means nothing: its row count then)."""
    if t.dim() == 1 or t.shape[1] == 1:
        return max(1, t.shape[0])
    return max(1, t.stride(1))


def _i64(vals):
    a = np.asarray(vals, dtype=np.int64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def _ptrs(ts):
    arr = (ctypes.c_void_p * max(len(ts), 1))()
    for q, t in enumerate(ts):
        arr[q] = None if t is None else t.data_ptr()
    return arr


def potrf(A):
    """Upper Cholesky factor of a dense SPD block (a new tensor)."""
    R = _copy(A)
    n = R.shape[0]
    _, np_ = _i64([n])
    _, lp = _i64([_ld(R)])
    info, ip = _i64([0])
    check(lib.hm_dense_potrf(1, np_, _ptrs([R]), lp, ip, _stream_ptr(None)))
    return R


def potrf_async(A, infos):
    """Upper Cholesky factor of a dense SPD block (a new tensor), without
    waiting: the pivot report (0 or the failing pivot + 1) is appended to
    `infos` as a device tensor, to be read once by check_infos."""
    torch = _torch()
    R = _copy(A)
    n = R.shape[0]
    _, np_ = _i64([n])
    _, lp = _i64([_ld(R)])
    info = torch.empty(1, dtype=torch.int64, device="cuda")
    check(lib.hm_dense_potrf_dev(1, np_, _ptrs([R]), lp, ctypes.c_void_p(info.data_ptr()),
                                 _stream_ptr(None)))
    infos.append(info)
    return R


def check_infos(infos):
    """Raise for the first leaf whose factorization met a non-positive pivot
    (Julia's PosDefException), in the order the leaves were factored."""
    if not infos:
        return
    torch = _torch()
    vals = torch.cat(infos).cpu().numpy()
    bad = np.flatnonzero(vals)
    if len(bad):
        t = int(bad[0])
        raise HMError(HM_ERR_INVALID, f"hierarchicalcholesky: leaf {t} is not positive definite "
                      f"(pivot {int(vals[t])})")


def trsm_t(R, B, out=None):
    """R^{-T} B for an upper dense R (B n x k): into `out` (column-major,
    may be B itself) or a new tensor."""
    if out is None:
        X = _copy(B)
    else:
        X = out
        if not _same(X, B):
            X.copy_(B)
    n, k = X.shape
    _, np_ = _i64([n])
    _, kp = _i64([k])
    _, lr = _i64([_ld(R)])
    _, lb = _i64([_ld(X)])
    check(lib.hm_dense_trsm_t(1, np_, kp, _ptrs([R]), lr, _ptrs([X]), lb, _stream_ptr(None)))
    return X


def gemm(ta, tb, alpha, A, B, beta, C, D=None):
    """C = beta C + alpha op(A) diag(D) op(B), in place on the column-major C."""
    A, B = _colmajor(A), _colmajor(B)
    m, n = C.shape
    k = A.shape[0] if ta else A.shape[1]
    _, mnk = _i64([m, n, k])
    _, la = _i64([_ld(A)])
    _, lb = _i64([_ld(B)])
    _, lc = _i64([_ld(C)])
    check(lib.hm_dense_gemm(int(ta), int(tb), 1, mnk, float(alpha), _ptrs([A]), la,
                            _ptrs([D]), _ptrs([B]), lb, float(beta), _ptrs([C]), lc,
                            _stream_ptr(None)))
    return C


# ------------------------------------------------ host <-> device trees

def to_device(A):
    torch = _torch()
    if isinstance(A, HierarchicalMatrix):
        M, N = A.assigned.shape
        D = DTree(M, N)
        for m in range(M):
            for n in range(N):
                if A.assigned[m, n]:
                    D.set(m, n, to_device(A.blocks[m][n]))
        return D
    if isinstance(A, LowRankMatrix):
        dev = lambda a: _colmajor(torch.from_numpy(np.array(a, dtype=np.float64)).cuda())
        return DLowRank(dev(A.U), torch.from_numpy(np.array(A.S)).cuda(), dev(A.V))
    return _colmajor(torch.from_numpy(np.array(A, dtype=np.float64, order="F")).cuda())


def to_host(D):
    if isinstance(D, DTree):
        M, N = D.codes.shape
        H = HierarchicalMatrix(M, N)
        for m in range(M):
            for n in range(N):
                if D.codes[m, n]:
                    H[Block(m + 1), Block(n + 1)] = to_host(D.blocks[m][n])
        return H
    if isinstance(D, DLowRank):
        return LowRankMatrix(D.U.cpu().numpy(), D.S.cpu().numpy(), D.V.cpu().numpy())
    return np.asfortranarray(D.cpu().numpy())


def densify(D, out=None):
    """The dense matrix of a device tree (unassigned blocks zero, as the
    reference's getindex gives them), column-major."""
    if not isinstance(D, (DTree, DLowRank)):
        if out is None:
            return _colmajor(D)
        out.copy_(D)
        return out
    m, n = D.shape
    if out is None:
        out = _new(m, n)
    if isinstance(D, DLowRank):
        gemm(0, 1, 1.0, D.U, D.V, 0.0, out, D=D.S)
        return out
    M, N = D.codes.shape
    p = 0
    for a in range(M):
        q = 0
        for b in range(N):
            c = D.blocks[a][b]
            if c is not None:
                r, s = c.shape
                densify(c, out[p:p + r, q:q + s])
            q += D.blocksize(0, b, 2)
        p += D.blocksize(a, N - 1, 1)
    return out


# ------------------------------------------------------- the recursion

def _codes(A):
    return int(A.codes[0, 0]), int(A.codes[0, 1]), int(A.codes[1, 1])


def _check_case(A, what):
    if A.codes.shape != (2, 2):
        raise AssertionError("blocksize(A) == (2, 2)")
    c11, c12, c22 = _codes(A)
    if c11 not in (1, 3) or c12 not in (2, 3) or c22 not in (1, 3):
        raise TypeError(f"{what}: no method for the block kinds ({c11}, {c12}, {c22})")


def solvetransposed_device(R, B, out=None):
    """R^{-T} B: src/cholesky.jl:97-228. B: a tensor (vector or n x k) or a
    DLowRank (its U is solved, S and V kept). The result goes to `out` (a
    column-major n x k tensor, which may be B itself) or a new tensor; the
    halves of a tree are solved in place in their rows of it."""
    if isinstance(B, DLowRank):
        return DLowRank(solvetransposed_device(R, B.U), B.S, B.V)
    vec = B.dim() == 1
    B2 = B.reshape(B.shape[0], -1)
    X = out if out is not None else _empty(*B2.shape)
    if not isinstance(R, DTree):
        trsm_t(R, B2, out=X)
    else:
        _check_case(R, "solvetransposed")
        n = R.size(2)
        s = R.blocks[0][0].shape[1]
        x1 = X[:s]
        solvetransposed_device(R.blocks[0][0], B2[:s], out=x1)
        rhs = X[s:n]
        if not _same(rhs, B2[s:n]):
            rhs.copy_(B2[s:n])
        R12 = R.blocks[0][1]
        if isinstance(R12, DLowRank):
            # rhs -= V diag(S) (U' x1)
            T1 = _empty(R12.U.shape[1], x1.shape[1])
            gemm(1, 0, 1.0, R12.U, x1, 0.0, T1)
            gemm(0, 0, -1.0, R12.V, T1, 1.0, rhs, D=R12.S)
        else:
            gemm(1, 0, -1.0, R12, x1, 1.0, rhs)
        solvetransposed_device(R.blocks[1][1], rhs, out=rhs)
    return X.reshape(-1) if vec else X


def _lowrank_slice(L, rows, cols):
    """L[pr, qr] = LowRankMatrix(U[pr, :], S, V[qr, :]), src/LowRankMatrix.jl:60-62."""
    return DLowRank(L.U[rows[0]:rows[1]], L.S, L.V[cols[0]:cols[1]])


def _sub_lowrank(A, L, pending):
    """A - L for a DTree or dense A and a DLowRank L (src/algebra.jl:460-491 for
    a tree, leaf by leaf; Matrix - LowRankMatrix densely). LowRank - LowRank
    leaves are queued in `pending` and recompressed in one batched call."""
    if isinstance(A, DTree):
        M, N = A.codes.shape
        G = DTree(M, N)
        p = 0
        for a in range(M):
            q = 0
            for b in range(N):
                c = A.blocks[a][b]
                if c is not None:
                    r, s = c.shape
                    G.set(a, b, _sub_lowrank(c, _lowrank_slice(L, (p, p + r), (q, q + s)),
                                             pending))
                q += A.blocksize(0, b, 2)
            p += A.blocksize(a, N - 1, 1)
        return G
    if isinstance(A, DLowRank):
        slot = [None]
        pending.append((A, L, slot))
        return slot
    C = _copy(A)
    gemm(0, 1, -1.0, L.U, L.V, 1.0, C, D=L.S)
    return C


def _resolve(G):
    """Replace queued LowRank - LowRank slots by their results."""
    if isinstance(G, list):
        return G[0]
    if isinstance(G, DTree):
        M, N = G.codes.shape
        for a in range(M):
            for b in range(N):
                c = G.blocks[a][b]
                if c is not None:
                    G.set(a, b, _resolve(c))
        return G
    return G


def _schur(A22, R12):
    """A22 - R12' R12 (src/cholesky.jl:25,34,...)."""
    if isinstance(R12, DLowRank):
        # R12' * R12: LowRankMatrix' * LowRankMatrix with recompression
        P = _lr.batch("adjmul", [{"U1": R12.U, "S1": R12.S, "V1": R12.V, "U2": R12.U,
                                   "S2": R12.S, "V2": R12.V}])[0]
        P = DLowRank(_colmajor(P[0]), P[1].contiguous(), _colmajor(P[2]))
        pending = []
        G = _sub_lowrank(A22, P, pending)
        if pending:
            res = _lr.batch("sub", [{"U1": a.U, "S1": a.S, "V1": a.V, "U2": b.U, "S2": b.S,
                                     "V2": b.V} for a, b, _ in pending])
            for (a, b, slot), (U, S, V) in zip(pending, res):
                slot[0] = DLowRank(_colmajor(U), S.contiguous(), _colmajor(V))
        return _resolve(G)
    # dense R12: the generic AbstractMatrix difference is dense
    n2 = R12.shape[1]
    C = densify(A22, _new(n2, n2)) if isinstance(A22, DTree) else _copy(A22)
    gemm(1, 0, -1.0, R12, R12, 1.0, C)
    return C


def _cholesky_rec(A, infos):
    if not isinstance(A, DTree):
        return potrf_async(A, infos)
    _check_case(A, "hierarchicalcholesky")
    R = DTree(2, 2)
    R11 = _cholesky_rec(A.blocks[0][0], infos)
    R.set(0, 0, R11)
    R12 = solvetransposed_device(R11, A.blocks[0][1])
    R.set(0, 1, R12)
    R.set(1, 1, _cholesky_rec(_schur(A.blocks[1][1], R12), infos))
    return R


def hierarchicalcholesky_device(A):
    """src/cholesky.jl:12-95 on a device tree (or a dense device block). The
    leaf factorizations do not wait for their pivot reports; those are read
    once at the end, and a non-positive pivot raises HMError then (the
    reference raises PosDefException at that leaf)."""
    infos = []
    R = _cholesky_rec(A, infos)
    check_infos(infos)
    return R


# ------------------------------------------------------------ host API

def hierarchicalcholesky(A):
    """hierarchicalcholesky(A) for a HierarchicalMatrix (2 x 2, upper blocks)
    or a dense SPD ndarray; returns R of the same kind (R' R = A)."""
    return to_host(hierarchicalcholesky_device(to_device(A)))


def solvetransposed(R, b):
    """R^{-T} b for R from hierarchicalcholesky; b a 1-D or 2-D ndarray or a
    LowRankMatrix."""
    Rd = to_device(R)
    if isinstance(b, LowRankMatrix):
        return to_host(solvetransposed_device(Rd, to_device(b)))
    torch = _torch()
    bt = torch.from_numpy(np.array(b, dtype=np.float64)).cuda()
    x = solvetransposed_device(Rd, bt)
    return x.cpu().numpy()
