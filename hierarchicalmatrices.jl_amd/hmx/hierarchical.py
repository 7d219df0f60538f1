"""The generic block tree of HierarchicalMatrices.jl: `HierarchicalMatrix` with
`LowRankMatrix` and dense leaves, mirrored on the host and run on the device.

  LowRankMatrix(U, S, V)          src/LowRankMatrix.jl:28-41 (U diag(S) V')
  HierarchicalMatrix(M, N)        src/hierarchical.jl:54-69 (@hierarchical
                                  HierarchicalMatrix LowRankMatrix Matrix,
                                  src/HierarchicalMatrix.jl:1)
  H[Block(m), Block(n)] = A       src/hierarchical.jl:149-172 (assigned codes
                                  1 = HierarchicalMatrix, 2 = LowRankMatrix,
                                  3 = Matrix; other types stay unassigned)
  blocksize(H, m, n, k), size(H)  src/hierarchical.jl:31-47,76-97
  mul_(y, H, x, istart, jstart, INCX, INCY)
                                  src/HierarchicalMatrix.jl:24-52 with the leaf
                                  loops of src/algebra.jl:37-48,110-131
  rmul_(H, b), lmul_(b, H)        rmul!(H, Diagonal(b)), lmul!(Diagonal(b), H):
                                  scale!, src/HierarchicalMatrix.jl:15-16,54-102
                                  with src/algebra.jl:282-321
  add_col_(H, u, j[, istart])     add_col!, src/HierarchicalMatrix.jl:21,104-146
                                  with src/algebra.jl:324-380

The tree is flattened on first use into the leaves' absolute offsets (the
traversal's istart + p, jstart + q) and handed to hm_plan_from_blocks; the
product then runs the same gfx950 kernels as the KernelMatrix product.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib
from .barycentric import BarycentricFactors, BarycentricMatrix2D
from .kernelmatrix import Block, KernelMatrix, _stream_ptr


class LowRankMatrix:
    """U * Diagonal(S) * V' (src/LowRankMatrix.jl:28-38)."""

    __array_ufunc__ = None  # numpy defers `ndarray @ LowRankMatrix` to __rmatmul__

    def __init__(self, U, S, V, dtype=np.float64):
        """dtype: the element type T (float64, or float32 for the factors of
        a Float32 barycentricmatrix; the device arithmetic below is Float64)."""
        self.U = np.asfortranarray(U, dtype=dtype)
        self.S = np.ascontiguousarray(np.diag(S) if np.ndim(S) == 2 else S, dtype=dtype)
        self.V = np.asfortranarray(V, dtype=dtype)
        if self.U.shape[1] != len(self.S) or self.V.shape[1] != len(self.S):
            raise ValueError("U, S, V ranks differ")

    @property
    def shape(self):
        return (self.U.shape[0], self.V.shape[0])

    def size(self, k=None):
        return self.shape if k is None else self.shape[k - 1]

    def rank(self):
        return len(self.S)

    # recompressing arithmetic on the device, src/LowRankMatrix.jl:95-158
    def __add__(self, o):
        from . import lowrank
        return lowrank.add(self, o) if isinstance(o, LowRankMatrix) else NotImplemented

    def __sub__(self, o):
        from . import lowrank
        return lowrank.sub(self, o) if isinstance(o, LowRankMatrix) else NotImplemented

    def __matmul__(self, o):
        from . import lowrank
        if isinstance(o, LowRankMatrix):
            return lowrank.mul(self, o)
        if isinstance(o, np.ndarray) and o.ndim == 2:
            return lowrank.muld(self, o)
        return NotImplemented

    def __rmatmul__(self, o):
        from . import lowrank
        if isinstance(o, np.ndarray) and o.ndim == 2:
            return lowrank.dmul(o, self)
        return NotImplemented

    def __getitem__(self, ij):
        """Entry (i, j), 0-based; sum over k from r down to 1 as
        src/LowRankMatrix.jl:50-58."""
        i, j = ij
        ret = 0.0
        for k in range(self.rank() - 1, -1, -1):
            ret += self.U[i, k] * self.S[k] * self.V[j, k]
        return ret


def _child_size(c, k):
    if c is None:
        return 0
    if isinstance(c, HierarchicalMatrix):
        return c.size(k)
    return c.shape[k - 1]


class UpperTriangular:
    """UpperTriangular(A): the upper triangle of a Matrix or LowRankMatrix
    (LinearAlgebra's wrapper; a leaf type of the reference's test tree
    @hierarchical UpperTriangularHierarchicalMatrix Matrix (UpperTriangular,
    LowRankMatrix) Diagonal, test/runtests.jl:13)."""

    def __init__(self, data):
        self.data = data

    @property
    def shape(self):
        return self.data.shape

    @property
    def dtype(self):
        return _leaf_dtype(self.data)

    def dense(self):
        return np.triu(_dense(self.data))

    def __getitem__(self, ij):
        i, j = ij
        return self.data[i, j] if i <= j else 0.0


class Diagonal:
    """Diagonal(d) (LinearAlgebra's), a leaf type of the reference's test
    tree."""

    def __init__(self, d):
        self.diag = np.ascontiguousarray(d)

    @property
    def shape(self):
        return (len(self.diag), len(self.diag))

    @property
    def dtype(self):
        return self.diag.dtype.type

    def dense(self):
        return np.diag(self.diag)

    def __getitem__(self, ij):
        i, j = ij
        return float(self.diag[i]) if i == j else 0.0


def _dense(c):
    if isinstance(c, np.ndarray):
        return c
    if isinstance(c, LowRankMatrix):
        return (c.U * c.S[None, :]) @ c.V.T
    if isinstance(c, (BarycentricMatrix2D, BarycentricFactors)):
        return c.U @ c.F @ c.V.T
    return c.dense()


def _leaf_matches(A, t):
    """isinstance for a leaf type: a class, a tuple of classes, or a
    parametric wrapper (Wrapper, Inner) such as (UpperTriangular,
    LowRankMatrix), the reference's UpperTriangular{T, LowRankMatrix{T}}."""
    if isinstance(t, tuple) and len(t) == 2 and t[0] in (UpperTriangular,):
        return isinstance(A, t[0]) and isinstance(A.data, t[1])
    if t is np.ndarray:
        return isinstance(A, np.ndarray) and A.ndim == 2 and A.dtype in (np.float64, np.float32)
    return isinstance(A, t)


def hierarchical(name, *leaf_types):
    """@hierarchical Name Types... (src/hierarchical.jl:4-172): a tree class
    whose children are trees of the same class (assigned code 1) or leaves
    of the given types (codes 2, 3, ... in order; anything else stays
    unassigned). A type is a class (np.ndarray standing for Matrix), or a
    pair (Wrapper, Inner) for the reference's parametric Wrapper{T, Inner{T}}.
    The fields <Type>blocks (M x N object arrays) mirror the reference's;
    products run on the device like HierarchicalMatrix's, with leaves that are
    neither Matrix nor LowRankMatrix formed densely."""
    types = tuple(leaf_types)

    def tname(t):
        t = t[0] if isinstance(t, tuple) else t
        return "Matrix" if t is np.ndarray else t.__name__

    field_types = {f"{name}blocks": name}
    field_types.update({f"{tname(t)}blocks": t for t in types})

    def blocks_field(self, code):
        M, N = self.assigned.shape
        out = np.empty((M, N), dtype=object)
        for m in range(M):
            for n in range(N):
                out[m, n] = self.blocks[m][n] if self.assigned[m, n] == code else None
        return out

    ns = {"leaf_types": types, "field_types": field_types,
          "__doc__": f"@hierarchical {name} " + " ".join(tname(t) for t in types)}
    cls = type(name, (HierarchicalMatrix,), ns)
    for code, f in enumerate(field_types, start=1):
        setattr(cls, f, property(lambda self, c=code: blocks_field(self, c)))
    return cls


class HierarchicalMatrix:
    """An M x N block tree whose children are HierarchicalMatrix,
    LowRankMatrix or dense (2-D float64 numpy) blocks."""

    def __init__(self, M, N):
        self.blocks = [[None] * N for _ in range(M)]
        self.assigned = np.zeros((M, N), dtype=np.int64)
        self._plan = None

    def blocksize(self, m=None, n=None, k=None):
        """blocksize(H) = size(H.assigned); blocksize(H, m, n, k) = size of
        child (m, n) (1-based) along k, 0 when unassigned."""
        if m is None:
            return self.assigned.shape
        return _child_size(self.blocks[m - 1][n - 1], k)

    def size(self, k=None):
        M, N = self.assigned.shape
        rows = sum(self.blocksize(m, N, 1) for m in range(1, M + 1))
        cols = sum(self.blocksize(1, n, 2) for n in range(1, N + 1))
        return (rows, cols) if k is None else (rows, cols)[k - 1]

    @property
    def shape(self):
        return self.size()

    # the leaf types after the tree itself, in code order (2, 3, ...): the
    # fields of @hierarchical HierarchicalMatrix LowRankMatrix Matrix
    leaf_types = ((LowRankMatrix, BarycentricMatrix2D, BarycentricFactors), np.ndarray)

    def _code_of(self, A):
        """The `assigned` code of a child: 1 for a tree of this type, 2 + l
        for the l-th leaf type, 0 (unassigned) for anything else."""
        if type(A) is type(self):
            return 1
        for l, t in enumerate(self.leaf_types):
            if _leaf_matches(A, t):
                return 2 + l
        return 0

    def __setitem__(self, key, A):
        b1, b2 = key
        m, n = (b1.K if isinstance(b1, Block) else int(b1)), (b2.K if isinstance(b2, Block) else int(b2))
        code = self._code_of(A)
        if code == 0:  # no matching field: the reference leaves the block unassigned
            return
        if isinstance(A, np.ndarray):
            A = np.asfortranarray(A)
        self.blocks[m - 1][n - 1] = A
        self.assigned[m - 1, n - 1] = code
        self._plan = None

    def __getitem__(self, key):
        b1, b2 = key
        if isinstance(b1, Block):
            return self.blocks[b1.K - 1][b2.K - 1]
        return self.getindex(int(b1), int(b2))

    def getindex(self, i, j):
        """Entry (i, j), 0-based, by the block walk of src/hierarchical.jl:120-147."""
        M, N = self.assigned.shape
        m = 0
        while m < M and i >= self.blocksize(m + 1, N, 1):
            i -= self.blocksize(m + 1, N, 1)
            m += 1
        n = 0
        while n < N and j >= self.blocksize(1, n + 1, 2):
            j -= self.blocksize(1, n + 1, 2)
            n += 1
        if m >= M or n >= N:
            raise IndexError("index out of range")
        c = self.blocks[m][n]
        if c is None:
            return 0.0
        if isinstance(c, HierarchicalMatrix):
            return c.getindex(i, j)
        return c[i, j] if not isinstance(c, np.ndarray) else float(c[i, j])

    def leaves(self, istart=0, jstart=0):
        """Leaves in the traversal order of src/HierarchicalMatrix.jl:24-52,
        as (block, row0, col0) with 0-based absolute offsets."""
        out = []
        M, N = self.assigned.shape
        p = 0
        for m in range(M):
            q = 0
            for n in range(N):
                c = self.blocks[m][n]
                if isinstance(c, HierarchicalMatrix):
                    out.extend(c.leaves(istart + p, jstart + q))
                elif c is not None:
                    out.append((c, istart + p, jstart + q))
                q += self.blocksize(1, n + 1, 2)
            p += self.blocksize(m + 1, N, 1)
        return out

    def plan(self, stream=None):
        """The device plan (built on first use)."""
        if self._plan is None:
            self._plan = _BlocksPlan(self, stream)
        return self._plan

    def _drop_plans(self):
        self._plan = None
        for row in self.blocks:
            for c in row:
                if isinstance(c, HierarchicalMatrix):
                    c._drop_plans()

    # products, as KernelMatrix
    def mul(self, u, v, istart=1, jstart=1, INCX=1, INCY=1, **kw):
        return self.plan().mul(u, v, istart, jstart, INCX, INCY, **kw)

    def matvec(self, v, **kw):
        return self.plan().matvec(v, **kw)

    def __matmul__(self, v):
        return self.matvec(v)

    __mul__ = __matmul__


class _BlocksPlan(KernelMatrix):
    """hm_plan_from_blocks (or its Float32 twin) over the flattened leaves of
    a HierarchicalMatrix. The element type is that of the leaves (all float64
    or all float32)."""

    def __init__(self, H, stream=None):
        rows, cols = H.size()
        leaves = H.leaves()
        nl = len(leaves)
        dts = {_leaf_dtype(c) for c, _, _ in leaves}
        if len(dts) > 1:
            raise TypeError("leaves of one HierarchicalMatrix must share one element type")
        dt = dts.pop() if dts else np.float64
        self.dtype = dt
        kind = np.zeros(max(nl, 1), dtype=np.int32)
        r0 = np.zeros(max(nl, 1), dtype=np.int64)
        c0 = np.zeros_like(r0)
        mm = np.zeros_like(r0)
        nn = np.zeros_like(r0)
        rk = np.zeros_like(r0)
        DP = ctypes.POINTER(ctypes.c_float if dt == np.float32 else ctypes.c_double)
        A = (DP * max(nl, 1))()
        U = (DP * max(nl, 1))()
        S = (DP * max(nl, 1))()
        V = (DP * max(nl, 1))()
        keep = []

        def put(arr, dst, t):
            a = np.ascontiguousarray(np.asarray(arr).ravel(order="F"), dtype=dt)
            keep.append(a)
            dst[t] = a.ctypes.data_as(DP)

        for t, (c, i0, j0) in enumerate(leaves):
            r0[t], c0[t] = i0, j0
            mm[t], nn[t] = c.shape
            if isinstance(c, LowRankMatrix):
                kind[t] = 2
                rk[t] = c.rank()
                put(c.U, U, t)
                put(c.S, S, t)
                put(c.V, V, t)
            elif isinstance(c, (BarycentricMatrix2D, BarycentricFactors)):
                kind[t] = 4  # U F V' with the core F in the S slot
                rk[t] = c.rank()
                put(c.U, U, t)
                put(c.F, S, t)
                put(c.V, V, t)
            else:  # Matrix, or another leaf type formed densely
                kind[t] = 3
                put(_dense(c), A, t)
        IP = ctypes.POINTER(ctypes.c_int64)
        h = ctypes.c_void_p()
        ctor = lib.hm_plan_from_blocks_f32 if dt == np.float32 else lib.hm_plan_from_blocks
        check(ctor(rows, cols, nl, kind.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                   r0.ctypes.data_as(IP), c0.ctypes.data_as(IP), mm.ctypes.data_as(IP),
                   nn.ctypes.data_as(IP), rk.ctypes.data_as(IP), A, U, S, V,
                   _stream_ptr(stream), ctypes.byref(h)))
        self._h = h
        self.f = None
        self.rank = int(rk.max()) if nl else 0
        self.blocksize_ = 0
        self.flags = 0
        self.x = np.zeros(rows, dtype=dt)
        self.y = np.zeros(cols, dtype=dt)


def _leaf_dtype(c):
    if isinstance(c, (LowRankMatrix, BarycentricMatrix2D, BarycentricFactors)):
        return np.dtype(c.U.dtype).type
    return np.dtype(c.dtype).type if isinstance(c, np.ndarray) else c.dtype


class _FlatLeaves:
    """A leaf list with the two methods _BlocksPlan reads."""

    def __init__(self, rows, cols, leaves):
        self._shape = (int(rows), int(cols))
        self._leaves = [(c, int(i0), int(j0)) for c, i0, j0 in leaves]

    def size(self):
        return self._shape

    def leaves(self):
        return self._leaves


def blocks_plan(rows, cols, leaves, stream=None):
    """A device plan of a rows x cols matrix given as leaves (block, row0,
    col0) with 0-based offsets: dense numpy blocks, LowRankMatrix,
    BarycentricMatrix2D or BarycentricFactors (hm_plan_from_blocks). The
    product adds every leaf, as the reference's traversal does."""
    return _BlocksPlan(_FlatLeaves(rows, cols, leaves), stream)


def _scale_host(H, b, start, side):
    """scale!(H, b, jstart) (side 0) / scale!(b, H, istart) (side 1) on the host
    leaves, in the reference's loop order (src/HierarchicalMatrix.jl:54-102);
    start is 1-based."""
    M, N = H.assigned.shape
    if side == 0:
        q = 0
        for n in range(N):
            for m in range(M):
                c = H.blocks[m][n]
                if isinstance(c, HierarchicalMatrix):
                    _scale_host(c, b, start + q, 0)
                elif isinstance(c, LowRankMatrix):  # scale!(C.V, b, A.V, jstart)
                    c.V *= b[start + q - 1:start + q - 1 + c.V.shape[0], None]
                elif c is not None:  # C[i, j] = A[i, j] * b[jshift + j]
                    c *= b[None, start + q - 1:start + q - 1 + c.shape[1]]
            q += H.blocksize(1, n + 1, 2)
    else:
        p = 0
        for m in range(M):
            for n in range(N):
                c = H.blocks[m][n]
                if isinstance(c, HierarchicalMatrix):
                    _scale_host(c, b, start + p, 1)
                elif isinstance(c, LowRankMatrix):  # scale!(C.U, b, A.U, istart)
                    c.U *= b[start + p - 1:start + p - 1 + c.U.shape[0], None]
                elif c is not None:  # C[i, j] = A[i, j] * b[ishift + i]
                    c *= b[start + p - 1:start + p - 1 + c.shape[0], None]
            p += H.blocksize(m + 1, N, 1)


def _scale(H, b, side):
    b_host = b.detach().cpu().numpy() if hasattr(b, "detach") else np.asarray(b, dtype=np.float64)
    b_host = np.ascontiguousarray(b_host, dtype=np.float64)
    n = H.size(2 if side == 0 else 1)
    if len(b_host) < n:
        raise ValueError("b is shorter than the scaled dimension of H")
    _scale_host(H, b_host, 1, side)
    plan = H._plan
    for row in H.blocks:  # nested plans would no longer match their leaves
        for c in row:
            if isinstance(c, HierarchicalMatrix):
                c._drop_plans()
    if plan is not None:  # the device factors are scaled in place, as the host ones
        from .kernelmatrix import _torch
        torch = _torch()
        bd = torch.from_numpy(b_host).cuda()
        check(lib.hm_plan_scale(plan._h, side, ctypes.c_void_p(bd.data_ptr()), 0,
                                _stream_ptr(None)))
        torch.cuda.synchronize()
    return H


def rmul_(H, b):
    """rmul!(H, Diagonal(b)): H := H * Diagonal(b) (scale!(H, b, 1),
    src/HierarchicalMatrix.jl:15,54-77): dense leaves scale their columns,
    LowRank leaves the rows of V. A device plan built for H is scaled in place
    (hm_plan_scale)."""
    return _scale(H, b, 0)


def lmul_(b, H):
    """lmul!(Diagonal(b), H): H := Diagonal(b) * H (scale!(b, H, 1),
    src/HierarchicalMatrix.jl:16,79-102): dense leaves scale their rows,
    LowRank leaves the rows of U."""
    return _scale(H, b, 1)


def _add_col_lowrank(L, u, istart, j):
    """add_col(L, u, istart, j) (src/algebra.jl:352-360): rank + 1 with the
    new term last, U1 = [U u[istart:...]], S1 = [S; 1], V1 = [V e_j]. (The
    Float64 method of :362-380 puts the term first but calls Vector{T}(r+1),
    which Julia no longer has; the matrix is the same.)"""
    m, n = L.shape
    U1 = np.concatenate([L.U, u[istart - 1:istart - 1 + m, None]], axis=1)
    S1 = np.concatenate([L.S, [1.0]])
    V1 = np.concatenate([L.V, np.zeros((n, 1))], axis=1)
    V1[j - 1, -1] = 1.0
    return LowRankMatrix(U1, S1, V1)


def _add_col_host(H, u, j, istart):
    """add_col! on the host leaves (src/HierarchicalMatrix.jl:104-146)."""
    M, N = H.assigned.shape
    n = 0
    q = H.blocksize(1, 1, 2)
    while q < j:
        n += 1
        q += H.blocksize(1, n + 1, 2)
    q -= H.blocksize(1, n + 1, 2)
    p = 0
    for m in range(M):
        c = H.blocks[m][n]
        if isinstance(c, HierarchicalMatrix):
            _add_col_host(c, u, j - q, istart + p)
        elif isinstance(c, LowRankMatrix):
            H.blocks[m][n] = _add_col_lowrank(c, u, istart + p, j - q)
        elif isinstance(c, np.ndarray):  # A[i + jm] += u[i + ishift]
            c[:, j - q - 1] += u[istart + p - 1:istart + p - 1 + c.shape[0]]
        elif c is not None:
            raise TypeError(f"add_col!: no method for a {type(c).__name__} leaf")
        p += H.blocksize(m + 1, N, 1)


def add_col_(H, u, j, istart=1):
    """add_col!(H, u, istart, j): column j (1-based) of H gains u[istart:...]
    (src/HierarchicalMatrix.jl:21,104-146): dense leaves add in place, LowRank
    leaves become their rank + 1 form, nested matrices recurse. A device plan
    built for H is updated in place (hm_plan_add_col) when every LowRank leaf
    holding column j has a free rank column (its rank below the plan's);
    otherwise it is rebuilt on the next product."""
    u_host = np.ascontiguousarray(u.detach().cpu().numpy() if hasattr(u, "detach") else u,
                                  dtype=np.float64)
    rows = H.size(1)
    if len(u_host) < istart - 1 + rows:
        raise ValueError("u is shorter than istart - 1 + size(H, 1)")
    _add_col_host(H, u_host, j, istart)
    plan = H._plan
    for row in H.blocks:  # nested plans would no longer match their leaves
        for c in row:
            if isinstance(c, HierarchicalMatrix):
                c._drop_plans()
    if plan is not None:
        from .kernelmatrix import _torch
        torch = _torch()
        ud = torch.from_numpy(u_host).cuda()
        rc = lib.hm_plan_add_col(plan._h, ctypes.c_void_p(ud.data_ptr()), istart - 1, j - 1,
                                 _stream_ptr(None))
        torch.cuda.synchronize()
        if rc == _lib.HM_ERR_UNSUPPORTED:
            H._plan = None  # no free rank column: rebuilt from the leaves when next used
        else:
            check(rc)
    return H
