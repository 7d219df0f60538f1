"""EvenBarycentricMatrix and barycentricmatrix on the device.

Mirrors src/BarycentricMatrix.jl:5-89 and the leaf mul! of
src/algebra.jl:168-239. Rows are the integer grid a..b, columns c..d. The
caller's f is evaluated on the host at the nodes t_k (f is arbitrary Python,
as it is arbitrary Julia in the reference); W is formed on the device by
hm_even_weights; products run on the device through hm_even_* plans.

f is called once, as f(T, t, j) with t of shape (1, n) and j of shape (p, 1)
(numpy broadcasting), and must return a (p, n) array: F[j, k] = f(T, t_k, j)
of src/BarycentricMatrix.jl:39-44, converted to T.

The element type T is np.float64 or np.float32 (the reference's
EvenBarycentricMatrix{T}); a Float32 matrix forms its nodes, W and products
in Float32, through the _f32 entry points. The nodes t_k are Float64 for
either type, as the reference's (a+b)/2 promotes them.
"""
import ctypes

import numpy as np

from ._lib import check, lib
from .kernelmatrix import BLOCKRANK, _Vec, _stream_ptr, _torch, chebyshevbarycentricweights, \
    chebyshevpoints
from .hierarchical import LowRankMatrix


def _f32(dtype):
    dt = np.dtype(dtype)
    if dt not in (np.float64, np.float32):
        raise TypeError("the element type must be float64 or float32")
    return dt == np.float32


def _torch_dtype(dtype):
    torch = _torch()
    return torch.float32 if _f32(dtype) else torch.float64


def _ptr(a):
    ct = ctypes.c_float if a.dtype == np.float32 else ctypes.c_double
    return a.ctypes.data_as(ctypes.POINTER(ct))


def even_nodes(a, b, n, dtype=np.float64):
    """t_k = (a+b)/2 + (b-a)*x_k/2, src/BarycentricMatrix.jl:41, for
    T = dtype (Float64 values either way)."""
    t = np.zeros(n)
    fn = lib.hm_even_nodes_f32 if _f32(dtype) else lib.hm_even_nodes
    check(fn(int(a), int(b), int(n), t.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return t


def _device_weights(a, b, n, layout, dtype=np.float64):
    torch = _torch()
    m = b - a + 1
    W = torch.empty(m * n, dtype=_torch_dtype(dtype), device="cuda")
    w = torch.empty(m, dtype=_torch_dtype(dtype), device="cuda")
    fn = lib.hm_even_weights_f32 if _f32(dtype) else lib.hm_even_weights
    check(fn(int(a), int(b), int(n), layout, ctypes.c_void_p(W.data_ptr()),
             ctypes.c_void_p(w.data_ptr()), _stream_ptr(None)))
    return W, w


def even_F(f, a, b, c, d, n, dtype=np.float64):
    """F[j, k] = f(T, t_k, j), (d-c+1) x n, src/BarycentricMatrix.jl:39-44."""
    t = even_nodes(a, b, n, dtype)
    j = np.arange(c, d + 1, dtype=np.int64)
    F = np.asarray(f(np.dtype(dtype).type, t[None, :], j[:, None]), dtype=dtype)
    return np.asfortranarray(np.broadcast_to(F, (d - c + 1, n)))


class EvenBarycentricMatrix:
    """EvenBarycentricMatrix(T, f, a, b, c, d), src/BarycentricMatrix.jl:5-46.

    Fields as the reference's: a, b, c, d, x (nodes), lam (weights), w (m),
    beta (n, scratch in the reference), W (n x m), F (p x n), all of the
    element type dtype (T; rank BLOCKRANK(T) unless given)."""

    def __init__(self, f, a, b, c, d, rank=None, F=None, dtype=np.float64):
        a, b, c, d = int(a), int(b), int(c), int(d)
        if b < a or d < c:
            raise ValueError("empty grid")
        _f32(dtype)
        self.dtype = np.dtype(dtype)
        n = int(rank) if rank is not None else BLOCKRANK(self.dtype)
        self.a, self.b, self.c, self.d = a, b, c, d
        self.x = chebyshevpoints(n, dtype=self.dtype)
        self.lam = chebyshevbarycentricweights(n, dtype=self.dtype)
        if F is None:
            F = even_F(f, a, b, c, d, n, self.dtype)
        self.F = np.asfortranarray(F, dtype=self.dtype)
        if self.F.shape != (d - c + 1, n):
            raise ValueError("F must be (d-c+1) x rank")
        Wd, wd = _device_weights(a, b, n, 0, self.dtype)
        m = b - a + 1
        self.W = np.asfortranarray(Wd.cpu().numpy().reshape(m, n).T)
        self.w = wd.cpu().numpy()
        self.beta = np.zeros(n, dtype=self.dtype)
        self._plan = None

    # size, src/BarycentricMatrix.jl:3
    @property
    def shape(self):
        return (self.b - self.a + 1, self.d - self.c + 1)

    def size(self, k=None):
        return self.shape if k is None else self.shape[k - 1]

    def rank(self):
        return len(self.x)

    def __getitem__(self, ij):
        """getindex(B, i, j), src/BarycentricMatrix.jl:49-58, with 0-based i, j:
        nonzero where size(B,1) + size(B,2) + (i+1) + (j+1) is even."""
        i, j = ij
        m, p = self.shape
        ret = self.dtype.type(0)
        if (m + p + i + j) % 2 == 0:
            for k in range(self.rank()):
                ret += self.F[j, k] * self.W[k, i]
        return ret

    def _copy_with_F(self, F):
        out = object.__new__(EvenBarycentricMatrix)
        out.__dict__.update({k: v for k, v in self.__dict__.items()})
        out.F = np.asfortranarray(F)
        out.beta = np.zeros(self.rank(), dtype=self.dtype)
        out._plan = None
        return out

    # unary and binary +/- on F, src/algebra.jl:8-23
    def __pos__(self):
        return self._copy_with_F(+self.F)

    def __neg__(self):
        return self._copy_with_F(-self.F)

    def _binop(self, o, sign):
        if not isinstance(o, EvenBarycentricMatrix):
            return NotImplemented
        if self.dtype != o.dtype:
            raise TypeError("both matrices must have one element type")
        if (self.a, self.b, self.c, self.d) != (o.a, o.b, o.c, o.d):
            raise AssertionError("B1.a == B2.a && B1.b == B2.b && B1.c == B2.c && B1.d == B2.d")
        if self.rank() != o.rank():
            raise AssertionError("length(B1.x) == length(B2.x)")
        return self._copy_with_F(self.F + sign * o.F)

    def __add__(self, o):
        return self._binop(o, 1.0)

    def __sub__(self, o):
        return self._binop(o, -1.0)

    def plan(self):
        if self._plan is None:
            m, p = self.shape
            self._plan = EvenBarycentricLeaves(m, p, [(0, 0, self)], dtype=self.dtype)
        return self._plan

    def mul_(self, u, v, istart=1, jstart=1):
        """mul!(u, B, v, istart, jstart), src/algebra.jl:168-239:
        u[istart-1+i] += B(i, j) v[jstart-1+j] where (istart-1+i) +
        (jstart-1+j) is even. u and v are CUDA tensors or numpy arrays (copied
        back) of the matrix's element type."""
        return self.plan().mul_(u, v, istart, jstart)

    def __matmul__(self, v):
        """B * v: the generic AbstractMatrix product over getindex (the
        reference defines no * for the type), i.e. the parity of
        size(B,1) + size(B,2) + i + j."""
        torch = _torch()
        m, p = self.shape
        host = not isinstance(v, torch.Tensor)
        vt = torch.from_numpy(np.ascontiguousarray(v, dtype=self.dtype)).cuda() if host else v
        u = torch.zeros((m,) + tuple(vt.shape[1:]), dtype=_torch_dtype(self.dtype),
                        device=vt.device)
        if (m + p) % 2 == 0:
            self.plan().mul_(u, vt, 1, 1)
        else:
            # shift the column parity by one without moving the data: v
            # behind one leading zero, read from offset 1
            vs = torch.cat([torch.zeros((1,) + tuple(vt.shape[1:]), dtype=vt.dtype,
                                        device=vt.device), vt])
            self.plan().mul_(u, vs, 1, 2)
        return u.cpu().numpy() if host else u

    def close(self):
        if self._plan is not None:
            self._plan.close()
            self._plan = None


def barycentricmatrix(f, a, b, c, d, rank=None, F=None, dtype=np.float64):
    """barycentricmatrix(T, f, a, b, c, d), src/BarycentricMatrix.jl:61-89:
    LowRankMatrix(W, I, F) with W (m x n) formed on the device, T = dtype."""
    a, b, c, d = int(a), int(b), int(c), int(d)
    _f32(dtype)
    n = int(rank) if rank is not None else BLOCKRANK(dtype)
    if F is None:
        F = even_F(f, a, b, c, d, n, dtype)
    Wd, _ = _device_weights(a, b, n, 1, dtype)
    m = b - a + 1
    W = Wd.cpu().numpy().reshape(n, m).T
    return LowRankMatrix(W, np.ones(n, dtype=dtype), F, dtype=dtype)


class EvenLeaf:
    """A leaf given by its grids and F only (W is formed on the device when a
    plan is built), for plans whose leaves the host never inspects."""

    def __init__(self, a, b, c, d, F, dtype=np.float64):
        self.a, self.b, self.c, self.d = int(a), int(b), int(c), int(d)
        _f32(dtype)
        self.dtype = np.dtype(dtype)
        self.F = np.asfortranarray(F, dtype=self.dtype)
        if self.F.shape[0] != self.d - self.c + 1:
            raise ValueError("F must have d-c+1 rows")

    @property
    def shape(self):
        return (self.b - self.a + 1, self.d - self.c + 1)

    def rank(self):
        return self.F.shape[1]


class EvenBarycentricLeaves:
    """Any number of EvenBarycentricMatrix leaves of one rank placed in an
    nrows x ncols matrix, applied together by one device plan (hm_even_*).
    leaves: iterable of (row0, col0, B) with 0-based placement, B an
    EvenBarycentricMatrix or an EvenLeaf, all of one element type (dtype,
    for a plan without leaves)."""

    def __init__(self, nrows, ncols, leaves, stream=None, dtype=None):
        leaves = list(leaves)
        ranks = {B.rank() for _, _, B in leaves}
        if len(ranks) > 1:
            raise ValueError("all leaves must have one rank")
        types = {np.dtype(B.dtype) for _, _, B in leaves}
        if dtype is not None:
            types.add(np.dtype(dtype))
        if len(types) > 1:
            raise TypeError("all leaves must have one element type")
        self.dtype = types.pop() if types else np.dtype(np.float64)
        f32 = _f32(self.dtype)
        n = ranks.pop() if ranks else BLOCKRANK(self.dtype)
        self.nrows, self.ncols, self.rank = int(nrows), int(ncols), n
        cnt = len(leaves)
        r0 = np.array([int(l[0]) for l in leaves], dtype=np.int64)
        c0 = np.array([int(l[1]) for l in leaves], dtype=np.int64)
        abcd = np.array([[B.a, B.b, B.c, B.d] for _, _, B in leaves], dtype=np.int64).reshape(-1)
        Fs = [np.asfortranarray(B.F, dtype=self.dtype) for _, _, B in leaves]
        ct = ctypes.c_float if f32 else ctypes.c_double
        Fp = (ctypes.POINTER(ct) * max(cnt, 1))(*[_ptr(F) for F in Fs])
        h = ctypes.c_void_p()
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        create = lib.hm_even_create_f32 if f32 else lib.hm_even_create
        check(create(self.nrows, self.ncols, cnt, ip(r0), ip(c0), ip(abcd), n, Fp,
                     _stream_ptr(stream), ctypes.byref(h)))
        self._h = h
        self._shapes = [B.shape for _, _, B in leaves]

    @property
    def info(self):
        keys = ["nrows", "ncols", "nleaves", "rank", "w_entries", "f_entries", "n_up",
                "n_reduce", "n_down", "list_entries", "f32", "w_blocks"]
        out = np.zeros(len(keys), dtype=np.int64)
        check(lib.hm_even_info(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                               len(keys)))
        return dict(zip(keys, (int(v) for v in out)))

    def leaf_w_entries(self):
        """m*n summed over the leaves: the W entries the leaves would hold
        each on their own (the plan stores w_entries, one W per row count)."""
        return sum(m * self.rank for m, _ in self._shapes)

    def bytes_per_product(self, k=1):
        """Algorithmic HBM bytes of one product: the stored W and F once, v
        read and u read and written, per right-hand side for the vectors."""
        i = self.info
        e = self.dtype.itemsize
        return e * (i["w_entries"] + i["f_entries"]) + e * k * (self.ncols + 2 * self.nrows)

    def _call(self, u, v, ioff, joff, beta, stream):
        U = _Vec(u, True, self.dtype.type)
        V = _Vec(v, False, self.dtype.type)
        if U.k != V.k:
            raise ValueError("u and v have different numbers of columns")
        fn = lib.hm_even_matvec_f32 if self.dtype == np.float32 else lib.hm_even_matvec
        check(fn(self._h, U.ptr(), V.ptr(), int(ioff), int(joff), U.inc, V.inc, U.k, U.ld, V.ld,
                 int(beta), _stream_ptr(stream)))
        U.writeback()
        return u

    def mul_(self, u, v, istart=1, jstart=1, stream=None):
        """u[istart-1+I] += sum over leaves (the reference's mul! of each leaf
        at its placement); 1-based istart, jstart."""
        return self._call(u, v, istart - 1, jstart - 1, 1, stream)

    def matvec(self, v, stream=None):
        """u = H v (beta = 0): a new vector, or (nrows, k) for (ncols, k) v."""
        torch = _torch()
        host = not isinstance(v, torch.Tensor)
        vt = torch.from_numpy(np.ascontiguousarray(v, dtype=self.dtype)).cuda() if host else v
        u = torch.empty((self.nrows,) + tuple(vt.shape[1:]), dtype=_torch_dtype(self.dtype),
                        device=vt.device)
        self._call(u, vt, 0, 0, 0, stream)
        return u.cpu().numpy() if host else u

    def leaf_data(self, t):
        """(W, F) of leaf t as the device holds them: W n x m, F p x n."""
        m, p = self._shapes[t]
        W = np.zeros((self.rank, m), dtype=self.dtype, order="F")
        F = np.zeros((p, self.rank), dtype=self.dtype, order="F")
        fn = lib.hm_even_leaf_data_f32 if self.dtype == np.float32 else lib.hm_even_leaf_data
        check(fn(self._h, int(t), _ptr(W), _ptr(F)))
        return W, F

    def close(self):
        if getattr(self, "_h", None):
            lib.hm_even_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
