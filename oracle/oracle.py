"""ctypes front-end of the CPU oracle (oracle/hmx_oracle.c).

TEST INFRASTRUCTURE ONLY. Imported by tests/, by the smoke check and by the
cpu_baseline leg of bench.py, always as the checker. The product package never
imports this module.

Index conventions follow the reference (src/KernelMatrix.jl, src/algebra.jl):
leaf-kernel offsets are 1-based, block records use 0-based row/col offsets.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
LIB_PATH_F32 = os.path.join(HERE, "liboracle_f32.so")

CAUCHY, COULOMB, COULOMB3, LOG = 0, 1, 2, 3
CODE_KM, CODE_BARY, CODE_DENSE = 1, 2, 3

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int64)
_lib = None
_lib32 = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _setup(L, _dp, real):
    L.or_kernel.restype = real
    L.or_kernel.argtypes = [ctypes.c_int, real, real]
    L.or_blockrank.restype = ctypes.c_int
    L.or_blockrank.argtypes = [ctypes.c_int]
    L.or_sinpi.restype = real
    L.or_sinpi.argtypes = [real]
    L.or_chebyshevpoints.argtypes = [_dp, ctypes.c_int, ctypes.c_int]
    L.or_chebyshevbarycentricweights.argtypes = [_dp, ctypes.c_int, ctypes.c_int]
    L.or_indsplit.restype = ctypes.c_int
    L.or_indsplit.argtypes = [_dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, real, real, _ip]
    L.or_tree_build.restype = ctypes.c_void_p
    L.or_tree_build.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int64, ctypes.c_int64,
                                real, real, real, real, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_int)]
    L.or_tree_free.argtypes = [ctypes.c_void_p]
    L.or_tree_stats.argtypes = [ctypes.c_void_p, _ip]
    L.or_tree_blocks.restype = ctypes.c_int64
    L.or_tree_blocks.argtypes = [ctypes.c_void_p, _ip, _dp, ctypes.c_int64]
    L.or_tree_matvec.restype = ctypes.c_int
    L.or_tree_matvec.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.or_tree_matvec_omp.restype = ctypes.c_int
    L.or_tree_matvec_omp.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
    L.or_tree_matvec_rows.restype = ctypes.c_int
    L.or_tree_matvec_rows.argtypes = [ctypes.c_void_p, _dp, _ip, ctypes.c_int64, _dp, ctypes.c_int]
    L.or_tree_getindex.restype = real
    L.or_tree_getindex.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64]
    L.or_tree_block_data.restype = ctypes.c_int
    L.or_tree_block_data.argtypes = [ctypes.c_void_p, ctypes.c_int64, _dp, _dp, _dp, _dp]
    for name in ("or_dense_mul", "or_dense_tmul"):
        getattr(L, name).argtypes = [_dp, _dp, ctypes.c_int64, ctypes.c_int64, _dp,
                                     ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_int64]
    for name in ("or_lowrank_mul", "or_lowrank_adjmul"):
        getattr(L, name).argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_int64, _dp, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_int64, ctypes.c_int64]
    L.or_bary_mul.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int64, ctypes.c_int64,
                              ctypes.c_int64, _dp, ctypes.c_int64, ctypes.c_int64]
    L.or_bary_evaluate.restype = real
    L.or_bary_evaluate.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, real, real]
    L.or_dense_matvec_rows.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int64, _dp, _ip,
                                       ctypes.c_int64, _dp, ctypes.c_int]
    L.or_even_nodes.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_double)]
    L.or_even_weights.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int, _dp, _dp]
    L.or_barycentricmatrix_weights.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int, _dp]
    L.or_even_getindex.restype = real
    L.or_even_getindex.argtypes = [_dp, _dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                   ctypes.c_int64, ctypes.c_int64]
    L.or_even_mul.argtypes = [_dp, _dp, _dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, _dp,
                              ctypes.c_int64, ctypes.c_int64]
    return L


def lib(f32=False):
    """The Float64 oracle, or its Float32 twin (same source, or_real = float)."""
    global _lib, _lib32
    if f32:
        if _lib32 is None:
            if not os.path.exists(LIB_PATH_F32):
                build()
            _lib32 = _setup(ctypes.CDLL(LIB_PATH_F32), _fp, ctypes.c_float)
        return _lib32
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _setup(ctypes.CDLL(LIB_PATH), _dp, ctypes.c_double)
    return _lib


def _d(a):
    return a.ctypes.data_as(_fp if a.dtype == np.float32 else _dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _dt(f32):
    return np.float32 if f32 else np.float64


def kernel(kid, x, y, f32=False):
    return lib(f32).or_kernel(kid, float(x), float(y))


def blockrank(mant_bits):
    return lib().or_blockrank(mant_bits)


def chebyshevpoints(n, kind=1, f32=False):
    out = np.zeros(n, dtype=_dt(f32))
    lib(f32).or_chebyshevpoints(_d(out), n, kind)
    return out


def chebyshevbarycentricweights(n, kind=1, f32=False):
    out = np.zeros(n, dtype=_dt(f32))
    lib(f32).or_chebyshevbarycentricweights(_d(out), n, kind)
    return out


def indsplit(x, first, length, a, b, f32=False):
    """Returns ((first1, len1), (first2, len2)) with 1-based firsts."""
    x = np.ascontiguousarray(x, dtype=_dt(f32))
    out = np.zeros(4, dtype=np.int64)
    rc = lib(f32).or_indsplit(_d(x), len(x), first, length, a, b, _i(out))
    if rc:
        raise IndexError("BoundsError in indsplit")
    return (int(out[0]), int(out[1])), (int(out[2]), int(out[3]))


class Tree:
    """KernelMatrix tree built by the oracle (src/KernelMatrix.jl:47-116)."""

    def __init__(self, kid, x, y, a, b, c, d, blocksize, rank, store=True, f32=False):
        self.f32 = f32
        self.dt = _dt(f32)
        self.L = lib(f32)
        self.x = np.ascontiguousarray(x, dtype=self.dt)
        self.y = np.ascontiguousarray(y, dtype=self.dt)
        self.kid, self.rank, self.blocksize = kid, rank, blocksize
        quirk = ctypes.c_int(0)
        h = self.L.or_tree_build(kid, _d(self.x), _d(self.y), len(self.x), len(self.y), a, b, c, d,
                                blocksize, rank, 1 if store else 0, ctypes.byref(quirk))
        if not h:
            raise ValueError("oracle tree build failed (bounds error or unbounded recursion)")
        self.h = h
        self.quirk = quirk.value
        self.store = store

    def __del__(self):
        try:
            if getattr(self, "h", None) and getattr(self, "L", None) is not None:
                self.L.or_tree_free(self.h)
        except Exception:
            pass
        self.h = None

    def stats(self):
        out = np.zeros(9, dtype=np.int64)
        self.L.or_tree_stats(self.h, _i(out))
        keys = ["d", "k1", "k2", "dense", "far", "dense_entries", "far_rows", "far_cols", "depth"]
        return dict(zip(keys, (int(v) for v in out)))

    def blocks(self):
        """Leaf blocks in traversal order: ints (nb,6) = code, depth, row0,
        col0, m, n; dbls (nb,4) = a, b, c, d."""
        nb = self.L.or_tree_blocks(self.h, None, None, 0)
        ints = np.zeros((max(nb, 1), 6), dtype=np.int64)
        dbls = np.zeros((max(nb, 1), 4), dtype=self.dt)
        self.L.or_tree_blocks(self.h, _i(ints), _d(dbls), nb)
        return ints[:nb], dbls[:nb]

    def block_data(self, k):
        ints, _ = self.blocks()
        code, _, _, _, m, n = ints[k]
        r = self.rank
        if code == CODE_DENSE:
            A = np.zeros(m * n, dtype=self.dt)
            self.L.or_tree_block_data(self.h, k, _d(A), None, None, None)
            return {"A": A.reshape((m, n), order="F")}
        U, V, F = (np.zeros(k, dtype=self.dt) for k in (m * r, n * r, r * r))
        self.L.or_tree_block_data(self.h, k, None, _d(U), _d(V), _d(F))
        return {"U": U.reshape((m, r), order="F"), "V": V.reshape((n, r), order="F"),
                "F": F.reshape((r, r), order="F")}

    def matvec(self, v, u=None):
        """Faithful mul!: returns u + H*v (u defaults to zeros)."""
        v = np.ascontiguousarray(v, dtype=self.dt)
        u = np.zeros(len(self.x), dtype=self.dt) if u is None else np.array(u, dtype=self.dt)
        if self.L.or_tree_matvec(self.h, _d(u), _d(v)):
            raise ValueError("tree has no stored factors")
        return u

    def matvec_omp(self, v, nthreads=0, u=None):
        v = np.ascontiguousarray(v, dtype=self.dt)
        u = np.zeros(len(self.x), dtype=self.dt) if u is None else np.array(u, dtype=self.dt)
        if self.L.or_tree_matvec_omp(self.h, _d(u), _d(v), nthreads):
            raise ValueError("tree has no stored factors")
        return u

    def matvec_rows(self, v, rows, nthreads=0):
        v = np.ascontiguousarray(v, dtype=self.dt)
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        out = np.zeros(len(rows), dtype=self.dt)
        self.L.or_tree_matvec_rows(self.h, _d(v), _i(rows), len(rows), _d(out), nthreads)
        return out

    def getindex(self, i, j):
        return self.L.or_tree_getindex(self.h, i, j)


def dense_matvec(kid, x, y, v, rows=None, nthreads=0):
    """sum_j f(x_i, y_j) v_j in long double, for rows (0-based) or all rows."""
    x, y, v = f64(x), f64(y), f64(v)
    if rows is None:
        out = np.zeros(len(x))
        lib().or_dense_matvec_rows(kid, _d(x), _d(y), len(y), _d(v), None, len(x), _d(out),
                                   nthreads)
        return out
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    out = np.zeros(len(rows))
    lib().or_dense_matvec_rows(kid, _d(x), _d(y), len(y), _d(v), _i(rows), len(rows), _d(out),
                               nthreads)
    return out


def dense_mul(y, A, x, istart, jstart, incx=1, incy=1):
    """src/algebra.jl:37-48 in the element type of y (Float64 or Float32)."""
    f32 = y.dtype == np.float32
    A = np.asfortranarray(A, dtype=_dt(f32))
    x = np.ascontiguousarray(x, dtype=_dt(f32))
    m, n = A.shape
    lib(f32).or_dense_mul(_d(y), _d(A), m, n, _d(x), istart, jstart, incx, incy)
    return y


def dense_tmul(y, A, x, istart, jstart, incx=1, incy=1):
    """src/algebra.jl:52-65 in the element type of y (Float64 or Float32)."""
    f32 = y.dtype == np.float32
    A = np.asfortranarray(A, dtype=_dt(f32))
    x = np.ascontiguousarray(x, dtype=_dt(f32))
    m, n = A.shape
    lib(f32).or_dense_tmul(_d(y), _d(A), m, n, _d(x), istart, jstart, incx, incy)
    return y


def lowrank_mul(y, U, S, V, x, istart, jstart, incx=1, incy=1, adjoint=False):
    U = np.asfortranarray(U, dtype=np.float64)
    V = np.asfortranarray(V, dtype=np.float64)
    S = f64(S)
    fn = lib().or_lowrank_adjmul if adjoint else lib().or_lowrank_mul
    fn(_d(y), U.ctypes.data_as(_dp), _d(S), V.ctypes.data_as(_dp), U.shape[0], V.shape[0],
       len(S), _d(x), istart, jstart, incx, incy)
    return y


def bary_mul(u, U, F, V, v, istart, jstart):
    U = np.asfortranarray(U, dtype=np.float64)
    V = np.asfortranarray(V, dtype=np.float64)
    F = np.asfortranarray(F, dtype=np.float64)
    lib().or_bary_mul(_d(u), U.ctypes.data_as(_dp), F.ctypes.data_as(_dp),
                      V.ctypes.data_as(_dp), U.shape[0], V.shape[0], F.shape[0], _d(v),
                      istart, jstart)
    return u


def bary_evaluate(tx, ty, lx, ly, F, x, y, f32=False):
    """evaluate(B::BarycentricPoly2D, x, y) at each pair (x[t], y[t])."""
    dt = _dt(f32)
    tx, ty, lx, ly = (np.ascontiguousarray(a, dtype=dt) for a in (tx, ty, lx, ly))
    F = np.asfortranarray(F, dtype=dt)
    fn = lib(f32).or_bary_evaluate
    xs = np.atleast_1d(np.asarray(x, dtype=dt))
    ys = np.atleast_1d(np.asarray(y, dtype=dt))
    out = np.array([fn(len(tx), _d(tx), _d(ty), _d(lx), _d(ly), _d(F), float(a), float(b))
                    for a, b in zip(xs, ys)], dtype=dt)
    return out


# ---------------------------------------------------- EvenBarycentricMatrix

def even_nodes(a, b, n, f32=False):
    """t_k = (a+b)/2 + (b-a)*x_k/2 where f is evaluated (always Float64)."""
    t = np.zeros(n, dtype=np.float64)
    lib(f32).or_even_nodes(a, b, n, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return t


def even_weights(a, b, n, f32=False):
    """(w, W) of EvenBarycentricMatrix: w has b-a+1 entries, W is n x (b-a+1)."""
    m = b - a + 1
    w = np.zeros(m, dtype=_dt(f32))
    W = np.zeros((n, m), dtype=_dt(f32), order="F")
    lib(f32).or_even_weights(a, b, n, _d(w), _d(W))
    return w, W


def barycentricmatrix_weights(a, b, n, f32=False):
    """W of barycentricmatrix: (b-a+1) x n."""
    W = np.zeros((b - a + 1, n), dtype=_dt(f32), order="F")
    lib(f32).or_barycentricmatrix_weights(a, b, n, _d(W))
    return W


def even_getindex(W, F, i, j):
    """getindex(B, i, j) with 1-based i, j."""
    f32 = W.dtype == np.float32
    W = np.asfortranarray(W)
    F = np.asfortranarray(F, dtype=W.dtype)
    n, m = W.shape
    return lib(f32).or_even_getindex(_d(W), _d(F), m, F.shape[0], n, i, j)


def even_mul(u, W, F, v, istart=1, jstart=1):
    """mul!(u, B, v, istart, jstart) of src/algebra.jl:168-239, in u's type."""
    f32 = u.dtype == np.float32
    W = np.asfortranarray(W, dtype=u.dtype)
    F = np.asfortranarray(F, dtype=u.dtype)
    v = np.ascontiguousarray(v, dtype=u.dtype)
    n, m = W.shape
    lib(f32).or_even_mul(_d(u), _d(W), _d(F), m, F.shape[0], n, _d(v), istart, jstart)
    return u
