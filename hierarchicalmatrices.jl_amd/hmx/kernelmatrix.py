"""Host-side mirror of HierarchicalMatrices.jl's KernelMatrix API, backed by libhmx.

Names, argument meaning and accumulate semantics follow the reference:

  KernelMatrix(f, x, y, a, b, c, d)        src/KernelMatrix.jl:47
  mul_(u, H, v, istart, jstart)            mul!, src/KernelMatrix.jl:14-45 (u += H v)
  H * v, H @ v                             src/KernelMatrix.jl:5-12 (zeros, then mul!)
  BLOCKRANK(T), BLOCKSIZE(T)               src/HierarchicalMatrices.jl:5-7
  chebyshevpoints, chebyshevbarycentricweights, indsplit
                                           src/BarycentricMatrix.jl:92-136,299-307
  blocksize(H), size(H), H[i, j]           src/hierarchical.jl:31-47,120-147

f is either one of the named kernels of examples/Kernel.jl:34-37
(cauchykernel, coulombkernel, coulombprimekernel, logkernel), which the device
evaluates itself (every plan form, gradients), or any host function of two
scalars (the reference's arbitrary f): then the device builds the tree and the
bases U, V, which do not depend on f, while f is evaluated on the host at the
dense leaves' point pairs and the far blocks' node pairs and handed over
(kernel HM_USER, hm_plan_set_leaf_data); such a matrix has the stored and
hybrid forms only. Element access H[i, j] uses 0-based Python indices; vectors
may be torch tensors on the GPU (used in place, stream-ordered on torch's
current stream) or numpy arrays (copied to and from the device).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib


class NamedKernel:
    """A kernel of examples/Kernel.jl:34-37, known to the device code by id."""

    def __init__(self, name, kid, fn):
        self.name, self.kid, self._fn = name, kid, fn

    def __call__(self, x, y):
        return self._fn(np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64))

    def __repr__(self):
        return self.name


cauchykernel = NamedKernel("cauchykernel", 0, lambda x, y: 1.0 / (x - y))
coulombkernel = NamedKernel("coulombkernel", 1, lambda x, y: 1.0 / ((x - y) * (x - y)))
coulombprimekernel = NamedKernel("coulombprimekernel", 2,
                                 lambda x, y: 1.0 / ((x - y) * (x - y) * (x - y)))
logkernel = NamedKernel("logkernel", 3, lambda x, y: np.log(np.abs(x - y)))
KERNELS = {k.name: k for k in (cauchykernel, coulombkernel, coulombprimekernel, logkernel)}

_MANT = {np.dtype(np.float16): 10, np.dtype(np.float32): 23, np.dtype(np.float64): 52}


def BLOCKRANK(T=np.float64):
    """2*round(0.5*log_{3+sqrt 8}(1/eps(T))), src/HierarchicalMatrices.jl:5."""
    return lib.hm_blockrank(_MANT[np.dtype(T)])


def BLOCKSIZE(T=np.float64):
    """4*BLOCKRANK(T), src/HierarchicalMatrices.jl:7."""
    return 4 * BLOCKRANK(T)


def _dptr(a):
    if a.dtype == np.float32:
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def chebyshevpoints(n, kind=1, dtype=np.float64):
    """chebyshevpoints(T, n; kind), src/BarycentricMatrix.jl:92-111."""
    out = np.zeros(n, dtype=dtype)
    fn = lib.hm_chebyshevpoints_f32 if out.dtype == np.float32 else lib.hm_chebyshevpoints
    check(fn(_dptr(out), n, kind))
    return out


def chebyshevbarycentricweights(n, kind=1, dtype=np.float64):
    """chebyshevbarycentricweights(T, n; kind), src/BarycentricMatrix.jl:114-136."""
    out = np.zeros(n, dtype=dtype)
    fn = (lib.hm_chebyshevbarycentricweights_f32 if out.dtype == np.float32
          else lib.hm_chebyshevbarycentricweights)
    check(fn(_dptr(out), n, kind))
    return out


def indsplit(x, ir, a, b):
    """indsplit(x, ir, a, b): ir is a range of 1-based indices (Julia's
    first:last); returns the two halves as ranges of 1-based indices."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(4, dtype=np.int64)
    check(lib.hm_indsplit(_dptr(x), len(x), ir.start, len(ir), a, b,
                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
    f1, l1, f2, l2 = (int(v) for v in out)
    return range(f1, f1 + l1), range(f2, f2 + l2)


class Block:
    """Block index, src/block.jl:8-10."""

    def __init__(self, K):
        self.K = int(K)

    def __eq__(self, o):
        return isinstance(o, Block) and o.K == self.K

    def __repr__(self):
        return f"Block({self.K})"


def _torch():
    import torch
    return torch


def _stream_ptr(stream):
    torch = _torch()
    if stream is None:
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if hasattr(stream, "cuda_stream"):
        return ctypes.c_void_p(stream.cuda_stream)
    return ctypes.c_void_p(int(stream))


class _Vec:
    """A vector or (n, k) matrix of the plan's element type on the device,
    with its strides."""

    def __init__(self, a, writable, dtype=np.float64):
        torch = _torch()
        tdt = torch.float32 if dtype == np.float32 else torch.float64
        name = np.dtype(dtype).name
        self.host = None
        if isinstance(a, torch.Tensor):
            if a.device.type != "cuda" or a.dtype != tdt:
                raise TypeError(f"device vectors must be {name} CUDA tensors (the eltype of H)")
            self.t = a
        else:
            arr = np.asarray(a)
            if arr.dtype != dtype:
                raise TypeError(f"vectors must be {name} (the eltype of H, src/KernelMatrix.jl:17)")
            self.host = arr
            self.t = torch.from_numpy(np.array(arr, copy=True)).cuda()
        self.writable = writable
        t = self.t
        if t.dim() == 1:
            self.n, self.k, self.inc, self.ld = t.shape[0], 1, t.stride(0), 0
        elif t.dim() == 2:
            self.n, self.k = t.shape
            self.inc, self.ld = t.stride(0), t.stride(1)
        else:
            raise ValueError("vectors must be 1-D or 2-D")

    def ptr(self):
        return ctypes.c_void_p(self.t.data_ptr())

    def writeback(self):
        if self.host is not None and self.writable:
            torch = _torch()
            torch.cuda.synchronize()
            self.host[...] = self.t.cpu().numpy()


_FAR = {"f64": 0, "f32": _lib.HM_MV_FAR_F32, "bf16x3": _lib.HM_MV_FAR_BF16X3,
        "bf16x6": _lib.HM_MV_FAR_BF16X6}


def _mfma_flag(mfma):
    if mfma is None:
        return 0
    return _lib.HM_MV_MFMA if mfma else _lib.HM_MV_NO_MFMA


class KernelMatrix:
    """KernelMatrix(f, x, y, a, b, c, d), src/KernelMatrix.jl:47-116.

    x and y must be non-increasing (the reference's indsplit assumes it,
    src/BarycentricMatrix.jl:299-307); (a, b), (c, d) are their intervals as
    (upper, lower). The element type T is that of x (float32 or float64; pass
    dtype to force one); blocksize and rank default to BLOCKSIZE(T) and
    BLOCKRANK(T) (80 and 20 for Float64, 40 and 10 for Float32). A Float32
    matrix forms its tree, factors and products in FP32, as the reference's
    generic code does for T = Float32. matrix_free=True keeps only the points
    and regenerates every factor in each product; hybrid=True keeps the dense
    blocks and F and applies the far-field basis in Chebyshev form (no U, V).
    """

    def __init__(self, f, x, y, a, b, c, d, *, blocksize=None, rank=None, matrix_free=False,
                 hybrid=False, tree_only=False, stream=None, dtype=None):
        if isinstance(f, str):
            f = KERNELS[f]
        if not isinstance(f, NamedKernel) and not callable(f):
            raise TypeError("f must be a named kernel or a function of two scalars")
        user = not isinstance(f, NamedKernel)
        if user and matrix_free:
            raise TypeError("a host function f cannot be regenerated on the device: "
                            "stored or hybrid form only")
        self.f = f
        if dtype is None:
            dtype = np.float32 if np.asarray(x).dtype == np.float32 else np.float64
        self.dtype = np.dtype(dtype).type
        if self.dtype not in (np.float32, np.float64):
            raise TypeError("element type must be float32 or float64")
        self.rank = BLOCKRANK(self.dtype) if rank is None else int(rank)
        self.blocksize_ = 4 * self.rank if blocksize is None else int(blocksize)
        self.x = np.ascontiguousarray(x, dtype=self.dtype)
        self.y = np.ascontiguousarray(y, dtype=self.dtype)
        self.intervals = (float(a), float(b), float(c), float(d))
        flags = 0
        if matrix_free:
            flags |= _lib.HM_PLAN_MATRIX_FREE
        if hybrid:
            flags |= _lib.HM_PLAN_HYBRID
        if tree_only:
            flags |= _lib.HM_PLAN_TREE_ONLY
        self.flags = flags
        h = ctypes.c_void_p()
        sp = ctypes.c_void_p(0) if tree_only else _stream_ptr(stream)
        ctor = lib.hm_plan_from_points_f32 if self.f32 else lib.hm_plan_from_points
        if self.f32:  # the intervals are Float32 values too
            a, b, c, d = (float(np.float32(v)) for v in (a, b, c, d))
        check(ctor(_lib.HM_USER if user else f.kid, _dptr(self.x), len(self.x), _dptr(self.y),
                   len(self.y), float(a), float(b), float(c), float(d), self.blocksize_,
                   self.rank, flags, sp, ctypes.byref(h)))
        self._h = h
        if user and not tree_only:
            self._set_user_leaves(stream)

    def _set_user_leaves(self, stream=None):
        """f on the host at every dense leaf's (x_i, y_j) and every far leaf's
        mapped node pairs (t_m, s_n) (src/KernelMatrix.jl:57-60,82,104;
        src/BarycentricMatrix.jl:147-178), in traversal order, into the plan."""
        from .barycentric import eval_pairs, mapped_nodes
        dt = self.dtype
        ints, dbls = self.blocks()
        dense, far = [], []
        r = self.rank
        for (code, _, i0, j0, m, n), (a, b, c, d) in zip(ints, dbls):
            if m == 0 or n == 0:
                continue
            if code == 3:
                blk = eval_pairs(self.f, self.x[i0:i0 + m, None], self.y[None, j0:j0 + n], dt)
                dense.append(np.asarray(blk, dtype=dt).ravel(order="F"))
            else:
                tx, ty = mapped_nodes(a, b, r, dt), mapped_nodes(c, d, r, dt)
                far.append((tx, ty))
        A = np.concatenate(dense) if dense else np.zeros(0, dtype=dt)
        if far:
            TX = np.stack([t for t, _ in far])
            TY = np.stack([t for _, t in far])
            Fb = eval_pairs(self.f, TX[:, :, None], TY[:, None, :], dt)  # (nb, m, n)
            F = np.ascontiguousarray(np.transpose(Fb, (0, 2, 1))).ravel()  # F[m + r n] per block
        else:
            F = np.zeros(0, dtype=dt)
        A = np.ascontiguousarray(A, dtype=dt)
        F = np.ascontiguousarray(F, dtype=dt)
        fn = lib.hm_plan_set_leaf_data_f32 if self.f32 else lib.hm_plan_set_leaf_data
        check(fn(self._h, _dptr(A), len(A), _dptr(F), len(F), _stream_ptr(stream)))

    @property
    def f32(self):
        return getattr(self, "dtype", np.float64) == np.float32

    def _tdtype(self):
        torch = _torch()
        return torch.float32 if self.f32 else torch.float64

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib.hm_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- structure (src/hierarchical.jl)
    @property
    def info(self):
        out = np.zeros(len(_lib.INFO_KEYS), dtype=np.int64)
        check(lib.hm_plan_info(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                               len(out)))
        return dict(zip(_lib.INFO_KEYS, (int(v) for v in out)))

    @property
    def shape(self):
        return (len(self.x), len(self.y))

    def size(self, k=None):
        return self.shape if k is None else self.shape[k - 1]

    def blocksize(self):
        """blocksize(H) = size(H.assigned) of the root node (2, 2)."""
        return (2, 2)

    def blocks(self):
        """Leaf blocks in the reference's traversal order: (ints (nb, 6) =
        code, depth, row0, col0, m, n; dbls (nb, 4) = a, b, c, d)."""
        cnt = ctypes.c_int64(0)
        check(lib.hm_plan_blocks(self._h, None, None, 0, ctypes.byref(cnt)))
        nb = cnt.value
        ints = np.zeros((max(nb, 1), 6), dtype=np.int64)
        dbls = np.zeros((max(nb, 1), 4))
        check(lib.hm_plan_blocks(self._h, ints.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                 _dptr(dbls), nb, ctypes.byref(cnt)))
        return ints[:nb], dbls[:nb]

    # -- products
    def _mv(self, u, v, istart, jstart, incy, incx, beta, matrix_free, transpose, stream,
            cheb_basis=False, tuned=False, far_precision="f64", mfma=None):
        dt = np.float32 if self.f32 else np.float64
        U = _Vec(u, True, dt)
        V = _Vec(v, False, dt)
        if U.k != V.k:
            raise ValueError("u and v hold different numbers of right-hand sides")
        m, n = self.shape
        if transpose:
            m, n = n, m
        ioff, joff = istart - 1, jstart - 1
        incy = U.inc * incy
        incx = V.inc * incx
        if ioff < 0 or joff < 0:
            raise ValueError("istart and jstart are 1-based")
        if ioff + (m - 1) * incy >= U.n * U.inc or joff + (n - 1) * incx >= V.n * V.inc:
            raise ValueError("vector too short for H at the given offset and stride")
        if self.f32:
            fn = lib.hm_matvec_t_f32 if transpose else lib.hm_matvec_f32
        else:
            fn = lib.hm_matvec_t if transpose else lib.hm_matvec
        flags = (_lib.HM_MV_MATRIX_FREE if matrix_free else 0) | \
            (_lib.HM_MV_CHEB_BASIS if cheb_basis else 0) | (_lib.HM_MV_TUNED if tuned else 0) | \
            _FAR[far_precision] | _mfma_flag(mfma)
        check(fn(self._h, U.ptr(), V.ptr(), ioff * U.inc, joff * V.inc, incy, incx, U.k,
                 U.ld, V.ld, beta, flags, _stream_ptr(stream)))
        U.writeback()
        return u

    def mul(self, u, v, istart=1, jstart=1, INCX=1, INCY=1, *, matrix_free=False, stream=None):
        """u[istart + (i-1)INCY] += (H v)_i: the reference's accumulating mul!."""
        return self._mv(u, v, istart, jstart, INCY, INCX, 1, matrix_free, False, stream)

    def tmul(self, u, v, istart=1, jstart=1, INCX=1, INCY=1, *, matrix_free=False, stream=None):
        """u += H^T v (accumulating)."""
        return self._mv(u, v, istart, jstart, INCY, INCX, 1, matrix_free, True, stream)

    def matvec(self, v, *, matrix_free=False, transpose=False, cheb_basis=False, tuned=False,
               far_precision="f64", mfma=None, stream=None):
        """H v (or H^T v) into a new vector on the device of v. mfma=True runs
        forward products in groups of 16 right-hand sides on the MFMA passes,
        False never does, None leaves the choice to the library (MFMA from 16
        right-hand sides on)."""
        torch = _torch()
        ishost = not isinstance(v, torch.Tensor)
        dt = np.float32 if self.f32 else np.float64
        vt = torch.from_numpy(np.asarray(v, dtype=dt)).cuda() if ishost else v
        m = self.shape[1] if transpose else self.shape[0]
        out = torch.empty((m,) + tuple(vt.shape[1:]), dtype=self._tdtype(), device=vt.device)
        self._mv(out, vt, 1, 1, 1, 1, 0, matrix_free, transpose, stream, cheb_basis, tuned,
                 far_precision, mfma)
        if ishost:
            torch.cuda.synchronize()
            return out.cpu().numpy()
        return out

    def update_points(self, x, y, check=True, stream=None):
        """New point values x, y (device tensors of the plan's element type,
        non-increasing) for the same tree (hm_plan_update_points): the dense
        blocks and stored bases are formed again on the device, F stays. The
        device checks that every split of the tree falls at the same index;
        with check=True the answer is read back (one small synchronisation)
        and returned, with check=False it stays on the device (read it later
        through tree_status()), so the call can be captured into a graph."""
        torch = _torch()
        m, n = self.shape
        x = self._grad_vec(x, m, "x")
        y = self._grad_vec(y, n, "y")
        if x.stride(0) != 1 or y.stride(0) != 1:
            raise ValueError("x and y must be contiguous")
        if getattr(self, "_status", None) is None:
            self._status = torch.zeros(1, dtype=torch.int32, device=x.device)
        fn = lib.hm_plan_update_points_f32 if self.f32 else lib.hm_plan_update_points
        _lib.check(fn(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                      ctypes.c_void_p(self._status.data_ptr()), _stream_ptr(stream)))
        return self.tree_status() if check else None

    def tree_status(self):
        """True when the points of the last update_points keep the tree."""
        return int(self._status.item()) == 0

    def _grad_vec(self, t, n, name):
        """A 1-D device tensor of the plan's element type with at least n
        entries at its stride (the gradient entry points read n of them)."""
        torch = _torch()
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda" or t.dtype != self._tdtype():
            raise TypeError(f"{name} must be a {np.dtype(self.dtype).name} CUDA tensor (the eltype of H)")
        if t.dim() != 1:
            raise ValueError(f"{name} must be 1-D")
        if t.shape[0] < n:
            raise ValueError(f"{name} has {t.shape[0]} entries, H needs {n}")
        return t

    def grad_x(self, b, s, stream=None):
        """g_i = s_i * d/dx_i (H(x, y) b)_i (device tensors, 1-D: b of length
        size(H, 2), s of length size(H, 1))."""
        torch = _torch()
        m, n = self.shape
        b = self._grad_vec(b, n, "b")
        s = self._grad_vec(s, m, "s")
        g = torch.empty(m, dtype=self._tdtype(), device=b.device)
        fn = lib.hm_matvec_grad_x_f32 if self.f32 else lib.hm_matvec_grad_x
        check(fn(self._h, ctypes.c_void_p(g.data_ptr()),
                                   ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(s.data_ptr()),
                                   0, 0, 1, b.stride(0), s.stride(0), 0, _stream_ptr(stream)))
        return g

    def matvec_grad(self, b, s=None, stream=None):
        """(H b, s .* d/dx_i (H b)_i) in one pass (device tensors, 1-D: b of
        length size(H, 2), s of length size(H, 1) or None for ones)."""
        torch = _torch()
        m, n = self.shape
        b = self._grad_vec(b, n, "b")
        if s is not None:
            s = self._grad_vec(s, m, "s")
        u = torch.empty(m, dtype=self._tdtype(), device=b.device)
        g = torch.empty_like(u)
        fn = lib.hm_matvec_grad_f32 if self.f32 else lib.hm_matvec_grad
        check(fn(self._h, ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(g.data_ptr()),
                 ctypes.c_void_p(b.data_ptr()),
                 ctypes.c_void_p(s.data_ptr() if s is not None else None), 1, 1, b.stride(0),
                 s.stride(0) if s is not None else 1, 0, _stream_ptr(stream)))
        return u, g

    def tgrad(self, r, s, stream=None):
        """(H^T r, s .* d/dy_j (H^T r)_j) for device tensors r (rows) and s (cols)."""
        torch = _torch()
        m, n = self.shape
        r = self._grad_vec(r, m, "r")
        s = self._grad_vec(s, n, "s")
        gb = torch.empty(n, dtype=self._tdtype(), device=r.device)
        gy = torch.empty_like(gb)
        fn = lib.hm_matvec_t_grad_f32 if self.f32 else lib.hm_matvec_t_grad
        check(fn(self._h, ctypes.c_void_p(gb.data_ptr()),
                                   ctypes.c_void_p(gy.data_ptr()), ctypes.c_void_p(r.data_ptr()),
                                   ctypes.c_void_p(s.data_ptr()), 1, 1, r.stride(0), s.stride(0),
                                   0, _stream_ptr(stream)))
        return gb, gy

    def graph(self, out, v, *, beta=0, matrix_free=False, cheb_basis=False, tuned=False,
              far_precision="f64", mfma=None):
        """Capture `out = beta*out + H v` (device tensors, 1-D or (n, k)) into an
        executable HIP graph; returns a ProductGraph whose launch() replays it."""
        dt = np.float32 if self.f32 else np.float64
        U, V = _Vec(out, True, dt), _Vec(v, False, dt)
        if U.host is not None or V.host is not None:
            raise TypeError("graphs need device tensors")
        flags = (_lib.HM_MV_MATRIX_FREE if matrix_free else 0) | \
            (_lib.HM_MV_CHEB_BASIS if cheb_basis else 0) | (_lib.HM_MV_TUNED if tuned else 0) | \
            _FAR[far_precision] | _mfma_flag(mfma)
        g = ctypes.c_void_p()
        fn = lib.hm_graph_matvec_f32 if self.f32 else lib.hm_graph_matvec
        check(fn(self._h, U.ptr(), V.ptr(), 0, 0, U.inc, V.inc, U.k, U.ld, V.ld, beta, flags,
                 ctypes.byref(g)))
        return ProductGraph(g, (out, v, self))

    def tune(self, k=1, stream=None):
        """Autotune launch shape and far-field form for k right-hand sides;
        returns (rhs per launch, mode, microseconds). Use tuned=True later."""
        torch = _torch()
        m, n = self.shape
        y = torch.empty(m * k, dtype=self._tdtype(), device="cuda")
        x = torch.ones(n * k, dtype=self._tdtype(), device="cuda")
        out = np.zeros(3, dtype=np.int64)
        check(lib.hm_plan_tune(self._h, ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                               k, _stream_ptr(stream),
                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return tuple(int(v) for v in out)

    def __matmul__(self, v):
        return self.matvec(v)

    __mul__ = __matmul__

    @property
    def T(self):
        return _Transposed(self)

    def __getitem__(self, ij):
        """H[i, j] with 0-based indices (getindex, src/hierarchical.jl:120-147)."""
        torch = _torch()
        i, j = ij
        ii = torch.as_tensor(np.atleast_1d(np.asarray(i, dtype=np.int64))).cuda()
        jj = torch.as_tensor(np.atleast_1d(np.asarray(j, dtype=np.int64))).cuda()
        ii, jj = torch.broadcast_tensors(ii, jj)
        ii, jj = ii.contiguous(), jj.contiguous()
        out = torch.empty(ii.shape, dtype=self._tdtype(), device=ii.device)
        fn = lib.hm_entries_f32 if self.f32 else lib.hm_entries
        check(fn(self._h, ii.numel(), ctypes.c_void_p(ii.data_ptr()),
                             ctypes.c_void_p(jj.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                             _stream_ptr(None)))
        res = out.cpu().numpy()
        return float(res.reshape(-1)[0]) if np.ndim(i) == 0 and np.ndim(j) == 0 else res


class ProductGraph:
    """An executable HIP graph of one product (KernelMatrix.graph)."""

    def __init__(self, handle, keep):
        self._g = handle
        self._keep = keep  # the captured buffers and plan must outlive the graph

    def launch(self, stream=None):
        check(lib.hm_graph_launch(self._g, _stream_ptr(stream)))

    def close(self):
        if self._g and self._g.value:
            lib.hm_graph_destroy(self._g)
            self._g = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class KernelMatrixBatch(KernelMatrix):
    """Many independent KernelMatrix instances (same kernel, blocksize, rank)
    in one grouped plan (hm_plan_batched): every pass of a product is one
    launch over all instances. Vectors are the instances' vectors
    concatenated; `row_offsets` / `col_offsets` give the instance boundaries.
    `instances` is a sequence of (x, y, (a, b, c, d))."""

    def __init__(self, f, instances, *, blocksize=None, rank=None, matrix_free=False,
                 hybrid=False, tree_only=False, stream=None):
        if isinstance(f, str):
            f = KERNELS[f]
        self.f = f
        self.rank = BLOCKRANK(np.float64) if rank is None else int(rank)
        self.blocksize_ = 4 * self.rank if blocksize is None else int(blocksize)
        xs = [np.ascontiguousarray(x, dtype=np.float64) for x, _, _ in instances]
        ys = [np.ascontiguousarray(y, dtype=np.float64) for _, y, _ in instances]
        iv = np.ascontiguousarray([list(map(float, v)) for _, _, v in instances], dtype=np.float64)
        cnt = len(xs)
        DP = ctypes.POINTER(ctypes.c_double)
        xp = (DP * cnt)(*[_dptr(a) for a in xs])
        yp = (DP * cnt)(*[_dptr(a) for a in ys])
        nx = np.array([len(a) for a in xs], dtype=np.int64)
        ny = np.array([len(a) for a in ys], dtype=np.int64)
        flags = (_lib.HM_PLAN_MATRIX_FREE if matrix_free else 0) | \
            (_lib.HM_PLAN_HYBRID if hybrid else 0) | (_lib.HM_PLAN_TREE_ONLY if tree_only else 0)
        self.flags = flags
        h = ctypes.c_void_p()
        sp = ctypes.c_void_p(0) if tree_only else _stream_ptr(stream)
        IP = ctypes.POINTER(ctypes.c_int64)
        check(lib.hm_plan_batched(f.kid, cnt, xp, nx.ctypes.data_as(IP), yp, ny.ctypes.data_as(IP),
                                  _dptr(iv), self.blocksize_, self.rank, flags, sp,
                                  ctypes.byref(h)))
        self._h = h
        self.x = np.concatenate(xs)
        self.y = np.concatenate(ys)
        ro = np.zeros(cnt + 1, dtype=np.int64)
        co = np.zeros(cnt + 1, dtype=np.int64)
        check(lib.hm_plan_batch_offsets(h, ro.ctypes.data_as(IP), co.ctypes.data_as(IP)))
        self.row_offsets, self.col_offsets = ro, co


class _Transposed:
    def __init__(self, H):
        self.H = H

    def __matmul__(self, v):
        return self.H.matvec(v, transpose=True)

    @property
    def shape(self):
        return self.H.shape[::-1]


def mul_(u, H, v, istart=1, jstart=1, INCX=1, INCY=1, **kw):
    """mul!(u, H, v[, istart, jstart[, INCX, INCY]]): u += H v in place,
    src/KernelMatrix.jl:14-45 (offsets 1-based, strides as in
    src/HierarchicalMatrix.jl:24-52)."""
    return H.mul(u, v, istart, jstart, INCX, INCY, **kw)


def blocksize(H):
    return H.blocksize()


def size(H, k=None):
    return H.size(k)
