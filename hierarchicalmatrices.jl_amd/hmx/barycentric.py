"""BarycentricPoly2D and BarycentricMatrix2D of HierarchicalMatrices.jl as
standalone objects (outside a KernelMatrix plan), backed by libhmx.

  BarycentricPoly2D(f, a, b, c, d)      src/BarycentricMatrix.jl:139-178
  evaluate(B, x, y), B(x, y)            src/BarycentricMatrix.jl:180-205
  BarycentricMatrix2D(B, x, y, ir, jr)  src/BarycentricMatrix.jl:208-246
  BarycentricMatrix2D(f, a, b, c, d, x, y, ir, jr)
  update_(B, x, y, ir, jr)              update!, src/BarycentricMatrix.jl:248-297
  B[i, j]                               getindex, src/BarycentricMatrix.jl:222-234
  mul_(u, B, v, istart, jstart)         src/algebra.jl:243-277 (u += B v)

f is any host function of two scalars, called on numpy arrays (broadcast over
the node pairs) when it accepts them, else entry by entry. The nodes, the
weights, F = f(t_m, s_n) and the mapped nodes follow the reference's formulas
and operation order (t = (a+b)/2 + (b-a)/2 * x_hat in T). evaluate and the
update! bases run on the device (hm_bary_evaluate, hm_bary_basis); a leaf's
product runs through a one-leaf plan of hm_plan_from_blocks (kind 4, U F V').
Ranges ir, jr are Python ranges of 1-based indices, as Julia's UnitRange.
"""
import ctypes

import numpy as np

from ._lib import check, lib
from .kernelmatrix import (BLOCKRANK, _stream_ptr, _torch, chebyshevbarycentricweights,
                           chebyshevpoints)


def _tdt(dtype):
    torch = _torch()
    return torch.float32 if np.dtype(dtype) == np.float32 else torch.float64


def _dev(a, dtype):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).cuda()


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def eval_pairs(f, X, Y, dtype):
    """f on every pair of the broadcast arrays X, Y: vectorized when f
    accepts arrays and returns their broadcast shape, else entry by entry."""
    X = np.asarray(X, dtype=dtype)
    Y = np.asarray(Y, dtype=dtype)
    shape = np.broadcast_shapes(X.shape, Y.shape)
    try:
        with np.errstate(all="ignore"):
            out = np.asarray(f(X, Y))
        if out.shape == shape:
            return out.astype(dtype, copy=False)
    except Exception:  # a scalar-only function: evaluate it entry by entry
        pass
    Xb, Yb = np.broadcast_to(X, shape), np.broadcast_to(Y, shape)
    out = np.empty(shape, dtype=dtype)
    for idx in np.ndindex(shape):
        out[idx] = f(Xb[idx], Yb[idx])
    return out


def mapped_nodes(a, b, r, dtype=np.float64):
    """ab2 + ba2 * x_hat with ab2 = (a+b)/2, ba2 = (b-a)/2 in T
    (src/BarycentricMatrix.jl:156-166)."""
    T = np.dtype(dtype).type
    a, b = T(a), T(b)
    ab2 = T(0.5) * (a + b)
    ba2 = T(0.5) * (b - a)
    return ab2 + ba2 * chebyshevpoints(r, dtype=dtype)


class BarycentricPoly2D:
    """BarycentricPoly2D(T, f, a, b, c, d): fields x, y (mapped first-kind
    Chebyshev nodes on [a, b] and [c, d]), lamx, lamy (barycentric weights)
    and F (r x r, F[m, n] = f(x[m], y[n])). rank defaults to BLOCKRANK(T)."""

    def __init__(self, f, a, b, c, d, rank=None, dtype=np.float64):
        self.dtype = np.dtype(dtype).type
        if self.dtype not in (np.float32, np.float64):
            raise TypeError("element type must be float32 or float64")
        r = BLOCKRANK(self.dtype) if rank is None else int(rank)
        if r < 1:
            raise ValueError("rank must be >= 1")
        self.x = mapped_nodes(a, b, r, self.dtype)
        self.y = mapped_nodes(c, d, r, self.dtype)
        self.lamx = chebyshevbarycentricweights(r, dtype=self.dtype)
        self.lamy = chebyshevbarycentricweights(r, dtype=self.dtype)
        self.F = np.asfortranarray(eval_pairs(f, self.x[:, None], self.y[None, :], self.dtype))
        self._d = None

    def rank(self):
        return len(self.x)

    def _device(self):
        if self._d is None:
            dt = self.dtype
            self._d = tuple(_dev(a, dt) for a in (self.x, self.y, self.lamx, self.lamy,
                                                  self.F.ravel(order="F")))
        return self._d

    def evaluate(self, x, y, stream=None):
        """evaluate(B, x, y) at every pair of the broadcast x, y (on the
        device; numpy in, numpy out; scalars give a scalar)."""
        torch = _torch()
        X, Y = np.broadcast_arrays(np.asarray(x, dtype=self.dtype), np.asarray(y, dtype=self.dtype))
        shape = X.shape
        px, py = _dev(X.ravel(), self.dtype), _dev(Y.ravel(), self.dtype)
        out = torch.empty(px.shape[0], dtype=_tdt(self.dtype), device=px.device)
        tx, ty, lx, ly, F = self._device()
        fn = lib.hm_bary_evaluate_f32 if self.dtype == np.float32 else lib.hm_bary_evaluate
        check(fn(self.rank(), _vp(tx), _vp(ty), _vp(lx), _vp(ly), _vp(F), px.shape[0], _vp(px),
                 _vp(py), _vp(out), _stream_ptr(stream)))
        res = out.cpu().numpy().reshape(shape)
        return res[()] if res.ndim == 0 else res

    __call__ = evaluate


def evaluate(B, x, y):
    """evaluate(B::BarycentricPoly2D, x, y), src/BarycentricMatrix.jl:180-205."""
    return B.evaluate(x, y)


def _basis(B, pts, nodes, lam, dtype, stream=None):
    """The update! basis of src/BarycentricMatrix.jl:256-287 on the device:
    len(pts) x r, column-major."""
    torch = _torch()
    n, r = len(pts), len(nodes)
    # every device buffer stays referenced until the result is copied back
    # (a temporary freed at once would hand its memory to the next one)
    p, t, lm = _dev(pts, dtype), _dev(nodes, dtype), _dev(lam, dtype)
    out = torch.empty(max(n, 1) * r, dtype=_tdt(dtype), device=p.device)
    fn = lib.hm_bary_basis_f32 if dtype == np.float32 else lib.hm_bary_basis
    check(fn(r, _vp(t), _vp(lm), n, _vp(p), _vp(out), max(n, 1), _stream_ptr(stream)))
    res = np.asfortranarray(out.cpu().numpy()[:n * r].reshape(r, n).T)
    del p, t, lm
    return res


class BarycentricMatrix2D:
    """BarycentricMatrix2D(B, x, y, ir, jr) or BarycentricMatrix2D(f, a, b, c,
    d, x, y, ir, jr): the far leaf U F V' of a KernelMatrix, with U (|ir| x r)
    and V (|jr| x r) the barycentric Lagrange bases of the points x[ir], y[jr]
    (update!) and F the core of B. x and y are the global point vectors, ir
    and jr ranges of 1-based indices into them."""

    def __init__(self, *args, rank=None, dtype=None):
        if len(args) == 5 and isinstance(args[0], BarycentricPoly2D):
            B, x, y, ir, jr = args
        elif len(args) == 9:
            f, a, b, c, d, x, y, ir, jr = args
            dt = dtype if dtype is not None else (np.float32 if np.asarray(x).dtype == np.float32
                                                  else np.float64)
            B = BarycentricPoly2D(f, a, b, c, d, rank=rank, dtype=dt)
        else:
            raise TypeError("BarycentricMatrix2D(B, x, y, ir, jr) or "
                            "BarycentricMatrix2D(f, a, b, c, d, x, y, ir, jr)")
        self.B = B
        self.dtype = B.dtype
        self.x = np.array(x, dtype=self.dtype)
        self.y = np.array(y, dtype=self.dtype)
        self.ir, self.jr = range(ir.start, ir.stop), range(jr.start, jr.stop)
        self.temp1 = np.zeros(B.rank(), dtype=self.dtype)
        self.temp2 = np.zeros(B.rank(), dtype=self.dtype)
        self.U = self.V = None
        self._plan = None
        self.update_(x, y, ir, jr)

    @property
    def F(self):
        return self.B.F

    @property
    def shape(self):
        return (len(self.ir), len(self.jr))

    def size(self, k=None):
        return self.shape if k is None else self.shape[k - 1]

    def rank(self):
        return self.B.rank()

    def update_(self, x, y, ir, jr, stream=None):
        """update!(B, x, y, ir, jr): new points for the same ranges' sizes
        (the reference asserts equal lengths); U and V are re-formed on the
        device and x[ir], y[jr] copied into the leaf's point vectors."""
        if len(ir) != len(self.ir) or len(jr) != len(self.jr):
            raise ValueError("update!: ranges must keep their lengths (src/BarycentricMatrix.jl:249-250)")
        x = np.asarray(x, dtype=self.dtype)
        y = np.asarray(y, dtype=self.dtype)
        xs, ys = x[ir.start - 1:ir.stop - 1], y[jr.start - 1:jr.stop - 1]
        self.U = _basis(self, xs, self.B.x, self.B.lamx, self.dtype, stream)
        self.V = _basis(self, ys, self.B.y, self.B.lamy, self.dtype, stream)
        self.x[ir.start - 1:ir.stop - 1] = xs
        self.y[jr.start - 1:jr.stop - 1] = ys
        self._plan = None
        return self

    def __getitem__(self, ij):
        """getindex(B, i, j) with 0-based i, j (src/BarycentricMatrix.jl:222-234):
        sum_k U[i, k] sum_l F[k, l] V[j, l], in that order."""
        i, j = ij
        T = self.dtype
        ret = T(0)
        for k in range(self.rank()):
            temp = T(0)
            for l in range(self.rank()):
                temp = T(temp + T(self.F[k, l] * self.V[j, l]))
            ret = T(ret + T(self.U[i, k] * temp))
        return ret

    def _leaf_plan(self):
        if self._plan is None:
            from .hierarchical import HierarchicalMatrix
            from .kernelmatrix import Block
            H = HierarchicalMatrix(1, 1)
            H[Block(1), Block(1)] = self
            self._plan = H.plan()
        return self._plan

    def mul_(self, u, v, istart=1, jstart=1, stream=None):
        """mul!(u, B, v, istart, jstart): u[istart + i - 1] += (B v[jstart:...])_i
        (src/algebra.jl:243-277) on the device."""
        return self._leaf_plan().mul(u, v, istart, jstart, stream=stream)

    def __matmul__(self, v):
        return self._leaf_plan().matvec(v)


class BarycentricFactors:
    """A far leaf U F V' given by its factors: the fields U, F (the core
    B.F) and V of a BarycentricMatrix2D built elsewhere, e.g. by the reference
    from an arbitrary f (src/BarycentricMatrix.jl:208-218). Its product is the
    leaf mul! of src/algebra.jl:243-277."""

    def __init__(self, U, F, V, dtype=np.float64):
        self.U = np.asfortranarray(U, dtype=dtype)
        self.F = np.asfortranarray(F, dtype=dtype)
        self.V = np.asfortranarray(V, dtype=dtype)
        r = self.F.shape[0]
        if self.F.shape != (r, r) or self.U.shape[1] != r or self.V.shape[1] != r:
            raise ValueError("U (m x r), F (r x r), V (n x r) expected")

    @property
    def shape(self):
        return (self.U.shape[0], self.V.shape[0])

    def rank(self):
        return self.F.shape[0]

    def __getitem__(self, ij):
        i, j = ij
        return float(self.U[i] @ (self.F @ self.V[j]))


def update_(B, x, y, ir, jr):
    """update!(B, x, y, ir, jr), src/BarycentricMatrix.jl:248-297."""
    return B.update_(x, y, ir, jr)
