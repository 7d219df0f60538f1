"""LowRankMatrix arithmetic with recompression on the device.

Mirrors src/LowRankMatrix.jl:64-171 over hm_lowrank_op:

  L1 + L2, L1 - L2         (+), (-)           :95-111
  L1 @ L2                  (*)(L1, L2)        :113-118
  adjmul(L1, L2)           L1' * L2           :120-126
  dmul(A, L), A @ L        A * L (dense or Diagonal A)  :128-134,144-150
  muld(L, A), L @ A        L * A                        :136-142,152-158
  svdtrunc(A)              convert(LowRankMatrix, A)    :82-86

`batch(op, items)` runs many items of one operation in one call on torch
device tensors (float64); the operator forms take host LowRankMatrix objects
and return one, computed on the device.
"""
import ctypes

import numpy as np

from ._lib import check, lib
from .kernelmatrix import _stream_ptr, _torch

ADD, SUB, MUL, ADJMUL, DMUL, MULD, SVDTRUNC = range(7)
_NAMES = {"add": ADD, "sub": SUB, "mul": MUL, "adjmul": ADJMUL, "dmul": DMUL, "muld": MULD,
          "svdtrunc": SVDTRUNC}


def _rmax(op, m, n, r1, r2):
    return {ADD: r1 + r2, SUB: r1 + r2, MUL: min(r1, r2), ADJMUL: min(r1, r2), DMUL: r2,
            MULD: r1, SVDTRUNC: min(m, n)}[op]


def _shape(op, it):
    """(m, n, k, r1, r2) of an item dict."""
    g = lambda key: it.get(key)
    r1 = g("U1").shape[1] if g("U1") is not None else 0
    r2 = g("U2").shape[1] if g("U2") is not None else 0
    if op in (ADD, SUB):
        return it["U1"].shape[0], it["V1"].shape[0], 0, r1, r2
    if op == MUL:
        return it["U1"].shape[0], it["V2"].shape[0], it["V1"].shape[0], r1, r2
    if op == ADJMUL:
        return it["V1"].shape[0], it["V2"].shape[0], it["U1"].shape[0], r1, r2
    if op == DMUL:
        return it["A"].shape[0], it["V2"].shape[0], it["A"].shape[1], 0, r2
    if op == MULD:
        return it["U1"].shape[0], it["A"].shape[1], it["A"].shape[0], r1, 0
    return it["A"].shape[0], it["A"].shape[1], 0, 0, 0


def _fortran_dev(a):
    """A float64 CUDA tensor holding a in column-major order (as the ABI
    reads it): the transpose of a contiguous (cols, rows) tensor."""
    torch = _torch()
    if a.dim() == 1:
        return a.contiguous()
    return a.t().contiguous().t()


def batch(op, items, stream=None):
    """items: dicts of float64 CUDA tensors with keys U1, S1, V1, U2, S2, V2, A
    (those the operation uses). Returns a list of (U, S, V) tensors, each
    truncated to its rank."""
    torch = _torch()
    op = _NAMES.get(op, op)
    cnt = len(items)
    dims = np.zeros(5 * max(cnt, 1), dtype=np.int64)
    keys = ("U1", "S1", "V1", "U2", "S2", "V2", "A")
    keep, outs = [], []
    inp = (ctypes.c_void_p * (7 * max(cnt, 1)))()
    outp = (ctypes.c_void_p * (3 * max(cnt, 1)))()
    for t, it in enumerate(items):
        m, n, k, r1, r2 = _shape(op, it)
        dims[5 * t:5 * t + 5] = (m, n, k, r1, r2)
        for q, key in enumerate(keys):
            a = it.get(key)
            if a is not None:
                if a.dtype != torch.float64 or a.device.type != "cuda":
                    raise TypeError("operands must be float64 CUDA tensors")
                a = _fortran_dev(a)
                keep.append(a)
                inp[7 * t + q] = a.data_ptr()
        rm = _rmax(op, m, n, r1, r2)
        dev = next(iter(v for v in it.values() if v is not None)).device
        U = torch.empty(max(rm, 1), m, dtype=torch.float64, device=dev)  # column-major m x rm
        S = torch.empty(max(rm, 1), dtype=torch.float64, device=dev)
        V = torch.empty(max(rm, 1), n, dtype=torch.float64, device=dev)
        outs.append((U, S, V))
        outp[3 * t], outp[3 * t + 1], outp[3 * t + 2] = U.data_ptr(), S.data_ptr(), V.data_ptr()
    rank = np.zeros(max(cnt, 1), dtype=np.int64)
    check(lib.hm_lowrank_op(op, cnt, dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), inp,
                            outp, rank.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                            _stream_ptr(stream)))
    res = []
    for t, (U, S, V) in enumerate(outs):
        r = int(rank[t])
        res.append((U[:r].t(), S[:r], V[:r].t()))
    return res


def _dev(a):
    torch = _torch()
    return torch.from_numpy(np.array(a, dtype=np.float64, order="C", copy=True)).cuda()


def _lr_items(L, idx):
    return {f"U{idx}": _dev(L.U), f"S{idx}": _dev(L.S), f"V{idx}": _dev(L.V)}


def _host(res):
    from .hierarchical import LowRankMatrix
    U, S, V = res
    return LowRankMatrix(U.cpu().numpy(), S.cpu().numpy(), V.cpu().numpy())


def add(L1, L2):
    return _host(batch(ADD, [{**_lr_items(L1, 1), **_lr_items(L2, 2)}])[0])


def sub(L1, L2):
    return _host(batch(SUB, [{**_lr_items(L1, 1), **_lr_items(L2, 2)}])[0])


def mul(L1, L2):
    return _host(batch(MUL, [{**_lr_items(L1, 1), **_lr_items(L2, 2)}])[0])


def adjmul(L1, L2):
    return _host(batch(ADJMUL, [{**_lr_items(L1, 1), **_lr_items(L2, 2)}])[0])


def dmul(A, L):
    A = np.diag(A) if np.ndim(A) == 1 else A  # Diagonal given by its diagonal
    return _host(batch(DMUL, [{"A": _dev(A), **_lr_items(L, 2)}])[0])


def muld(L, A):
    A = np.diag(A) if np.ndim(A) == 1 else A
    return _host(batch(MULD, [{"A": _dev(A), **_lr_items(L, 1)}])[0])


def svdtrunc(A):
    return _host(batch(SVDTRUNC, [{"A": _dev(A)}])[0])
