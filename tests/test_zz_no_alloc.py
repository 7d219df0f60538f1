"""hm_dense_gemm_noalloc, the batched GEMM of the Cholesky steps that
allocates nothing (its item table lives in a fixed device array), against
hm_dense_gemm on the same inputs: two calls in a row, the second with fewer
and smaller items, each bit for bit equal to hm_dense_gemm's result (the same
kernel on the same items). Also: one item (passed by value), and a count
beyond the array's capacity is refused with HM_ERR_UNSUPPORTED."""
import ctypes

import numpy as np
import pytest
import torch

from hmx._lib import HM_DENSE_NOALLOC_MAX, HM_ERR_UNSUPPORTED, check, lib

pytestmark = pytest.mark.gpu

CAP = HM_DENSE_NOALLOC_MAX


def _arr(vals, ct):
    a = (ct * len(vals))(*vals)
    return a


def ld(t):
    """The leading dimension of a column-major matrix (torch reports the
    stride of a one-column matrix as 1)."""
    return t.stride(1) if t.shape[1] > 1 else max(1, t.shape[0])


def call(fn, ta, tb, items, alpha, beta):
    """items: list of (A, D or None, B, C) column-major device tensors."""
    cnt = len(items)
    mnk, ptrs = [], {"A": [], "D": [], "B": [], "C": []}
    lda, ldb, ldc = [], [], []
    for A, D, B, C in items:
        m, n = C.shape
        k = A.shape[0] if ta else A.shape[1]
        mnk += [m, n, k]
        for key, t in (("A", A), ("D", D), ("B", B), ("C", C)):
            ptrs[key].append(ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0))
        lda.append(ld(A))
        ldb.append(ld(B))
        ldc.append(ld(C))
    I64 = ctypes.c_int64
    VP = ctypes.c_void_p
    return fn(ta, tb, cnt, _arr(mnk, I64), alpha, _arr(ptrs["A"], VP), _arr(lda, I64),
              _arr(ptrs["D"], VP), _arr(ptrs["B"], VP), _arr(ldb, I64), beta,
              _arr(ptrs["C"], VP), _arr(ldc, I64), ctypes.c_void_p(0))


def colmajor(rng, r, c, dev):
    return torch.from_numpy(np.asfortranarray(rng.standard_normal((r, c)))).to(dev).T.contiguous().T


def make_items(rng, count, smax, dev, ta, tb):
    items = []
    for _ in range(count):
        m, n, k = (int(v) for v in rng.integers(1, smax, 3))
        A = colmajor(rng, k, m, dev) if ta else colmajor(rng, m, k, dev)
        B = colmajor(rng, n, k, dev) if tb else colmajor(rng, k, n, dev)
        D = torch.from_numpy(rng.standard_normal(k)).to(dev) if rng.random() < 0.5 else None
        C = colmajor(rng, m, n, dev)
        items.append((A, D, B, C))
    return items


def run_pair(items, ta, tb, alpha, beta):
    """(result of hm_dense_gemm, result of hm_dense_gemm_noalloc) on copies."""
    ref = [(A, D, B, C.clone()) for A, D, B, C in items]
    out = [(A, D, B, C.clone()) for A, D, B, C in items]
    check(call(lib.hm_dense_gemm, ta, tb, ref, alpha, beta))
    check(call(lib.hm_dense_gemm_noalloc, ta, tb, out, alpha, beta))
    torch.cuda.synchronize()
    return [c[3] for c in ref], [c[3] for c in out]


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1)])
def test_noalloc_gemm_matches_allocating_gemm(cuda, ta, tb):
    rng = np.random.default_rng(10 * ta + tb)
    # first call: 40 items up to 150 x 150; second: 7 items up to 40 x 40
    for count, smax in ((40, 150), (7, 40), (1, 90)):
        print(f"case: noalloc gemm ta={ta} tb={tb} count={count}", flush=True)
        items = make_items(rng, count, smax, cuda, ta, tb)
        want, got = run_pair(items, ta, tb, -1.0, 1.0)
        for t, (w, g) in enumerate(zip(want, got)):
            assert torch.equal(w, g), (count, t)
        # and the values are the product itself
        A, D, B, C = items[0]
        Am = A.T if ta else A
        Bm = B.T if tb else B
        Dm = torch.diag(D) if D is not None else torch.eye(Am.shape[1], dtype=A.dtype, device=cuda)
        exp = C - Am @ Dm @ Bm
        assert float(torch.linalg.norm(got[0] - exp) / torch.linalg.norm(exp)) <= 1e-13


def test_noalloc_gemm_rejects_too_many_items(cuda):
    rng = np.random.default_rng(3)
    it = make_items(rng, 1, 4, cuda, 0, 0)
    items = it * (CAP + 1)
    rc = call(lib.hm_dense_gemm_noalloc, 0, 0, items, 1.0, 0.0)
    assert rc == HM_ERR_UNSUPPORTED
    assert b"at most 8192" in lib.hm_last_error()
