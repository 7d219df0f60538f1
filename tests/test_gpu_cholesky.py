"""Hierarchical Cholesky factorization on the GPU (hmx.hierarchicalcholesky,
hmx.solvetransposed; dense leaves by hm_dense_potrf / hm_dense_trsm_t /
hm_dense_gemm, low-rank steps by hm_lowrank_op) against the oracle's numpy
restatement of src/cholesky.jl:12-228 (pinned in test_oracle_cholesky.py).

Bars: the same block kinds and low-rank ranks as the oracle's R; the dense
matrix of R within 1e-12 relative of the oracle's; solves within 1e-10 of the
oracle's; the dense kernels against numpy / scipy to round-off."""
import numpy as np
import pytest
import scipy.linalg as sla
import torch

import cholesky as OC
import hmx
from test_oracle_cholesky import ALL, build_case, case_matrix, to_oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def same_structure(Rh, Ro):
    if isinstance(Rh, hmx.HierarchicalMatrix):
        assert isinstance(Ro, OC.OTree)
        assert np.array_equal(Rh.assigned, Ro.codes)
        M, N = Rh.assigned.shape
        for m in range(M):
            for n in range(N):
                if Rh.assigned[m, n]:
                    same_structure(Rh.blocks[m][n], Ro.blocks[m][n])
    elif isinstance(Rh, hmx.LowRankMatrix):
        assert isinstance(Ro, tuple) and Rh.rank() == len(Ro[1]), (Rh.rank(), len(Ro[1]))
    else:
        assert not isinstance(Ro, (tuple, OC.OTree))


@pytest.mark.parametrize("kinds", ALL)
def test_factor_vs_oracle(cuda, kinds):
    K, H = build_case(kinds)
    Ro = OC.chol(to_oracle(H))
    Rh = hmx.hierarchicalcholesky(H)
    same_structure(Rh, Ro)
    assert rel(OC.densify(to_oracle(Rh)), OC.densify(Ro)) <= 1e-12


def test_deep_tree(cuda):
    K, H = case_matrix(2048, 64)
    Ro = OC.chol(to_oracle(H))
    Rh = hmx.hierarchicalcholesky(H)
    same_structure(Rh, Ro)
    Rd = OC.densify(to_oracle(Rh))
    assert rel(Rd, OC.densify(Ro)) <= 1e-12
    Ad = OC.symmetric_upper(OC.densify(to_oracle(H)))
    assert np.linalg.norm(Rd.T @ Rd - Ad) <= 1e-12 * np.linalg.norm(Ad)


@pytest.mark.parametrize("kinds", [(1, 2, 1), (3, 3, 3), (1, 3, 1), (3, 2, 1)])
def test_solvetransposed_vs_oracle(cuda, kinds):
    K, H = build_case(kinds)
    Rh = hmx.hierarchicalcholesky(H)
    Ro = to_oracle(Rh)
    rng = np.random.default_rng(2)
    b = rng.standard_normal(K.shape[0])
    assert rel(hmx.solvetransposed(Rh, b), OC.solvet(Ro, b)) <= 1e-10
    B = rng.standard_normal((K.shape[0], 5))
    assert rel(hmx.solvetransposed(Rh, B), OC.solvet(Ro, B)) <= 1e-10
    L = hmx.LowRankMatrix(rng.standard_normal((K.shape[0], 3)), np.array([2.0, 1.0, 0.5]),
                          rng.standard_normal((9, 3)))
    Lx = hmx.solvetransposed(Rh, L)
    assert rel(Lx.U, OC.solvet(Ro, L.U)) <= 1e-10
    assert np.array_equal(Lx.S, L.S) and np.array_equal(Lx.V, L.V)


def test_dense_kernels(cuda):
    rng = np.random.default_rng(3)
    for n in (1, 7, 33, 64, 255, 256, 300, 600):
        M = rng.standard_normal((n, n))
        A = M @ M.T + n * np.eye(n)
        A_up = np.triu(A) + np.tril(rng.standard_normal((n, n)), -1)  # lower part ignored
        R = hmx.cholesky.potrf(torch.from_numpy(np.asfortranarray(A_up)).cuda()).cpu().numpy()
        Rc = np.linalg.cholesky(A).T
        assert np.allclose(R, Rc, rtol=0, atol=1e-12 * np.abs(Rc).max())
        assert np.array_equal(np.tril(R, -1), np.zeros((n, n)))
        B = rng.standard_normal((n, 70))
        X = hmx.cholesky.trsm_t(torch.from_numpy(np.asfortranarray(Rc)).cuda(),
                                torch.from_numpy(B).cuda()).cpu().numpy()
        assert rel(X, sla.solve_triangular(Rc, B, trans="T")) <= 1e-12
        b = rng.standard_normal((n, 1))
        x = hmx.cholesky.trsm_t(torch.from_numpy(np.asfortranarray(Rc)).cuda(),
                                torch.from_numpy(b).cuda()).cpu().numpy()
        assert rel(x, sla.solve_triangular(Rc, b, trans="T")) <= 1e-12
    for ta in (0, 1):
        for tb in (0, 1):
            m, n, k = (130, 70, 45) if ta == tb else (12, 9, 300)
            A = rng.standard_normal((k, m) if ta else (m, k))
            B = rng.standard_normal((n, k) if tb else (k, n))
            C = rng.standard_normal((m, n))
            d = rng.random(k) + 0.5
            Ct = hmx.cholesky._copy(torch.from_numpy(C).cuda())
            hmx.cholesky.gemm(ta, tb, -0.5, torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(),
                              2.0, Ct, D=torch.from_numpy(d).cuda())
            want = 2.0 * C - 0.5 * ((A.T if ta else A) * d[None, :]) @ (B.T if tb else B)
            assert rel(Ct.cpu().numpy(), want) <= 1e-14


def test_not_positive_definite(cuda):
    with pytest.raises(hmx.HMError) as e:
        hmx.cholesky.potrf(torch.tensor([[1.0, 2.0], [2.0, 1.0]], dtype=torch.float64).cuda())
    assert e.value.code == 1 and b"positive definite" in str(e.value).encode()


def test_indefinite_leaf_raises_after_the_recursion(cuda):
    """The leaf factorizations report their pivots without waiting; a tree
    whose last diagonal leaf is indefinite raises HMError (the reference's
    PosDefException) once the recursion has ended, and the steps that ran on
    the NaN factors in between end normally."""
    K, H = case_matrix(1024, 128)
    leaf = H
    while isinstance(leaf.blocks[1][1], hmx.HierarchicalMatrix):
        leaf = leaf.blocks[1][1]
    D = np.array(leaf.blocks[1][1])
    D[0, 0] = -1.0
    leaf.blocks[1][1] = D
    with pytest.raises(hmx.HMError) as e:
        hmx.hierarchicalcholesky(H)
    assert b"not positive definite" in str(e.value).encode()
    # the device is usable afterwards
    K2, H2 = case_matrix(512, 128)
    Rh = hmx.hierarchicalcholesky(H2)
    Rd = OC.densify(to_oracle(Rh))
    Ad = OC.symmetric_upper(OC.densify(to_oracle(H2)))
    assert np.linalg.norm(Rd.T @ Rd - Ad) <= 1e-12 * np.linalg.norm(Ad)


def _dev_cm(a):
    return torch.from_numpy(np.asfortranarray(a)).cuda().t().contiguous().t()


def test_batched_dense_entry_points(cuda):
    """Several items per call: hm_dense_potrf_dev (pivot reports in device
    memory, no wait; one item indefinite) and hm_dense_trsm_t with sizes on
    both sides of the LDS path's limit (256 rows)."""
    import ctypes
    from hmx import _lib
    L = _lib.lib
    rng = np.random.default_rng(11)
    ns = [40, 256, 100, 3]
    As, Rc = [], []
    for t, n in enumerate(ns):
        M = rng.standard_normal((n, n))
        A = M @ M.T + n * np.eye(n)
        if t == 3:
            A = -np.eye(n)  # first pivot negative
        else:
            Rc.append(np.linalg.cholesky(A).T)
        As.append(_dev_cm(A))
    p64 = lambda v: np.asarray(v, dtype=np.int64)
    n_ = p64(ns)
    ld = p64([a.stride(1) for a in As])
    ptrs = (ctypes.c_void_p * len(As))(*[a.data_ptr() for a in As])
    info = torch.full((len(ns),), -7, dtype=torch.int64, device="cuda")
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    assert L.hm_dense_potrf_dev(len(ns), ip(n_), ptrs, ip(ld), ctypes.c_void_p(info.data_ptr()),
                                None) == 0
    got = info.cpu().numpy()
    assert list(got[:3]) == [0, 0, 0] and got[3] == 1
    for t in range(3):
        R = As[t].cpu().numpy()
        assert np.allclose(R, Rc[t], rtol=0, atol=1e-12 * np.abs(Rc[t]).max()), t
    for sizes in ([(40, 3), (256, 20), (100, 1)], [(40, 3), (300, 17), (7, 2)]):
        Rs, Bs, want = [], [], []
        for n, k in sizes:
            M = rng.standard_normal((n, n))
            R = np.linalg.cholesky(M @ M.T + n * np.eye(n)).T
            B = rng.standard_normal((n, k))
            Rs.append(_dev_cm(R))
            Bs.append(_dev_cm(B))
            want.append(sla.solve_triangular(R, B, trans="T"))
        nn = p64([n for n, _ in sizes])
        kk = p64([k for _, k in sizes])
        lr = p64([r.stride(1) for r in Rs])
        lb = p64([b.stride(1) if b.shape[1] > 1 else b.shape[0] for b in Bs])
        rp = (ctypes.c_void_p * 3)(*[r.data_ptr() for r in Rs])
        bp = (ctypes.c_void_p * 3)(*[b.data_ptr() for b in Bs])
        assert L.hm_dense_trsm_t(3, ip(nn), ip(kk), rp, ip(lr), bp, ip(lb), None) == 0
        for b, w in zip(Bs, want):
            assert rel(b.cpu().numpy(), w) <= 1e-12
