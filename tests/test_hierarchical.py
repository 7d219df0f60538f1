"""The generic HierarchicalMatrix container (src/HierarchicalMatrix.jl,
src/hierarchical.jl, src/LowRankMatrix.jl): host structure on the CPU; the
strided product on the GPU against a restatement of the traversal
(src/HierarchicalMatrix.jl:24-52) that calls the oracle's leaf kernels
(src/algebra.jl:37-48,110-131)."""
import numpy as np
import pytest

import hmx
import oracle as O
from hmx import Block, HierarchicalMatrix, LowRankMatrix


def dense_of(H):
    rows, cols = H.size()
    D = np.zeros((rows, cols))
    for c, i0, j0 in H.leaves():
        if isinstance(c, LowRankMatrix):
            blk = c.U @ np.diag(c.S) @ c.V.T
        else:
            blk = c
        D[i0:i0 + blk.shape[0], j0:j0 + blk.shape[1]] += blk
    return D


def oracle_traversal(H, y, x, istart, jstart, incx, incy):
    """src/HierarchicalMatrix.jl:24-52 restated: y[...] += H x[...] by walking
    the children with the oracle's leaf loops."""
    M, N = H.blocksize()
    p = 0
    for m in range(1, M + 1):
        q = 0
        for n in range(1, N + 1):
            c = H[Block(m), Block(n)]
            if isinstance(c, HierarchicalMatrix):
                oracle_traversal(c, y, x, istart + p, jstart + q, incx, incy)
            elif isinstance(c, LowRankMatrix):
                O.lowrank_mul(y, c.U, c.S, c.V, x, istart + p, jstart + q, incx, incy)
            elif c is not None:
                O.dense_mul(y, c, x, istart + p, jstart + q, incx, incy)
            q += incx * H.blocksize(1, n, 2)
        p += incy * H.blocksize(m, N, 1)
    return y


def consistent_tree(rng):
    """2x2 of 2x2 with matching block sizes (rows r1, r2; cols c1, c2)."""
    def leaf(mm, nn):
        if rng.random() < 0.5:
            r = int(rng.integers(1, 5))
            return LowRankMatrix(rng.standard_normal((mm, r)), rng.standard_normal(r),
                                 rng.standard_normal((nn, r)))
        return rng.standard_normal((mm, nn))

    def node(rows, cols, depth):
        H = HierarchicalMatrix(2, 2)
        for m in range(2):
            for n in range(2):
                if depth < 2 and m == n and rows[m] >= 8 and cols[n] >= 8:
                    rs = [rows[m] // 2, rows[m] - rows[m] // 2]
                    cs = [cols[n] // 2, cols[n] - cols[n] // 2]
                    H[Block(m + 1), Block(n + 1)] = node(rs, cs, depth + 1)
                else:
                    H[Block(m + 1), Block(n + 1)] = leaf(rows[m], cols[n])
        return H

    return node([int(rng.integers(20, 60)), int(rng.integers(20, 60))],
                [int(rng.integers(20, 60)), int(rng.integers(20, 60))], 0)


def test_structure_and_getindex():
    rng = np.random.default_rng(0)
    H = consistent_tree(rng)
    D = dense_of(H)
    rows, cols = H.size()
    assert D.shape == (rows, cols)
    for i, j in zip(rng.integers(0, rows, 200), rng.integers(0, cols, 200)):
        assert abs(H.getindex(int(i), int(j)) - D[i, j]) <= 1e-12 * (1 + abs(D[i, j]))
    # an unknown block type stays unassigned (src/hierarchical.jl:149-172)
    H2 = HierarchicalMatrix(2, 2)
    H2[Block(1), Block(1)] = "not a matrix"
    assert H2.assigned[0, 0] == 0 and H2.blocksize() == (2, 2)


def test_oracle_traversal_equals_dense():
    rng = np.random.default_rng(1)
    H = consistent_tree(rng)
    D = dense_of(H)
    rows, cols = H.size()
    x = rng.standard_normal(3 * cols + 9)
    y = np.zeros(3 * rows + 9)
    oracle_traversal(H, y, x, 4, 2, 2, 3)
    np.testing.assert_allclose(y[3:3 + 3 * rows:3], D @ x[1:1 + 2 * cols:2], rtol=1e-12, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_gpu_strided_product_vs_oracle(cuda, seed):
    import torch
    rng = np.random.default_rng(seed)
    H = consistent_tree(rng)
    rows, cols = H.size()
    istart, jstart, incx, incy = 4, 2, 2, 3
    x = rng.standard_normal(incx * cols + 9)
    y0 = rng.standard_normal(incy * rows + 9)
    want = oracle_traversal(H, y0.copy(), x, istart, jstart, incx, incy)
    yd = torch.from_numpy(y0.copy()).to(cuda)
    hmx.mul_(yd, H, torch.from_numpy(x).to(cuda), istart, jstart, incx, incy)
    got = yd.cpu().numpy()
    assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.norm(want)
    # transpose against the dense assembly
    D = dense_of(H)
    r = rng.standard_normal(rows)
    g = H.plan().matvec(torch.from_numpy(r).to(cuda), transpose=True).cpu().numpy()
    assert np.linalg.norm(g - D.T @ r) <= 1e-13 * np.linalg.norm(D.T @ r)
    # several right-hand sides
    B = rng.standard_normal((cols, 5))
    Y = H.plan().matvec(torch.from_numpy(B).to(cuda)).cpu().numpy()
    assert np.linalg.norm(Y - D @ B) <= 1e-13 * np.linalg.norm(D @ B)
    # on the matrix cores: the leaves' odd sizes start chunks at odd rows of
    # the dense blocks around them; 32 right-hand sides (copies of row pairs
    # and of value pairs of x) and 17 (single values)
    for k in (32, 17):
        B = rng.standard_normal((cols, k))
        Y = H.plan().matvec(torch.from_numpy(B).to(cuda), mfma=True).cpu().numpy()
        assert np.linalg.norm(Y - D @ B) <= 1e-13 * np.linalg.norm(D @ B), k


# ------------------------------------- scale! and add_col! (src/HierarchicalMatrix.jl)

def test_scale_and_add_col_host_identities():
    """rmul!/lmul! and add_col! on the host leaves against the dense
    assembly: H D_b, D_b H and H + u e_j^T (tolerance: rounding of U S V')."""
    rng = np.random.default_rng(11)
    H = consistent_tree(rng)
    rows, cols = H.size()
    D0 = dense_of(H)
    b = rng.standard_normal(cols)
    hmx.rmul_(H, b)
    assert np.allclose(dense_of(H), D0 * b[None, :], rtol=1e-13, atol=1e-13)
    c = rng.standard_normal(rows)
    D1 = dense_of(H)
    hmx.lmul_(c, H)
    assert np.allclose(dense_of(H), c[:, None] * D1, rtol=1e-13, atol=1e-13)
    for j in (1, cols // 3, cols):
        u = rng.standard_normal(rows)
        D2 = dense_of(H)
        ranks = sorted(l.rank() for l, _, _ in H.leaves() if isinstance(l, LowRankMatrix))
        hmx.add_col_(H, u, j)
        want = D2.copy()
        want[:, j - 1] += u
        assert np.allclose(dense_of(H), want, rtol=1e-13, atol=1e-13)
        # the LowRank leaves in column j gained one rank, the new term first
        after = sorted(l.rank() for l, _, _ in H.leaves() if isinstance(l, LowRankMatrix))
        assert sum(after) >= sum(ranks)


def ranked_tree(rng, free):
    """2x2 of 2x2; LowRank leaves of ranks 2-5, and (free=True) one leaf of
    rank 8 so that every other leaf has free rank columns in the plan."""
    def leaf(mm, nn, r):
        return LowRankMatrix(rng.standard_normal((mm, r)), rng.standard_normal(r),
                             rng.standard_normal((nn, r)))
    rows, cols = [37, 29, 33, 41], [31, 40, 26, 38]
    top = HierarchicalMatrix(2, 2)
    for m in range(2):
        for n in range(2):
            H = HierarchicalMatrix(2, 2)
            for a in range(2):
                for b in range(2):
                    mm, nn = rows[2 * m + a], cols[2 * n + b]
                    if (m, n, a, b) == (0, 1, 0, 0):
                        H[Block(a + 1), Block(b + 1)] = leaf(mm, nn, 8 if free else 5)
                    elif (a + b) % 2:
                        H[Block(a + 1), Block(b + 1)] = leaf(mm, nn, int(rng.integers(2, 6)))
                    else:
                        H[Block(a + 1), Block(b + 1)] = rng.standard_normal((mm, nn))
            top[Block(m + 1), Block(n + 1)] = H
    return top


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("free", [True, False])
def test_gpu_scale_and_add_col_vs_oracle(cuda, seed, free):
    """rmul!/lmul! (hm_plan_scale) and add_col! (hm_plan_add_col) applied in
    place to the device plan, against the oracle's restatement of scale! and
    add_col! (oracle/hierarchical.py, pinned by dense assembly) and its
    traversal mul!: products within 1e-13. With free rank columns add_col!
    updates the plan in place; without (free=False: every LowRank leaf of the
    column at the plan's rank), the plan is rebuilt and the product agrees
    all the same."""
    import torch
    import hierarchical as OH
    from hmx._lib import lib
    rng = np.random.default_rng(seed)
    H = ranked_tree(rng, free)
    T = OH.from_tree(H)
    rows, cols = H.size()
    x = rng.standard_normal(cols)
    xd = torch.from_numpy(x).to(cuda)

    def check_product(T):
        want = OH.mul(T, np.zeros(rows), x, 1, 1)
        got = H.matvec(xd).cpu().numpy()
        assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.norm(want)

    check_product(T)                      # builds the plan
    plan = H._plan
    b = rng.standard_normal(cols)
    hmx.rmul_(H, torch.from_numpy(b).to(cuda))
    T = OH.scale_right(T, b, 1)
    assert H._plan is plan                # scaled in place
    check_product(T)
    c = rng.standard_normal(rows)
    hmx.lmul_(c, H)
    T = OH.scale_left(c, T, 1)
    check_product(T)
    for j in (3, cols // 2 + 1, cols - 2):
        u = rng.standard_normal(rows + 3)
        hmx.add_col_(H, u, j, istart=2)
        T = OH.add_col(T, u, 2, j)
        if free:
            assert H._plan is plan        # updated in place on the device
        check_product(T)
    # a NULL vector is refused
    assert lib.hm_plan_add_col(H.plan()._h, None, 0, 0, None) != 0


@pytest.mark.gpu
def test_gpu_scale_rejects_kernel_matrix(cuda):
    import ctypes
    import torch
    from hmx._lib import HM_ERR_UNSUPPORTED, lib
    w = hmx.workloads.config_plain(n=1024, blocksize=64, rank=16)
    K = hmx.KernelMatrix("cauchykernel", w["x"], w["y"], *w["intervals"], blocksize=64, rank=16)
    b = torch.ones(1024, dtype=torch.float64, device=cuda)
    assert lib.hm_plan_scale(K._h, 0, ctypes.c_void_p(b.data_ptr()), 0, None) == HM_ERR_UNSUPPORTED
