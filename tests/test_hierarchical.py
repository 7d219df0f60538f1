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
