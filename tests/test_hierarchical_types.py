"""@hierarchical with other leaf types (src/hierarchical.jl:4-172), as the
reference's test tree @hierarchical UpperTriangularHierarchicalMatrix Matrix
(UpperTriangular, LowRankMatrix) Diagonal (test/runtests.jl:13,54-56): the
assigned codes follow the leaf-type order, a leaf of another type (or another
parameter, UpperTriangular of a Matrix) stays unassigned, size, blocksize and
getindex follow the reference's block walk; the strided product on the device
(the non-Matrix leaves formed densely) against the oracle's leaf loops over
the same dense leaves."""
import numpy as np
import pytest

import hmx
import oracle as O
from hmx import Block, Diagonal, LowRankMatrix, UpperTriangular
from hmx.hierarchical import _dense

UTHM = hmx.hierarchical("UpperTriangularHierarchicalMatrix", np.ndarray,
                        (UpperTriangular, LowRankMatrix), Diagonal)


def lowrank(rng, m, n, r):
    return LowRankMatrix(rng.standard_normal((m, r)), rng.standard_normal(r),
                         rng.standard_normal((n, r)))


def build(rng):
    H = UTHM(2, 2)
    inner = UTHM(2, 2)
    inner[Block(1), Block(1)] = Diagonal(rng.standard_normal(12))
    inner[Block(1), Block(2)] = UpperTriangular(lowrank(rng, 12, 12, 3))
    inner[Block(2), Block(2)] = rng.standard_normal((12, 12))
    H[Block(1), Block(1)] = inner
    H[Block(1), Block(2)] = UpperTriangular(lowrank(rng, 24, 24, 4))
    H[Block(2), Block(1)] = rng.standard_normal((24, 24))
    H[Block(2), Block(2)] = Diagonal(rng.standard_normal(24))
    return H


def dense(H):
    rows, cols = H.size()
    D = np.zeros((rows, cols))
    for c, i0, j0 in H.leaves():
        blk = _dense(c)
        D[i0:i0 + blk.shape[0], j0:j0 + blk.shape[1]] = blk
    return D


def test_codes_fields_and_structure():
    rng = np.random.default_rng(0)
    H = build(rng)
    assert H.field_types["UpperTriangularblocks"] == (UpperTriangular, LowRankMatrix)
    assert list(H.field_types) == ["UpperTriangularHierarchicalMatrixblocks", "Matrixblocks",
                                   "UpperTriangularblocks", "Diagonalblocks"]
    assert H.assigned.tolist() == [[1, 3], [2, 4]]
    assert H.UpperTriangularblocks[0, 1] is H.blocks[0][1] and H.Matrixblocks[0, 1] is None
    assert H.size() == (48, 48) and H.blocksize() == (2, 2)
    # UpperTriangular of a Matrix is not a field type: unassigned
    G = UTHM(2, 2)
    G[Block(1), Block(1)] = UpperTriangular(np.ones((3, 3)))
    G[Block(2), Block(2)] = hmx.HierarchicalMatrix(2, 2)  # another tree type
    assert G.assigned.tolist() == [[0, 0], [0, 0]]
    D = dense(H)
    for i, j in zip(rng.integers(0, 48, 100), rng.integers(0, 48, 100)):
        assert abs(H.getindex(int(i), int(j)) - D[i, j]) <= 1e-12 * (1 + abs(D[i, j]))
    # the upper triangle of the UpperTriangular leaves, the diagonal of the
    # Diagonal ones
    assert np.all(np.tril(D[:24, 24:], -1) == 0) and np.all(D[24:, 24:] == np.diag(np.diag(D[24:, 24:])))


@pytest.mark.gpu
def test_device_product_vs_oracle_leaf_loops(cuda):
    import torch
    rng = np.random.default_rng(1)
    H = build(rng)
    rows, cols = H.size()
    istart, jstart, incx, incy = 3, 2, 2, 3
    x = rng.standard_normal(incx * cols + 5)
    y0 = rng.standard_normal(incy * rows + 5)
    want = y0.copy()
    for c, i0, j0 in H.leaves():  # the leaf loops of src/algebra.jl:37-48
        O.dense_mul(want, np.asfortranarray(_dense(c)), x,
                    istart + incy * i0, jstart + incx * j0, incx, incy)
    yd = torch.from_numpy(y0.copy()).to(cuda)
    hmx.mul_(yd, H, torch.from_numpy(x).to(cuda), istart, jstart, incx, incy)
    got = yd.cpu().numpy()
    assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.norm(want)
