"""Pins of the oracle's scale! and add_col! restatements (oracle/hierarchical.py,
src/HierarchicalMatrix.jl:54-146, src/algebra.jl:282-380) by the dense
matrices they must produce, which the restatement never forms:

  scale!(H, b, jstart)   dense(H) Diagonal(b[jstart:...])
  scale!(b, H, istart)   Diagonal(b[istart:...]) dense(H)
  add_col!(H, u, i, j)   dense(H) + u[i:...] e_j'

on nested trees with LowRank and dense leaves, unassigned blocks, and offsets;
the oracle's strided traversal mul! against the dense product. A planted
mistake (the child's row offset dropped) fails the add_col! pin. CPU only."""
import copy

import numpy as np
import pytest

import hierarchical as OH
from cholesky import OTree, densify


def tree(rng, depth=0, rows=(30, 26), cols=(28, 35), holes=False):
    T = OTree(2, 2)
    for m in range(2):
        for n in range(2):
            if holes and (m, n) == (1, 0) and depth == 0:
                continue
            if depth < 2 and m == n:
                rs = (rows[m] // 2, rows[m] - rows[m] // 2)
                cs = (cols[n] // 2, cols[n] - cols[n] // 2)
                T.set(m, n, tree(rng, depth + 1, rs, cs))
            elif rng.random() < 0.5:
                r = int(rng.integers(1, 5))
                T.set(m, n, (rng.standard_normal((rows[m], r)), rng.standard_normal(r),
                             rng.standard_normal((cols[n], r))))
            else:
                T.set(m, n, rng.standard_normal((rows[m], cols[n])))
    return T


def test_traversal_mul_is_dense_product():
    rng = np.random.default_rng(0)
    T = tree(rng)
    D = densify(T)
    rows, cols = D.shape
    x = rng.standard_normal(2 * cols + 7)
    y = rng.standard_normal(3 * rows + 9)
    want = y.copy()
    want[4:4 + 3 * rows:3] += D @ x[2:2 + 2 * cols:2]
    got = OH.mul(T, y.copy(), x, 5, 3, 2, 3)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("holes", [False, True])
def test_scale_right_left(seed, holes):
    rng = np.random.default_rng(seed)
    T = tree(rng, holes=holes)
    D = densify(T)
    rows, cols = D.shape
    b = rng.standard_normal(cols + 9)
    for js in (1, 4):
        got = densify(OH.scale_right(T, b, js))
        np.testing.assert_allclose(got, D * b[None, js - 1:js - 1 + cols], rtol=1e-13, atol=1e-13)
    c = rng.standard_normal(rows + 5)
    for is_ in (1, 6):
        got = densify(OH.scale_left(c, T, is_))
        np.testing.assert_allclose(got, c[is_ - 1:is_ - 1 + rows, None] * D, rtol=1e-13, atol=1e-13)
    assert np.array_equal(densify(T), D)  # the input is untouched


@pytest.mark.parametrize("seed", [4, 5])
def test_add_col(seed):
    rng = np.random.default_rng(seed)
    T = tree(rng)
    D = densify(T)
    rows, cols = D.shape
    u = rng.standard_normal(rows + 6)
    for istart in (1, 4):
        for j in (1, 2, cols // 2, cols - 1, cols):
            got = densify(OH.add_col(T, u, istart, j))
            want = D.copy()
            want[:, j - 1] += u[istart - 1:istart - 1 + rows]
            np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)


def test_add_col_grows_lowrank_rank_last():
    rng = np.random.default_rng(6)
    T = OTree(1, 2)
    T.set(0, 0, rng.standard_normal((9, 4)))
    L = (rng.standard_normal((9, 3)), rng.standard_normal(3), rng.standard_normal((5, 3)))
    T.set(0, 1, L)
    u = rng.standard_normal(9)
    T2 = OH.add_col(T, u, 1, 6)  # column 2 of the LowRank leaf
    U1, S1, V1 = T2.blocks[0][1]
    assert U1.shape == (9, 4) and np.array_equal(U1[:, :3], L[0]) and np.array_equal(U1[:, 3], u)
    assert np.array_equal(S1, np.concatenate([L[1], [1.0]]))
    assert np.array_equal(V1[:, 3], np.eye(5)[1])


def test_planted_mistake_fails_the_add_col_pin(monkeypatch):
    """The child's row offset (istart + p) replaced by istart: the dense pin
    no longer holds."""
    rng = np.random.default_rng(7)
    T = tree(rng)
    D = densify(T)
    rows, cols = D.shape
    u = rng.standard_normal(rows)
    orig = OH._add_col_inplace

    def wrong(Tr, uu, istart, j):
        M, N = Tr.codes.shape
        n = 0
        q = Tr.blocksize(0, n, 2)
        while q < j:
            n += 1
            q += Tr.blocksize(0, n, 2)
        q -= Tr.blocksize(0, n, 2)
        for m in range(M):
            code, c = Tr.codes[m, n], Tr.blocks[m][n]
            if code == 1:
                wrong(c, uu, istart, j - q)
            elif code == 2:
                Tr.blocks[m][n] = OH._add_col_lowrank(c, uu, istart, j - q)
            elif code == 3:
                c = np.array(c, copy=True)
                c[:, j - q - 1] += uu[istart - 1:istart - 1 + c.shape[0]]
                Tr.blocks[m][n] = c

    monkeypatch.setattr(OH, "_add_col_inplace", wrong)
    got = densify(OH.add_col(T, u, 1, 3))
    monkeypatch.setattr(OH, "_add_col_inplace", orig)
    want = D.copy()
    want[:, 2] += u
    assert np.max(np.abs(got - want)) > 1e-3
