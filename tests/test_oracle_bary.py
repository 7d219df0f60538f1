"""Pins of the oracle's evaluate(B::BarycentricPoly2D, x, y)
(src/BarycentricMatrix.jl:180-205) by facts its code does not restate:

- the 2-D barycentric interpolant on r x r first-kind Chebyshev nodes
  reproduces every polynomial of degree < r in each variable, so with
  F = p(t_m, s_n) it returns p(x, y) (a transposed F, a dropped weight or an
  unmapped node breaks this for the non-symmetric p used);
- at a point pair of a far block it equals that block's entry
  sum_k U[i, k] sum_l F[k, l] V[j, l] (getindex, src/BarycentricMatrix.jl:
  222-234), which factors the same interpolant through the update! bases.
CPU only."""
import numpy as np

import oracle as O


def nodes(a, b, r):
    xh = O.chebyshevpoints(r)
    return 0.5 * (a + b) + 0.5 * (b - a) * xh


def p(x, y):
    return 3.0 * x ** 3 * y ** 2 - x * y + 2.0 + 0.5 * x ** 5


def test_reproduces_polynomials():
    r = 8
    tx, ty = nodes(2.0, 1.0, r), nodes(0.5, -0.5, r)
    lam = O.chebyshevbarycentricweights(r)
    F = np.asfortranarray(p(tx[:, None], ty[None, :]))
    rng = np.random.default_rng(0)
    x = 1.0 + rng.random(50)
    y = -0.5 + rng.random(50)
    got = O.bary_evaluate(tx, ty, lam, lam, F, x, y)
    assert np.max(np.abs(got - p(x, y)) / np.abs(p(x, y))) <= 1e-13
    # a planted mistake fails it: F transposed
    bad = O.bary_evaluate(tx, ty, lam, lam, np.asfortranarray(F.T), x, y)
    assert np.max(np.abs(bad - p(x, y)) / np.abs(p(x, y))) > 1e-3
    # and so does a degree-r term (outside the interpolation space)
    Fr = np.asfortranarray(F + tx[:, None] ** r)
    off = O.bary_evaluate(tx, ty, lam, lam, Fr, x, y) - (p(x, y) + x ** r)
    assert np.max(np.abs(off)) > 1e-6


def test_equals_far_block_entry():
    n, bs, r = 1024, 64, 16
    i = np.arange(n, 0, -1, dtype=np.float64)
    x, y = (i - 0.25) / n, (i - 0.75) / n
    t = O.Tree(O.CAUCHY, x, y, 1.0, 0.0, 1.0, 0.0, bs, r)
    ints, dbls = t.blocks()
    lam = O.chebyshevbarycentricweights(r)
    rng = np.random.default_rng(1)
    checked = 0
    for k in range(len(ints)):
        code, _, i0, j0, m, nn = ints[k]
        if code != 2:
            continue
        a, b, c, d = dbls[k]
        tx, ty = nodes(a, b, r), nodes(c, d, r)
        F = t.block_data(k)["F"]
        ii = i0 + rng.integers(0, m, 4)
        jj = j0 + rng.integers(0, nn, 4)
        got = O.bary_evaluate(tx, ty, lam, lam, F, x[ii], y[jj])
        want = np.array([t.getindex(int(a_), int(b_)) for a_, b_ in zip(ii, jj)])
        assert np.max(np.abs(got - want) / np.abs(want)) <= 1e-13
        checked += 1
    assert checked > 100
