"""Pins the oracle's EvenBarycentricMatrix restatement (src/BarycentricMatrix.jl:
5-89, mul! src/algebra.jl:168-239) with checks independent of its formulas:
the Lagrange basis evaluated in mpmath at 50 digits, brute-force dense products
with the parity mask written out entry by entry, and interpolation of a smooth
f on the integer grid."""
import numpy as np
import pytest
from mpmath import mp, mpf

import oracle as O

mp.dps = 50


def mp_nodes(a, b, n):
    """t_k = (a+b)/2 + (b-a)/2 * sin(pi (n-2k+1)/(2n)), k = 1..n (descending)."""
    return [mpf(a + b) / 2 + mpf(b - a) / 2 * mp.sin(mp.pi * mpf(n - 2 * k + 1) / (2 * n))
            for k in range(1, n + 1)]


def lagrange(ts, i, k):
    out = mpf(1)
    for l, tl in enumerate(ts):
        if l != k:
            out *= (mpf(i) - tl) / (ts[k] - tl)
    return out


@pytest.mark.parametrize("a,b,n", [(1, 40, 20), (3, 40, 20), (-7, 120, 16), (100, 163, 12),
                                   (0, 1, 8)])
def test_weights_are_the_lagrange_basis(a, b, n):
    """W[k, i] = l_k(i), the Lagrange basis of the nodes t_k at the integer i,
    and w_i = sum_k lambda_k / (2(i - t_k))."""
    w, W = O.even_weights(a, b, n)
    ts = mp_nodes(a, b, n)
    lam = O.chebyshevbarycentricweights(n)
    assert W.shape == (n, b - a + 1)
    worst = 0.0
    for i in range(a, b + 1):
        col = W[:, i - a]
        for k in range(n):
            want = lagrange(ts, i, k)
            worst = max(worst, abs(float(col[k] - want)))
        wi = sum(mpf(float(lam[k])) / (2 * (mpf(i) - ts[k])) for k in range(n))
        assert abs(float(w[i - a]) - float(wi)) <= 1e-13 * abs(float(wi))
    assert worst < 1e-12, worst


def test_single_row_grid_is_not_finite():
    """a == b puts every node on the one grid point: 2i-a-b-(b-a)x_k = 0, so
    inv gives Inf and the weights NaN, as in the reference's Float64 code."""
    w, W = O.even_weights(5, 5, 20)
    assert np.isnan(w[0]) and np.all(np.isnan(W))


def test_nodes_closed_form():
    for a, b, n in [(1, 40, 20), (2, 9, 7), (-3, 1000, 20)]:
        t = O.even_nodes(a, b, n)
        scale = abs(a + b) / 2 + abs(b - a) / 2  # the terms summed
        for k, tk in enumerate(mp_nodes(a, b, n)):
            assert abs(t[k] - float(tk)) <= 2 * np.spacing(scale), (a, b, k)
    # (a+b)/2 is a Float64 division: half-integers survive for odd a+b
    t = O.even_nodes(2, 9, 7)
    assert t[3] == 5.5


def test_barycentricmatrix_weights_are_the_same_values():
    _, W = O.even_weights(4, 77, 20)
    W2 = O.barycentricmatrix_weights(4, 77, 20)
    assert W2.shape == (74, 20)
    assert np.array_equal(W2, W.T)


def test_weights_float32():
    w, W = O.even_weights(1, 30, 10, f32=True)
    assert W.dtype == np.float32
    ts = mp_nodes(1, 30, 10)
    for i in (1, 7, 30):
        for k in range(10):
            assert abs(float(W[k, i - 1]) - float(lagrange(ts, i, k))) < 1e-5


def masked_dense(W, F, ishift, jshift):
    """M[i, j] = sum_k F[j, k] W[k, i] where ishift+i + jshift+j is even."""
    n, m = W.shape
    p = F.shape[0]
    M = np.zeros((m, p))
    for i in range(1, m + 1):
        for j in range(1, p + 1):
            if (ishift + i + jshift + j) % 2 == 0:
                M[i - 1, j - 1] = float(np.dot(F[j - 1, :], W[:, i - 1]))
    return M


def smooth_F(a, b, c, d, n, shift=10.0):
    t = O.even_nodes(a, b, n)
    j = np.arange(c, d + 1, dtype=np.float64)
    return np.asfortranarray(1.0 / (t[None, :] + j[:, None] + shift))


@pytest.mark.parametrize("istart,jstart", [(1, 1), (2, 1), (1, 2), (4, 6), (3, 8)])
@pytest.mark.parametrize("a,b,c,d", [(1, 40, 1, 33), (3, 40, 7, 40), (10, 30, 2, 2)])
def test_mul_is_the_globally_even_masked_product(istart, jstart, a, b, c, d):
    """u[istart-1+i] += sum_j B(i, j) v[jstart-1+j] over (istart-1+i) +
    (jstart-1+j) even: the parity of the global indices, whatever the offsets."""
    n = 20
    _, W = O.even_weights(a, b, n)
    F = smooth_F(a, b, c, d, n)
    m, p = b - a + 1, d - c + 1
    rng = np.random.default_rng(istart * 7 + jstart)
    v = rng.standard_normal(p + jstart + 3)
    u0 = rng.standard_normal(m + istart + 3)
    u = O.even_mul(u0.copy(), W, F, v, istart, jstart)
    M = masked_dense(W, F, istart - 1, jstart - 1)
    want = u0.copy()
    want[istart - 1:istart - 1 + m] += M @ v[jstart - 1:jstart - 1 + p]
    assert np.max(np.abs(u - want)) <= 1e-14 * (1 + np.max(np.abs(want)))
    # rows outside the block are untouched
    assert np.array_equal(u[:istart - 1], u0[:istart - 1])
    assert np.array_equal(u[istart - 1 + m:], u0[istart - 1 + m:])


def test_mul_unit_vectors_give_masked_columns():
    a, b, c, d, n = 1, 12, 1, 9, 8
    _, W = O.even_weights(a, b, n)
    F = smooth_F(a, b, c, d, n)
    for J in range(1, 10):
        v = np.zeros(9)
        v[J - 1] = 1.0
        u = O.even_mul(np.zeros(12), W, F, v)
        for i in range(1, 13):
            if (i + J) % 2:
                assert u[i - 1] == 0.0
            else:
                assert u[i - 1] == pytest.approx(float(np.dot(F[J - 1], W[:, i - 1])), rel=1e-15)


@pytest.mark.parametrize("m,p", [(12, 10), (12, 9)])
def test_getindex_parity_uses_the_block_size(m, p):
    """getindex is nonzero where size(B,1)+size(B,2)+i+j is even. With m+p even
    that is the mask of mul!(u, B, v); with m+p odd it is the complement."""
    n = 8
    _, W = O.even_weights(1, m, n)
    F = smooth_F(1, m, 1, p, n)
    G = np.array([[O.even_getindex(W, F, i, j) for j in range(1, p + 1)] for i in range(1, m + 1)])
    M = masked_dense(W, F, 0, 0)
    full = np.asarray(W).T @ np.asarray(F).T
    if (m + p) % 2 == 0:
        assert np.allclose(G, M, rtol=1e-15, atol=0)
    else:
        assert np.allclose(G + M, full, rtol=1e-14, atol=0)
        assert np.count_nonzero(G * M) == 0


def test_interpolates_a_smooth_kernel_on_the_grid():
    """B(i, j) = sum_k f(t_k, j) l_k(i) = f(i, j) up to the interpolation error
    of a function analytic near [a, b]: f = 1/(t + j + s) has its pole at
    t = -j - s, whose Bernstein ellipse gives an error ~ rho^-20: about 4e-8
    for s = 10 (rho = 2.33) and round-off for s = 100 (rho = 8.35)."""
    a, b, c, d, n = 1, 64, 1, 50, 20
    _, W = O.even_weights(a, b, n)
    i = np.arange(a, b + 1, dtype=np.float64)[:, None]
    j = np.arange(c, d + 1, dtype=np.float64)[None, :]
    for s, lo, hi in ((100.0, 0, 1e-13), (10.0, 1e-9, 2e-7)):
        F = smooth_F(a, b, c, d, n, shift=s)
        full = np.asarray(W).T @ np.asarray(F).T
        exact = 1.0 / (i + j + s)
        err = np.max(np.abs(full - exact) / np.abs(exact))
        assert lo <= err < hi, (s, err)


# ---------------------------------------------------------------- Float32
def test_chebyshevpoints_f32_are_sinpi_of_the_f32_quotient_rounded_once():
    """chebyshevpoints(Float32, n): q = (n-2k+1)/2n formed in Float32, then
    sinpi(q) rounded once to Float32 (mpmath at 50 digits)."""
    for n in (2, 7, 10, 16, 33, 64):
        x = O.chebyshevpoints(n, 1, f32=True)
        lam = O.chebyshevbarycentricweights(n, 1, f32=True)
        for k in range(1, n // 2 + 1):
            q = np.float32(np.float32(n - 2 * k) + np.float32(1)) / np.float32(2 * n)
            assert x[k - 1] == np.float32(float(mp.sin(mp.pi * mpf(float(q)))))
            assert x[n - k] == -x[k - 1]
        for k in range(1, n // 2 + 2):
            if k > n:
                break
            q = (np.float32(2 * k) - np.float32(1)) / np.float32(2 * n)
            want = np.float32(float(mp.sin(mp.pi * mpf(float(q)))))
            assert abs(lam[k - 1]) == want and np.sign(lam[k - 1]) == (-1) ** (k - 1)


def test_nodes_f32_closed_form():
    """(a+b)/2 exact in Float64, (b-a)*x_k/2 within one Float32 rounding of
    the Float32 node times (b-a)/2."""
    for a, b, n in [(1, 40, 10), (2, 9, 7), (-3, 1000, 10), (1, 1 << 22, 10)]:
        t = O.even_nodes(a, b, n, f32=True)
        x = O.chebyshevpoints(n, 1, f32=True)
        for k in range(n):
            h = float(b - a) * float(x[k]) / 2
            assert abs((t[k] - (a + b) / 2) - h) <= abs(h) * 2.0 ** -23 + 1e-300, (a, b, k)
        assert np.all(np.diff(t) < 0)


@pytest.mark.parametrize("istart,jstart", [(1, 1), (2, 1), (4, 6)])
def test_mul_f32_is_the_masked_product(istart, jstart):
    """Float32 mul! against the masked dense product of the same Float32
    factors in Float64: within a few eps(Float32) of the sum of |terms|."""
    a, b, c, d, n = 3, 200, 7, 300, 10
    _, W = O.even_weights(a, b, n, f32=True)
    F = smooth_F(a, b, c, d, n).astype(np.float32)
    m, p = b - a + 1, d - c + 1
    rng = np.random.default_rng(istart + 5 * jstart)
    v = rng.standard_normal(p + jstart + 3).astype(np.float32)
    u0 = rng.standard_normal(m + istart + 3).astype(np.float32)
    u = O.even_mul(u0.copy(), W, F, v, istart, jstart)
    assert u.dtype == np.float32
    M = masked_dense(W.astype(np.float64), F.astype(np.float64), istart - 1, jstart - 1)
    vs = v[jstart - 1:jstart - 1 + p].astype(np.float64)
    want = u0.astype(np.float64)
    want[istart - 1:istart - 1 + m] += M @ vs
    scale = np.abs(M) @ np.abs(vs) + 1.0
    err = np.abs(u[istart - 1:istart - 1 + m] - want[istart - 1:istart - 1 + m])
    assert np.all(err <= 4 * p * 2.0 ** -24 * scale)
    assert np.array_equal(u[:istart - 1], u0[:istart - 1])
