"""Pins the Float32 twin of the oracle (oracle/liboracle_f32.so, the same
source built with or_real = float). The reference's KernelMatrix path is
generic in T (src/KernelMatrix.jl:47-116, src/BarycentricMatrix.jl:92-136,
147-178,248-307, src/algebra.jl:34-48,243-277), so a Float32 tree splits at
Float32 midpoints, forms its nodes, weights, F, U, V and dense blocks in Float32
arithmetic and accumulates mul! in Float32. Each check below uses arithmetic
independent of the oracle: numpy's IEEE float32 operations, mpmath, brute
force, and the dense long-double product.
"""
import numpy as np
import pytest
from mpmath import mp, mpf

import hmx
import oracle as O

mp.dps = 50
f32 = np.float32


def ulp32(v):
    return float(np.spacing(np.abs(f32(v)))) if v != 0 else float(np.finfo(np.float32).tiny)


def _grid32(n):
    x, y, iv = hmx.workloads.shifted_grid(n)
    return x.astype(f32), y.astype(f32), iv


@pytest.mark.parametrize("n", [2, 4, 10, 16, 17, 24, 32])
def test_f32_chebyshev_points_and_weights(n):
    """x_k = sinpi(q_k) and lambda_k = sinpi(p_k) with the arguments
    q_k = (n-2k+1)/(2n), p_k = (2k-1)/(2n) formed in Float32
    (src/BarycentricMatrix.jl:97,121), correctly rounded to Float32."""
    x = O.chebyshevpoints(n, f32=True)
    lam = O.chebyshevbarycentricweights(n, f32=True)
    assert x.dtype == f32 and lam.dtype == f32
    for k in range(1, n // 2 + 1):
        q = (f32(n - 2 * k) + f32(1)) / f32(2 * n)
        want = mp.sin(mp.pi * mpf(float(q)))
        assert abs(float(x[k - 1]) - float(want)) <= 0.5 * ulp32(float(want)) * (1 + 1e-6), (n, k)
        assert x[n - k] == -x[k - 1]
    for k in range(1, n // 2 + 2):
        if k > n:
            break
        p = (f32(2 * k) - f32(1)) / f32(2 * n)
        want = mp.sin(mp.pi * mpf(float(p))) * (-1 if k % 2 == 0 else 1)
        assert abs(float(lam[k - 1]) - float(want)) <= 0.5 * ulp32(float(want)) * (1 + 1e-6), (n, k)


def test_f32_indsplit_midpoint_is_float32():
    """half(T)*(a+b) in Float32: with a = 1, b = 2^-24 the float sum rounds to
    1 (ties to even), so the Float32 midpoint is 0.5 while the Float64 one is
    0.5 + 2^-25; the point 0.5 goes to the first half only in Float32."""
    x = np.array([0.75, 0.5, 0.25], dtype=f32)
    b = 2.0 ** -24
    (f1, l1), (f2, l2) = O.indsplit(x, 1, 3, 1.0, b, f32=True)
    assert (f1, l1, f2, l2) == (1, 2, 3, 1)
    (f1, l1), (f2, l2) = O.indsplit(x.astype(np.float64), 1, 3, 1.0, b)
    assert (f1, l1, f2, l2) == (1, 1, 2, 2)


def test_f32_indsplit_brute_force():
    rng = np.random.default_rng(11)
    for _ in range(200):
        n = int(rng.integers(1, 40))
        x = np.sort(rng.random(n).astype(f32))[::-1].copy()
        a, b = f32(rng.random() * 0.3 + 0.7), f32(rng.random() * 0.3)
        mid = f32(0.5) * (a + b)
        first = int(rng.integers(1, n + 1))
        length = int(rng.integers(1, n - first + 2))
        seg = x[first - 1:first - 1 + length]
        want = int(np.sum(seg >= mid))
        (g1, m1), (g2, m2) = O.indsplit(x, first, length, float(a), float(b), f32=True)
        assert (g1, m1, g2, m2) == (first, want, first + want, length - want)


def test_f32_tree_on_exact_grid_equals_f64_tree():
    """The shifted grid of n = 4096 is exact in Float32 and every midpoint of
    (1, 0) is dyadic, so the Float32 tree is the Float64 tree."""
    x, y, iv = _grid32(4096)
    t32 = O.Tree(O.CAUCHY, x, y, *iv, 64, 16, store=False, f32=True)
    t64 = O.Tree(O.CAUCHY, x.astype(np.float64), y.astype(np.float64), *iv, 64, 16, store=False)
    i32, d32 = t32.blocks()
    i64, d64 = t64.blocks()
    assert np.array_equal(i32, i64)
    assert np.array_equal(d32.astype(np.float64), d64)
    assert t32.stats() == t64.stats()


def test_f32_tiling_and_admissibility_random_points():
    x, y, iv = hmx.workloads.sorted_uniform_points(900, 3)
    x, y = x.astype(f32), y.astype(f32)
    t = O.Tree(O.CAUCHY, x, y, *iv, 16, 8, store=False, f32=True)
    ints, dbls = t.blocks()
    cover = np.zeros((len(x), len(y)), dtype=np.int32)
    for (code, depth, r0, c0, m, n), (a, b, c, d) in zip(ints, dbls):
        cover[r0:r0 + m, c0:c0 + n] += 1
        if code == O.CODE_BARY and m and n:
            xs, ys = x[r0:r0 + m], y[c0:c0 + n]
            assert np.all(xs <= max(a, b)) and np.all(xs >= min(a, b))
            assert np.all(ys <= max(c, d)) and np.all(ys >= min(c, d))
        # every interval end is a Float32 value
        assert all(f32(v) == v for v in (a, b, c, d))
    assert np.all(cover == 1)


def _mapped32(a, b, xhat):
    a, b = f32(a), f32(b)
    return f32(0.5) * (a + b) + (f32(0.5) * (b - a)) * xhat


def test_f32_factors_closed_form_and_invariants():
    x, y, iv = _grid32(1024)
    r = 16
    t = O.Tree(O.CAUCHY, x, y, *iv, 64, r, f32=True)
    ints, dbls = t.blocks()
    xhat = O.chebyshevpoints(r, f32=True)
    lam = O.chebyshevbarycentricweights(r, f32=True)
    far = dense = 0
    for k in range(len(ints)):
        code, depth, r0, c0, m, n = ints[k]
        if not (m and n):
            continue
        f = t.block_data(k)
        if code == O.CODE_DENSE:
            want = f32(1) / (x[r0:r0 + m, None] - y[None, c0:c0 + n])
            assert f["A"].dtype == f32 and np.array_equal(f["A"], want)
            dense += 1
            continue
        a, b, c, d = dbls[k]
        U, V, F = f["U"], f["V"], f["F"]
        tx, ty = _mapped32(a, b, xhat), _mapped32(c, d, xhat)
        assert np.array_equal(F, f32(1) / (tx[:, None] - ty[None, :]))
        # update! in Float32: lam_m * inv(x_i - t_m), divided by the sum
        # accumulated m = 1..r
        row = lam[None, :] * (f32(1) / (x[r0:r0 + m, None] - tx[None, :]))
        s = np.zeros(m, dtype=f32)
        for q in range(r):
            s = s + row[:, q]
        assert np.array_equal(U, row / s[:, None])
        assert np.max(np.abs(U.astype(np.float64).sum(1) - 1)) < 4 * r * 6e-8
        # polynomial reproduction to Float32 accuracy
        sx = (x[r0:r0 + m].astype(np.float64) - 0.5 * (a + b)) / (0.5 * (a - b))
        st = (tx.astype(np.float64) - 0.5 * (a + b)) / (0.5 * (a - b))
        for q in (0, 1, 5):
            assert np.max(np.abs(U.astype(np.float64) @ st ** q - sx ** q)) < 1e-4
        far += 1
    assert far > 50 and dense > 20


@pytest.mark.parametrize("kid", [O.CAUCHY, O.COULOMB, O.LOG])
@pytest.mark.parametrize("bs,r", [(64, 16), (40, 10)])
def test_f32_traversal_vs_dense(kid, bs, r):
    """Float32 mul! against the dense long-double product of the same Float32
    points: interpolation error at rank 10 (~1e-7) plus Float32 rounding."""
    x, y, iv = _grid32(4096)
    b = hmx.workloads.normal(0, 4096).astype(f32)
    t = O.Tree(kid, x, y, *iv, bs, r, f32=True)
    u = t.matvec(b)
    assert u.dtype == f32
    d = O.dense_matvec(kid, x.astype(np.float64), y.astype(np.float64), b.astype(np.float64))
    err = np.linalg.norm(u - d) / np.linalg.norm(d)
    assert err < 2e-5, err


def test_f32_traversal_rounds_in_float32():
    """A Float32 tree carried out in Float64 would agree with the Float64
    oracle to ~1e-13; the Float32 one differs at the Float32 level."""
    x, y, iv = _grid32(2048)
    b = hmx.workloads.normal(0, 2048).astype(f32)
    u32 = O.Tree(O.CAUCHY, x, y, *iv, 64, 16, f32=True).matvec(b)
    u64 = O.Tree(O.CAUCHY, x.astype(np.float64), y.astype(np.float64), *iv, 64, 16).matvec(
        b.astype(np.float64))
    diff = np.linalg.norm(u32 - u64) / np.linalg.norm(u64)
    assert 1e-9 < diff < 2e-5, diff
    assert np.array_equal(O.Tree(O.CAUCHY, x, y, *iv, 64, 16, f32=True).matvec_omp(b), u32)


@pytest.mark.parametrize("exact", [False, True])
def test_f32_leaf_known_answers(exact):
    """test/runtests.jl:15-33 at T = Float32: the four offset/stride cases of
    the dense leaf mul! within eps(Float32)*|ref|, and exact on small
    integers (the BigFloat twins of :35-52)."""
    rng = np.random.default_rng(0)
    if exact:
        A = rng.integers(-8, 9, (10, 5)).astype(f32)
        x = rng.integers(-8, 9, 40).astype(f32)
    else:
        A = hmx.workloads.uniform(100, 50).reshape(10, 5, order="F").astype(f32)
        x = hmx.workloads.uniform(101, 40).astype(f32)
    Ad, xd = A.astype(np.float64), x.astype(np.float64)
    eps = float(np.finfo(np.float32).eps)

    def check(got, want):
        if exact:
            assert np.array_equal(got.astype(np.float64), want)
        else:
            assert np.linalg.norm(got.astype(np.float64) - want) <= eps * np.linalg.norm(want)

    y = np.zeros(40, dtype=f32)
    O.dense_mul(y, A, x, 1, 1)
    check(y[0:10], Ad @ xd[0:5])
    y = np.zeros(40, dtype=f32)
    O.dense_mul(y, A, x, 5, 5, 2, 2)
    check(y[4:23:2], Ad @ xd[4:14:2])
    y = np.zeros(40, dtype=f32)
    O.dense_tmul(y, A, x, 1, 5, 2, 1)
    check(y[0:5], Ad.T @ xd[4:23:2])
    y = np.zeros(40, dtype=f32)
    O.dense_tmul(y, A, x, 6, 3, 1, 3)
    check(y[5:18:3], Ad.T @ xd[2:12])
