"""Pins the CPU oracle (oracle/hmx_oracle.c) before it is trusted as the parity
reference. The reference (Julia) cannot run here, and its own tests hold no
golden vectors for the KernelMatrix path (only the leaf offset/stride
known-answer tests of test/runtests.jl:15-52), so each oracle function is pinned
by what the mathematics fixes: closed forms checked in 50-digit arithmetic,
invariants, brute force on small inputs, the dense product, and those
known-answer tests restated.
"""
import numpy as np
import pytest
from mpmath import mp, mpf

import hmx
import oracle as O

mp.dps = 50


def ulp_err(a, b):
    """|a - b| in units of the spacing of doubles at b."""
    return abs(float(a) - float(b)) / np.spacing(abs(float(b)) if b != 0 else 1e-300)


# ----------------------------------------------------------- A5, A6: nodes

@pytest.mark.parametrize("n", [2, 3, 4, 10, 16, 17, 20, 24, 32, 33, 64, 100])
def test_chebyshev_points_closed_form(n):
    """First kind: cos((2k-1)pi/(2n)); second kind: cos((k-1)pi/(n-1)),
    k = 1..n (src/BarycentricMatrix.jl:92-111), to <= 1 ulp."""
    for kind in (1, 2):
        if kind == 2 and n < 2:
            continue
        x = O.chebyshevpoints(n, kind)
        for k in range(1, n + 1):
            if kind == 1:
                want = mp.cos((2 * k - 1) * mp.pi / (2 * n))
            else:
                want = mp.cos((k - 1) * mp.pi / (n - 1))
            if abs(want) < mpf(10) ** -30:
                assert abs(x[k - 1]) < 1e-15
            else:
                assert ulp_err(x[k - 1], want) <= 1.0, (n, kind, k)


@pytest.mark.parametrize("n", [3, 4, 10, 16, 17, 20, 24, 32])
def test_barycentric_weights_are_proportional_to_lagrange_weights(n):
    """lambda_k must be proportional to 1/prod_{j != k}(x_k - x_j) for the
    nodes they belong to (the definition of barycentric weights)."""
    for kind in (1, 2):
        x = [mp.cos((2 * k - 1) * mp.pi / (2 * n)) if kind == 1 else mp.cos((k - 1) * mp.pi / (n - 1))
             for k in range(1, n + 1)]
        lam = O.chebyshevbarycentricweights(n, kind)
        true = []
        for k in range(n):
            p = mpf(1)
            for j in range(n):
                if j != k:
                    p *= x[k] - x[j]
            true.append(1 / p)
        ratio = [mpf(lam[k]) / true[k] for k in range(n)]
        for k in range(n):
            assert abs(ratio[k] / ratio[0] - 1) < 1e-14, (n, kind, k)


def test_sinpi_correctly_rounded_on_grid():
    for q in np.linspace(0, 0.5, 257):
        assert ulp_err(O.lib().or_sinpi(q), mp.sin(mp.pi * mpf(q))) <= 0.5 + 1e-9


# ------------------------------------------------------------ A1: BLOCKRANK

def test_blockrank_table():
    """Values of src/HierarchicalMatrices.jl:5 for Float16/32/64 and 256-bit
    BigFloat, all even (the assertion of test/runtests.jl:5-7)."""
    assert O.blockrank(10) == 4
    assert O.blockrank(23) == 10
    assert O.blockrank(52) == 20
    assert O.blockrank(255) == 100
    for bits in range(5, 120):
        assert O.blockrank(bits) % 2 == 0


# -------------------------------------------------------------- A2: indsplit

def test_indsplit_brute_force():
    rng = np.random.default_rng(0)
    for _ in range(300):
        n = int(rng.integers(1, 40))
        x = np.sort(rng.uniform(-1, 1, n))[::-1]
        f = int(rng.integers(1, n + 1))
        ln = int(rng.integers(1, n - f + 2))
        a, b = 1.0, -1.0 + rng.uniform(0, 1)
        (f1, l1), (f2, l2) = O.indsplit(x, f, ln, a, b)
        mid = 0.5 * (a + b)
        cnt = 0
        while cnt < ln and x[f - 1 + cnt] >= mid:
            cnt += 1
        assert (f1, l1, f2, l2) == (f, cnt, f + cnt, ln - cnt)


def test_indsplit_empty_range_quirk():
    """An empty range whose first point lies above the midpoint comes back as a
    one-point range in the reference (src/BarycentricMatrix.jl:302-306)."""
    x = np.array([0.9, 0.8, 0.1])
    assert O.indsplit(x, 2, 0, 1.0, 0.0) == ((2, 1), (3, 0))
    assert O.indsplit(x, 3, 0, 1.0, 0.0) == ((3, 0), (3, 0))
    with pytest.raises(IndexError):
        O.indsplit(x, 4, 0, 1.0, 0.0)


# ------------------------------------------------------------ A3: the tree

def _grid(n):
    return hmx.workloads.shifted_grid(n)


@pytest.mark.parametrize("n,bs,r,counts", [
    (4096, 64, 16, (127, 120, 120, 382, 720)),
    (1 << 20, 256, 24, (8191, 8178, 8178, 24574, 49068)),
])
def test_block_counts(n, bs, r, counts):
    x, y, iv = _grid(n)
    st = O.Tree(O.CAUCHY, x, y, *iv, bs, r, store=False).stats()
    assert (st["d"], st["k1"], st["k2"], st["dense"], st["far"]) == counts


@pytest.mark.parametrize("case", ["grid", "random", "chebyshev"])
def test_tiling_and_admissibility(case):
    if case == "grid":
        x, y, iv = _grid(700)
        bs, r = 16, 8
    elif case == "random":
        x, y, iv = hmx.workloads.sorted_uniform_points(900, 3)
        bs, r = 16, 8
    else:
        x = O.chebyshevpoints(800)
        y = O.chebyshevpoints(800, 2)
        iv = (1.0, -1.0, 1.0, -1.0)
        bs, r = 20, 8
    t = O.Tree(O.CAUCHY, x, y, *iv, bs, r, store=False)
    ints, dbls = t.blocks()
    cover = np.zeros((len(x), len(y)), dtype=np.int32)
    for (code, depth, r0, c0, m, n), (a, b, c, d) in zip(ints, dbls):
        cover[r0:r0 + m, c0:c0 + n] += 1
        if code == O.CODE_BARY and m and n:
            xs, ys = x[r0:r0 + m], y[c0:c0 + n]
            lo_x, hi_x, lo_y, hi_y = min(a, b), max(a, b), min(c, d), max(c, d)
            assert np.all(xs <= hi_x) and np.all(xs >= lo_x)
            assert np.all(ys <= hi_y) and np.all(ys >= lo_y)
            gap = max(lo_x - hi_y, lo_y - hi_x)
            # x and y trees split their own intervals; they differ by `shift`
            shift = abs(iv[0] - iv[2]) + abs(iv[1] - iv[3])
            assert gap >= (hi_x - lo_x) * (1 - 1e-12) - 1e-15 - shift
            assert abs((hi_x - lo_x) - (hi_y - lo_y)) <= 1e-12 * (hi_x - lo_x) or case == "random"
    assert np.all(cover == 1)


def test_traversal_offsets_match_ranges():
    """The offsets of the traversal (src/KernelMatrix.jl:38,40, sizes from
    src/hierarchical.jl:33-47) equal the blocks' own index ranges."""
    x, y, iv = _grid(4096)
    t = O.Tree(O.CAUCHY, x, y, *iv, 64, 16, store=False)
    ints, _ = t.blocks()
    assert ints[0, 2] == 0 and ints[0, 3] == 0
    # rows and columns of each block row are contiguous and start where the
    # previous block row ended
    starts = sorted(set(ints[:, 2]))
    assert starts[0] == 0


# --------------------------------------------- A4, A7, A9: generated factors

def test_far_factors_invariants_and_closed_form():
    x, y, iv = _grid(1024)
    r = 16
    t = O.Tree(O.CAUCHY, x, y, *iv, 64, r)
    ints, dbls = t.blocks()
    xhat = O.chebyshevpoints(r)
    checked = 0
    for k in range(len(ints)):
        code, depth, r0, c0, m, n = ints[k]
        if code != O.CODE_BARY:
            continue
        a, b, c, d = dbls[k]
        f = t.block_data(k)
        U, V, F = f["U"], f["V"], f["F"]
        # partition of unity
        assert np.max(np.abs(U.sum(1) - 1)) < 1e-14
        assert np.max(np.abs(V.sum(1) - 1)) < 1e-14
        # polynomial reproduction: U p(t) = p(x_i) for deg < r
        tx = (0.5 * (a + b)) + (0.5 * (b - a)) * xhat
        ty = (0.5 * (c + d)) + (0.5 * (d - c)) * xhat
        s = (x[r0:r0 + m] - 0.5 * (a + b)) / (0.5 * (a - b))
        st = (tx - 0.5 * (a + b)) / (0.5 * (a - b))
        for q in (0, 1, 5, r - 1):
            assert np.max(np.abs(U @ st ** q - s ** q)) < 1e-12
        # F is the kernel sampled at the mapped node pairs, bit for bit
        assert np.array_equal(F, 1.0 / (tx[:, None] - ty[None, :]))
        checked += 1
    assert checked > 50


def test_dense_blocks_closed_form():
    x, y, iv = _grid(512)
    for kid, fn in [(O.CAUCHY, lambda d: 1.0 / d), (O.COULOMB, lambda d: 1.0 / (d * d)),
                    (O.COULOMB3, lambda d: 1.0 / ((d * d) * d))]:
        t = O.Tree(kid, x, y, *iv, 64, 16)
        ints, _ = t.blocks()
        for k in range(len(ints)):
            code, _, r0, c0, m, n = ints[k]
            if code == O.CODE_DENSE and m and n:
                A = t.block_data(k)["A"]
                want = fn(x[r0:r0 + m, None] - y[None, c0:c0 + n])
                assert np.array_equal(A, want)


# ------------------------------------------ A13-A17: leaf kernels (known answers)

def _kat_data(dtype_exact):
    rng = np.random.default_rng(0)
    if dtype_exact:
        A = rng.integers(-8, 9, (10, 5)).astype(np.float64)
        x = rng.integers(-8, 9, 40).astype(np.float64)
    else:
        A = hmx.workloads.uniform(100, 50).reshape(10, 5, order="F")
        x = hmx.workloads.uniform(101, 40)
    return A, x


@pytest.mark.parametrize("exact", [False, True])
def test_leaf_known_answers(exact):
    """The four cases of test/runtests.jl:15-33 (tolerance eps*|ref|) and their
    exact-arithmetic twins (:35-52, BigFloat there; small integers here)."""
    A, x = _kat_data(exact)
    eps = np.finfo(np.float64).eps

    def check(got, want):
        if exact:
            assert np.array_equal(got, want)
        else:
            assert np.linalg.norm(got - want) <= eps * np.linalg.norm(want)

    y = np.zeros(40)
    O.dense_mul(y, A, x, 1, 1)
    check(y[0:10], A @ x[0:5])
    y = np.zeros(40)
    O.dense_mul(y, A, x, 5, 5, 2, 2)
    check(y[4:23:2], A @ x[4:14:2])
    y = np.zeros(40)
    O.dense_tmul(y, A, x, 1, 5, 2, 1)
    check(y[0:5], A.T @ x[4:23:2])
    y = np.zeros(40)
    O.dense_tmul(y, A, x, 6, 3, 1, 3)
    check(y[5:18:3], A.T @ x[2:12])


def test_lowrank_and_barycentric_leaves():
    """LowRankMatrix mul! and its adjoint (src/algebra.jl:110-159) and the
    BarycentricMatrix2D mul! (src/algebra.jl:243-277) against GEMV chains."""
    rng = np.random.default_rng(5)
    m, n, r = 13, 9, 4
    U, V, S = rng.standard_normal((m, r)), rng.standard_normal((n, r)), rng.standard_normal(r)
    x = rng.standard_normal(60)
    y = rng.standard_normal(60)
    y0 = y.copy()
    O.lowrank_mul(y, U, S, V, x, 3, 7, 2, 3)
    want = y0.copy()
    want[2:2 + 3 * m:3] += U @ (S * (V.T @ x[6:6 + 2 * n:2]))
    assert np.allclose(y, want, rtol=1e-14, atol=1e-14)
    y = y0.copy()
    O.lowrank_mul(y, U, S, V, x, 2, 1, 1, 1, adjoint=True)
    want = y0.copy()
    want[1:1 + n] += V @ (S * (U.T @ x[0:m]))
    assert np.allclose(y, want, rtol=1e-14, atol=1e-14)
    F = rng.standard_normal((r, r))
    u = y0.copy()
    O.bary_mul(u, U, F, V, x, 4, 11)
    want = y0.copy()
    want[3:3 + m] += U @ (F @ (V.T @ x[10:10 + n]))
    assert np.allclose(u, want, rtol=1e-14, atol=1e-14)


# ------------------------------------------------ A12: the traversal mul!

@pytest.mark.parametrize("kid", [O.CAUCHY, O.COULOMB, O.COULOMB3, O.LOG])
def test_traversal_matvec_vs_dense(kid):
    """Whole path against the dense long-double product: the error is the
    Chebyshev interpolation error, ~1e-14 at rank 16 for these kernels."""
    w = hmx.workloads.config_plain()
    t = O.Tree(kid, w["x"], w["y"], *w["intervals"], 64, 16)
    u = t.matvec(w["b"])
    d = O.dense_matvec(kid, w["x"], w["y"], w["b"])
    assert np.linalg.norm(u - d) / np.linalg.norm(d) < 2e-13


def test_variants_bit_identical():
    w = hmx.workloads.config_plain()
    t = O.Tree(O.CAUCHY, w["x"], w["y"], *w["intervals"], 64, 16)
    u = t.matvec(w["b"])
    assert np.array_equal(t.matvec_omp(w["b"]), u)
    rows = np.array([0, 1, 31, 32, 2047, 2048, 4095])
    assert np.array_equal(t.matvec_rows(w["b"], rows), u[rows])
    ts = O.Tree(O.CAUCHY, w["x"], w["y"], *w["intervals"], 64, 16, store=False)
    assert np.array_equal(ts.matvec_rows(w["b"], rows), u[rows])
    # accumulate semantics: u0 + H b
    u0 = np.arange(4096.0)
    assert np.array_equal(t.matvec(w["b"], u=u0), u0 + u) or \
        np.linalg.norm(t.matvec(w["b"], u=u0) - (u0 + u)) < 1e-12 * np.linalg.norm(u0)


def test_getindex_matches_blocks():
    w = hmx.workloads.config_plain(n=1024)
    t = O.Tree(O.CAUCHY, w["x"], w["y"], *w["intervals"], 64, 16)
    K = 1.0 / (w["x"][:, None] - w["y"][None, :])
    ints, _ = t.blocks()
    kind = np.zeros((1024, 1024), dtype=np.int64)
    for code, _, r0, c0, m, n in ints:
        kind[r0:r0 + m, c0:c0 + n] = code
    rng = np.random.default_rng(3)
    for i, j in zip(rng.integers(0, 1024, 400), rng.integers(0, 1024, 400)):
        g = t.getindex(int(i), int(j))
        if kind[i, j] == O.CODE_DENSE:
            assert g == K[i, j]
        else:  # entrywise Chebyshev interpolation error at rank 16
            assert abs(g - K[i, j]) <= 1e-10 * abs(K[i, j])


def test_planted_mistakes_are_detected():
    """A transposed F, an unnormalised basis or unmapped nodes in a far block
    all move the product far outside the 1e-12 bar the dense check uses."""
    w = hmx.workloads.config_plain(n=1024)
    x, y, b = w["x"], w["y"], w["b"]
    t = O.Tree(O.CAUCHY, x, y, *w["intervals"], 64, 16)
    ints, dbls = t.blocks()
    dense = O.dense_matvec(O.CAUCHY, x, y, b)

    def product(mutate):
        u = np.zeros(len(x))
        for k in range(len(ints)):
            code, _, r0, c0, m, n = ints[k]
            f = t.block_data(k)
            if code == O.CODE_DENSE:
                u[r0:r0 + m] += f["A"] @ b[c0:c0 + n]
            else:
                U, F, V = mutate(f["U"], f["F"], f["V"], dbls[k])
                u[r0:r0 + m] += U @ (F @ (V.T @ b[c0:c0 + n]))
        return np.linalg.norm(u - dense) / np.linalg.norm(dense)

    assert product(lambda U, F, V, iv: (U, F, V)) < 1e-13
    assert product(lambda U, F, V, iv: (U, F.T, V)) > 1e-6
    assert product(lambda U, F, V, iv: (U * 1.001, F, V)) > 1e-6
    xhat = O.chebyshevpoints(16)

    def unmapped(U, F, V, iv):
        return U, 1.0 / (xhat[:, None] - (xhat[None, :] + 2.5)), V
    assert product(unmapped) > 1e-6
