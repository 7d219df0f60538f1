"""CPU checks of libhmx: every symbol of include/hmx.h is exported, the host-side
functions agree with the oracle, the block tree (built with HM_PLAN_TREE_ONLY,
no device needed) equals the oracle's leaf for leaf, and bad input is turned
into status codes."""
import ctypes
import os
import re

import numpy as np
import pytest

import hmx
import oracle as O
from hmx import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "hmx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hm_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 14
    L = ctypes.CDLL(_lib.LIB_PATH)
    for s in syms:
        assert hasattr(L, s), s


def test_host_functions_match_oracle():
    for bits, want in [(10, 4), (23, 10), (52, 20), (255, 100)]:
        assert _lib.lib.hm_blockrank(bits) == want == O.blockrank(bits)
    assert hmx.BLOCKRANK(np.float64) == 20 and hmx.BLOCKSIZE(np.float64) == 80
    assert hmx.BLOCKRANK(np.float32) == 10 and hmx.BLOCKRANK(np.float16) == 4
    for n in (2, 3, 10, 16, 17, 20, 24, 32, 64):
        for kind in (1, 2):
            assert np.array_equal(hmx.chebyshevpoints(n, kind), O.chebyshevpoints(n, kind))
            assert np.array_equal(hmx.chebyshevbarycentricweights(n, kind),
                                  O.chebyshevbarycentricweights(n, kind))


def test_indsplit_matches_oracle_exactly():
    rng = np.random.default_rng(2)
    for _ in range(300):
        n = int(rng.integers(1, 50))
        x = np.sort(rng.uniform(-1, 1, n))[::-1].copy()
        f = int(rng.integers(1, n + 1))
        ln = int(rng.integers(1, n - f + 2))
        a, b = 1.0, float(rng.uniform(-1, 1))
        r1, r2 = hmx.indsplit(x, range(f, f + ln), a, b)
        (f1, l1), (f2, l2) = O.indsplit(x, f, ln, a, b)
        assert (r1.start, len(r1), r2.start, len(r2)) == (f1, l1, f2, l2)


def tree(kernel, x, y, iv, bs, r):
    return hmx.KernelMatrix(kernel, x, y, *iv, blocksize=bs, rank=r, tree_only=True)


def compare_structure(H, x, y, iv, bs, r, kid=O.CAUCHY):
    t = O.Tree(kid, x, y, *iv, bs, r, store=False)
    ib, db = H.blocks()
    ob, odb = t.blocks()
    assert np.array_equal(ib, ob)
    assert np.array_equal(db, odb)
    st = t.stats()
    info = H.info
    assert (info["n_d"], info["n_k1"], info["n_k2"], info["n_dense"], info["n_far"]) == \
        (st["d"], st["k1"], st["k2"], st["dense"], st["far"])
    assert info["dense_entries"] == st["dense_entries"]
    assert info["depth"] == st["depth"]
    return info, st


def test_structure_plain_and_headline():
    w = hmx.workloads.config_plain()
    info, _ = compare_structure(tree("cauchykernel", w["x"], w["y"], w["intervals"], 64, 16),
                                w["x"], w["y"], w["intervals"], 64, 16)
    assert info["bytes_ref_layout"] == 20447232
    x, y, iv = hmx.workloads.shifted_grid(1 << 20)
    info, st = compare_structure(tree("cauchykernel", x, y, iv, 256, 24), x, y, iv, 256, 24)
    assert info["bytes_ref_layout"] == 16751695872
    # one U per row cluster and one V per column cluster: 12 basis rows per
    # point instead of the reference's 33 per-block copies
    assert info["basis_rows_u"] == info["basis_rows_v"] == 12 * (1 << 20)
    assert st["far_rows"] == st["far_cols"] == 34603776
    assert info["bytes_stored"] == 8295684096


@pytest.mark.parametrize("seed,n", [(0, 1034), (3, 20000), (7, 65000)])
def test_structure_random_points(seed, n):
    x, y, iv = hmx.workloads.sorted_uniform_points(n, seed)
    compare_structure(tree("cauchykernel", x, y, iv, 64, 16), x, y, iv, 64, 16)


def test_structure_log_scale_family():
    x, y, iv = hmx.workloads.sorted_uniform_points(1 << 18, 1)
    compare_structure(tree("logkernel", x, y, iv, 512, 32), x, y, iv, 512, 32, O.LOG)


@pytest.mark.parametrize("N", [1000, 10000])
def test_structure_example_families(N):
    x = hmx.chebyshevpoints(N)
    y = hmx.chebyshevpoints(N, kind=2)
    compare_structure(tree("cauchykernel", x, y, (1.0, -1.0, 1.0, -1.0), 80, 20), x, y,
                      (1.0, -1.0, 1.0, -1.0), 80, 20)
    i = np.arange(N, 0, -1, dtype=np.float64)
    x, y = i * (i + 1), (i + 0.5) * (i + 1.5)
    iv = (x.max(), x.min(), y.max(), y.min())
    compare_structure(tree("cauchykernel", x, y, iv, 80, 20), x, y, iv, 80, 20)


def status_of(fn):
    with pytest.raises(hmx.HMError) as e:
        fn()
    return e.value.code


def test_error_statuses():
    x, y, iv = hmx.workloads.shifted_grid(256)
    # ascending points make the reference recurse without bound
    assert status_of(lambda: tree("cauchykernel", x[::-1].copy(), y, iv, 16, 8)) == \
        _lib.HM_ERR_NOT_DESCENDING
    xn = x.copy()
    xn[5] = np.nan
    assert status_of(lambda: tree("cauchykernel", xn, y, iv, 16, 8)) == _lib.HM_ERR_NOT_DESCENDING
    assert status_of(lambda: tree("cauchykernel", x, y, iv, 16, 0)) == _lib.HM_ERR_UNSUPPORTED
    assert status_of(lambda: tree("cauchykernel", x, y, iv, 16, 65)) == _lib.HM_ERR_UNSUPPORTED
    assert status_of(lambda: tree("cauchykernel", x, y, iv, 0, 8)) == _lib.HM_ERR_INVALID
    assert status_of(lambda: tree("cauchykernel", x[:0], y, iv, 16, 8)) == _lib.HM_ERR_DEGENERATE
    h = ctypes.c_void_p()
    xd = x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    yd = y.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert _lib.lib.hm_plan_from_points(9, xd, 256, yd, 256, 1.0, 0.0, 1.0, 0.0, 16, 8,
                                        _lib.HM_PLAN_TREE_ONLY, None, ctypes.byref(h)) == \
        _lib.HM_ERR_INVALID
    assert b"kernel" in _lib.lib.hm_last_error()
    assert _lib.lib.hm_plan_from_points(0, xd, 256, yd, 256, 1.0, 0.0, 1.0, 0.0, 16, 8,
                                        64, None, ctypes.byref(h)) == _lib.HM_ERR_INVALID
    H = tree("cauchykernel", x, y, iv, 16, 8)
    assert _lib.lib.hm_matvec(H._h, None, None, 0, 0, 1, 1, 1, 0, 0, 0, 0, None) == \
        _lib.HM_ERR_INVALID
    # the one-pass product and x-derivative sums: NULL outputs, then a tree-only
    # plan (rejected on the host before any pointer is used)
    assert _lib.lib.hm_matvec_grad(H._h, None, None, None, None, 1, 1, 1, 1, 0, None) == \
        _lib.HM_ERR_INVALID
    buf = np.zeros(256)
    p = ctypes.c_void_p(buf.ctypes.data)
    assert _lib.lib.hm_matvec_grad(H._h, p, p, p, None, 1, 1, 1, 1, 0, None) == \
        _lib.HM_ERR_INVALID
    assert b"tree-only" in _lib.lib.hm_last_error()


def test_empty_cluster_is_reported():
    """x entirely in the upper half while y fills both halves: the root's lower
    x range is empty (first index n+1) but its D child still recurses because y
    needs splitting, and the reference's indsplit then reads x[n+1]
    (BoundsError). The oracle raises; libhmx returns HM_ERR_DEGENERATE."""
    n = 300
    x = np.linspace(0.99, 0.6, n)
    y = np.linspace(0.99, 0.01, n)
    with pytest.raises(ValueError):
        O.Tree(O.CAUCHY, x, y, 1.0, 0.0, 1.0, 0.0, 16, 8, store=False)
    assert status_of(lambda: tree("cauchykernel", x, y, (1.0, 0.0, 1.0, 0.0), 16, 8)) == \
        _lib.HM_ERR_DEGENERATE
    # points on both sides of every split build the same tree in both
    x2 = np.linspace(0.99, 0.01, n)
    compare_structure(tree("cauchykernel", x2, y, (1.0, 0.0, 1.0, 0.0), 16, 8), x2, y,
                      (1.0, 0.0, 1.0, 0.0), 16, 8)


# ------------------------------------------------------------ Float32 plans

def compare_structure_f32(x, y, iv, bs, r, kid=O.CAUCHY, kernel="cauchykernel"):
    x, y = x.astype(np.float32), y.astype(np.float32)
    H = hmx.KernelMatrix(kernel, x, y, *iv, blocksize=bs, rank=r, tree_only=True)
    assert H.dtype == np.float32 and H.info["float32"] == 1
    t = O.Tree(kid, x, y, *iv, bs, r, store=False, f32=True)
    ib, db = H.blocks()
    ob, odb = t.blocks()
    assert np.array_equal(ib, ob)
    assert np.array_equal(db, odb.astype(np.float64))
    st = t.stats()
    info = H.info
    assert (info["n_d"], info["n_k1"], info["n_k2"], info["n_dense"], info["n_far"]) == \
        (st["d"], st["k1"], st["k2"], st["dense"], st["far"])
    return info


def test_structure_f32_families():
    """Float32 trees (midpoints rounded to Float32) equal the Float32
    oracle's leaf for leaf: grid, random points, the example families."""
    w = hmx.workloads.config_plain()
    info = compare_structure_f32(w["x"], w["y"], w["intervals"], 64, 16)
    assert info["bytes_ref_layout"] == 20447232 // 2  # 4-byte elements
    for seed, n in [(0, 1034), (3, 20000)]:
        x, y, iv = hmx.workloads.sorted_uniform_points(n, seed)
        compare_structure_f32(x, y, iv, 40, 10)
    x = hmx.chebyshevpoints(1000)
    y = hmx.chebyshevpoints(1000, kind=2)
    compare_structure_f32(x, y, (1.0, -1.0, 1.0, -1.0), 40, 10)
    x, y, iv = hmx.workloads.sorted_uniform_points(1 << 16, 1)
    compare_structure_f32(x, y, iv, 512, 32, O.LOG, "logkernel")


def test_f32_defaults_and_errors():
    x, y, iv = hmx.workloads.shifted_grid(256)
    x32, y32 = x.astype(np.float32), y.astype(np.float32)
    H = hmx.KernelMatrix("cauchykernel", x32, y32, *iv, tree_only=True)
    assert (H.rank, H.blocksize_) == (hmx.BLOCKRANK(np.float32), hmx.BLOCKSIZE(np.float32)) == (10, 40)
    # Float32 plans take rank <= 32
    assert status_of(lambda: hmx.KernelMatrix("cauchykernel", x32, y32, *iv, blocksize=16,
                                              rank=33, tree_only=True)) == _lib.HM_ERR_UNSUPPORTED
    assert status_of(lambda: hmx.KernelMatrix("cauchykernel", x32[::-1].copy(), y32, *iv,
                                              tree_only=True)) == _lib.HM_ERR_NOT_DESCENDING
    # a Float64 product entry point on a Float32 plan fails before any GPU work
    assert _lib.lib.hm_matvec_f32(H._h, None, None, 0, 0, 1, 1, 1, 0, 0, 0, 0, None) == \
        _lib.HM_ERR_INVALID


def test_wide_intervals_are_reported():
    """Points in [0.1, 0.9] with intervals (1.5, -0.5): deep in the tree one
    side's range is empty while the other still splits, so the reference's
    indsplit reads outside its range. The oracle raises; libhmx returns
    HM_ERR_DEGENERATE."""
    n = 1250
    x = np.repeat((np.arange(n, 0, -1) - 0.25) / n, 2) * 0.8 + 0.1
    y = np.repeat((np.arange(n, 0, -1) - 0.75) / n, 2) * 0.8 + 0.1
    iv = (1.5, -0.5, 1.5, -0.5)
    with pytest.raises(ValueError):
        O.Tree(O.CAUCHY, x, y, *iv, 40, 12, store=False)
    assert status_of(lambda: tree("cauchykernel", x, y, iv, 40, 12)) == _lib.HM_ERR_DEGENERATE
