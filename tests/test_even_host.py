"""CPU checks of the EvenBarycentricMatrix entry points: the nodes at which
the caller evaluates f (hm_even_nodes, host) equal the oracle's restatement of
src/BarycentricMatrix.jl:41 bit for bit, and bad input is turned into status
codes before any device work."""
import ctypes

import numpy as np
import pytest

import hmx
import oracle as O
from hmx import _lib

L = _lib.lib
_ip = ctypes.POINTER(ctypes.c_int64)
_dp = ctypes.POINTER(ctypes.c_double)


@pytest.mark.parametrize("a,b,n", [(1, 40, 20), (2, 9, 7), (-3, 1000, 20), (0, 1, 64), (7, 7, 8),
                                   (-(1 << 40), 1 << 40, 20)])
def test_nodes_match_oracle(a, b, n):
    assert np.array_equal(hmx.even.even_nodes(a, b, n), O.even_nodes(a, b, n))


def test_nodes_reject_bad_input():
    t = np.zeros(70)
    tp = t.ctypes.data_as(_dp)
    assert L.hm_even_nodes(1, 10, 0, tp) == _lib.HM_ERR_INVALID
    assert L.hm_even_nodes(1, 10, 65, tp) == _lib.HM_ERR_INVALID
    assert L.hm_even_nodes(10, 1, 8, tp) == _lib.HM_ERR_INVALID
    assert b"empty grid" in L.hm_last_error()
    assert L.hm_even_nodes(0, 1 << 51, 8, tp) == _lib.HM_ERR_INVALID


def create(nrows, ncols, row0, col0, abcd, n, Fs):
    r0 = np.array(row0, dtype=np.int64)
    c0 = np.array(col0, dtype=np.int64)
    g = np.array(abcd, dtype=np.int64).reshape(-1)
    Fp = (_dp * max(len(Fs), 1))(*[F.ctypes.data_as(_dp) for F in Fs])
    h = ctypes.c_void_p()
    rc = L.hm_even_create(nrows, ncols, len(Fs), r0.ctypes.data_as(_ip), c0.ctypes.data_as(_ip),
                          g.ctypes.data_as(_ip), n, Fp, None, ctypes.byref(h))
    return rc, h


def test_create_rejects_bad_leaves():
    F = np.zeros((10, 8), order="F")
    # rank out of range
    assert create(20, 20, [0], [0], [1, 10, 1, 10], 0, [F])[0] == _lib.HM_ERR_INVALID
    assert create(20, 20, [0], [0], [1, 10, 1, 10], 65, [F])[0] == _lib.HM_ERR_INVALID
    # leaf outside the matrix (10 rows at row 15 of 20)
    rc, h = create(20, 20, [15], [0], [1, 10, 1, 10], 8, [F])
    assert rc == _lib.HM_ERR_INVALID and not h.value
    assert b"outside" in L.hm_last_error()
    # empty column grid
    assert create(20, 20, [0], [0], [1, 10, 5, 4], 8, [F])[0] == _lib.HM_ERR_INVALID
    # negative placement
    assert create(20, 20, [-1], [0], [1, 10, 1, 10], 8, [F])[0] == _lib.HM_ERR_INVALID


def test_matvec_rejects_bad_arguments():
    assert L.hm_even_matvec(None, None, None, 0, 0, 1, 1, 1, 0, 0, 1, None) == _lib.HM_ERR_INVALID
    assert L.hm_even_weights(1, 10, 8, 2, None, None, None) == _lib.HM_ERR_INVALID
    assert L.hm_even_weights(1, 10, 8, 0, None, None, None) == _lib.HM_ERR_INVALID


@pytest.mark.parametrize("a,b,n", [(1, 40, 10), (2, 9, 7), (-3, 1000, 10), (0, 1, 64), (7, 7, 8),
                                   (-(1 << 30), 1 << 30, 10), (1, 1 << 22, 10)])
def test_nodes_f32_match_oracle(a, b, n):
    """Float32 nodes: x_k and (b-a)*x_k/2 in Float32, the rest in Float64."""
    t = hmx.even.even_nodes(a, b, n, np.float32)
    assert np.array_equal(t, O.even_nodes(a, b, n, f32=True))
    if n > 1 and b > a + 8:
        assert not np.array_equal(t, O.even_nodes(a, b, n))


@pytest.mark.parametrize("n", [1, 2, 7, 10, 16, 33, 64])
@pytest.mark.parametrize("kind", [1, 2])
def test_chebyshev_f32_match_oracle(n, kind):
    if kind == 2 and n < 2:
        return
    x = hmx.chebyshevpoints(n, kind, dtype=np.float32)
    lam = hmx.chebyshevbarycentricweights(n, kind, dtype=np.float32)
    assert x.dtype == np.float32 and lam.dtype == np.float32
    assert np.array_equal(x, O.chebyshevpoints(n, kind, f32=True))
    assert np.array_equal(lam, O.chebyshevbarycentricweights(n, kind, f32=True))


def test_create_f32_rejects_bad_leaves():
    _fp = ctypes.POINTER(ctypes.c_float)
    F = np.zeros((10, 8), dtype=np.float32, order="F")
    r0 = np.array([15], dtype=np.int64)
    c0 = np.array([0], dtype=np.int64)
    g = np.array([1, 10, 1, 10], dtype=np.int64)
    Fp = (_fp * 1)(F.ctypes.data_as(_fp))
    h = ctypes.c_void_p()
    rc = L.hm_even_create_f32(20, 20, 1, r0.ctypes.data_as(_ip), c0.ctypes.data_as(_ip),
                              g.ctypes.data_as(_ip), 8, Fp, None, ctypes.byref(h))
    assert rc == _lib.HM_ERR_INVALID and not h.value
    assert b"outside" in L.hm_last_error()
    assert L.hm_even_matvec_f32(None, None, None, 0, 0, 1, 1, 1, 0, 0, 1, None) == \
        _lib.HM_ERR_INVALID
    assert L.hm_even_weights_f32(1, 10, 8, 2, None, None, None) == _lib.HM_ERR_INVALID
    assert L.hm_even_nodes_f32(1, 10, 65, None) == _lib.HM_ERR_INVALID
