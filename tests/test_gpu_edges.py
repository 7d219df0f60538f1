"""Edge cases of the KernelMatrix product against the oracle's mul!: a single
point, fewer points than the leaf threshold (the root's four children are all
dense), rectangular matrices (nx != ny), tied points, k = 0 right-hand sides,
and the largest supported rank on a small problem. (Intervals much wider than
the points make the reference's recursion read out of range; that error path
is checked on the CPU in test_abi.py.) Bar: <= 1e-12 relative 2-norm (exact equality where the product is a
single dense block)."""
import numpy as np
import pytest
import torch

import hmx
import oracle as O

pytestmark = pytest.mark.gpu
MODES = {"stored": {}, "hybrid": {"hybrid": True}, "free": {"matrix_free": True}}


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb else float(np.linalg.norm(a - b))


def check(x, y, iv, bs, r, mode, k=1, seed=0):
    H = hmx.KernelMatrix("cauchykernel", x, y, *iv, blocksize=bs, rank=r, **MODES[mode])
    B = hmx.workloads.normal(seed, len(y) * k).reshape(len(y), k)
    Y = H.matvec(torch.from_numpy(B).cuda()).cpu().numpy()
    t = O.Tree(O.CAUCHY, x, y, *iv, bs, r)
    for q in range(k):
        ref = t.matvec(np.ascontiguousarray(B[:, q]))
        assert rel(Y[:, q], ref) <= 1e-12, (mode, q, rel(Y[:, q], ref))
    return H, Y


@pytest.mark.parametrize("mode", list(MODES))
def test_single_point(cuda, mode):
    x = np.array([0.75])
    y = np.array([0.25])
    H, Y = check(x, y, (1.0, 0.0, 1.0, 0.0), 64, 16, mode)
    assert Y[0, 0] == pytest.approx(1.0 / 0.5 * hmx.workloads.normal(0, 1)[0], rel=1e-15)
    assert H.info["n_far"] == 0


@pytest.mark.parametrize("mode", list(MODES))
def test_below_leaf_threshold_is_all_dense(cuda, mode):
    x, y, iv = hmx.workloads.shifted_grid(50)
    H, Y = check(x, y, iv, 64, 16, mode, k=3)
    info = H.info
    assert info["n_far"] == 0 and info["n_dense"] == 4 and info["dense_entries"] == 50 * 50


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("k", [1, 9])
def test_rectangular(cuda, mode, k):
    """nx = 3000, ny = 5000 on the same interval: the two trees split their own
    points, and the product is 3000 x 5000."""
    x = (np.arange(3000, 0, -1) - 0.3) / 3000
    y = (np.arange(5000, 0, -1) - 0.6) / 5000
    H, Y = check(x, y, (1.0, 0.0, 1.0, 0.0), 64, 16, mode, k=k)
    assert Y.shape == (3000, k)


@pytest.mark.parametrize("mode", list(MODES))
def test_tied_points(cuda, mode):
    """Every point doubled (non-increasing, not strictly): ties stay on one
    side of every split."""
    n = 1250
    x = np.repeat((np.arange(n, 0, -1) - 0.25) / n, 2)
    y = np.repeat((np.arange(n, 0, -1) - 0.75) / n, 2)
    check(x, y, (1.0, 0.0, 1.0, 0.0), 40, 12, mode, k=2)


def test_zero_rhs_is_a_no_op(cuda):
    """k = 0 returns at once, before looking at the (here absent) vectors."""
    from hmx._lib import HM_OK, lib
    x, y, iv = hmx.workloads.shifted_grid(1000)
    H = hmx.KernelMatrix("cauchykernel", x, y, *iv, blocksize=64, rank=16)
    assert lib.hm_matvec(H._h, None, None, 0, 0, 1, 1, 0, 0, 0, 0, 0, None) == HM_OK
    assert lib.hm_matvec_t(H._h, None, None, 0, 0, 1, 1, 0, 0, 0, 0, 0, None) == HM_OK


@pytest.mark.parametrize("mode", list(MODES))
def test_rank_64_small_problem(cuda, mode):
    x, y, iv = hmx.workloads.shifted_grid(3000)
    check(x, y, iv, 256, 64, mode, k=1)


def test_leaf_threshold_larger_than_n(cuda):
    """BLOCKSIZE beyond n: the root is a single 2 x 2 grid of dense blocks, and
    the product equals the dense product computed in the reference's order."""
    x, y, iv = hmx.workloads.shifted_grid(700)
    b = hmx.workloads.normal(1, 700)
    H = hmx.KernelMatrix("cauchykernel", x, y, *iv, blocksize=4096, rank=16)
    u = H.matvec(torch.from_numpy(b).cuda()).cpu().numpy()
    ref = O.Tree(O.CAUCHY, x, y, *iv, 4096, 16).matvec(b)
    assert rel(u, ref) <= 1e-13
