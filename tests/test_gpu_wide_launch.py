"""Launch shapes whose per-workgroup work exceeds one entry per thread:
8 right-hand sides per launch on the vector units at ranks above 32 (the
upward pass's partial sums of 8 x r entries exceed its 256 threads), and
EvenBarycentricMatrix leaves whose reduced coefficients (2 parities x right-
hand sides x rank) exceed the reduce kernel's 256 threads. Every product is
checked against the oracle's restatement of the reference's mul! per column.
Bars: <= 1e-12 relative 2-norm for KernelMatrix products (rank 48/64 makes the
far field exact to round-off, so only the summation order differs), <= 1e-13
for HierarchicalMatrix leaves (dense assembly), <= 1e-14 for the even leaves
(same factors as the oracle)."""
import numpy as np
import pytest
import torch

import hmx
import oracle as O
from hmx import Block, HierarchicalMatrix, LowRankMatrix

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("rank,bs", [(48, 192), (64, 256)])
@pytest.mark.parametrize("mode", ["stored", "hybrid", "free"])
@pytest.mark.parametrize("k", [8, 13, 16])
def test_eight_rhs_per_launch_high_rank(cuda, rank, bs, mode, k):
    """k >= 8 on the vector units (mfma=False): launches of 8 right-hand
    sides; at rank 48 and 64 the partial-sum write covers 384 and 512
    entries."""
    n = 12000
    x, y, iv = hmx.workloads.sorted_uniform_points(n, 2)
    print(f"case: wide kb=8 rank={rank} mode={mode} k={k}", flush=True)
    H = hmx.KernelMatrix("cauchykernel", x, y, *iv, blocksize=bs, rank=rank,
                         matrix_free=mode == "free", hybrid=mode == "hybrid")
    B = hmx.workloads.normal(11, n * k).reshape(n, k)
    Y = H.matvec(torch.from_numpy(B).to(cuda), mfma=False).cpu().numpy()
    t = O.Tree(O.CAUCHY, x, y, *iv, bs, rank)
    for q in sorted({0, 7, k - 1}):
        ref = t.matvec(np.ascontiguousarray(B[:, q]))
        assert rel(Y[:, q], ref) <= 1e-12, (q, rel(Y[:, q], ref))
    # the transpose (one right-hand side per launch) of the same plan against
    # the dense product: sum_j f(y_i, x_j) v_j = -(K^T v)_i for Cauchy, and
    # rank 48/64 interpolates to round-off
    Yt = H.matvec(torch.from_numpy(np.ascontiguousarray(B[:, :2])).to(cuda),
                  transpose=True).cpu().numpy()
    for q in range(2):
        want = -O.dense_matvec(O.CAUCHY, y, x, np.ascontiguousarray(B[:, q]))
        assert rel(Yt[:, q], want) <= 1e-12, (q, rel(Yt[:, q], want))


def test_blocks_plan_rank40_leaf_eight_rhs(cuda):
    """A HierarchicalMatrix with a rank-40 LowRank leaf, 8 and 11 right-hand
    sides (one launch of 8, then 2 and 1), against its dense assembly."""
    rng = np.random.default_rng(40)
    H = HierarchicalMatrix(2, 2)
    H[Block(1), Block(1)] = rng.standard_normal((150, 130))
    H[Block(1), Block(2)] = LowRankMatrix(rng.standard_normal((150, 40)), rng.standard_normal(40),
                                          rng.standard_normal((170, 40)))
    H[Block(2), Block(1)] = LowRankMatrix(rng.standard_normal((90, 7)), rng.standard_normal(7),
                                          rng.standard_normal((130, 7)))
    H[Block(2), Block(2)] = rng.standard_normal((90, 170))
    D = np.zeros((240, 300))
    for c, i0, j0 in H.leaves():
        blk = c.U @ np.diag(c.S) @ c.V.T if isinstance(c, LowRankMatrix) else c
        D[i0:i0 + blk.shape[0], j0:j0 + blk.shape[1]] = blk
    for k in (8, 11):
        X = rng.standard_normal((300, k))
        Y = H.plan().matvec(torch.from_numpy(X).to(cuda)).cpu().numpy()
        assert rel(Y, D @ X) <= 1e-13, (k, rel(Y, D @ X))


def test_tune_on_blocks_plan(cuda):
    """hm_plan_tune on a plan built from blocks: stored form only; the tuned
    product equals the untuned one."""
    rng = np.random.default_rng(41)
    H = HierarchicalMatrix(2, 2)
    H[Block(1), Block(1)] = rng.standard_normal((200, 200))
    H[Block(1), Block(2)] = LowRankMatrix(rng.standard_normal((200, 12)), rng.standard_normal(12),
                                          rng.standard_normal((180, 12)))
    H[Block(2), Block(1)] = LowRankMatrix(rng.standard_normal((160, 12)), rng.standard_normal(12),
                                          rng.standard_normal((200, 12)))
    H[Block(2), Block(2)] = rng.standard_normal((160, 180))
    P = H.plan()
    kb, mode, us = P.tune(k=8)
    assert kb in (1, 2, 4, 8) and mode == 0 and us > 0
    X = torch.from_numpy(rng.standard_normal((380, 8))).to(cuda)
    assert rel(P.matvec(X, tuned=True).cpu().numpy(), P.matvec(X).cpu().numpy()) <= 1e-14


def f_osc(T, t, j):
    return np.cos(0.01 * t * j) / (1.0 + 0.001 * (t - j) ** 2)


@pytest.mark.parametrize("rank", [48, 64])
@pytest.mark.parametrize("k", [3, 4])
def test_even_wide_leaf_high_rank_reduce(cuda, rank, k):
    """p = 9000 columns (three upward tasks) at rank 48/64 with 3-4
    right-hand sides per launch: 288-512 reduced entries per leaf."""
    B = hmx.EvenBarycentricMatrix(f_osc, 1, 700, 1, 9000, rank=rank)
    H = hmx.EvenBarycentricLeaves(700, 9000, [(0, 0, B)])
    assert H.info["n_reduce"] == 1
    rng = np.random.default_rng(rank + k)
    V = rng.standard_normal((9000, k))
    U = H.matvec(torch.from_numpy(V).to(cuda)).cpu().numpy()
    for q in range(k):
        want = O.even_mul(np.zeros(700), B.W, B.F, np.ascontiguousarray(V[:, q]), 1, 1)
        assert rel(U[:, q], want) <= 1e-14, (q, rel(U[:, q], want))
