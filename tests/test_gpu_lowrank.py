"""LowRankMatrix arithmetic on the GPU (hm_lowrank_op) against the oracle's
numpy restatement of src/LowRankMatrix.jl:64-171 (pinned in
test_oracle_lowrank.py).

Singular vectors are fixed only up to sign, so the bars are on what is
unique: the rank (getrank), the singular values (<= 1e-13 * S[1]), the
reconstruction U diag(S) V^T (<= 1e-13 relative) and the orthonormality of
U and V (<= 1e-13)."""
import numpy as np
import pytest
import torch

import hmx
import lowrank as LR
from hmx import lowrank as DL

pytestmark = pytest.mark.gpu


def rand_lr(rng, m, n, r, decay=False):
    U, _ = np.linalg.qr(rng.standard_normal((m, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)))
    S = (np.sort(rng.random(r) + 0.1)[::-1] if not decay
         else 3.0 * 10.0 ** (-12.0 * np.arange(r) / max(r - 1, 1)))
    return U, S, V


def dev(a):
    return torch.from_numpy(np.array(a, dtype=np.float64, order="C", copy=True)).cuda()


def items_of(**kw):
    return {k: dev(v) for k, v in kw.items()}


def agree(got, want):
    U, S, V = (x.cpu().numpy() if torch.is_tensor(x) else x for x in got)
    Uw, Sw, Vw = want
    assert len(S) == len(Sw), (len(S), len(Sw))
    r = len(S)
    if r == 0:
        return
    assert np.max(np.abs(S - Sw)) <= 1e-13 * Sw[0]
    assert np.allclose(U.T @ U, np.eye(r), atol=1e-13)
    assert np.allclose(V.T @ V, np.eye(r), atol=1e-13)
    D, Dw = LR.dense((U, S, V)), LR.dense((Uw, Sw, Vw))
    assert np.linalg.norm(D - Dw) <= 1e-13 * np.linalg.norm(Dw)


@pytest.mark.parametrize("op", ["add", "sub"])
def test_sums_batched(cuda, op):
    rng = np.random.default_rng(1)
    shapes = [(200, 150, 12, 9), (64, 64, 32, 32), (1000, 77, 5, 20), (40, 33, 1, 1),
              (300, 300, 16, 16)]
    items, want = [], []
    for m, n, r1, r2 in shapes:
        L1, L2 = rand_lr(rng, m, n, r1), rand_lr(rng, m, n, r2)
        items.append(items_of(U1=L1[0], S1=L1[1], V1=L1[2], U2=L2[0], S2=L2[1], V2=L2[2]))
        want.append(LR.add(L1, L2, 1.0 if op == "add" else -1.0))
    got = DL.batch(op, items)
    for g, w in zip(got, want):
        agree(g, w)


def test_sum_with_shared_columns_truncates(cuda):
    rng = np.random.default_rng(2)
    L = rand_lr(rng, 120, 100, 10)
    got = DL.batch("add", [items_of(U1=L[0], S1=L[1], V1=L[2], U2=L[0], S2=L[1], V2=L[2])])[0]
    agree(got, LR.add(L, L))
    assert len(got[1]) == 10


def test_sum_of_decaying_spectra(cuda):
    """Singular values spanning 1 .. 1e-12, all above getrank's tolerance:
    every one is kept, and the small ones agree to the absolute accuracy of
    the factorization (~1e-14 of the largest)."""
    rng = np.random.default_rng(3)
    L1, L2 = rand_lr(rng, 256, 200, 12, decay=True), rand_lr(rng, 256, 200, 8, decay=True)
    want = LR.add(L1, L2)
    got = DL.batch("add", [items_of(U1=L1[0], S1=L1[1], V1=L1[2], U2=L2[0], S2=L2[1],
                                    V2=L2[2])])[0]
    U, S, V = (x.cpu().numpy() for x in got)
    assert len(S) == len(want[1])
    assert np.allclose(S, want[1], rtol=1e-9, atol=1e-14 * want[1][0])
    assert np.linalg.norm(LR.dense((U, S, V)) - LR.dense(want)) <= 1e-13 * want[1][0] * 10


@pytest.mark.parametrize("r1,r2", [(8, 11), (11, 8), (16, 16), (1, 5)])
def test_products(cuda, r1, r2):
    rng = np.random.default_rng(r1 * 7 + r2)
    L1, L2 = rand_lr(rng, 90, 70, r1), rand_lr(rng, 70, 60, r2)
    agree(DL.batch("mul", [items_of(U1=L1[0], S1=L1[1], V1=L1[2], U2=L2[0], S2=L2[1],
                                    V2=L2[2])])[0], LR.mul(L1, L2))
    L3 = rand_lr(rng, 90, 40, r2)
    agree(DL.batch("adjmul", [items_of(U1=L1[0], S1=L1[1], V1=L1[2], U2=L3[0], S2=L3[1],
                                       V2=L3[2])])[0], LR.adjmul(L1, L3))


def test_dense_times_lowrank(cuda):
    rng = np.random.default_rng(4)
    L = rand_lr(rng, 80, 50, 7)
    A = rng.standard_normal((60, 80))
    agree(DL.batch("dmul", [items_of(A=A, U2=L[0], S2=L[1], V2=L[2])])[0], LR.dmul(A, L))
    B = rng.standard_normal((50, 45))
    agree(DL.batch("muld", [items_of(A=B, U1=L[0], S1=L[1], V1=L[2])])[0], LR.muld(L, B))
    d = rng.random(80) + 0.5
    agree(DL.batch("dmul", [items_of(A=np.diag(d), U2=L[0], S2=L[1], V2=L[2])])[0],
          LR.dmul(np.diag(d), L))


@pytest.mark.parametrize("m,n", [(64, 48), (48, 64), (500, 30), (30, 500), (20, 20)])
def test_svdtrunc(cuda, m, n):
    rng = np.random.default_rng(m + n)
    A = LR.dense(rand_lr(rng, m, n, 6)) + 1e-3 * rng.standard_normal((m, n))
    agree(DL.batch("svdtrunc", [items_of(A=A)])[0], LR.svdtrunc(A))


def test_operator_forms_on_host_objects(cuda):
    rng = np.random.default_rng(5)
    L1 = hmx.LowRankMatrix(*rand_lr(rng, 100, 80, 6))
    L2 = hmx.LowRankMatrix(*rand_lr(rng, 100, 80, 4))
    agree(((L1 + L2).U, (L1 + L2).S, (L1 + L2).V), LR.add((L1.U, L1.S, L1.V), (L2.U, L2.S, L2.V)))
    D = L1 - L2
    agree((D.U, D.S, D.V), LR.add((L1.U, L1.S, L1.V), (L2.U, L2.S, L2.V), -1.0))
    L3 = hmx.LowRankMatrix(*rand_lr(rng, 80, 30, 5))
    P = L1 @ L3
    agree((P.U, P.S, P.V), LR.mul((L1.U, L1.S, L1.V), (L3.U, L3.S, L3.V)))
    A = rng.standard_normal((25, 100))
    Q = A @ L1
    agree((Q.U, Q.S, Q.V), LR.dmul(A, (L1.U, L1.S, L1.V)))


def test_unsupported_sizes_raise(cuda):
    rng = np.random.default_rng(6)
    L1, L2 = rand_lr(rng, 100, 100, 40), rand_lr(rng, 100, 100, 30)
    with pytest.raises(hmx.HMError) as e:
        DL.batch("add", [items_of(U1=L1[0], S1=L1[1], V1=L1[2], U2=L2[0], S2=L2[1], V2=L2[2])])
    assert e.value.code == 6


def test_factors_beyond_lds_in_one_batch(cuda):
    """Tall factors too large for the in-LDS factorization (5000 x 32, and
    3000 x 48 for a product's dense side) run the global-memory QR; a batch
    mixing both sizes gives each item its own result."""
    rng = np.random.default_rng(7)
    shapes = [(5000, 3000, 16, 16), (200, 150, 12, 9), (4096, 64, 30, 2)]
    items, want = [], []
    for m, n, r1, r2 in shapes:
        L1, L2 = rand_lr(rng, m, n, r1), rand_lr(rng, m, n, r2)
        items.append(items_of(U1=L1[0], S1=L1[1], V1=L1[2], U2=L2[0], S2=L2[1], V2=L2[2]))
        want.append(LR.add(L1, L2))
    for g, w in zip(DL.batch("add", items), want):
        agree(g, w)
    L = rand_lr(rng, 3000, 100, 48)
    A = rng.standard_normal((2500, 3000))
    agree(DL.batch("dmul", [items_of(A=A, U2=L[0], S2=L[1], V2=L[2])])[0], LR.dmul(A, L))
