"""Multi-RHS products on the FP64 MFMA passes (16 right-hand sides per
launch, far field in Chebyshev form) against the oracle's mul! per column and
against the vector-unit passes (up to 8 right-hand sides per launch). Bars:
<= 1e-12 relative 2-norm per column against the oracle (the reference's
summation order versus the MFMA one, and barycentric versus Chebyshev form),
<= 1e-13 against the vector-unit product."""
import numpy as np
import pytest
import torch

import hmx
import oracle as O

pytestmark = pytest.mark.gpu
KIDS = {"cauchykernel": O.CAUCHY, "coulombkernel": O.COULOMB, "coulombprimekernel": O.COULOMB3,
        "logkernel": O.LOG}


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rhs(seed, n, k):
    return hmx.workloads.normal(seed, n * k).reshape(n, k)


@pytest.mark.parametrize("kernel", list(KIDS))
@pytest.mark.parametrize("mode", ["stored", "hybrid", "free"])
@pytest.mark.parametrize("k", [16, 20, 32, 45])
def test_mfma_vs_oracle_and_vector_units(cuda, kernel, mode, k):
    w = hmx.workloads.config_plain(n=4096, blocksize=64, rank=16)
    print(f"case: mfma kernel={kernel} mode={mode} k={k}", flush=True)
    H = hmx.KernelMatrix(kernel, w["x"], w["y"], *w["intervals"], blocksize=64, rank=16,
                         matrix_free=mode == "free", hybrid=mode == "hybrid")
    B = rhs(3, 4096, k)
    Bt = torch.from_numpy(B).to(cuda)
    Ym = H.matvec(Bt, mfma=True).cpu().numpy()
    Yv = H.matvec(Bt, mfma=False).cpu().numpy()
    assert rel(Ym, Yv) <= 1e-13
    t = O.Tree(KIDS[kernel], w["x"], w["y"], *w["intervals"], 64, 16)
    # launches of 32 right-hand sides (two groups of 16) and the rest
    for q in sorted({0, 7, 15, min(16, k - 1), min(31, k - 1), k - 1}):
        ref = t.matvec(np.ascontiguousarray(B[:, q]))
        assert rel(Ym[:, q], ref) <= 1e-12, (q, rel(Ym[:, q], ref))


@pytest.mark.parametrize("rank,bs", [(8, 32), (24, 96), (40, 160), (64, 256)])
def test_mfma_rank_buckets(cuda, rank, bs):
    x, y, iv = hmx.workloads.sorted_uniform_points(12000, 2)
    print(f"case: mfma rank={rank}", flush=True)
    H = hmx.KernelMatrix("cauchykernel", x, y, *iv, blocksize=bs, rank=rank, hybrid=True)
    B = rhs(4, 12000, 32)
    Ym = H.matvec(torch.from_numpy(B).to(cuda), mfma=True).cpu().numpy()
    t = O.Tree(O.CAUCHY, x, y, *iv, bs, rank)
    for q in (0, 9, 15, 16, 31):
        ref = t.matvec(np.ascontiguousarray(B[:, q]))
        assert rel(Ym[:, q], ref) <= 1e-12, (q, rel(Ym[:, q], ref))


def test_mfma_large_clusters_multi_task(cuda):
    """n = 2^17 at leaf 256: the top clusters hold more than 8192 points, so
    they have several upward tasks and go through the reduce."""
    n = 1 << 17
    x, y, iv = hmx.workloads.shifted_grid(n)
    H = hmx.KernelMatrix("cauchykernel", x, y, *iv, blocksize=256, rank=24, hybrid=True)
    B = rhs(5, n, 16)
    Bt = torch.from_numpy(B).to(cuda)
    Ym = H.matvec(Bt, mfma=True).cpu().numpy()
    Yv = H.matvec(Bt, mfma=False).cpu().numpy()
    assert rel(Ym, Yv) <= 1e-13
    t = O.Tree(O.CAUCHY, x, y, *iv, 256, 24, store=False)
    rows = np.arange(0, n, 997)
    for q in (0, 15):
        ref = t.matvec_rows(np.ascontiguousarray(B[:, q]), rows)
        assert rel(Ym[rows, q], ref) <= 1e-12


def test_mfma_batched_vs_oracle(cuda):
    w = hmx.workloads.config_batched(count=12, nmin_log2=10, nmax_log2=13, k=32)
    inst = w["instances"]
    H = hmx.KernelMatrixBatch("cauchykernel", [(i["x"], i["y"], i["intervals"]) for i in inst],
                              blocksize=64, rank=16)
    Bcat = np.concatenate([i["b"] for i in inst])
    Y = H.matvec(torch.from_numpy(Bcat).to(cuda)).cpu().numpy()  # k = 32: MFMA by default
    ro = H.row_offsets
    for t_, i in enumerate(inst):
        tr = O.Tree(O.CAUCHY, i["x"], i["y"], *i["intervals"], 64, 16)
        for q in (0, 16, 31):
            ref = tr.matvec(np.ascontiguousarray(i["b"][:, q]))
            assert rel(Y[ro[t_]:ro[t_ + 1], q], ref) <= 1e-12, (t_, q)


def test_mfma_strides_and_accumulate(cuda):
    """y[:, q] = y + H x[:, q] with row strides and padded leading dimensions."""
    w = hmx.workloads.config_plain(n=4096, blocksize=64, rank=16)
    H = hmx.KernelMatrix("cauchykernel", w["x"], w["y"], *w["intervals"], blocksize=64, rank=16)
    B = rhs(6, 4096, 17)
    big = torch.zeros(2 * 4096 + 3, 17, dtype=torch.float64, device=cuda)
    big[1::2][:4096] = torch.from_numpy(B).to(cuda)
    X = big[1:1 + 2 * 4096:2]  # row stride 2 x 17
    Y0 = torch.from_numpy(rhs(7, 4096, 17)).to(cuda)
    Y = Y0.clone()
    H._mv(Y, X, 1, 1, 1, 1, 1, False, False, None, mfma=True)
    want = Y0 + H.matvec(torch.from_numpy(B).to(cuda), mfma=False)
    assert rel(Y.cpu().numpy(), want.cpu().numpy()) <= 1e-13
    # column-major output (each right-hand side contiguous, ldy = 4099)
    Yc = torch.zeros(17, 4099, dtype=torch.float64, device=cuda)
    Yt = Yc[:, :4096].t()
    Yt.copy_(Y0)
    H._mv(Yt, X, 1, 1, 1, 1, 1, False, False, None, mfma=True)
    assert Yt.stride() == (1, 4099)
    assert rel(Yt.cpu().numpy(), want.cpu().numpy()) <= 1e-13
    assert torch.count_nonzero(Yc[:, 4096:]) == 0


def test_mfma_graph_and_tuner(cuda):
    w = hmx.workloads.config_plain(n=4096, blocksize=64, rank=16)
    H = hmx.KernelMatrix("cauchykernel", w["x"], w["y"], *w["intervals"], blocksize=64, rank=16,
                         hybrid=True)
    B = torch.from_numpy(rhs(8, 4096, 32)).to(cuda)
    eager = H.matvec(B, mfma=True)
    Y = torch.zeros_like(eager)
    g = H.graph(Y, B, mfma=True)
    for _ in range(3):
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(Y, eager)
        Y.zero_()
        _ = H.matvec(B[:, :3].contiguous())
        torch.cuda.synchronize()
    g.close()
    kb, mode, us = H.tune(k=32)
    assert kb in (1, 2, 4, 8, 16) and mode in (1, 2) and us > 0
    Yt = H.matvec(B, tuned=True).cpu().numpy()
    assert rel(Yt, eager.cpu().numpy()) <= 1e-13
