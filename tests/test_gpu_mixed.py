"""Reduced-precision far field (HM_MV_FAR_F32, HM_MV_FAR_BF16X3,
HM_MV_FAR_BF16X6): far-field products on MFMA with FP32 accumulation per
step and FP64 across steps, FP64 for point differences, the Chebyshev
recurrence, the near field and F. Stated bound against the FP64 product and
the oracle: <= 1e-6 relative 2-norm for every form on these Cauchy matrices
(the near field, exact in FP64, carries most of each row; the log kernel's
far-field-dominated bounds are in test_gpu_fullsize.py)."""
import numpy as np
import pytest
import torch

import hmx
import oracle as O

pytestmark = pytest.mark.gpu
BOUND = {"f32": 1e-6, "bf16x3": 1e-6, "bf16x6": 1e-6}


def rel(a, b):
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


@pytest.mark.parametrize("prec", ["f32", "bf16x3", "bf16x6"])
@pytest.mark.parametrize("mode", ["stored", "hybrid", "free"])
@pytest.mark.parametrize("k", [1, 7, 16, 32])
def test_mixed_vs_fp64(cuda, prec, mode, k):
    w = hmx.workloads.config_plain(n=8192, blocksize=64, rank=16)
    H = hmx.KernelMatrix("cauchykernel", w["x"], w["y"], *w["intervals"], blocksize=64, rank=16,
                         matrix_free=mode == "free", hybrid=mode == "hybrid")
    B = torch.from_numpy(hmx.workloads.normal(8, 8192 * k).reshape(8192, k)).to(cuda)
    ref = H @ B
    got = H.matvec(B, far_precision=prec)
    e = rel(got, ref)
    print(f"mixed {prec} mode={mode} k={k}: rel err {e:.3e}", flush=True)
    assert e <= BOUND[prec]


@pytest.mark.parametrize("prec", ["f32", "bf16x3", "bf16x6"])
@pytest.mark.parametrize("rank,bs", [(16, 64), (24, 96), (32, 128)])
def test_mixed_ranks_batched_vs_oracle(cuda, prec, rank, bs):
    w = hmx.workloads.config_batched(count=6, nmin_log2=11, nmax_log2=13, k=16, rank=rank,
                                     blocksize=bs)
    inst = w["instances"]
    H = hmx.KernelMatrixBatch("cauchykernel", [(i["x"], i["y"], i["intervals"]) for i in inst],
                              blocksize=bs, rank=rank, hybrid=True)
    Bcat = torch.from_numpy(np.concatenate([i["b"] for i in inst])).to(cuda)
    Y = H.matvec(Bcat, far_precision=prec).cpu().numpy()
    ro = H.row_offsets
    for t, i in enumerate(inst):
        tr = O.Tree(O.CAUCHY, i["x"], i["y"], *i["intervals"], bs, rank)
        for q in (0, 9, 15):
            ref = tr.matvec(np.ascontiguousarray(i["b"][:, q]))
            e = np.linalg.norm(Y[ro[t]:ro[t + 1], q] - ref) / np.linalg.norm(ref)
            assert e <= BOUND[prec], (t, q, e)


def test_mixed_graph_replays(cuda):
    """A captured reduced-precision product replays to the eager result, with
    eager work and synchronisation between replays."""
    w = hmx.workloads.config_plain(n=4096, blocksize=64, rank=16)
    H = hmx.KernelMatrix("cauchykernel", w["x"], w["y"], *w["intervals"], blocksize=64, rank=16,
                         hybrid=True)
    B = torch.from_numpy(hmx.workloads.normal(3, 4096 * 16).reshape(4096, 16)).to(cuda)
    eager = H.matvec(B, far_precision="f32")
    Y = torch.zeros_like(eager)
    g = H.graph(Y, B, far_precision="f32")
    for _ in range(3):
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(Y, eager)
        Y.zero_()
        _ = H.matvec(B)  # eager FP64 work between replays
        torch.cuda.synchronize()
    g.close()


def test_mixed_rejects_transpose_and_blocks(cuda):
    w = hmx.workloads.config_plain(n=1024, blocksize=64, rank=16)
    H = hmx.KernelMatrix("cauchykernel", w["x"], w["y"], *w["intervals"], blocksize=64, rank=16)
    b = torch.ones(1024, dtype=torch.float64, device=cuda)
    with pytest.raises(hmx.HMError):
        H.matvec(b, transpose=True, far_precision="f32")
