"""Multi-RHS and batched (grouped) products against the oracle, per instance
and per right-hand side."""
import numpy as np
import pytest
import torch

import hmx
import oracle as O

pytestmark = pytest.mark.gpu
MODES = ["stored", "hybrid", "free"]


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def mk(mode):
    return dict(matrix_free=mode == "free", hybrid=mode == "hybrid")


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("k", [2, 5, 8, 13, 32])
def test_multi_rhs_matches_single(cuda, mode, k):
    w = hmx.workloads.config_plain(n=4096, blocksize=64, rank=16)
    H = hmx.KernelMatrix("cauchykernel", w["x"], w["y"], *w["intervals"], blocksize=64, rank=16,
                         **mk(mode))
    B = torch.from_numpy(hmx.workloads.normal(31, 4096 * k).reshape(4096, k)).to(cuda)
    Y = H @ B
    # from 8 right-hand sides on the product runs on the MFMA passes (far
    # field in Chebyshev form): equal to the single-RHS product up to rounding
    tol = 1e-13 if k >= 8 else 1e-14
    for q in range(k):
        yq = H @ B[:, q].contiguous()
        assert rel(Y[:, q].cpu().numpy(), yq.cpu().numpy()) <= tol, q
    # column-major right-hand sides (ld > 1 between RHS, unit stride inside)
    Bt = B.T.contiguous().T
    assert rel((H @ Bt).cpu().numpy(), Y.cpu().numpy()) <= 1e-15


@pytest.mark.parametrize("mode", MODES)
def test_batched_vs_oracle(cuda, mode):
    w = hmx.workloads.config_batched(count=10, nmin_log2=10, nmax_log2=13, k=6)
    inst = w["instances"]
    print(f"case: batched 10 instances mode={mode}", flush=True)
    H = hmx.KernelMatrixBatch("cauchykernel", [(i["x"], i["y"], i["intervals"]) for i in inst],
                              blocksize=64, rank=16, **mk(mode))
    assert H.info["n_instances"] == 10
    Bcat = np.concatenate([i["b"] for i in inst])
    Y = (H @ torch.from_numpy(Bcat).to(cuda)).cpu().numpy()
    ro = H.row_offsets
    for t, i in enumerate(inst):
        tr = O.Tree(O.CAUCHY, i["x"], i["y"], *i["intervals"], 64, 16)
        for q in range(6):
            ref = tr.matvec(np.ascontiguousarray(i["b"][:, q]))
            assert rel(Y[ro[t]:ro[t + 1], q], ref) <= 1e-12, (t, q)


def test_batched_config_sampled_rows(cuda):
    """256 instances, n log-uniform in [2^10, 2^16], rank 16, leaf 64, 32 RHS
    (hybrid plan): sampled rows of several instances against the oracle."""
    w = hmx.workloads.config_batched()
    inst = w["instances"]
    H = hmx.KernelMatrixBatch("cauchykernel", [(i["x"], i["y"], i["intervals"]) for i in inst],
                              blocksize=64, rank=16, hybrid=True)
    info = H.info
    print("case: batched config", info["nrows"], info["n_instances"], flush=True)
    Bcat = torch.from_numpy(np.concatenate([i["b"] for i in inst])).to(cuda)
    Y = (H @ Bcat).cpu().numpy()
    assert np.all(np.isfinite(Y))
    ro = H.row_offsets
    rng = np.random.default_rng(4)
    for t in [0, 17, 100, 255] + list(rng.integers(0, 256, 3)):
        i = inst[t]
        tr = O.Tree(O.CAUCHY, i["x"], i["y"], *i["intervals"], 64, 16, store=False)
        rows = np.unique(rng.integers(0, len(i["x"]), 24))
        for q in (0, 13, 31):
            ref = tr.matvec_rows(np.ascontiguousarray(i["b"][:, q]), rows)
            assert rel(Y[ro[t] + rows, q], ref) <= 1e-12, (t, q)


def test_graph_replays_match_eager(cuda):
    """A captured product replayed several times, with eager work and
    synchronisations between replays, equals the eager product each time."""
    w = hmx.workloads.config_batched(count=8, nmin_log2=10, nmax_log2=12, k=8)
    inst = w["instances"]
    H = hmx.KernelMatrixBatch("cauchykernel", [(i["x"], i["y"], i["intervals"]) for i in inst],
                              blocksize=64, rank=16, hybrid=True)
    B = torch.from_numpy(np.concatenate([i["b"] for i in inst])).to(cuda)
    Y = torch.zeros(H.shape[0], 8, dtype=torch.float64, device=cuda)
    g = H.graph(Y, B)
    for rep in range(4):
        B.mul_(1.5).add_(0.25)  # eager work between replays
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        want = H @ B
        assert torch.equal(Y, want), rep
    g.close()


def test_tuned_products_match(cuda):
    w = hmx.workloads.config_plain(n=8192, blocksize=64, rank=16)
    H = hmx.KernelMatrix("cauchykernel", w["x"], w["y"], *w["intervals"], blocksize=64, rank=16)
    kb, mode, us = H.tune(k=16)
    # 1, 2, 4, 8 right-hand sides per launch on the vector units, 16 on MFMA
    assert kb in (1, 2, 4, 8, 16) and mode in (0, 1, 2) and us > 0
    B = torch.from_numpy(hmx.workloads.normal(2, 8192 * 16).reshape(8192, 16)).to(cuda)
    # the default for 16 right-hand sides is the MFMA passes; any tuned choice
    # agrees with them to rounding
    assert rel(H.matvec(B, tuned=True).cpu().numpy(), (H @ B).cpu().numpy()) <= 1e-13
