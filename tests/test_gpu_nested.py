"""The nested far field of the MFMA passes: the moments of each column leaf,
combined up the column interval tree by the exact re-expansion of a half's
Chebyshev series on its parent's interval, and the translated coefficients
pushed down the row tree to each row's leaf. The plans of points (grid,
random points, the log kernel, batched instances) use it (info "nested");
plans built from blocks have no interval trees and do not. Bars: <= 1e-13
against the vector-unit product, which applies every covering cluster per
level, and <= 1e-12 per column against the oracle's mul!."""
import numpy as np
import pytest
import torch

import hmx
import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


CASES = {
    "grid": lambda: (*hmx.workloads.shifted_grid(4096), 64, 16, "cauchykernel", O.CAUCHY),
    "random": lambda: (*hmx.workloads.sorted_uniform_points(12000, 7), 96, 24, "cauchykernel",
                       O.CAUCHY),
    "log": lambda: (*hmx.workloads.sorted_uniform_points(16384, 3), 128, 32, "logkernel", O.LOG),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("mode", ["stored", "hybrid", "free"])
def test_nested_mfma_vs_vector_units_and_oracle(cuda, case, mode):
    x, y, iv, bs, r, kern, kid = CASES[case]()
    print(f"case: nested {case} {mode}", flush=True)
    H = hmx.KernelMatrix(kern, x, y, *iv, blocksize=bs, rank=r, hybrid=mode == "hybrid",
                         matrix_free=mode == "free")
    assert H.info["nested"] == 1
    n = len(x)
    B = hmx.workloads.normal(11, n * 32).reshape(n, 32)
    Bt = torch.from_numpy(B).to(cuda)
    Ym = H.matvec(Bt, mfma=True).cpu().numpy()
    Yv = H.matvec(Bt, mfma=False).cpu().numpy()
    assert rel(Ym, Yv) <= 1e-13, rel(Ym, Yv)
    t = O.Tree(kid, x, y, *iv, bs, r)
    for q in (0, 17, 31):
        ref = t.matvec(np.ascontiguousarray(B[:, q]))
        assert rel(Ym[:, q], ref) <= 1e-12, (q, rel(Ym[:, q], ref))
    # 16 right-hand sides: one group
    Y16 = H.matvec(Bt[:, :16].contiguous(), mfma=True).cpu().numpy()
    assert rel(Y16, Yv[:, :16]) <= 1e-13


def test_nested_batched(cuda):
    w = hmx.workloads.config_batched(count=10, nmin_log2=10, nmax_log2=13, k=32)
    inst = w["instances"]
    H = hmx.KernelMatrixBatch("cauchykernel", [(i["x"], i["y"], i["intervals"]) for i in inst],
                              blocksize=64, rank=16, hybrid=True)
    assert H.info["nested"] == 1
    n = H.shape[1]
    B = hmx.workloads.normal(12, n * 32).reshape(n, 32)
    Bt = torch.from_numpy(B).to(cuda)
    Ym = H.matvec(Bt, mfma=True).cpu().numpy()
    Yv = H.matvec(Bt, mfma=False).cpu().numpy()
    assert rel(Ym, Yv) <= 1e-13
    H.close()


def test_plan_from_blocks_is_not_nested(cuda):
    rng = np.random.default_rng(0)
    L = hmx.LowRankMatrix(rng.standard_normal((40, 3)), rng.standard_normal(3),
                          rng.standard_normal((30, 3)))
    P = hmx.blocks_plan(70, 60, [(np.asfortranarray(rng.standard_normal((30, 60))), 0, 0),
                                 (L, 30, 30)])
    assert P.info["nested"] == 0
