"""tune_blocksize: the leaf threshold and plan form chosen by building and
timing a plan per candidate.

- Every candidate is reported with its time and its error on sampled rows
  against the dense FP64 product; the chosen one is the fastest within the
  tolerance.
- The error reported for a candidate is that of a plan built with its
  parameters (checked again here on the same rows), and every candidate of a
  rank-16 Cauchy plan is within 1e-11 of the dense product.
- A batch of uneven instances and a single matrix."""
import numpy as np
import pytest
import torch

import hmx

pytestmark = pytest.mark.gpu


def small_batch(count=6, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for t in range(count):
        n = int(2 ** rng.uniform(10, 12.5))
        x, y, iv = hmx.workloads.sorted_uniform_points(n, 100 + t)
        out.append((x, y, iv))
    return out


@pytest.mark.parametrize("batched", [True, False])
def test_tune_blocksize(cuda, batched):
    spec = small_batch() if batched else [hmx.workloads.shifted_grid(1 << 13)]
    arg = spec if batched else spec[0]
    best, res = hmx.tune_blocksize("cauchykernel", arg, 16, k=32, candidates=(32, 64, 128),
                                   forms=("hybrid", "matrix_free"), reps=2, check_rows=48,
                                   tol=1e-11)
    print(f"tune_blocksize batched={batched}: " + ", ".join(
        f"{d['blocksize']}/{d['form']} {d['ms']:.3f} ms {d['rel_err']:.1e}" for d in res),
        flush=True)
    assert len(res) == 6
    assert all(d["rel_err"] <= 1e-11 for d in res)
    assert [d["ms"] for d in res] == sorted(d["ms"] for d in res)
    assert best is res[0]
    # the chosen parameters rebuild a plan of the same accuracy
    kw = dict(blocksize=best["blocksize"], rank=16, hybrid=best["form"] == "hybrid",
              matrix_free=best["form"] == "matrix_free")
    H = hmx.KernelMatrixBatch("cauchykernel", spec, **kw) if batched else \
        hmx.KernelMatrix("cauchykernel", spec[0][0], spec[0][1], *spec[0][2], **kw)
    x, y, _ = spec[0]
    rng = np.random.default_rng(9)
    B = torch.from_numpy(rng.standard_normal((H.shape[1], 4))).to(cuda)
    Y = H.matvec(B)
    rows = np.arange(0, len(x), 37)
    xt, yt = torch.from_numpy(x).to(cuda), torch.from_numpy(y).to(cuda)
    K = 1.0 / (xt[rows][:, None] - yt[None, :])
    want = K @ B[:len(y)]
    e = float(torch.linalg.norm(Y[rows] - want) / torch.linalg.norm(want))
    assert e <= 1e-11
    H.close()
