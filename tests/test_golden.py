"""Golden fixtures (tests/golden/, written by tests/golden/make_fixtures.py from
the oracle's traversal mul! and the dense long-double product)."""
import glob
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = sorted(glob.glob(os.path.join(HERE, "*.npz")))
KNAMES = {O.CAUCHY: "cauchykernel", O.COULOMB: "coulombkernel", O.COULOMB3: "coulombprimekernel",
          O.LOG: "logkernel"}


def load(f):
    z = np.load(f, allow_pickle=False)
    return {k: z[k] for k in z.files}


def test_fixtures_present():
    assert len(FILES) >= 5


@pytest.mark.parametrize("f", FILES, ids=os.path.basename)
def test_oracle_reproduces_fixture(f):
    g = load(f)
    t = O.Tree(int(g["kernel"]), g["x"], g["y"], *g["intervals"], int(g["blocksize"]),
               int(g["rank"]))
    assert np.array_equal(t.matvec(g["b"]), g["u"])
    # interpolation error of rank r on admissible blocks: O(rho^-r), rho = 3 + sqrt 8
    bound = 10.0 * (3.0 + np.sqrt(8.0)) ** -float(g["rank"]) + 1e-14
    assert np.linalg.norm(g["u"] - g["dense"]) <= bound * np.linalg.norm(g["dense"])


@pytest.mark.gpu
@pytest.mark.parametrize("f", FILES, ids=os.path.basename)
def test_gpu_matches_fixture(cuda, f):
    import torch
    import hmx
    g = load(f)
    for mode in ("stored", "hybrid", "free"):
        H = hmx.KernelMatrix(KNAMES[int(g["kernel"])], g["x"], g["y"], *g["intervals"],
                             blocksize=int(g["blocksize"]), rank=int(g["rank"]),
                             matrix_free=mode == "free", hybrid=mode == "hybrid")
        u = (H @ torch.from_numpy(g["b"]).to(cuda)).cpu().numpy()
        assert np.linalg.norm(u - g["u"]) <= 1e-12 * np.linalg.norm(g["u"])
