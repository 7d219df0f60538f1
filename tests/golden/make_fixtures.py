"""Writes the golden fixtures of tests/golden/ (run from the repository root:
python tests/golden/make_fixtures.py).

The reference (Julia) is not runnable in this environment and its tests hold no
vectors for the KernelMatrix path, so these fixtures come from the oracle's
restatement of the reference mul! (oracle/hmx_oracle.c, traversal order) and
from the dense long-double product, on inputs fixed by the documented RNG. They
pin the oracle against accidental change and give the GPU tests fixed data.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "hierarchicalmatrices.jl_amd")]
import oracle as O  # noqa: E402
from hmx import workloads as W  # noqa: E402

CASES = [
    ("cauchy_n64", O.CAUCHY, 64, 16, 4),
    ("cauchy_n256", O.CAUCHY, 256, 16, 8),
    ("cauchy_n4096", O.CAUCHY, 4096, 64, 16),
    ("log_n4096", O.LOG, 4096, 64, 16),
    ("coulomb_n1024", O.COULOMB, 1024, 64, 16),
]


def make(name, kid, n, bs, r):
    x, y, iv = W.shifted_grid(n)
    b = W.normal(0, n)
    t = O.Tree(kid, x, y, *iv, bs, r)
    u = t.matvec(b)
    d = O.dense_matvec(kid, x, y, b)
    return dict(x=x, y=y, b=b, u=u, dense=d, intervals=np.array(iv), kernel=np.array(kid),
                blocksize=np.array(bs), rank=np.array(r))


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    for name, kid, n, bs, r in CASES:
        np.savez(os.path.join(here, name + ".npz"), **make(name, kid, n, bs, r))
        print("wrote", name)
