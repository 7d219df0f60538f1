"""Leaf-size autotuning for KernelMatrix and KernelMatrixBatch products.

hm_plan_tune picks the launch shape of one plan (right-hand sides per launch,
vector units or matrix cores, far-field form). The leaf threshold
(`blocksize`, src/KernelMatrix.jl:47-116: a block is dense below it) and the
plan's form (stored, hybrid, matrix-free) are fixed when a plan is built, and
they change the matrix's near/far split, so they are chosen here by building
a plan per candidate and measuring it:

- every candidate (blocksize, form) plan runs hm_plan_tune for k right-hand
  sides, then `reps` timed products with its tuned launch shape (HIP events
  on the current stream);
- its product is compared on `check_rows` sampled rows of up to
  `check_instances` instances with the dense FP64 product (the named kernel
  on the device);
- the fastest candidate whose relative error is within `tol` wins.

A different leaf threshold is a different H-matrix (the reference's tree
for that blocksize): the results report the error so that the caller
chooses with it, and nothing here changes a plan the caller already holds.
"""
import numpy as np

from .kernelmatrix import KERNELS, KernelMatrix, KernelMatrixBatch

# the named kernels of examples/Kernel.jl:34-37 on device differences d = x - y
_TORCH_F = {0: lambda d: 1.0 / d, 1: lambda d: 1.0 / (d * d), 2: lambda d: 1.0 / (d * d * d),
            3: lambda d: d.abs().log()}


def _dense_rows(kid, x, y, B, rows, chunk=1 << 15):
    import torch
    f = _TORCH_F[kid]
    xr = x[rows]
    out = torch.zeros(len(rows), B.shape[1], dtype=torch.float64, device=B.device)
    for j0 in range(0, y.numel(), chunk):
        out += f(xr[:, None] - y[None, j0:j0 + chunk]) @ B[j0:j0 + chunk]
    return out


def tune_blocksize(kernel, spec, rank, *, k=1, candidates=(32, 64, 128, 256),
                   forms=("stored", "hybrid", "matrix_free"), reps=5, check_rows=64,
                   check_instances=4, tol=1e-10, seed=0):
    """spec: (x, y, intervals) for one KernelMatrix, or a sequence of them for
    a KernelMatrixBatch. Returns (best, results): results is a list of dicts
    (blocksize, form, rhs_per_launch, launch_mode, ms, rel_err), fastest
    first; best is the fastest entry with rel_err <= tol (None if none)."""
    import torch
    name = kernel if isinstance(kernel, str) else kernel.name
    kid = KERNELS[name].kid
    # a batch is a sequence of (x, y, intervals); one matrix is (x, y, intervals)
    batch = isinstance(spec, (list, tuple)) and len(spec) > 0 and isinstance(spec[0], (list, tuple))
    insts = list(spec) if batch else [spec]
    dev = torch.device("cuda")
    n_cols = sum(len(i[1]) for i in insts)
    rng = np.random.default_rng(seed)
    B = torch.from_numpy(rng.standard_normal((n_cols, k))).to(dev)
    # reference rows (dense FP64 on the device) of the first instances
    checks = []
    r0 = c0 = 0
    for t, (x, y, _) in enumerate(insts):
        if t < check_instances:
            rows = np.sort(rng.choice(len(x), min(check_rows, len(x)), replace=False))
            xt, yt = torch.from_numpy(np.asarray(x, np.float64)).to(dev), \
                torch.from_numpy(np.asarray(y, np.float64)).to(dev)
            want = _dense_rows(kid, xt, yt, B[c0:c0 + len(y)], torch.from_numpy(rows).to(dev))
            checks.append((r0 + rows, want))
        r0 += len(x)
        c0 += len(y)
    stream = torch.cuda.current_stream()
    results = []
    for bs in candidates:
        for form in forms:
            kw = dict(blocksize=bs, rank=rank, hybrid=form == "hybrid",
                      matrix_free=form == "matrix_free")
            H = KernelMatrixBatch(name, insts, **kw) if batch else \
                KernelMatrix(name, insts[0][0], insts[0][1], *insts[0][2], **kw)
            kb, mode, _ = H.tune(k=k)
            Y = H.matvec(B, tuned=True)  # warm-up
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            t0.record(stream)
            for _ in range(reps):
                Y = H.matvec(B, tuned=True)
            t1.record(stream)
            torch.cuda.synchronize()
            ms = t0.elapsed_time(t1) / reps
            num = den = 0.0
            for rows, want in checks:
                got = Y[torch.from_numpy(rows).to(dev)]
                num += float(torch.linalg.norm(got - want)) ** 2
                den += float(torch.linalg.norm(want)) ** 2
            results.append(dict(blocksize=bs, form=form, rhs_per_launch=int(kb),
                                launch_mode=int(mode), ms=ms, rel_err=(num / den) ** 0.5))
            H.close()
            del Y
    results.sort(key=lambda d: d["ms"])
    best = next((d for d in results if d["rel_err"] <= tol), None)
    return best, results
