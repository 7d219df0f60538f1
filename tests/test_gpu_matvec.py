"""GPU parity of the KernelMatrix compression and product against the oracle.

The oracle (oracle/hmx_oracle.c) restates src/KernelMatrix.jl, src/BarycentricMatrix.jl
and src/algebra.jl loop for loop; its traversal mul! is "the reference's mul!"
here. Bars: stored factors bit-exact (IEEE kernels); products within 1e-12
relative 2-norm of the oracle in FP64.
"""
import numpy as np
import pytest
import torch

import hmx
import oracle as O

pytestmark = pytest.mark.gpu

KID = {"cauchykernel": O.CAUCHY, "coulombkernel": O.COULOMB,
       "coulombprimekernel": O.COULOMB3, "logkernel": O.LOG}
TOL = 1e-12
MODES = ["stored", "hybrid", "free"]


def plan(kernel, w, bs, r, mode, x=None, y=None):
    x = w["x"] if x is None else x
    y = w["y"] if y is None else y
    return hmx.KernelMatrix(kernel, x, y, *w["intervals"], blocksize=bs, rank=r,
                            matrix_free=mode == "free", hybrid=mode == "hybrid")


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def plain(n=4096, bs=64, r=16):
    return hmx.workloads.config_plain(n=n, blocksize=bs, rank=r)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("kernel", list(KID))
def test_plain_vs_oracle(cuda, kernel, mode):
    w = plain()
    print(f"case: plain n=4096 kernel={kernel} mode={mode}", flush=True)
    ref = O.Tree(KID[kernel], w["x"], w["y"], *w["intervals"], 64, 16).matvec(w["b"])
    H = plan(kernel, w, 64, 16, mode)
    assert H.info["nonfinite"] == 0
    b = torch.from_numpy(w["b"]).to(cuda)
    u = (H @ b).cpu().numpy()
    assert rel(u, ref) <= TOL
    if mode == "stored":  # Chebyshev-form basis on the stored plan
        assert rel(H.matvec(b, cheb_basis=True).cpu().numpy(), ref) <= TOL


@pytest.mark.parametrize("kernel", ["cauchykernel", "coulombkernel", "coulombprimekernel"])
def test_compression_bit_exact(cuda, kernel):
    """Entries of the stored plan (dense blocks and U F V' at far blocks,
    src/BarycentricMatrix.jl:222-234) equal the oracle's bit for bit."""
    w = plain(n=2048)
    t = O.Tree(KID[kernel], w["x"], w["y"], *w["intervals"], 64, 16)
    H = hmx.KernelMatrix(kernel, w["x"], w["y"], *w["intervals"], blocksize=64, rank=16)
    rng = np.random.default_rng(1)
    ii = rng.integers(0, 2048, 3000)
    jj = rng.integers(0, 2048, 3000)
    got = H[ii, jj]
    want = np.array([t.getindex(i, j) for i, j in zip(ii, jj)])
    assert np.array_equal(got, want)


def test_log_entries_close(cuda):
    w = plain(n=2048)
    t = O.Tree(O.LOG, w["x"], w["y"], *w["intervals"], 64, 16)
    H = hmx.KernelMatrix("logkernel", w["x"], w["y"], *w["intervals"], blocksize=64, rank=16)
    ii = np.arange(0, 2048, 7)
    jj = (ii * 13) % 2048
    got = H[ii, jj]
    want = np.array([t.getindex(i, j) for i, j in zip(ii, jj)])
    assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-300)) <= 4e-15


@pytest.fixture(scope="module")
def headline_ref():
    """The oracle's traversal mul! of the whole n = 2^20 product (leaf 256,
    rank 24): a stored-factor oracle tree (about 17 GB of host memory, built
    in about 15 s), its OpenMP product, and the tree freed."""
    w = hmx.workloads.config_headline()
    t = O.Tree(O.CAUCHY, w["x"], w["y"], *w["intervals"], 256, 24)
    ref = t.matvec_omp(w["b"])
    del t
    return w, ref


@pytest.mark.parametrize("mode", MODES)
def test_headline_size_full_vector(cuda, headline_ref, mode):
    """n = 2^20, leaf 256, rank 24: all 2^20 entries against the oracle's
    traversal mul! (relative 2-norm and entry by entry), and sampled rows
    against the dense product."""
    w, ref = headline_ref
    print(f"case: headline n=2^20 mode={mode}", flush=True)
    H = plan("cauchykernel", w, 256, 24, mode)
    u = (H @ torch.from_numpy(w["b"]).to(cuda)).cpu().numpy()
    H.close()
    assert np.all(np.isfinite(u))
    e = rel(u, ref)
    worst = float(np.max(np.abs(u - ref)) / np.max(np.abs(ref)))
    print(f"headline {mode}: rel. 2-norm {e:.3e}, max |diff| / max |ref| {worst:.3e}", flush=True)
    assert e <= TOL and worst <= TOL
    rng = np.random.default_rng(2)
    rows = np.unique(np.concatenate([rng.integers(0, 1 << 20, 32), [0, 1, 127, 128, (1 << 20) - 1]]))
    dense = O.dense_matvec(O.CAUCHY, w["x"], w["y"], w["b"], rows=rows)
    assert rel(u[rows], dense) <= 1e-11


def test_stored_equals_free_full_vector(cuda):
    w = hmx.workloads.config_plain(n=1 << 16, blocksize=256, rank=24)
    b = torch.from_numpy(w["b"]).to(cuda)
    H = hmx.KernelMatrix("cauchykernel", w["x"], w["y"], *w["intervals"], blocksize=256, rank=24)
    u1 = H @ b
    u2 = H.matvec(b, matrix_free=True)
    assert rel(u2.cpu().numpy(), u1.cpu().numpy()) <= 1e-14
    u4 = H.matvec(b, cheb_basis=True)
    assert rel(u4.cpu().numpy(), u1.cpu().numpy()) <= 1e-14
    # linearity
    b2 = torch.from_numpy(hmx.workloads.normal(5, 1 << 16)).to(cuda)
    u3 = H @ (2.5 * b + b2)
    assert rel(u3.cpu().numpy(), (2.5 * u1 + H @ b2).cpu().numpy()) <= 1e-13


def test_accumulate_offsets_strides(cuda):
    """mul!(u, H, v, istart, jstart, INCX, INCY) accumulates (src/KernelMatrix.jl:14,
    src/HierarchicalMatrix.jl:24-52); offsets 1-based."""
    w = plain(n=1024)
    n = 1024
    t = O.Tree(O.CAUCHY, w["x"], w["y"], *w["intervals"], 64, 16)
    H = hmx.KernelMatrix("cauchykernel", w["x"], w["y"], *w["intervals"], blocksize=64, rank=16)
    v = hmx.workloads.normal(3, 2 * n + 10)
    u0 = hmx.workloads.normal(4, 3 * n + 10)
    istart, jstart, incx, incy = 5, 3, 2, 3
    vs = v[jstart - 1: jstart - 1 + incx * n: incx]
    want = u0.copy()
    want[istart - 1: istart - 1 + incy * n: incy] += t.matvec(vs)
    ud = torch.from_numpy(u0.copy()).to(cuda)
    hmx.mul_(ud, H, torch.from_numpy(v).to(cuda), istart, jstart, incx, incy)
    got = ud.cpu().numpy()
    mask = np.zeros(len(u0), bool)
    mask[istart - 1: istart - 1 + incy * n: incy] = True
    assert np.array_equal(got[~mask], u0[~mask])
    assert rel(got[mask], want[mask]) <= TOL
    # host arrays are updated in place
    uh = u0.copy()
    hmx.mul_(uh, H, v, istart, jstart, incx, incy)
    assert rel(uh[mask], want[mask]) <= TOL


def test_multiple_rhs_columns(cuda):
    w = plain(n=2048)
    H = hmx.KernelMatrix("cauchykernel", w["x"], w["y"], *w["intervals"], blocksize=64, rank=16)
    B = torch.from_numpy(hmx.workloads.normal(9, 2048 * 5).reshape(2048, 5)).to(cuda)
    Y = H @ B
    for q in range(5):  # several right-hand sides per launch: same values up to rounding
        assert rel(Y[:, q].cpu().numpy(), (H @ B[:, q].contiguous()).cpu().numpy()) <= 1e-14


def test_transpose(cuda):
    w = plain(n=4096)
    x, y = w["x"], w["y"]
    H = hmx.KernelMatrix("cauchykernel", x, y, *w["intervals"], blocksize=64, rank=16)
    r = torch.from_numpy(hmx.workloads.normal(21, 4096)).to(cuda)
    b = torch.from_numpy(w["b"]).to(cuda)
    for free, cheb in ((False, False), (False, True), (True, False)):
        g = H.matvec(r, transpose=True, matrix_free=free, cheb_basis=cheb)
        # adjoint identity <H^T r, b> = <r, H b>
        lhs = float(torch.dot(g, b))
        rhs = float(torch.dot(r, H @ b))
        assert abs(lhs - rhs) <= 1e-13 * float(torch.linalg.norm(g) * torch.linalg.norm(b))
        # Cauchy on equal intervals: H(x, y)^T = -H(y, x) block for block
        Hs = hmx.KernelMatrix("cauchykernel", y, x, *w["intervals"], blocksize=64, rank=16)
        want = -(Hs @ r)
        assert rel(g.cpu().numpy(), want.cpu().numpy()) <= 1e-14
        # and the dense transpose within the interpolation error
        K = 1.0 / (x[:, None] - y[None, :])
        assert rel(g.cpu().numpy(), K.T @ r.cpu().numpy()) <= 1e-11


@pytest.mark.parametrize("family", ["chebyshev", "quadratic"])
@pytest.mark.parametrize("kernel", list(KID))
@pytest.mark.parametrize("N", [1000, 10000])
def test_example_families(cuda, family, kernel, N):
    """The two point families of examples/Kernel.jl:58-110 with the default
    BLOCKRANK(Float64) = 20 and BLOCKSIZE = 80."""
    if family == "chebyshev":
        x = hmx.chebyshevpoints(N)
        y = hmx.chebyshevpoints(N, kind=2)
        iv = (1.0, -1.0, 1.0, -1.0)
    else:
        i = np.arange(N, 0, -1, dtype=np.float64)
        x = i * (i + 1)
        y = (i + 0.5) * (i + 1.5)
        iv = (x.max(), x.min(), y.max(), y.min())
    b = hmx.workloads.normal(N + KID[kernel], N)
    t = O.Tree(KID[kernel], x, y, *iv, 80, 20)
    ref = t.matvec(b)
    for mode in MODES:
        H = hmx.KernelMatrix(kernel, x, y, *iv, matrix_free=mode == "free",
                             hybrid=mode == "hybrid")
        assert H.rank == 20 and H.blocksize_ == 80
        u = H.matvec(b)
        assert rel(u, ref) <= TOL, (mode, rel(u, ref))


@pytest.mark.parametrize("mode", MODES)
def test_side_stream_near_field_matches_one_stream(cuda, mode):
    """Hybrid and matrix-free plans with at least 2^24 near-field entries run the
    near field on a side stream beside the far-field chain and add the far field
    afterwards (stored plans stay on one stream; their case checks that). Pass
    profiling keeps a product on one stream, so the same plan gives both forms:
    forward, transpose, accumulate (beta = 1) and four right-hand sides per
    launch on the vector units, on a stream other than the default, over
    repeated products. 20 right-hand sides on the MFMA passes stay on one
    stream (both halves would compete for the matrix cores) and must agree."""
    import ctypes
    from hmx._lib import lib
    w = hmx.workloads.config_plain(n=1 << 16, blocksize=256, rank=24)
    print(f"case: side-stream split n=2^16 leaf 256 mode={mode}", flush=True)
    H = plan("cauchykernel", w, 256, 24, mode)
    assert H.info["dense_entries"] >= 1 << 24
    b = torch.from_numpy(w["b"]).to(cuda)
    B4 = torch.from_numpy(np.random.default_rng(4).standard_normal((1 << 16, 4))).to(cuda)
    # column-major (20 columns of n): the MFMA passes pack each group of rows
    B20 = torch.from_numpy(np.random.default_rng(6).standard_normal((20, 1 << 16))).to(cuda).t()
    u0 = torch.from_numpy(hmx.workloads.normal(9, 1 << 16)).to(cuda)
    s = torch.cuda.Stream()

    def products():
        with torch.cuda.stream(s):
            out = [H.matvec(b, stream=s), H.matvec(b, transpose=True, stream=s),
                   H.mul(u0.clone(), b, stream=s), H.matvec(B4, mfma=False, stream=s),
                   H.matvec(B20, mfma=True, stream=s)]
            out += [H.matvec(b, stream=s) for _ in range(3)]
        s.synchronize()
        return [o.cpu().numpy() for o in out]

    split = products()
    assert lib.hm_plan_profile(H._h, 1) == 0
    one = products()
    assert lib.hm_plan_profile(H._h, 0) == 0
    H.close()
    for a, e in zip(split, one):
        assert np.all(np.isfinite(a))
        assert rel(a, e) <= 1e-14
    # repeated products on the side stream are identical
    for q in (6, 7):
        assert np.array_equal(split[q], split[5])


def test_side_stream_plan_captured_on_one_stream(cuda):
    """A matrix-free plan whose products split onto its side stream, captured
    into a graph after an eager product: the captured product stays on one
    stream, so every replay equals the one-stream eager product bit for bit
    (pass profiling keeps an eager product on one stream)."""
    from hmx._lib import lib
    w = hmx.workloads.config_plain(n=1 << 16, blocksize=256, rank=24)
    print("case: side-stream plan under graph capture, n=2^16 leaf 256 matrix-free", flush=True)
    H = plan("cauchykernel", w, 256, 24, "free")
    assert H.info["dense_entries"] >= 1 << 24
    b = torch.from_numpy(w["b"]).to(cuda)
    split = H @ b  # makes the side stream
    out = torch.zeros_like(b)
    g = H.graph(out, b)
    for rep in range(3):
        b.mul_(1.25).add_(0.5)  # eager work between replays
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        assert lib.hm_plan_profile(H._h, 1) == 0
        one = H @ b
        torch.cuda.synchronize()
        assert lib.hm_plan_profile(H._h, 0) == 0
        assert torch.equal(out, one), rep
        assert rel((H @ b).cpu().numpy(), one.cpu().numpy()) <= 1e-14
    assert torch.all(torch.isfinite(split))
    g.close()
    H.close()
