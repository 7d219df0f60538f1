"""The reduced-precision and Float32 legs at the full sizes of their workloads.

1. Batched (256 Cauchy instances, n log-uniform in [2^10, 2^16], rank 16,
   leaf 64, 32 right-hand sides, hybrid plan): the far field in FP32, split
   BF16 hi+lo and split BF16 hi+mid+lo, against the oracle's mul! on sampled
   rows of several instances and against the FP64 product of the whole
   batch. Bound <= 1e-6 relative 2-norm for every form.
2. Scale (log|x - y|, n = 2^22 sorted uniform points, y = x + 0.5/n, rank 32,
   leaf 512, 8 right-hand sides; hybrid and matrix-free plans): FP64 MFMA,
   FP32, split BF16 hi+mid+lo and hi+lo far fields against 4096 sampled dense
   FP64 rows (computed on the device by plain tiles of K times B, the tiles'
   log|x - y| formed in FP64), the dense rows pinned on 64 rows against the
   oracle's long-double dense product (to 1e-12). Bounds: FP64 <= 1e-12; FP32 and
   BF16 hi+mid+lo <= 1e-6; BF16 hi+lo <= 2e-5 (operands carry 16 bits).
3. Differentiation in Float32 at n = 2^16 (Float32 points, b, t; Float32
   plans): against the dense FP64 gradient of the same Float32 inputs
   (<= 1e-4 relative 2-norm) and against central differences of the FP64
   loss at the same points (h = 1e-6 of the point spacing)."""
import numpy as np
import pytest
import torch

import hmx
import oracle as O
from hmx.autodiff import kernel_matvec

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ---------------------------------------------------------------- batched

BATCH_BOUND = {"f32": 1e-6, "bf16x3": 1e-6, "bf16x6": 1e-6}


@pytest.fixture(scope="module")
def batched(cuda):
    w = hmx.workloads.config_batched()
    inst = w["instances"]
    H = hmx.KernelMatrixBatch("cauchykernel", [(i["x"], i["y"], i["intervals"]) for i in inst],
                              blocksize=64, rank=16, hybrid=True)
    B = torch.from_numpy(np.concatenate([i["b"] for i in inst])).to(cuda)
    ref = H @ B  # FP64 (MFMA passes at 32 right-hand sides)
    return inst, H, B, ref


@pytest.mark.parametrize("prec", ["f32", "bf16x3", "bf16x6"])
def test_batched_full_reduced_precision(cuda, batched, prec):
    inst, H, B, ref = batched
    print(f"case: batched 256 instances far={prec}", flush=True)
    Y = H.matvec(B, far_precision=prec)
    e_all = float(torch.linalg.norm(Y - ref) / torch.linalg.norm(ref))
    Yc = Y.cpu().numpy()
    ro = H.row_offsets
    rng = np.random.default_rng(12)
    worst = 0.0
    for t in [0, 31, 128, 255] + list(rng.integers(0, 256, 4)):
        i = inst[t]
        tr = O.Tree(O.CAUCHY, i["x"], i["y"], *i["intervals"], 64, 16, store=False)
        rows = np.unique(rng.integers(0, len(i["x"]), 32))
        for q in (0, 15, 31):
            want = tr.matvec_rows(np.ascontiguousarray(i["b"][:, q]), rows)
            worst = max(worst, rel(Yc[ro[t] + rows, q], want))
    print(f"batched far={prec}: vs FP64 product {e_all:.3e}, vs oracle rows {worst:.3e}",
          flush=True)
    assert e_all <= BATCH_BOUND[prec] and worst <= BATCH_BOUND[prec]


# ------------------------------------------------------------------ scale

SCALE_BOUND = {"f64": 1e-12, "f32": 1e-6, "bf16x6": 1e-6, "bf16x3": 2e-5}


def dense_rows_log(x, y, B, rows, chunk=1 << 16):
    """sum_j log|x_i - y_j| B[j, :] for the sampled rows, FP64 on the device
    in column tiles."""
    xr = x[rows]
    out = torch.zeros(len(rows), B.shape[1], dtype=torch.float64, device=x.device)
    for j0 in range(0, y.numel(), chunk):
        K = torch.log(torch.abs(xr[:, None] - y[None, j0:j0 + chunk]))
        out += K @ B[j0:j0 + chunk]
    return out


@pytest.fixture(scope="module")
def scale(cuda):
    w = hmx.workloads.config_scale()
    n, k = len(w["x"]), w["k"]
    x = torch.from_numpy(w["x"]).to(cuda)
    y = torch.from_numpy(w["y"]).to(cuda)
    B = torch.from_numpy(w["b"]).to(cuda)
    rows = np.sort(np.random.default_rng(41).choice(n, 4096, replace=False))
    ref = dense_rows_log(x, y, B, torch.from_numpy(rows).to(cuda))
    # pin the dense rows on 64 of them against the oracle (long double)
    sub = np.arange(0, 4096, 64)
    for q in (0, k - 1):
        want = O.dense_matvec(O.LOG, w["x"], w["y"], np.ascontiguousarray(w["b"][:, q]),
                              rows=rows[sub])
        got = ref[sub, q].cpu().numpy()
        assert rel(got, want) <= 1e-12, rel(got, want)
    return w, B, rows, ref


@pytest.mark.parametrize("mode", ["hybrid", "free"])
def test_scale_log_kernel_n2p22(cuda, scale, mode):
    w, B, rows, ref = scale
    print(f"case: scale n=2^22 log kernel mode={mode}", flush=True)
    H = hmx.KernelMatrix("logkernel", w["x"], w["y"], *w["intervals"], blocksize=w["blocksize"],
                         rank=w["rank"], hybrid=mode == "hybrid", matrix_free=mode == "free")
    rt = torch.from_numpy(rows).to(B.device)
    errs = {}
    for prec in ("f64", "f32", "bf16x6", "bf16x3"):
        Y = H.matvec(B, far_precision=prec)
        assert bool(torch.all(torch.isfinite(Y)))
        errs[prec] = float(torch.linalg.norm(Y[rt] - ref) / torch.linalg.norm(ref))
        del Y
    print(f"scale {mode}: rel. errors vs dense FP64 rows " +
          ", ".join(f"{p} {e:.3e}" for p, e in errs.items()), flush=True)
    for p, e in errs.items():
        assert e <= SCALE_BOUND[p], (p, e)
    H.close()


# ------------------------------------------------------ Float32 gradients

def test_f32_gradient_n2p16_dense_and_fd(cuda):
    n = 1 << 16
    w = hmx.workloads.config_gradient(n=n)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    x, y, b, t = (T(w[k]).float() for k in ("x", "y", "b", "t"))
    print("case: f32 gradient n=2^16", flush=True)
    xg, yg, bg = (v.clone().requires_grad_(True) for v in (x, y, b))
    u = kernel_matvec("cauchykernel", xg, yg, bg, w["intervals"], 64, 16)
    assert u.dtype == torch.float32
    L = 0.5 * torch.sum((u.double() - t.double()) ** 2)
    L.backward()
    gx, gy, gb = xg.grad.double(), yg.grad.double(), bg.grad.double()
    # dense FP64 gradient of the same (Float32) inputs, row tiles
    xd, yd, bd, td = (v.double() for v in (x, y, b, t))
    Kb = torch.zeros(n, dtype=torch.float64, device=cuda)
    Kxb = torch.zeros_like(Kb)
    for i0 in range(0, n, 4096):
        d = xd[i0:i0 + 4096, None] - yd[None, :]
        Kb[i0:i0 + 4096] = (1.0 / d) @ bd
        Kxb[i0:i0 + 4096] = (-1.0 / (d * d)) @ bd
    r = Kb - td
    KTr = torch.zeros(n, dtype=torch.float64, device=cuda)
    KyTr = torch.zeros_like(KTr)
    for i0 in range(0, n, 4096):
        d = xd[i0:i0 + 4096, None] - yd[None, :]
        KTr += (1.0 / d).T @ r[i0:i0 + 4096]
        KyTr += (1.0 / (d * d)).T @ r[i0:i0 + 4096]
    errs = dict(u=rel(u.detach().double().cpu(), Kb.cpu()), b=rel(gb.cpu(), KTr.cpu()),
                x=rel(gx.cpu(), (r * Kxb).cpu()), y=rel(gy.cpu(), (bd * KyTr).cpu()))
    print("f32 n=2^16 gradient relative errors vs dense FP64", errs, flush=True)
    assert all(v <= 1e-4 for v in errs.values()), errs

    # central differences of the FP64 loss at the Float32 points (exact in FP64)
    iv = w["intervals"]

    def prod(xx, yy, bb):
        H = hmx.KernelMatrix("cauchykernel", xx.cpu().numpy(), yy.cpu().numpy(), *iv,
                             blocksize=64, rank=16, matrix_free=True)
        return H @ bb

    def dL(up, um):
        return float(torch.sum(0.5 * (up - um) * (up + um - 2 * td)))

    h = 1e-6 / n
    rng = np.random.default_rng(8)
    scale_x = float(torch.sqrt(torch.mean(gx * gx)))
    scale_y = float(torch.sqrt(torch.mean(gy * gy)))
    worst = 0.0
    for i in rng.integers(0, n, 4):
        e = torch.zeros(n, dtype=torch.float64, device=cuda)
        e[i] = h
        fd = dL(prod(xd + e, yd, bd), prod(xd - e, yd, bd)) / (2 * h)
        worst = max(worst, abs(fd - float(gx[i])) / max(abs(fd), scale_x))
        fd = dL(prod(xd, yd + e, bd), prod(xd, yd - e, bd)) / (2 * h)
        worst = max(worst, abs(fd - float(gy[i])) / max(abs(fd), scale_y))
    print(f"f32 n=2^16 gradient vs central differences of the FP64 loss: {worst:.3e}", flush=True)
    assert worst <= 1e-4
