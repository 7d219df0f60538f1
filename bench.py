"""Headline benchmark: Cauchy KernelMatrix product at n = 2^20 (FP64).

Workload (one "step" = one product y = H b over the whole matrix):
  x_i = (i - 1/4)/n, y_j = (j - 3/4)/n (non-increasing), K_ij = 1/(x_i - y_j),
  leaf threshold 256, Chebyshev rank 24, one right-hand side b ~ N(0, 1)
  (splitmix64 seed 0). The plan (tree + compressed factors) is built and
  resident in HBM before timing; each timed step is one hm_matvec.

Prints one JSON line. `value` is matvecs/s of the stored-factor product; the
matrix-free variant is reported beside it. `roofline` describes the dominant
kernel (the downward pass) from HIP events recorded around each pass;
`cpu_baseline` times the oracle's restatement of the reference mul! on the
host (bounded sample, see its `sample` field).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hierarchicalmatrices.jl_amd"))

N = 1 << 20
BS = 256
RANK = 24
HBM_PEAK_GBS = 8000.0
# FP64 vector issue: 1024 SIMDs x 2.4 GHz, one wave64 FP64 instruction per SIMD per 4 cycles
VALU_F64_PEAK_GINSTR = 614.4


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU leg threads (default: the affinity set, capped by the cgroup quota)")
    ap.add_argument("--cpu-n-log2", type=int, default=None,
                    help="CPU leg instance size (default: n=2^20 when 24 GB of RAM are free)")
    ap.add_argument("--variants", default=None,
                    help="headline: extra variants timed beside the stored product (default "
                         "hybrid,matrix_free,f32_hybrid; f32_stored on request); batched: plan "
                         "modes to time (default stored,hybrid,matrix_free)")
    ap.add_argument("--no-workloads", action="store_true",
                    help="headline: skip the batched / gradient / scale summaries")
    ap.add_argument("--no-peaks", action="store_true", help="skip the torch bandwidth probes")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run torch copies of 2^27 doubles (for PMC byte calibration)")
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "batched", "gradient", "scale", "even", "lowrank",
                             "cholesky"],
                    help="headline = the n=2^20 product (the bench line); the others time "
                         "the batched, gradient, scale, EvenBarycentricMatrix, LowRankMatrix "
                         "and hierarchical Cholesky workloads")
    ap.add_argument("--no-tuning", action="store_true",
                    help="batched workload: time the products only (no autotuner, graph "
                         "replays or leaf-size tuning), for kernel traces and counters")
    args = ap.parse_args()
    args.variants_given = args.variants is not None
    if args.variants is None:
        args.variants = "hybrid,matrix_free,f32_hybrid"
    return args


def time_steps(fn, steps, stream):
    import torch
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for _ in range(steps):
        fn()
    t1.record(stream)
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / steps


def pass_times(H, fn, steps):
    import ctypes
    import numpy as np
    import torch
    from hmx._lib import check, lib
    torch.cuda.synchronize()
    check(lib.hm_plan_profile(H._h, 1))
    for _ in range(steps):
        fn()
    ms = np.zeros(4)
    cnt = ctypes.c_int64(0)
    check(lib.hm_plan_profile_read(H._h, ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   ctypes.byref(cnt)))
    check(lib.hm_plan_profile(H._h, 0))
    return ms / max(cnt.value, 1)


def pmc_traffic():
    """HBM bytes per launch of the stored downward pass, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (scaled by a copy-kernel
    calibration); None if the summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_down_cauchy_n2p20_pmc_traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch k_down_stored:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def pmc_valu():
    """FP64 VALU wave instructions per product of the matrix-free downward
    pass, from the committed rocprofv3 SQ counter summary; None if absent."""
    path = os.path.join(ROOT, "profiles", "k_down_matrix_free_n2p20_valu_counters.md")
    try:
        for line in open(path):
            if line.startswith("valu_f64_instr_per_product k_down_free:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def stream_peaks(stream):
    """Measured HBM bandwidth of torch's copy (read + write) and sum (read) on a
    4 GiB float64 buffer, for context beside the nominal 8 TB/s."""
    import torch
    n = 1 << 29
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
        src.sum()
    ms_copy = time_steps(lambda: dst.copy_(src), 10, stream)
    ms_sum = time_steps(lambda: src.sum(), 10, stream)
    del src, dst
    torch.cuda.empty_cache()
    return {"copy_gbs": round(2 * 8 * n / (ms_copy * 1e-3) / 1e9, 1),
            "read_gbs": round(8 * n / (ms_sum * 1e-3) / 1e9, 1)}


def host_facts():
    """The host the CPU leg runs on: CPU model, logical CPUs, the affinity set,
    the cgroup CPU quota (cores), total and available RAM."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    mem = {}
    try:
        for line in open("/proc/meminfo"):
            k, v = line.split(":", 1)
            if k in ("MemTotal", "MemAvailable"):
                mem[k] = int(v.split()[0]) * 1024
    except OSError:
        pass
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = int(q) / int(per)
    except (OSError, ValueError):
        pass
    aff = len(os.sched_getaffinity(0))
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": aff,
            "cgroup_cpu_quota": quota, "ram_total_gb": round(mem.get("MemTotal", 0) / 1e9, 1),
            "ram_available_gb": round(mem.get("MemAvailable", 0) / 1e9, 1)}


def cpu_baseline(threads=0, n_log2=None, min_seconds=10.0):
    """Oracle traversal mul! (ref layout, bit-identical to the single-thread
    restatement) on the n = 2^20 headline instance itself when the host has
    room for its 16.75 GB of reference-layout factors (24 GB available),
    else on the largest 2^k instance of the same family that fits, scaled to
    n = 2^20 by the ratio of reference-layout bytes. Threads: every CPU of the
    affinity set, capped by the cgroup CPU quota when there is one. The OpenMP
    product is repeated until min_seconds of timed work and the median run is
    reported; the single-thread figure is the median of 3 runs. Tree
    construction is timed apart."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import hmx
    import oracle as O
    facts = host_facts()
    if not threads:
        threads = facts["affinity_cpus"]
        if facts["cgroup_cpu_quota"]:
            threads = max(1, min(threads, int(facts["cgroup_cpu_quota"])))
    if n_log2 is None:
        avail = facts["ram_available_gb"]
        n_log2 = 20
        while n_log2 > 14 and avail < 24.0 / (1 << (20 - n_log2)):
            n_log2 -= 1
    n = 1 << n_log2
    w = hmx.workloads.config_plain(n=n, blocksize=BS, rank=RANK)
    t0 = time.perf_counter()
    t = O.Tree(O.CAUCHY, w["x"], w["y"], *w["intervals"], BS, RANK, store=True)
    build = time.perf_counter() - t0
    st = t.stats()
    ref_bytes = lambda s, nn: 8 * (s["dense_entries"] + RANK * (s["far_rows"] + s["far_cols"])
                                   + s["far"] * RANK * RANK) + 16 * nn
    b_sample = ref_bytes(st, n)
    t.matvec_omp(w["b"], nthreads=threads)
    runs = []
    while sum(runs) < min_seconds and len(runs) < 2000:
        t0 = time.perf_counter()
        t.matvec_omp(w["b"], nthreads=threads)
        runs.append(time.perf_counter() - t0)
    sec = float(np.median(runs))
    runs1 = []
    for _ in range(3):
        t0 = time.perf_counter()
        t.matvec(w["b"])
        runs1.append(time.perf_counter() - t0)
    sec1 = float(np.median(runs1))
    del t
    if n_log2 == 20:
        b_full = b_sample
    else:
        b_full = hmx.KernelMatrix("cauchykernel", *hmx.workloads.shifted_grid(N)[:2],
                                  *hmx.workloads.shifted_grid(N)[2], blocksize=BS, rank=RANK,
                                  tree_only=True).info["bytes_ref_layout"]
    scale = b_sample / b_full
    how = ("the n=2^20 instance itself" if n_log2 == 20 else
           f"the n=2^{n_log2} instance of the same family, scaled to n=2^20 by ref-layout bytes "
           f"{b_sample}/{b_full}")
    return dict(value=(1.0 / sec) * scale, unit="matvecs/s", cores=threads, kind="port",
                sample=(f"oracle traversal mul! (OpenMP, {threads} threads, ref layout, "
                        f"{b_sample / 1e9:.2f} GB of factors) on {how} (leaf 256, rank 24): "
                        f"median {sec * 1e3:.1f} ms of {len(runs)} runs ({sum(runs):.1f} s "
                        f"timed); single thread median {sec1 * 1e3:.0f} ms of 3; tree and "
                        f"factors built in {build:.1f} s"),
                single_thread_value=(1.0 / sec1) * scale, n_log2=n_log2, host=facts)


def dense_rows(f, xr, y, B, chunk=1 << 16):
    """sum_j f(x_r - y_j) B[j, :] for the given rows, FP64 on the GPU."""
    import torch
    out = torch.zeros(xr.numel(), B.shape[1], dtype=torch.float64, device=xr.device)
    for j0 in range(0, y.numel(), chunk):
        out += f(xr[:, None] - y[None, j0:j0 + chunk]) @ B[j0:j0 + chunk]
    return out


def timed(fn, steps, warmup, stream):
    for _ in range(warmup):
        fn()
    return time_steps(fn, steps, stream)


def rel_err(a, b):
    import torch
    return float(torch.linalg.norm(a.double() - b.double()) / torch.linalg.norm(b.double()))


def summary_batched(steps, warmup, stream):
    """Config 2 in the headline run: 256 Cauchy instances (n log-uniform in
    [2^10, 2^16]), rank 16, leaf 64, 32 right-hand sides, hybrid plan; the
    plain FP64 MFMA product against FP32, split-BF16 hi+lo and hi+mid+lo far
    fields (ms, batched products/s, speed-up and error against the FP64 GPU
    product, error of instance 0 against its dense FP64 product)."""
    import numpy as np
    import torch
    import hmx
    dev = torch.device("cuda:0")
    w = hmx.workloads.config_batched()
    inst = w["instances"]
    t0 = time.perf_counter()
    H = hmx.KernelMatrixBatch("cauchykernel", [(i["x"], i["y"], i["intervals"]) for i in inst],
                              blocksize=64, rank=16, hybrid=True)
    torch.cuda.synchronize()
    build = time.perf_counter() - t0
    B = torch.from_numpy(np.concatenate([i["b"] for i in inst])).to(dev)
    Y = torch.empty(H.shape[0], 32, dtype=torch.float64, device=dev)
    n0 = len(inst[0]["x"])
    d0 = dense_rows(lambda d: 1.0 / d, torch.from_numpy(inst[0]["x"]).to(dev),
                    torch.from_numpy(inst[0]["y"]).to(dev), B[:len(inst[0]["y"])])
    out = {"plan": "hybrid", "instances": len(inst), "points_total": int(H.shape[0]), "rhs": 32,
           "rank": 16, "leaf_blocksize": 64, "build_s": round(build, 3)}
    ref = None
    for prec in ("f64", "f32", "bf16x3", "bf16x6"):
        ms = timed(lambda: H._mv(Y, B, 1, 1, 1, 1, 0, False, False, stream, False, False, prec),
                   steps, warmup, stream)
        if ref is None:
            ref, ms64 = Y.clone(), ms
        out[prec] = {"ms_per_step": round(ms, 4), "value": round(1000.0 / ms, 3),
                     "unit": "batched products/s",
                     "speedup_vs_fp64": round(ms64 / ms, 3),
                     "rel_err_vs_fp64_gpu": rel_err(Y, ref),
                     "rel_err_instance0_vs_dense_fp64": rel_err(Y[:n0], d0)}
    H.close()
    del Y
    out["leaf_tuning"] = leaf_tuning([(i["x"], i["y"], i["intervals"]) for i in inst])
    return out


def leaf_tuning(spec, reps=3):
    """hmx.tune_blocksize over leaf thresholds 32/64/128 and the hybrid and
    matrix-free forms for the batch, 32 right-hand sides, FP64: each
    candidate's tuned time and its error on sampled rows of four instances
    against the dense FP64 product."""
    import hmx
    best, res = hmx.tune_blocksize("cauchykernel", spec, 16, k=32, candidates=(32, 64, 128),
                                   forms=("hybrid", "matrix_free"), reps=reps, tol=1e-10)
    return {"best": None if best is None else {
                "blocksize": best["blocksize"], "form": best["form"],
                "ms_per_step": round(best["ms"], 4), "rel_err_vs_dense_fp64": best["rel_err"],
                "rhs_per_launch": best["rhs_per_launch"]},
            "candidates": [{"blocksize": d["blocksize"], "form": d["form"],
                            "ms_per_step": round(d["ms"], 4), "rel_err": d["rel_err"]}
                           for d in res]}


def summary_gradient(steps, warmup, stream):
    """Config 3 in the headline run: L = 1/2 ||H(x, y) b - t||^2, n = 2^16,
    leaf 64, rank 16; one step = forward product + pullback to x, y and b
    (plans reused and refitted on the device from step to step), Float64 and
    Float32, against the plain FP64 forward product of the same matrix-free
    plan; errors against the dense FP64 gradient on 256 sampled entries."""
    import numpy as np
    import torch
    import hmx
    from hmx.autodiff import kernel_matvec
    dev = torch.device("cuda:0")
    w = hmx.workloads.config_gradient()
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    n = len(w["x"])
    H = hmx.KernelMatrix("cauchykernel", w["x"], w["y"], *w["intervals"], blocksize=64, rank=16,
                         matrix_free=True)
    b64 = T(w["b"])
    u = torch.empty_like(b64)
    ms_plain = timed(lambda: H._mv(u, b64, 1, 1, 1, 1, 0, False, False, stream), steps, warmup,
                     stream)
    H.close()
    out = {"n": n, "rank": 16, "leaf_blocksize": 64,
           "plain_fp64_forward_ms": round(ms_plain, 4)}
    idx = torch.arange(0, n, n // 256, device=dev)
    for name, dtype in (("f64", torch.float64), ("f32", torch.float32)):
        x, y, b, t = (T(w[k]).to(dtype) for k in ("x", "y", "b", "t"))
        x.requires_grad_(True)
        y.requires_grad_(True)
        b.requires_grad_(True)

        def step():
            x.grad = y.grad = b.grad = None
            uu = kernel_matvec("cauchykernel", x, y, b, w["intervals"], 64, 16)
            L = 0.5 * torch.sum((uu - t) ** 2)
            L.backward()

        ms = timed(step, steps, warmup, stream)
        with torch.no_grad():
            xd, yd, bd, td = (v.detach().double() for v in (x, y, b, t))
            Kb = dense_rows(lambda d: 1.0 / d, xd, yd, bd[:, None])[:, 0]
            r = Kb - td
            gx = r[idx] * dense_rows(lambda d: -1.0 / (d * d), xd[idx], yd, bd[:, None])[:, 0]
            gb = dense_rows(lambda d: 1.0 / d, -yd[idx], -xd, r[:, None])[:, 0]
            gy = bd[idx] * dense_rows(lambda d: 1.0 / (d * d), -yd[idx], -xd, r[:, None])[:, 0]
        step()
        out[name] = {"ms_per_step": round(ms, 4), "value": round(1000.0 / ms, 3),
                     "unit": "gradient steps/s",
                     "step_vs_plain_forward": round(ms / ms_plain, 3),
                     "rel_err_vs_dense_fp64": {"grad_x": rel_err(x.grad[idx], gx),
                                               "grad_y": rel_err(y.grad[idx], gy),
                                               "grad_b": rel_err(b.grad[idx], gb)}}
        out[name].update(held_plan_steps(x, y, b, t, w["intervals"], steps, warmup, stream))
    return out


def held_plan_steps(x0, y0, b0, t, intervals, steps, warmup, stream):
    """The same step through hmx.autodiff.KernelProduct (one held plan, the
    tree check left on the device): eager, and captured into a CUDA graph and
    replayed; the replay's values and gradients against kernel_matvec's eager
    ones. Each phase differentiates its own leaf tensors, and no autograd
    graph of an eager phase is alive during the capture (its gradient
    accumulators would run on the stream they were made on)."""
    import torch
    from hmx.autodiff import KernelProduct, kernel_matvec

    def leaves():
        return [v.detach().clone().requires_grad_(True) for v in (x0, y0, b0)]

    def make_step(fn, x, y, b):
        def step():
            x.grad = y.grad = b.grad = None
            u = fn(x, y, b)
            (0.5 * torch.sum((u - t) ** 2)).backward()
            return u.detach()
        return step

    x, y, b = leaves()
    u = make_step(lambda *a: kernel_matvec("cauchykernel", *a, intervals, 64, 16), x, y, b)()
    want = [u, x.grad, y.grad, b.grad]
    del x, y, b
    kp = KernelProduct("cauchykernel", x0.detach(), y0.detach(), intervals, 64, 16)
    res = {"held_plan_eager_ms_per_step": round(timed(make_step(kp, *leaves()), steps, warmup,
                                                      stream), 4)}
    x, y, b = leaves()
    kstep = make_step(kp, x, y, b)
    side = torch.cuda.Stream()
    side.wait_stream(stream)
    with torch.cuda.stream(side):
        for _ in range(3):
            kstep()
    stream.wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        u_static = kstep()
    ms_g = timed(g.replay, steps, warmup, stream)
    got = [u_static, x.grad, y.grad, b.grad]
    res["graph_ms_per_step"] = round(ms_g, 4)
    res["graph_value"] = round(1000.0 / ms_g, 3)
    res["graph_equals_eager"] = all(bool(torch.equal(a, c)) for a, c in zip(got, want))
    res["graph_tree_kept"] = kp.tree_ok()
    del g
    kp.close()
    return res


def summary_scale(steps, warmup, stream):
    """Config 4 in the headline run: log|x - y|, n = 2^22 sorted uniform points
    (seed 1), y = x + 0.5/n, leaf 512, rank 32, 8 right-hand sides, hybrid
    plan: FP64 MFMA against FP32, split-BF16 hi+mid+lo and hi+lo far fields;
    errors against 4096 dense FP64 rows (torch on the GPU)."""
    import numpy as np
    import torch
    import hmx
    dev = torch.device("cuda:0")
    w = hmx.workloads.config_scale()
    B = torch.from_numpy(w["b"]).to(dev)
    n = len(w["x"])
    rows = torch.from_numpy(np.sort(np.random.default_rng(5).choice(n, 4096, replace=False))).to(dev)
    dense = dense_rows(lambda d: torch.log(torch.abs(d)), torch.from_numpy(w["x"]).to(dev)[rows],
                       torch.from_numpy(w["y"]).to(dev), B)
    t0 = time.perf_counter()
    H = hmx.KernelMatrix("logkernel", w["x"], w["y"], *w["intervals"], blocksize=512, rank=32,
                         hybrid=True)
    torch.cuda.synchronize()
    build = time.perf_counter() - t0
    Y = torch.empty(n, 8, dtype=torch.float64, device=dev)
    out = {"plan": "hybrid", "n": n, "rank": 32, "leaf_blocksize": 512, "rhs": 8,
           "build_s": round(build, 3), "hbm_bytes_stored": H.info["bytes_stored"]}
    ref = None
    for prec in ("f64", "f32", "bf16x6", "bf16x3"):
        ms = timed(lambda: H._mv(Y, B, 1, 1, 1, 1, 0, False, False, stream, False, False, prec),
                   steps, warmup, stream)
        if ref is None:
            ref, ms64 = Y.clone(), ms
        out[prec] = {"ms_per_step": round(ms, 4), "value": round(1000.0 / ms, 3),
                     "unit": "products/s (8 RHS)", "speedup_vs_fp64": round(ms64 / ms, 3),
                     "rel_err_vs_fp64_gpu": rel_err(Y, ref),
                     "rel_err_vs_dense_fp64_4096rows": rel_err(Y[rows], dense)}
    H.close()
    del Y, ref
    torch.cuda.empty_cache()
    return out


def run_batched(args):
    """256 Cauchy instances, n log-uniform in [2^10, 2^16], rank 16, leaf 64,
    32 right-hand sides; one step = the product over every instance and right-
    hand side. Variants: stored / hybrid / matrix-free plans on the FP64 MFMA
    passes (the default for 32 right-hand sides) and on the vector units, the
    FP32 and split-BF16 far field, the autotuned choice, and HIP-graph
    replays."""
    import numpy as np
    import torch
    import hmx
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    w = hmx.workloads.config_batched()
    inst = w["instances"]
    spec = [(i["x"], i["y"], i["intervals"]) for i in inst]
    B = torch.from_numpy(np.concatenate([i["b"] for i in inst])).to(dev)
    out = {"metric": "batched Cauchy H-matvecs (256 instances x 32 RHS) per second",
           "unit": "batched products/s", "dtype": "f64", "steps": args.steps,
           "config": {"workload": "batched", "instances": len(inst),
                      "points_total": int(sum(len(i["x"]) for i in inst)), "rank": 16,
                      "leaf_blocksize": 64, "rhs": 32}, "variants": {}}
    ref = None
    wanted = set(v for v in args.variants.split(",") if v) if args.variants_given else None
    for name, kw in (("stored", {}), ("hybrid", dict(hybrid=True)), ("matrix_free", dict(matrix_free=True))):
        if wanted is not None and name not in wanted:
            continue
        H = hmx.KernelMatrixBatch("cauchykernel", spec, blocksize=64, rank=16, **kw)
        Y = torch.empty(H.shape[0], 32, dtype=torch.float64, device=dev)
        ms = timed(lambda: H._mv(Y, B, 1, 1, 1, 1, 0, False, False, stream), args.steps,
                   args.warmup, stream)
        if ref is None:
            ref = Y.clone()
            # dense FP64 rows of instance 0 (torch, GPU) for the error
            i0 = inst[0]
            xr = torch.from_numpy(i0["x"]).to(dev)
            d = dense_rows(lambda d: 1.0 / d, xr, torch.from_numpy(i0["y"]).to(dev),
                           B[:len(i0["y"])])
            out["rel_err_instance0_vs_dense"] = float(torch.linalg.norm(Y[:len(i0["x"])] - d) /
                                                      torch.linalg.norm(d))
        err = float(torch.linalg.norm(Y - ref) / torch.linalg.norm(ref))
        info = H.info
        Yv = torch.empty_like(Y)
        ms_v = timed(lambda: H._mv(Yv, B, 1, 1, 1, 1, 0, False, False, stream, mfma=False),
                     args.steps, args.warmup, stream)
        out["variants"][name] = {"ms_per_step": round(ms, 4), "value": round(1000 / ms, 3),
                                 "hbm_bytes_stored": info["bytes_stored"],
                                 "rel_diff_vs_stored": err,
                                 "vector_units": {"ms_per_step": round(ms_v, 4),
                                                  "rel_diff_vs_mfma": float(
                                                      torch.linalg.norm(Yv - Y) / torch.linalg.norm(Y))}}
        del Yv
        for prec in ("f32", "bf16x3", "bf16x6"):
            Ym = torch.empty_like(Y)
            ms_m = timed(lambda: H._mv(Ym, B, 1, 1, 1, 1, 0, False, False, stream, False, False,
                                       prec), args.steps, args.warmup, stream)
            out["variants"][name][f"far_{prec}"] = {
                "ms_per_step": round(ms_m, 4), "value": round(1000 / ms_m, 3),
                "rel_err_vs_fp64": float(torch.linalg.norm(Ym - ref) / torch.linalg.norm(ref))}
            del Ym
        if (name == "matrix_free" or name == "hybrid") and not args.no_tuning:
            kb, mode, us = H.tune(k=32)
            ms_t = timed(lambda: H.matvec(B, tuned=True), args.steps, args.warmup, stream)
            g = H.graph(Y, B, tuned=True)
            ms_g = timed(lambda: g.launch(), args.steps, args.warmup, stream)
            out["variants"][name].update(tuned={"rhs_per_launch": kb, "mode": mode,
                                                "ms_per_step": round(ms_t, 4)},
                                         graph_ms_per_step=round(ms_g, 4))
            g.close()
            gm = H.graph(Y, B, far_precision="f32")
            ms_gm = timed(lambda: gm.launch(), args.steps, args.warmup, stream)
            out["variants"][name]["far_f32"]["graph_ms_per_step"] = round(ms_gm, 4)
            out["variants"][name]["far_f32"]["graph_rel_err_vs_fp64"] = float(
                torch.linalg.norm(Y - ref) / torch.linalg.norm(ref))
            gm.close()
        H.close()
        del Y
        torch.cuda.empty_cache()
    best = min(v["ms_per_step"] for v in out["variants"].values())
    out["ms_per_step"] = best
    out["value"] = round(1000 / best, 3)
    if not args.no_tuning:
        out["leaf_tuning"] = leaf_tuning(spec)
    return out


def run_gradient(args):
    """L = 1/2 ||H(x, y) b - t||^2, n = 2^16, leaf 64, rank 16: one step =
    forward product + pullback to x, y and b, with Float64 tensors (the bench
    line) and with Float32 tensors (Float32 plans)."""
    import numpy as np
    import torch
    import hmx
    from hmx.autodiff import kernel_matvec
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    w = hmx.workloads.config_gradient()
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)

    def measure(dtype):
        x, y, b, t = (T(w[k]).to(dtype) for k in ("x", "y", "b", "t"))
        x.requires_grad_(True)
        y.requires_grad_(True)
        b.requires_grad_(True)

        def step():
            x.grad = y.grad = b.grad = None
            u = kernel_matvec("cauchykernel", x, y, b, w["intervals"], 64, 16)
            L = 0.5 * torch.sum((u - t) ** 2)
            L.backward()

        ms = timed(step, args.steps, args.warmup, stream)
        held = held_plan_steps(x, y, b, t, w["intervals"], args.steps, args.warmup, stream)
        # dense FP64 gradient of the same (possibly Float32) points on 256
        # sampled rows / columns (torch, GPU)
        n = x.numel()
        idx = torch.arange(0, n, n // 256, device=dev)
        with torch.no_grad():
            xd, yd, bd, td = (v.detach().double() for v in (x, y, b, t))
            Kb = dense_rows(lambda d: 1.0 / d, xd, yd, bd[:, None])[:, 0]
            r = Kb - td
            gx = r[idx] * dense_rows(lambda d: -1.0 / (d * d), xd[idx], yd, bd[:, None])[:, 0]
            # column sums: K^T r and K_y^T r at the sampled columns
            gb = dense_rows(lambda d: 1.0 / d, -yd[idx], -xd, r[:, None])[:, 0]
            gy = bd[idx] * dense_rows(lambda d: 1.0 / (d * d), -yd[idx], -xd, r[:, None])[:, 0]
        e = lambda a, c: float(torch.linalg.norm(a.double() - c) / torch.linalg.norm(c))
        step()
        return n, ms, {"grad_x": e(x.grad[idx], gx), "grad_y": e(y.grad[idx], gy),
                       "grad_b": e(b.grad[idx], gb)}, held

    n, ms, err, held = measure(torch.float64)
    _, ms32, err32, held32 = measure(torch.float32)
    return {"metric": "gradient steps/s (forward + pullback to x, y, b), n=2^16",
            "unit": "steps/s", "value": round(1000 / ms, 3), "ms_per_step": round(ms, 4),
            "dtype": "f64", "steps": args.steps,
            "config": {"workload": "gradient", "n": n, "rank": 16, "leaf_blocksize": 64},
            "rel_err_vs_dense_fp64": err, **held,
            "variants": {"f32": {"value": round(1000 / ms32, 3), "ms_per_step": round(ms32, 4),
                                 "dtype": "f32", "rel_err_vs_dense_fp64": err32, **held32}}}


def run_scale(args):
    """log|x - y|, n = 2^22 sorted uniform points (seed 1), y = x + 0.5/n, leaf
    512, rank 32, 8 right-hand sides; error against dense FP64 rows (4096 rows,
    torch on the GPU)."""
    import numpy as np
    import torch
    import hmx
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    w = hmx.workloads.config_scale()
    B = torch.from_numpy(w["b"]).to(dev)
    out = {"metric": "log-kernel H-matvecs/s, n=2^22, 8 RHS", "unit": "products/s",
           "dtype": "f64", "steps": args.steps,
           "config": {"workload": "scale", "n": len(w["x"]), "rank": 32, "leaf_blocksize": 512,
                      "rhs": 8}, "variants": {}}
    rows = torch.from_numpy(np.sort(np.random.default_rng(5).choice(len(w["x"]), 4096, replace=False))).to(dev)
    dense = dense_rows(lambda d: torch.log(torch.abs(d)), torch.from_numpy(w["x"]).to(dev)[rows],
                       torch.from_numpy(w["y"]).to(dev), B)
    for name, kw in (("hybrid", dict(hybrid=True)), ("matrix_free", dict(matrix_free=True))):
        t0 = time.perf_counter()
        H = hmx.KernelMatrix("logkernel", w["x"], w["y"], *w["intervals"], blocksize=512,
                             rank=32, **kw)
        torch.cuda.synchronize()
        build = time.perf_counter() - t0
        Y = torch.empty(H.shape[0], 8, dtype=torch.float64, device=dev)
        ms = timed(lambda: H._mv(Y, B, 1, 1, 1, 1, 0, False, False, stream), args.steps,
                   args.warmup, stream)
        err = float(torch.linalg.norm(Y[rows] - dense) / torch.linalg.norm(dense))
        out["variants"][name] = {"ms_per_step": round(ms, 4), "value": round(1000 / ms, 3),
                                 "build_s": round(build, 3), "hbm_bytes_stored": H.info["bytes_stored"],
                                 "rel_err_vs_dense_4096rows": err}
        ms_m = timed(lambda: H._mv(Y, B, 1, 1, 1, 1, 0, False, False, stream, mfma=True),
                     args.steps, args.warmup, stream)
        out["variants"][name]["mfma_fp64"] = {
            "ms_per_step": round(ms_m, 4), "value": round(1000 / ms_m, 3),
            "rel_err_vs_dense_4096rows":
                float(torch.linalg.norm(Y[rows] - dense) / torch.linalg.norm(dense))}
        for prec in ("f32", "bf16x3", "bf16x6"):
            ms_m = timed(lambda: H._mv(Y, B, 1, 1, 1, 1, 0, False, False, stream, False, False,
                                       prec), args.steps, args.warmup, stream)
            out["variants"][name][f"far_{prec}"] = {
                "ms_per_step": round(ms_m, 4), "value": round(1000 / ms_m, 3),
                "rel_err_vs_dense_4096rows":
                    float(torch.linalg.norm(Y[rows] - dense) / torch.linalg.norm(dense))}
        H.close()
        del Y
        torch.cuda.empty_cache()
    best = min(v["ms_per_step"] for v in out["variants"].values())
    out["ms_per_step"], out["value"] = best, round(1000 / best, 3)
    return out


def run_even(args):
    """EvenBarycentricMatrix leaves (src/BarycentricMatrix.jl:5-46, mul!
    src/algebra.jl:168-239): N = 2^20, the 2046 off-diagonal leaves of a
    10-level dyadic partition, rank 20, Gaussian f; one step = the product of
    every leaf with one vector (and with 4 vectors). Error against the dense
    parity-masked product of f on 256 sampled rows (torch FP64 on the GPU);
    CPU: the oracle's leaf mul! on the N = 2^16 instance, scaled by bytes."""
    import numpy as np
    import torch
    import hmx
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    w = hmx.workloads.config_even()
    N, n, levels = w["N"], w["rank"], w["levels"]
    t0 = time.perf_counter()
    specs = [(r0, c0, hmx.EvenLeaf(a, b, c, d, hmx.even.even_F(w["f"], a, b, c, d, n)))
             for (r0, c0, a, b, c, d) in w["leaves"]]
    t1 = time.perf_counter()
    H = hmx.EvenBarycentricLeaves(N, N, specs)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    del specs
    info = H.info
    v = torch.from_numpy(w["v"]).to(dev)
    u = torch.empty(N, dtype=torch.float64, device=dev)
    ms = timed(lambda: H._call(u, v, 0, 0, 0, stream), args.steps, args.warmup, stream)
    nbytes = H.bytes_per_product(1)
    # the bytes with a W per leaf, as the leaves hold it on their own (the
    # plan stores one W per leaf size): the CPU leg is scaled by this
    nbytes_leaf_w = nbytes + 8 * (H.leaf_w_entries() - info["w_entries"])
    # dense check: row I pairs with columns J outside its finest diagonal block
    # with I + J even
    s = N >> levels
    rows = torch.from_numpy(np.sort(np.random.default_rng(7).choice(N, 256, replace=False))).to(dev)
    J = torch.arange(N, device=dev, dtype=torch.float64)
    Ji = torch.arange(N, device=dev)
    dense = torch.empty(256, dtype=torch.float64, device=dev)
    for q in range(0, 256, 32):
        I = rows[q:q + 32]
        keep = (((I[:, None] + Ji[None, :]) % 2) == 0) & ((I[:, None] // s) != (Ji[None, :] // s))
        M = torch.exp(-((I[:, None].double() - J[None, :]) / N) ** 2) * keep
        dense[q:q + 32] = M @ v
        del M, keep
    err = float(torch.linalg.norm(u[rows] - dense) / torch.linalg.norm(dense))
    V4 = torch.from_numpy(hmx.workloads.normal(22, 4 * N).reshape(4, N)).to(dev).T
    U4 = torch.empty(4, N, dtype=torch.float64, device=dev).T
    ms4 = timed(lambda: H._call(U4, V4, 0, 0, 0, stream), args.steps, args.warmup, stream)
    err4 = float(torch.linalg.norm(U4[:, 0] - H.matvec(V4[:, 0].contiguous())) /
                 torch.linalg.norm(U4[:, 0]))
    out = {"metric": "EvenBarycentricMatrix leaf products/s, N=2^20, 2046 leaves, rank 20",
           "value": round(1000.0 / ms, 3), "unit": "products/s", "ms_per_step": round(ms, 4),
           "steps": args.steps, "warmup": args.warmup, "dtype": "f64",
           "data": "synthetic (Gaussian f, v from splitmix64 seed 21)",
           "config": {"workload": "even_barycentric", "N": N, "levels": levels, "rank": n,
                      "leaves": info["nleaves"], "rhs": 1},
           "bytes_per_product": nbytes, "hbm_gbs": round(nbytes / (ms * 1e-3) / 1e9, 1),
           "bytes_with_w_per_leaf": nbytes_leaf_w,
           "plan": info, "f_eval_s": round(t1 - t0, 3), "build_s": round(t2 - t1, 3),
           "rel_err_vs_dense_256rows": err,
           "rhs4": {"ms_per_step": round(ms4, 4), "ms_per_rhs": round(ms4 / 4, 4),
                    "rel_diff_vs_single": err4,
                    "hbm_gbs": round(H.bytes_per_product(4) / (ms4 * 1e-3) / 1e9, 1)}}
    H.close()
    del H
    # EvenBarycentricMatrix{Float32} of the same leaves: rank BLOCKRANK(Float32)
    n32 = hmx.BLOCKRANK(np.float32)
    specs = [(r0, c0, hmx.EvenLeaf(a, b, c, d,
                                   hmx.even.even_F(w["f"], a, b, c, d, n32, np.float32),
                                   dtype=np.float32))
             for (r0, c0, a, b, c, d) in w["leaves"]]
    H32 = hmx.EvenBarycentricLeaves(N, N, specs)
    del specs
    v32 = v.float()
    u32 = torch.empty(N, dtype=torch.float32, device=dev)
    ms32 = timed(lambda: H32._call(u32, v32, 0, 0, 0, stream), args.steps, args.warmup, stream)
    err32 = float(torch.linalg.norm(u32[rows].double() - dense) / torch.linalg.norm(dense))
    out["f32"] = {"rank": n32, "ms_per_step": round(ms32, 4),
                  "products_per_s": round(1000.0 / ms32, 3),
                  "bytes_per_product": H32.bytes_per_product(1),
                  "hbm_gbs": round(H32.bytes_per_product(1) / (ms32 * 1e-3) / 1e9, 1),
                  "rel_err_vs_dense_256rows": err32}
    H32.close()
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        Nc = 1 << 16
        wc = hmx.workloads.config_even(N=Nc, levels=6)
        facs = []
        for (r0, c0, a, b, c, d) in wc["leaves"]:
            _, W = O.even_weights(a, b, n)
            facs.append((r0, c0, W, hmx.even.even_F(wc["f"], a, b, c, d, n)))
        cbytes = 8 * sum(W.size + F.size for _, _, W, F in facs) + 8 * 3 * Nc
        runs = []
        while sum(runs) < 10.0 and len(runs) < 500:
            uc = np.zeros(Nc)
            tc = time.perf_counter()
            for r0, c0, W, F in facs:
                O.even_mul(uc, W, F, wc["v"], r0 + 1, c0 + 1)
            runs.append(time.perf_counter() - tc)
        sec = float(np.median(runs))
        out["cpu_baseline"] = {
            "value": round((cbytes / sec) / nbytes_leaf_w, 4), "unit": "products/s", "cores": 1,
            "kind": "port",
            "sample": (f"oracle leaf mul! (src/algebra.jl:168-239 restated, one thread) over the "
                       f"{len(facs)} leaves of the N=2^16, 6-level instance: median "
                       f"{sec * 1e3:.1f} ms of {len(runs)}; scaled to N=2^20 by factor bytes "
                       f"{cbytes}/{nbytes_leaf_w} (W per leaf on both sides)")}
    return out


def run_lowrank(args):
    """LowRankMatrix recompression, L1 + L2 (src/LowRankMatrix.jl:95-102): 2048
    sums of 512 x 512 rank-16 matrices in one hm_lowrank_op call (QR of the
    two 512 x 32 stacked factors, SVD of the 32 x 32 core, getrank, Q times
    the singular vectors). One step = the whole call, workspace and
    synchronization included. Checked on 8 items against the dense sum of
    the operands; CPU = the oracle's numpy restatement (LAPACK) on 64 items."""
    import ctypes
    import numpy as np
    import torch
    import hmx
    from hmx._lib import check, lib
    dev = torch.device("cuda:0")
    count, m, r = 2048, 512, 16
    g = torch.Generator(device=dev).manual_seed(0)
    mk = lambda *shape: torch.randn(*shape, generator=g, dtype=torch.float64, device=dev)
    U1, V1, U2, V2 = mk(count, r, m), mk(count, r, m), mk(count, r, m), mk(count, r, m)
    S1 = torch.rand(count, r, generator=g, dtype=torch.float64, device=dev) + 0.1
    S2 = torch.rand(count, r, generator=g, dtype=torch.float64, device=dev) + 0.1
    Uo = torch.empty(count, 2 * r, m, dtype=torch.float64, device=dev)
    Vo = torch.empty(count, 2 * r, m, dtype=torch.float64, device=dev)
    So = torch.empty(count, 2 * r, dtype=torch.float64, device=dev)
    dims = np.tile(np.array([m, m, 0, r, r], dtype=np.int64), count)
    inp = (ctypes.c_void_p * (7 * count))()
    outp = (ctypes.c_void_p * (3 * count))()
    for t in range(count):
        inp[7 * t:7 * t + 6] = [U1[t].data_ptr(), S1[t].data_ptr(), V1[t].data_ptr(),
                                U2[t].data_ptr(), S2[t].data_ptr(), V2[t].data_ptr()]
        outp[3 * t:3 * t + 3] = [Uo[t].data_ptr(), So[t].data_ptr(), Vo[t].data_ptr()]
    rank = np.zeros(count, dtype=np.int64)
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    call = lambda: check(lib.hm_lowrank_op(0, count, ip(dims), inp, outp, ip(rank), stream))
    for _ in range(max(1, args.warmup)):
        call()
    steps = max(3, min(args.steps, 20))
    t0 = time.perf_counter()
    for _ in range(steps):
        call()
    sec = (time.perf_counter() - t0) / steps
    # check on 8 items: U diag(S) V^T against the dense sum of the operands
    # (torch FP64); random factors give full rank 2r
    errs = []
    for t in range(0, count, count // 8):
        rt = int(rank[t])
        dense = lambda U, S, V: (U.T * S[None, :]) @ V
        want = dense(U1[t], S1[t], V1[t]) + dense(U2[t], S2[t], V2[t])
        got = dense(Uo[t][:rt], So[t][:rt], Vo[t][:rt])
        errs.append(float(torch.linalg.norm(got - want) / torch.linalg.norm(want)))
    host = lambda x, t: x[t].cpu().numpy()
    # algorithmic bytes: read U1, V1, U2, V2 (m x r each) and write U, V (m x 2r)
    nbytes = count * 8 * (4 * m * r + 2 * m * 2 * r)
    # flops: two Householder QRs of m x 2r (4 m c^2 - 4/3 c^3 each, c = 2r), the
    # core (c^3 / 3), two Q applications (4 m c^2 each)
    c = 2 * r
    flops = count * (2 * (4 * m * c * c - 4 * c ** 3 / 3) + c ** 3 / 3 + 2 * 4 * m * c * c)
    out = {"metric": "LowRankMatrix sums recompressed/s (2048 x 512x512 rank 16+16)",
           "value": round(count / sec, 1), "unit": "sums/s", "ms_per_step": round(sec * 1e3, 3),
           "steps": steps, "warmup": args.warmup, "dtype": "f64",
           "data": "synthetic (torch normal factors, seed 0)",
           "config": {"workload": "lowrank_add", "count": count, "m": m, "n": m, "r1": r, "r2": r},
           "ranks": sorted(set(int(x) for x in rank)),
           "rel_err_vs_dense_sum_8items": max(errs),
           "hbm_gbs_algorithmic": round(nbytes / sec / 1e9, 1),
           "gflops_model": round(flops / sec / 1e9, 1)}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import lowrank as LR
        sample = 64
        Ls = [((host(U1, t).T, host(S1, t), host(V1, t).T), (host(U2, t).T, host(S2, t),
               host(V2, t).T)) for t in range(sample)]
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < 10.0:
            for L1, L2 in Ls:
                LR.add(L1, L2)
            reps += 1
        per = (time.perf_counter() - t0) / (reps * sample)
        out["cpu_baseline"] = {
            "value": round(1.0 / per, 1), "unit": "sums/s", "kind": "port",
            "cores": int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or None,
            "sample": (f"oracle numpy restatement (numpy.linalg.qr/svd = LAPACK geqrf/gesdd, "
                       f"as the reference's qr!/svd!) on {sample} of the items, {reps} passes "
                       f"in {time.perf_counter() - t0:.1f} s")}
    return out


def run_cholesky(args):
    """Hierarchical Cholesky factorization (src/cholesky.jl:12-95) of an SPD
    Gaussian-kernel matrix exp(-((x_i - x_j)/0.25)^2) + I, n = 8192 sorted
    uniform points, stored as the reference's 2 x 2 upper-block tree (leaf
    256, off-diagonal blocks LowRankMatrix truncated at 1e-13 sigma_1). One
    step = the whole factorization on the device tree (the host walks the
    recursion; leaves and updates run in the hm_dense_* and hm_lowrank_op
    kernels). Error: ||R'R - A|| / ||A||, dense, on the GPU. CPU = the
    oracle's numpy/LAPACK restatement of the same recursion."""
    import numpy as np
    import torch
    import hmx
    from hmx import cholesky as HC
    n, leaf = 8192, 256
    x = hmx.workloads.spd_points(n, 0)
    K = hmx.workloads.spd_kernel(x)
    H = hmx.workloads.spd_hierarchical(K, leaf)
    D = HC.to_device(H)
    torch.cuda.synchronize()
    for _ in range(max(1, args.warmup)):
        R = HC.hierarchicalcholesky_device(D)
    steps = max(2, min(args.steps, 10))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        R = HC.hierarchicalcholesky_device(D)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / steps
    Rd = HC.densify(R)
    Ad = HC.densify(D)
    Ad = torch.triu(Ad) + torch.triu(Ad, 1).t()
    err = float(torch.linalg.norm(Rd.t() @ Rd - Ad) / torch.linalg.norm(Ad))
    ranks = []

    def walk(T):
        for row in T.blocks:
            for c in row:
                if isinstance(c, HC.DTree):
                    walk(c)
                elif isinstance(c, HC.DLowRank):
                    ranks.append(int(c.S.shape[0]))
    walk(R)
    out = {"metric": "hierarchical Cholesky factorizations/s, n=8192 Gaussian-kernel SPD",
           "value": round(1.0 / sec, 3), "unit": "factorizations/s",
           "ms_per_step": round(sec * 1e3, 2), "steps": steps, "warmup": args.warmup,
           "dtype": "f64", "data": "synthetic (sorted uniform points, splitmix64 seed 31)",
           "config": {"workload": "hierarchical_cholesky", "n": n, "leaf": leaf,
                      "kernel": "exp(-((x-y)/0.25)^2) + I", "truncation": "1e-13 sigma_1"},
           "rel_err_RtR_vs_A": err, "lowrank_ranks_of_R": [min(ranks), max(ranks)],
           "n_lowrank_blocks_of_R": len(ranks)}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import cholesky as OC

        def to_oracle(Hh):
            if isinstance(Hh, hmx.HierarchicalMatrix):
                T = OC.OTree(*Hh.assigned.shape)
                for m in range(Hh.assigned.shape[0]):
                    for q in range(Hh.assigned.shape[1]):
                        if Hh.assigned[m, q]:
                            T.set(m, q, to_oracle(Hh.blocks[m][q]))
                return T
            if isinstance(Hh, hmx.LowRankMatrix):
                return (Hh.U, Hh.S, Hh.V)
            return np.array(Hh)
        A = to_oracle(H)
        runs = []
        while sum(runs) < 10.0 and len(runs) < 20:
            tc = time.perf_counter()
            OC.chol(A)
            runs.append(time.perf_counter() - tc)
        out["cpu_baseline"] = {
            "value": round(1.0 / float(np.median(runs)), 3), "unit": "factorizations/s",
            "kind": "port", "cores": int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or None,
            "sample": (f"oracle numpy/LAPACK restatement of the recursion on the same tree: "
                       f"median {np.median(runs) * 1e3:.0f} ms of {len(runs)}")}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box here "
              f"has one GPU, and the product needs no collectives).", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import hmx

    if args.copy_calibration:
        n = 1 << 27
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"doubles": n, "bytes_read": 8 * n,
                                               "bytes_written": 8 * n, "launches": 5}}))
        return

    if args.workload != "headline":
        fn = {"batched": run_batched, "gradient": run_gradient, "scale": run_scale,
              "even": run_even, "lowrank": run_lowrank, "cholesky": run_cholesky}[args.workload]
        print(json.dumps(fn(args)), flush=True)
        return

    dev = torch.device("cuda:0")
    w = hmx.workloads.config_headline()
    stream = torch.cuda.current_stream()
    b = torch.from_numpy(w["b"]).to(dev)

    # stored-factor plan (compression timed separately)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    H = hmx.KernelMatrix("cauchykernel", w["x"], w["y"], *w["intervals"], blocksize=BS,
                         rank=RANK)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    info = H.info
    u = torch.empty_like(b)

    def step():
        H._mv(u, b, 1, 1, 1, 1, 0, False, False, stream)

    for _ in range(args.warmup):
        step()
    ms = time_steps(step, args.steps, stream)
    passes = pass_times(H, step, args.steps)

    # sanity against the dense FP64 product on 64 rows (torch, on the GPU)
    rows = torch.arange(0, N, N // 64, device=dev)
    dense = dense_rows(lambda d: 1.0 / d, torch.from_numpy(w["x"]).to(dev)[rows],
                       torch.from_numpy(w["y"]).to(dev), b[:, None])[:, 0]
    err_dense = float(torch.linalg.norm(u[rows] - dense) / torch.linalg.norm(dense))

    # hybrid (stored A, F; Chebyshev-form basis) and matrix-free variants
    variants = {}
    wanted = [v for v in args.variants.split(",") if v]
    import numpy as np
    b32 = b.float()
    for name, kw in (("hybrid", dict(hybrid=True)), ("matrix_free", dict(matrix_free=True)),
                     ("f32_stored", dict(dtype=np.float32)),
                     ("f32_hybrid", dict(dtype=np.float32, hybrid=True))):
        if name not in wanted:
            continue
        is32 = name.startswith("f32")
        # Float32 plans: the same points (exact in Float32 at this n), the
        # tree, factors and products in FP32
        Hv = hmx.KernelMatrix("cauchykernel", w["x"], w["y"], *w["intervals"], blocksize=BS,
                              rank=RANK, **kw)
        bv = b32 if is32 else b
        uv = torch.empty_like(bv)

        def step_v():
            Hv._mv(uv, bv, 1, 1, 1, 1, 0, False, False, stream)

        for _ in range(args.warmup):
            step_v()
        ms_v = time_steps(step_v, args.steps, stream)
        pv = pass_times(Hv, step_v, args.steps)
        # a Float32 stored plan reproduces the reference's Float32 update!,
        # which divides by x - t and gives Inf where a Float32 node equals a
        # point (many do at this n); the count is reported, not an error
        err_v = float(torch.linalg.norm(uv.double() - u) / torch.linalg.norm(u))
        variants[name] = dict(ms=ms_v, passes=pv, bytes=Hv.info["bytes_stored"],
                              err=err_v if np.isfinite(err_v) else None,
                              nonfinite=Hv.info["nonfinite"])
        Hv.close()
        del uv
        torch.cuda.empty_cache()

    names = ["up", "reduce", "translate", "down"]
    r = RANK
    bytes_down = 8 * (info["dense_entries"] + r * info["basis_rows_u"]) + 8 * (2 * N)
    bytes_up = 8 * r * info["basis_rows_v"] + 8 * N
    ach_down = bytes_down / (passes[3] * 1e-3) / 1e9
    roof = dict(bound="hbm", kernel="k_down (near field + U z)", achieved=round(ach_down, 1),
                peak=HBM_PEAK_GBS, unit="GB/s", frac=round(ach_down / HBM_PEAK_GBS, 4),
                traffic=pmc_traffic(), algorithmic_bytes=bytes_down,
                traffic_source="profiles/k_down_cauchy_n2p20_pmc_traffic.md (rocprofv3 --pmc)",
                avg_ms=round(float(passes[3]), 4),
                up_pass=dict(achieved=round(bytes_up / (passes[0] * 1e-3) / 1e9, 1),
                             avg_ms=round(float(passes[0]), 4)))

    def valu_roof(down_ms):
        """The matrix-free downward pass against the FP64 vector-unit issue
        rate: FP64 VALU wave instructions per product (committed PMC counts)
        over the live pass time."""
        instr = pmc_valu()
        if instr is None or not down_ms:
            return None
        ach = instr / (float(down_ms) * 1e-3) / 1e9
        return dict(bound="valu", kernel="k_down matrix-free (near field regenerated + far field)",
                    achieved=round(ach, 1), peak=VALU_F64_PEAK_GINSTR,
                    unit="G FP64 wave-instructions/s", frac=round(ach / VALU_F64_PEAK_GINSTR, 4),
                    instructions_per_product=instr, avg_ms=round(float(down_ms), 4),
                    source="profiles/k_down_matrix_free_n2p20_valu_counters.md (rocprofv3 --pmc)")

    workloads = None
    if not args.no_workloads:
        wsteps, wwarm = max(3, min(args.steps, 10)), max(1, min(args.warmup, 2))
        workloads = {}
        for wname, fn, st, wu in (("batched", summary_batched, wsteps, wwarm),
                                  ("gradient", summary_gradient, wsteps, wwarm),
                                  ("scale", summary_scale, max(2, wsteps // 2), 1)):
            try:  # an error in one workload is reported in its place, not lost with the line
                workloads[wname] = fn(st, wu, stream)
            except Exception as e:  # noqa: BLE001
                workloads[wname] = {"error": f"{type(e).__name__}: {e}"}
            torch.cuda.synchronize()
            torch.cuda.empty_cache()
    peaks = None if args.no_peaks else stream_peaks(stream)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_threads, args.cpu_n_log2)

    out = {
        "metric": "Cauchy H-matvecs/s + HBM GB/s, n=2^20 fp64",
        "value": round(1000.0 / ms, 3),
        "unit": "matvecs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (shifted grid, b from splitmix64 seed 0)",
        "config": {"workload": "cauchy_kernelmatrix_matvec", "n": N, "leaf_blocksize": BS,
                   "rank": RANK, "rhs": 1, "layout": "stored, U/V one per cluster",
                   "parallelism": "single GPU"},
        "hbm_gbs_whole_matvec": round(info["bytes_stored"] / (ms * 1e-3) / 1e9, 1),
        "bytes_per_matvec": info["bytes_stored"],
        "bytes_ref_layout": info["bytes_ref_layout"],
        "pass_ms": {k: round(float(v), 4) for k, v in zip(names, passes)},
        "pass_ms_method": ("each pass timed with HIP events around its own launches, the "
                           "passes back to back on one stream (no side-stream overlap); "
                           "ms_per_step is the whole product as run"),
        "rel_err_vs_dense_64rows": err_dense,
        "build_s": round(build_s, 3),
        "variants": {k: {"value": (round(1000.0 / v["ms"], 3) if not v["nonfinite"] else None),
                         "invalid": (f"{v['nonfinite']} non-finite factors: Float32 nodes "
                                     f"coincide with points (the reference's Float32 update! "
                                     f"divides by zero there)" if v["nonfinite"] else None),
                         "unit": "matvecs/s",
                         "ms_per_step": round(v["ms"], 4),
                         "hbm_bytes_per_matvec": v["bytes"],
                         "pass_ms": {q: round(float(t), 4) for q, t in zip(names, v["passes"])},
                         "rel_diff_vs_stored": v["err"],
                         "nonfinite_factors": v["nonfinite"],
                         **({"valu_roofline": valu_roof(v["passes"][3])}
                            if k == "matrix_free" else {})} for k, v in variants.items()},
        "roofline": roof,
        "workloads": workloads,
        "measured_hbm_peaks": peaks,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
