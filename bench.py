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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-n-log2", type=int, default=17)
    ap.add_argument("--variants", default="hybrid,matrix_free",
                    help="extra variants timed beside the stored product")
    ap.add_argument("--no-peaks", action="store_true", help="skip the torch bandwidth probes")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run torch copies of 2^27 doubles (for PMC byte calibration)")
    return ap.parse_args()


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


def cpu_baseline(threads, n_log2):
    """Oracle traversal mul! (ref layout, bit-identical to the single-thread
    restatement) on the n = 2^n_log2 instance of the same family; matvecs/s
    scaled to n = 2^20 by the ratio of reference-layout bytes."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import hmx
    import oracle as O
    n = 1 << n_log2
    w = hmx.workloads.config_plain(n=n, blocksize=BS, rank=RANK)
    t = O.Tree(O.CAUCHY, w["x"], w["y"], *w["intervals"], BS, RANK, store=True)
    st = t.stats()
    ref_bytes = lambda s, nn: 8 * (s["dense_entries"] + RANK * (s["far_rows"] + s["far_cols"])
                                   + s["far"] * RANK * RANK) + 16 * nn
    b_sample = ref_bytes(st, n)
    t.matvec_omp(w["b"], nthreads=threads)
    runs = []
    for _ in range(5):
        t0 = time.perf_counter()
        t.matvec_omp(w["b"], nthreads=threads)
        runs.append(time.perf_counter() - t0)
    sec = float(np.median(runs))
    t0 = time.perf_counter()
    t.matvec(w["b"])
    sec1 = time.perf_counter() - t0
    del t
    big = hmx.KernelMatrix("cauchykernel", *hmx.workloads.shifted_grid(N)[:2],
                           *hmx.workloads.shifted_grid(N)[2], blocksize=BS, rank=RANK,
                           tree_only=True).info
    b_full = big["bytes_ref_layout"]
    scale = b_sample / b_full
    return dict(value=(1.0 / sec) * scale, unit="matvecs/s", cores=threads, kind="port",
                sample=(f"oracle traversal mul! (OpenMP, {threads} threads, ref layout) on the "
                        f"n=2^{n_log2} instance of the same family (leaf 256, rank 24): median "
                        f"{sec * 1e3:.1f} ms of 5; scaled to n=2^20 by ref-layout bytes "
                        f"{b_sample}/{b_full}"),
                single_thread_value=(1.0 / sec1) * scale)


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

    # quick sanity against the dense product on a few rows
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    rows = np.arange(0, N, N // 16)
    dense = O.dense_matvec(O.CAUCHY, w["x"], w["y"], w["b"], rows=rows)
    err_dense = float(np.linalg.norm(u.cpu().numpy()[rows] - dense) / np.linalg.norm(dense))

    # hybrid (stored A, F; Chebyshev-form basis) and matrix-free variants
    variants = {}
    wanted = [v for v in args.variants.split(",") if v]
    for name, kw in (("hybrid", dict(hybrid=True)), ("matrix_free", dict(matrix_free=True))):
        if name not in wanted:
            continue
        Hv = hmx.KernelMatrix("cauchykernel", w["x"], w["y"], *w["intervals"], blocksize=BS,
                              rank=RANK, **kw)
        uv = torch.empty_like(b)

        def step_v():
            Hv._mv(uv, b, 1, 1, 1, 1, 0, False, False, stream)

        for _ in range(args.warmup):
            step_v()
        ms_v = time_steps(step_v, args.steps, stream)
        pv = pass_times(Hv, step_v, args.steps)
        variants[name] = dict(ms=ms_v, passes=pv, bytes=Hv.info["bytes_stored"],
                              err=float(torch.linalg.norm(uv - u) / torch.linalg.norm(u)))
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

    peaks = None if args.no_peaks else stream_peaks(stream)
    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        cpu = cpu_baseline(threads, args.cpu_n_log2)

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
        "rel_err_vs_dense_16rows": err_dense,
        "build_s": round(build_s, 3),
        "variants": {k: {"value": round(1000.0 / v["ms"], 3), "unit": "matvecs/s",
                         "ms_per_step": round(v["ms"], 4),
                         "hbm_bytes_per_matvec": v["bytes"],
                         "pass_ms": {q: round(float(t), 4) for q, t in zip(names, v["passes"])},
                         "rel_diff_vs_stored": v["err"]} for k, v in variants.items()},
        "roofline": roof,
        "measured_hbm_peaks": peaks,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
