#!/usr/bin/env python3
"""Gray-Scott 256^2 IMEX-Strang throughput on one MI355X.

Workload (one "step" = one IMEX-Strang time step of every instance in the
batch): Grid2D 256x256 on [0,2.5]^2, periodic, Du=2e-5, Dv=1e-5, F=0.035,
k=0.065, dt=1.0; B instances with seeded ICs (seeds 0..B-1), FP32 state,
resident in HBM before the timed region.  Also reported in the same line:
the FP64 single-instance GPU form, the FP32-vs-FP64 relative L2 error curve,
per-step latency percentiles, the roofline of the dominant kernel (HIP events
on the launch stream) and the CPU oracle timed on this host.

Prints ONE JSON line.  Only the cpu_baseline leg touches the oracle.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gray-Scott 256² grid-pt-steps/s FP32-batched vs FP64 @1e-5 err; p99 latency"
N = 256
P = dict(Du=2e-5, Dv=1e-5, F=0.035, k=0.065)
DT = 1.0
HBM_PEAK = 8.0e12
KT_ROW_GS, KT_COL = 0, 1


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def setup(dtype, B, seeds):
    import torch
    from moljax_amd.core import Grid2D, create_gray_scott_periodic_fft
    from moljax_amd.core._arrays import inst_table, spectral_work
    from moljax_amd.core.stepping import _gs_rows
    from moljax_amd.workloads import gray_scott_batch
    grid = Grid2D.uniform(N, N, 0.0, 2.5, 0.0, 2.5)
    model, cache, diff = create_gray_scott_periodic_fft(grid, dtype=dtype, **P)
    U, V = gray_scott_batch(N, N, seeds, dtype=dtype)
    rows = inst_table(_gs_rows(model, DT, diff, B))
    work = spectral_work(dtype, B, N, N)
    return cache.laplacian_symbol, U, V, rows, work


def run(U, V, lam, rows, work, steps):
    from moljax_amd.core.stepping import gs_imex_strang_inplace
    gs_imex_strang_inplace(U, V, lam, rows, steps, work=work)


def time_form(dtype, B, steps, warmup, kernel_timing=False):
    import torch
    from moljax_amd import _hip
    lam, U, V, rows, work = setup(dtype, B, list(range(B)))
    run(U, V, lam, rows, work, warmup)
    torch.cuda.synchronize()
    if kernel_timing:
        _hip.lib().mjx_kernel_timing_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(U, V, lam, rows, work, steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kt = None
    if kernel_timing:
        kt = {k: _hip.kernel_timing(k) for k in (KT_ROW_GS, KT_COL)}
        _hip.lib().mjx_kernel_timing_enable(0)
    finite = bool(torch.isfinite(U).all() and torch.isfinite(V).all())
    del U, V, work
    return t1 - t0, kt, finite


def error_curve(checkpoints, seed=0):
    """FP32 vs FP64 (both on the GPU path; the FP64 path is pinned to the
    CPU oracle by the test suite) relative L2 of (u, v) at each checkpoint."""
    import torch
    out = {}
    l32, U32, V32, r32, w32 = setup(torch.float32, 1, [seed])
    l64, U64, V64, r64, w64 = setup(torch.float64, 1, [seed])
    done = 0
    for c in checkpoints:
        run(U32, V32, l32, r32, w32, c - done)
        run(U64, V64, l64, r64, w64, c - done)
        done = c
        num = torch.sqrt(((U32.double() - U64) ** 2).sum() + ((V32.double() - V64) ** 2).sum())
        den = torch.sqrt((U64 ** 2).sum() + (V64 ** 2).sum())
        out[str(c)] = float(num / den)
    return out


def cpu_baseline(target_s=10.0):
    """The float64 NumPy oracle (reference dataflow, numpy.fft, one thread)
    on one 256^2 instance, timed on this host for a bounded number of steps."""
    from oracle import gray_scott as ogs
    from oracle import spectral as osp
    y0 = ogs.initial_condition(N, N, seed=0)
    lam = osp.laplacian_symbol_2d(N, N, 2.5 / N, 2.5 / N)
    t0 = time.perf_counter()
    y, _ = ogs.integrate(y0, DT, 3, P["F"], P["k"], P["Du"], P["Dv"], lam)
    per = (time.perf_counter() - t0) / 3
    steps = int(max(10, min(3000, target_s / per)))
    t0 = time.perf_counter()
    y, _ = ogs.integrate(y0, DT, steps, P["F"], P["k"], P["Du"], P["Dv"], lam)
    el = time.perf_counter() - t0
    return {"value": N * N * steps / el, "unit": "grid-point-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle IMEX-Strang FP64 (numpy.fft, 1 thread), 1 instance 256^2, {steps} steps, {el:.1f} s",
            "cpu_model": cpu_model(), "os_cpu_count": os.cpu_count()}


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--fp64-steps", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-curve", action="store_true")
    ap.add_argument("--sweep", type=str, default="", help="comma-separated batch sizes to time as well")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        if int(os.environ.get("RANK", "0")) == 0:
            print("bench.py measures one GPU: every instance is independent and the multi-GPU layout "
                  "(instances partitioned by HBM footprint, no communication) is not built in this "
                  "single-GPU build; run with --gpus 1", flush=True)
        sys.exit(2)

    import torch
    import moljax_amd
    moljax_amd.load_library(True)
    B, K, W = args.batch, args.steps, args.warmup
    if args.calib_copy_mib:
        n = args.calib_copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    log(f"fp32 batched B={B} steps={K}")
    t32, kt, fin32 = time_form(torch.float32, B, K, W, kernel_timing=True)
    gps32 = B * N * N * K / t32
    ms_step = 1e3 * t32 / K

    log(f"fp64 single steps={args.fp64_steps}")
    t64, _, fin64 = time_form(torch.float64, 1, args.fp64_steps, W)
    gps64 = N * N * args.fp64_steps / t64

    # roofline of the dominant kernel: the step's compulsory 16 B/GPS (FP32
    # u, v read + written once) is split evenly over its two kernels.
    (ms_gs, n_gs), (ms_col, n_col) = kt[KT_ROW_GS], kt[KT_COL]
    dom = "row_gs" if ms_gs >= ms_col else "col_apply"
    ms_dom, n_dom = (ms_gs, n_gs) if dom == "row_gs" else (ms_col, n_col)
    avg_s = ms_dom / n_dom / 1e3
    algo_bytes = 8.0 * B * N * N
    achieved = algo_bytes / avg_s
    step_latency = {"row_gs_avg_ms": ms_gs / n_gs, "col_apply_avg_ms": ms_col / n_col}

    sweep = {}
    for b in [int(x) for x in args.sweep.split(",") if x]:
        log(f"sweep B={b}")
        t, _, _ = time_form(torch.float32, b, K, W)
        sweep[str(b)] = b * N * N * K / t

    curve = None
    if not args.no_curve:
        log("error curve")
        curve = error_curve([100, 1000, 2000, 3000, 5000, 10000])

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline()

    horizon = None
    if curve:
        ok = [int(c) for c, e in curve.items() if e <= 1e-5]
        horizon = max(ok) if ok else 0
    line = {
        "metric": METRIC,
        "value": gps32,
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded Gray-Scott ICs (u=1,v=0, centred 20x20 square u=0.5,v=0.25, U(-0.01,0.01) noise, "
                "numpy default_rng(seed), seeds 0..B-1)",
        "config": {"workload": "gray_scott_imex_strang_periodic_256x256_fp32_batched", "grid": [N, N],
                   "domain": [0.0, 2.5], "batch": B, "dt": DT, **P, "state_bytes": 2 * 4 * B * (N + 2) ** 2},
        "fp64_single_instance": {"value": gps64, "unit": "grid-point-steps/s", "steps": args.fp64_steps,
                                 "ms_per_step": 1e3 * t64 / args.fp64_steps},
        "speedup_fp32_batched_vs_fp64_single": gps32 / gps64,
        "rel_l2_fp32_vs_fp64": curve,
        "steps_within_1e-5": horizon,
        "finite": fin32 and fin64,
        "step_latency_ms": step_latency,
        "batch_sweep_gps": sweep or None,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None, "kernel": dom,
                     "algorithmic_bytes_per_launch": algo_bytes, "avg_launch_ms": avg_s * 1e3,
                     "step_frac": 16.0 * gps32 / HBM_PEAK},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
