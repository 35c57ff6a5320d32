#!/usr/bin/env python3
"""Gray-Scott 256^2 IMEX-Strang throughput on one MI355X.

Workload (one "step" = one IMEX-Strang time step of every instance in the
batch): Grid2D 256x256 on [0,2.5]^2, periodic, Du=2e-5, Dv=1e-5, F=0.035,
k=0.065, dt=1.0; B instances with seeded ICs (seeds 0..B-1), FP32 state,
resident in HBM before the timed region.  Also reported in the same line:
the FP64 single-instance GPU form, the FP64 batched form (the same batch in
FP64, the one form that stays within 1e-5 relative L2 error of an FP64
reference at 10000 steps), the FP32 increment form (the same operator with
only the diffusion increments through the FP32 transforms: within 1e-5 to
3000 steps, against 2000 for the plain FP32 chain), the FP32-vs-FP64
relative L2 error curves of both FP32 forms,
per-step latency percentiles, the roofline of the dominant kernel (HIP events
on the launch stream) and the CPU oracle timed on this host.  Secondary
workloads in the same line: the uneven-size batched ETDRK4 plan (1024
instances of {64,128,256,512}^2), backward-Euler Newton-Krylov on a 512^2
Dirichlet grid (iteration counts per step), the serving loop (Poisson
arrivals at 200/s, p50/p99 request latency, throughput vs sequential FP64),
and the same IMEX-Strang batch on 250^2 (mixed-radix Stockham kernels) and
251^2 (chirp-z kernels) grids.

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


def setup(dtype, B, seeds, n=N):
    import torch
    from moljax_amd.core import Grid2D, create_gray_scott_periodic_fft
    from moljax_amd.core._arrays import inst_table, spectral_work
    from moljax_amd.core.stepping import _gs_rows
    from moljax_amd.workloads import gray_scott_batch
    grid = Grid2D.uniform(n, n, 0.0, 2.5, 0.0, 2.5)
    model, cache, diff = create_gray_scott_periodic_fft(grid, dtype=dtype, **P)
    U, V = gray_scott_batch(n, n, seeds, dtype=dtype)
    rows = inst_table(_gs_rows(model, DT, diff, B))
    work = spectral_work(dtype, B, n, n)
    return cache.laplacian_symbol, U, V, rows, work


def run(U, V, lam, rows, work, steps, scheme=0):
    """IMEX-Strang steps in place: scheme 0 the plain chain, 4 the increment
    form (GS_STRANG_INC; the same operator, only the diffusion increments go
    through the transforms)."""
    from moljax_amd.core.stepping import gs_imex_inplace
    gs_imex_inplace(U, V, lam, rows, steps, scheme, work=work)


def time_form(dtype, B, steps, warmup, kernel_timing=False, n=N, keep_first=False, scheme=0):
    """Seconds for ``steps`` steps of a B-instance batch after ``warmup``
    untimed ones.  With kernel_timing, every launch of the chain is also
    bracketed by HIP events on its stream (a separate pass: the events cost
    time of their own).  keep_first returns instance 0's final (u, v)."""
    import torch
    from moljax_amd import _hip
    lam, U, V, rows, work = setup(dtype, B, list(range(B)), n)
    run(U, V, lam, rows, work, warmup, scheme)
    torch.cuda.synchronize()
    if kernel_timing:
        _hip.lib().mjx_kernel_timing_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(U, V, lam, rows, work, steps, scheme)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kt = kts = None
    if kernel_timing:
        kt = {k: _hip.kernel_timing(k) for k in (KT_ROW_GS, KT_COL)}
        kts = {k: _hip.kernel_timing_samples(k) for k in (KT_ROW_GS, KT_COL)}
        _hip.lib().mjx_kernel_timing_enable(0)
    finite = bool(torch.isfinite(U).all() and torch.isfinite(V).all())
    first = (U[0].double().clone(), V[0].double().clone()) if keep_first else None
    del U, V, work
    if keep_first:
        return t1 - t0, kt, finite, kts, first
    return t1 - t0, kt, finite, kts


def rel_l2_pair(a, b):
    """Relative L2 of the (u, v) pair a against the pair b."""
    import torch
    num = torch.sqrt(((a[0] - b[0]) ** 2).sum() + ((a[1] - b[1]) ** 2).sum())
    return float(num / torch.sqrt((b[0] ** 2).sum() + (b[1] ** 2).sum()))


def fp64_reference_state(n, steps, seed=0):
    """Instance ``seed`` advanced ``steps`` steps alone in FP64 on the GPU
    (the FP64 chain is held to the CPU oracle at 1e-12 by the test suite)."""
    import torch
    lam, U, V, rows, work = setup(torch.float64, 1, [seed], n)
    run(U, V, lam, rows, work, steps)
    return U[0].clone(), V[0].clone()


def error_curve(checkpoints, seed=0):
    """FP32 vs FP64 (both on the GPU path; the FP64 path is pinned to the
    CPU oracle by the test suite) relative L2 of (u, v) at each checkpoint,
    for the plain FP32 chain and for the FP32 increment form."""
    import torch
    from moljax_amd.core.stepping import GS_STRANG, GS_STRANG_INC
    out, out_inc = {}, {}
    forms = {"f32": (torch.float32, GS_STRANG), "inc": (torch.float32, GS_STRANG_INC),
             "f64": (torch.float64, GS_STRANG)}
    st = {k: setup(dt, 1, [seed]) for k, (dt, _) in forms.items()}
    done = 0
    for c in checkpoints:
        for k, (_, sc) in forms.items():
            lam, U, V, rows, work = st[k]
            run(U, V, lam, rows, work, c - done, sc)
        done = c
        U64, V64 = st["f64"][1], st["f64"][2]
        den = torch.sqrt((U64 ** 2).sum() + (V64 ** 2).sum())
        for k, dst in (("f32", out), ("inc", out_inc)):
            U, V = st[k][1], st[k][2]
            num = torch.sqrt(((U.double() - U64) ** 2).sum() + ((V.double() - V64) ** 2).sum())
            dst[str(c)] = float(num / den)
    return out, out_inc


def _cpu_variant(workers, target_s):
    """One warm-up, then the median and IQR of 5 timed repetitions of the
    oracle's IMEX-Strang on one 256^2 FP64 instance (the protocol of
    benchmarks/benchmark_utils.py:112-149)."""
    from oracle import gray_scott as ogs
    from oracle import spectral as osp
    osp.set_fft_workers(workers)
    try:
        y0 = ogs.initial_condition(N, N, seed=0)
        lam = osp.laplacian_symbol_2d(N, N, 2.5 / N, 2.5 / N)
        ogs.integrate(y0, DT, 2, P["F"], P["k"], P["Du"], P["Dv"], lam)  # warm-up (FFT plans, threads)
        t0 = time.perf_counter()
        ogs.integrate(y0, DT, 3, P["F"], P["k"], P["Du"], P["Dv"], lam)
        per = (time.perf_counter() - t0) / 3
        steps = int(max(5, min(2000, target_s / (5 * per))))
        times = []
        for _ in range(5):
            t0 = time.perf_counter()
            ogs.integrate(y0, DT, steps, P["F"], P["k"], P["Du"], P["Dv"], lam)
            times.append(time.perf_counter() - t0)
    finally:
        osp.set_fft_workers(None)
    med = float(np.median(times))
    q25, q75 = np.percentile(times, [25, 75])
    return N * N * steps / med, steps, med, float(q75 - q25)


def cpu_baseline(target_s=10.0):
    """The float64 NumPy oracle (the reference's IMEX-Strang dataflow) on one
    256^2 instance, timed on this host in two variants: numpy.fft on one
    thread, and scipy.fft real-to-complex with the box's CPU share of
    threads; the faster is the reported value."""
    workers = max(1, min(16, os.cpu_count() or 1))
    v1, s1, m1, i1 = _cpu_variant(None, target_s / 2)
    vw, sw, mw, iw = _cpu_variant(workers, target_s / 2)
    variants = [
        {"name": "numpy.fft complex, 1 thread", "value": v1, "cores": 1,
         "sample": f"{s1} steps x 5 reps, median {m1:.2f} s, IQR {i1:.3f} s"},
        {"name": f"scipy.fft rfft2, workers={workers}", "value": vw, "cores": workers,
         "sample": f"{sw} steps x 5 reps, median {mw:.2f} s, IQR {iw:.3f} s"},
    ]
    best = max(variants, key=lambda v: v["value"])
    return {"value": best["value"], "unit": "grid-point-steps/s", "cores": best["cores"], "kind": "port",
            "sample": f"oracle IMEX-Strang FP64, 1 instance 256^2, {best['name']}: {best['sample']}",
            "variants": variants, "cpu_model": cpu_model(), "os_cpu_count": os.cpu_count()}


def step_latency(kt_samples):
    """Per-step device latency of the fused chain (row_gs + col_apply launch
    durations of the same step, HIP events) and its percentiles."""
    row, col = kt_samples[KT_ROW_GS], kt_samples[KT_COL]
    steps = np.array(row[:len(col)]) + np.array(col[:len(row)])
    return {"row_gs_avg_ms": float(np.mean(row)), "col_apply_avg_ms": float(np.mean(col)),
            "step_p50_ms": float(np.percentile(steps, 50)), "step_p99_ms": float(np.percentile(steps, 99)),
            "step_max_ms": float(np.max(steps)), "samples": int(len(steps))}


def _profile_value(name, *keys):
    """The number on the line ``<keys...> <value>`` of a committed rocprofv3
    summary under profiles/, or None when there is no such line."""
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            f = line.split()
            if len(f) == len(keys) + 1 and tuple(f[:-1]) == keys:
                return float(f[-1])
    except OSError:
        pass
    return None


def measured_traffic(kernel, B):
    """HBM bytes per launch of ``kernel`` at batch B from the committed
    rocprofv3 --pmc summary (FETCH_SIZE x2 + WRITE_SIZE x1, copy-kernel
    calibrated), or None when the summary has no matching line."""
    return _profile_value("gray_scott_256_imex_traffic.md", "traffic_bytes_per_launch", kernel, f"B={B}")


def c4_batched_etdrk4(n_inst, steps, seed=0):
    """Uneven-size batched ETDRK4 (1024 instances of {64..512}^2, FP32 state,
    FP64 contour coefficients, hipGraph replay, autotuned tiles)."""
    import torch
    from moljax_amd.batch import GrayScottBatch, c4_workload
    from moljax_amd.workloads import gray_scott_ic
    sizes, F, k = c4_workload(n_inst, seed)
    t0 = time.perf_counter()
    b = GrayScottBatch(sizes, F, k, 2e-5, 1e-5, 1.0, scheme="etdrk4", dtype=torch.float32, steps_per_graph=16)
    setup_s = time.perf_counter() - t0
    for i, n in enumerate(sizes):
        u, v = gray_scott_ic(n, n, i, square=max(2, round(20 * n / 256)))
        b.set_state(i, torch.from_numpy(u[1:-1, 1:-1]), torch.from_numpy(v[1:-1, 1:-1]))
    b.run(2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run(steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    finite = all(bool(torch.isfinite(t).all()) for t in b.u + b.v)
    # FP32 against FP64: the first instance of every size group, advanced
    # the same steps in an FP64 plan of its own
    pick = [b.members[n][0] for n in b.order]
    ref = GrayScottBatch([sizes[i] for i in pick], [F[i] for i in pick], [k[i] for i in pick], 2e-5, 1e-5, 1.0,
                         scheme="etdrk4", dtype=torch.float64, steps_per_graph=16, tune=False)
    for j, i in enumerate(pick):
        n = sizes[i]
        u, v = gray_scott_ic(n, n, i, square=max(2, round(20 * n / 256)))
        ref.set_state(j, torch.from_numpy(u[1:-1, 1:-1]), torch.from_numpy(v[1:-1, 1:-1]))
    ref.run(2 + steps)
    torch.cuda.synchronize()
    errs = {}
    for j, i in enumerate(pick):
        a = tuple(t.double() for t in b.state(i))
        errs[str(sizes[i])] = rel_l2_pair(a, ref.state(j))
    ref.close()
    from moljax_amd import _hip
    tiles = {str(n): [_hip.lib().mjx_tune_get(2, 0, n), _hip.lib().mjx_tune_get(3, 0, n)] for n in b.order}
    gps = b.n_points * steps / el
    # Roofline: the state's 16 B per grid-point step, and the chain's own
    # compulsory traffic (FP32: four row passes of 2 planes, stages of 3, 3,
    # 6 and 3 planes, 8 B per point a plane), beside the bytes per point the
    # committed counter summary measured for the 512^2 group.
    chain = 8.0 * (4 * 2 + 3 + 3 + 6 + 3)
    meas = _profile_value("batched_etdrk4_uneven_counters.md", "traffic_bytes_per_point", "etdrk4_512")
    roof = {"bound": "hbm", "peak": HBM_PEAK, "unit": "B/s", "algorithmic_bytes_per_gps": 16.0,
            "frac": 16.0 * gps / HBM_PEAK, "chain_bytes_per_point": chain, "chain_frac": chain * gps / HBM_PEAK,
            "measured_bytes_per_point_512": meas,
            "traffic_ratio_512": None if meas is None else meas / chain}
    out = {"value": gps, "unit": "grid-point-steps/s", "instances": n_inst, "roofline": roof,
           "points": b.n_points, "state_bytes": b.state_bytes(), "steps": steps, "ms_per_step": 1e3 * el / steps,
           "sizes": {str(n): len(b.members[n]) for n in b.order}, "tiles_row_col": tiles,
           "setup_s": setup_s, "finite": finite,
           "rel_l2_fp32_vs_fp64": errs, "rel_l2_fp32_vs_fp64_max": max(errs.values()),
           "note": "2000-step workload timed on its first steps (per-step cost is constant)"}
    b.close()
    return out


def c3_newton_krylov(n, steps, op_dtype):
    """Backward-Euler Newton-Krylov steps of u_t = D lap u - u^3 + u on an
    n x n Dirichlet vertex grid, GMRES(30) rtol 1e-8, DST preconditioner."""
    import torch
    from moljax_amd.core.newton_krylov import NKParams
    from moljax_amd.core.preconditioners import DSTDiffusionPreconditioner
    from moljax_amd.core.reaction_diffusion import be_step, create_reaction_diffusion_model, sine_mode_ic
    model = create_reaction_diffusion_model(n, D=1e-3, reaction=(-1.0, 0.0, 1.0, 0.0), bc="dirichlet")
    y = {"u": torch.from_numpy(sine_mode_ic(n, seed=0)).cuda()}
    P = DSTDiffusionPreconditioner(field_diffusivity_keys={"u": "D"})
    nkp = NKParams(krylov_tol=1e-8, gmres_restart=30, operator_dtype=op_dtype)
    be_step(model, y, 0.0, 1e-2, P, nkp)  # warm-up
    torch.cuda.synchronize()
    newton, gmres, res = [], [], []
    t0 = time.perf_counter()
    for s in range(steps):
        y, st = be_step(model, y, s * 1e-2, 1e-2, P, nkp)
        newton.append(st.newton_iters)
        gmres.append(st.lin_iters)
        res.append(st.final_res_norm)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    arnoldi = sum(gmres)
    return {"ms_per_step": 1e3 * el / steps, "newton_iters": newton, "gmres_iters": gmres,
            "arnoldi_steps_per_be_step": arnoldi / steps,
            "us_per_arnoldi_step": 1e6 * el / arnoldi if arnoldi else None,
            "max_final_residual": max(res), "operator_dtype": str(op_dtype).replace("torch.", ""),
            "_u": y["u"].double()}


def c5_serving(n_req, rate, seed=0):
    import torch
    from moljax_amd.serving import Server, attach_inputs, latency_stats, poisson_stream, sequential
    reqs = attach_inputs(poisson_stream(n_req, rate=rate, seed=seed), torch.float32)
    srv = Server(dtype=torch.float32, chunk=32)
    srv.run(poisson_stream(8, rate=rate, seed=seed + 1, sizes=(64, 128)), realtime=True)  # warm-up
    srv = Server(dtype=torch.float32, chunk=32)
    t0 = time.perf_counter()
    srv.run(reqs, realtime=True)
    makespan = time.perf_counter() - t0
    gps = sum(r.points * r.steps for r in reqs)
    # saturation: the same requests all waiting at t = 0 (continuous batching
    # at full occupancy) -> serving capacity, against the same requests run
    # one at a time in FP64
    sat = attach_inputs(poisson_stream(n_req, rate=rate, seed=seed), torch.float32)
    for r in sat:
        r.arrival = 0.0
    srv2 = Server(dtype=torch.float32, chunk=32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    srv2.run(sat, realtime=False)
    sat_s = time.perf_counter() - t0
    seq_reqs = attach_inputs(poisson_stream(n_req, rate=rate, seed=seed), torch.float64)
    seq = sequential(seq_reqs, torch.float64, keep=True)
    # every served FP32 result against the same request run alone in FP64:
    # relative to the FP64 result where it kept at least 1e-6 of the input
    # norm, and relative to the input norm for all (a strongly damped linear
    # solve can decay faster than FP32 rounding noise in its slowest mode)
    errs, errs_in = [], []
    for a, b in zip(sorted(reqs, key=lambda r: r.rid), sorted(seq_reqs, key=lambda r: r.rid)):
        num = sum(float(((x.double() - y) ** 2).sum()) for x, y in zip(a.state, b.state)) ** 0.5
        den = sum(float((y ** 2).sum()) for y in b.state) ** 0.5
        inp = sum(float((np.asarray(f, dtype=np.float64) ** 2).sum()) for f in b.ic) ** 0.5
        errs_in.append(num / inp)
        if den >= 1e-6 * inp:
            errs.append(num / den)
    return {"requests": n_req, "rate_per_s": rate, **latency_stats(reqs), "makespan_s": makespan,
            "offered_gps": gps / makespan, "saturated_s": sat_s, "saturated_gps": gps / sat_s,
            "saturated_iterations": srv2.iterations, "saturated_max_group": srv2.max_group,
            "sequential_fp64_s": seq, "sequential_fp64_gps": gps / seq,
            "saturated_vs_sequential_fp64": seq / sat_s, "iterations": srv.iterations,
            "rel_l2_fp32_served_vs_fp64_alone_max": max(errs),
            "rel_l2_fp32_served_vs_fp64_alone_p50": float(np.median(errs)),
            "rel_l2_counted_requests": len(errs),
            "l2_fp32_served_vs_fp64_alone_rel_to_input_max": max(errs_in),
            "requests_per_s_saturated": n_req / sat_s,
            "peak_active": srv.peak_active, "max_group": srv.max_group,
            "mix": {k: sum(r.kind == k for r in reqs) for k in ("fft_cn", "dst", "dct", "imex_gs")},
            "note": "p50/p99: open-loop Poisson arrivals, arrival to result on the device; "
                    "saturated: all requests queued at t=0; gps counts request points x steps, and the "
                    "linear kinds (fft_cn, dst, dct) take each chunk of steps as one spectral power "
                    "multiply, in the served and the sequential runs alike"}


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor()


def kernel_legs(B=512, n=N, reps=50):
    """Per-kernel HBM rates of the row-independent operators at the headline
    shape (B padded FP32 fields of (n+2)^2): the FD Laplacian, central and
    upwind first derivatives (each reads the field once and writes the
    interior, 8 B per interior point), the fused stencil + reaction explicit
    stages (Euler and RK4, ten steps per call), the periodic ghost fill (reads the
    two interior rings it copies and writes the ghost ring) and the interior
    sum of squares (reads 4 B per interior point, FP64 accumulation).  Each
    leg: torch events on the launch stream around `reps` back-to-back calls."""
    import torch
    from moljax_amd.core import Grid2D, laplacian_2d, d1_central_2d, d1_upwind_2d, tree_norm2
    from moljax_amd.core.bc import FieldBCSpec, BCType
    from moljax_amd.core.bc_fill import fill_ghosts
    grid = Grid2D.uniform(n, n, 0.0, 2.5, 0.0, 2.5)
    gen = torch.Generator(device="cuda").manual_seed(0)
    f = torch.rand((B, n + 2, n + 2), generator=gen, device="cuda", dtype=torch.float32)
    pts = B * n * n
    ring = B * 4 * (n + 1) * 4  # ghost values per field ring, FP32
    spec = FieldBCSpec(kind=BCType.PERIODIC)
    from moljax_amd import _hip
    red_out = torch.empty(B, dtype=torch.float64, device="cuda")
    red_work = torch.empty(_hip.lib().mjx_reduce_workspace_elems(B), dtype=torch.float64, device="cuda")
    fi = f[:, 1:, 1:]

    def reduce_kernel():  # the C entry point alone: per-instance sums of squares of the interiors
        _hip.check(_hip.lib().mjx_reduce2d(0, 0, B, n, n, _hip.ptr(fi), n + 2, (n + 2) ** 2, None, 0, 0, 0.0, 0.0,
                                           _hip.ptr(red_out), 0, _hip.ptr(red_work), _hip.stream_ptr()), "reduce2d")
    # fused FD stencil + Gray-Scott reaction stages (mjx_gs_explicit, periodic,
    # no ghost fill), `ksteps` steps per call.  Euler reads and writes both
    # fields once per step (16 B per point); an RK4 step moves 136 B per point
    # (stage 0: 24, stages 1-2: 40 each, stage 3: 32).
    g1, g2 = (torch.rand((B, n + 2, n + 2), generator=gen, device="cuda", dtype=torch.float32) for _ in range(2))
    from moljax_amd.core._arrays import inst_table
    xrows = inst_table([[0.0, 0.0, P["Du"], P["Dv"], P["F"], P["k"], 0.5]] * B)
    xwork = torch.empty(_hip.lib().mjx_gs_explicit_workspace_bytes(0, B, n, n), dtype=torch.uint8, device="cuda")
    ksteps = 10

    def explicit(method):
        _hip.check(_hip.lib().mjx_gs_explicit(0, method, B, n, n, _hip.ptr(g1[:, 1:, 1:]), _hip.ptr(g2[:, 1:, 1:]), n + 2,
                                              (n + 2) ** 2, 0, grid.dy, grid.dx, _hip.ptr(xrows), ksteps,
                                              _hip.ptr(xwork), _hip.stream_ptr()), "gs_explicit")
    legs = {
        "gs_explicit_euler_x10": (lambda: explicit(0), 16 * ksteps * pts),
        "gs_explicit_rk4_x10": (lambda: explicit(2), 136 * ksteps * pts),
        "fd_laplacian_2d": (lambda: laplacian_2d(f, grid), 8 * pts),
        "fd_d1_central_2d_x": (lambda: d1_central_2d(f, grid, 1), 8 * pts),
        "fd_d1_upwind_2d_y": (lambda: d1_upwind_2d(f, 0.7, grid, 0), 8 * pts),
        "bc_fill_periodic": (lambda: fill_ghosts(f, grid, spec, 0.0, None), 2 * ring),
        "tree_norm2_interior": (lambda: tree_norm2({"u": f}, grid), 4 * pts),
        "reduce2d_sumsq_entry_point": (reduce_kernel, 4 * pts),
        # reference: a device copy of the same padded array (reads + writes it)
        "device_copy_same_shape": (lambda: f.clone(), 8 * B * (n + 2) ** 2),
    }
    out = {}
    for name, (fn, nbytes) in legs.items():
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        s = e0.elapsed_time(e1) / 1e3 / reps
        out[name] = {"us_per_call": s * 1e6, "algorithmic_bytes": nbytes, "GB_per_s": nbytes / s / 1e9,
                     "roofline_frac": nbytes / s / HBM_PEAK}
    return {"shape": [B, n + 2, n + 2], "dtype": "f32", "reps": reps, "legs": out}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--fp64-steps", type=int, default=2000)
    ap.add_argument("--fp64-batch", type=int, default=512,
                    help="batch of the FP64-batched leg (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-curve", action="store_true")
    ap.add_argument("--sweep", type=str, default="1,8,64,2048",
                    help="comma-separated batch sizes to time as well ('' to skip)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (byte-counter calibration)")
    ap.add_argument("--anylen-n", type=str, default="250,251",
                    help="comma-separated sides of the non-power-of-two IMEX-Strang legs at the headline batch "
                         "(250: mixed-radix Stockham kernels, 251: chirp-z kernels; '' or 0 to skip)")
    ap.add_argument("--c4-steps", type=int, default=100, help="steps of the uneven batched ETDRK4 workload (0: skip)")
    ap.add_argument("--c4-instances", type=int, default=1024)
    ap.add_argument("--c3-steps", type=int, default=10, help="Newton-Krylov backward-Euler steps (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=200, help="serving-loop requests (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--kernel-reps", type=int, default=50,
                    help="calls per leg of the FD / ghost-fill / reduction kernel rates (0: skip)")
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
    t32, _, fin32, _ = time_form(torch.float32, B, K, W)
    gps32 = B * N * N * K / t32
    ms_step = 1e3 * t32 / K
    # per-launch HIP events in a pass of their own (the events cost time)
    ke = max(10, min(K, 100))
    log(f"fp32 batched B={B}: kernel events over {ke} steps")
    _, kt, _, kts = time_form(torch.float32, B, ke, W, kernel_timing=True)

    log(f"fp64 single steps={args.fp64_steps}")
    t64, _, fin64, _ = time_form(torch.float64, 1, args.fp64_steps, W)
    gps64 = N * N * args.fp64_steps / t64

    fp64b = None
    if args.fp64_batch > 0:
        b64, k64 = args.fp64_batch, max(10, K // 2)
        log(f"fp64 batched B={b64} steps={k64}")
        tb, _, finb, _ = time_form(torch.float64, b64, k64, W)
        fp64b = {"value": b64 * N * N * k64 / tb, "unit": "grid-point-steps/s", "batch": b64, "steps": k64,
                 "ms_per_step": 1e3 * tb / k64, "finite": finb}

    log(f"fp32 increment form batched B={B} steps={K}")
    tinc, _, fininc, _ = time_form(torch.float32, B, K, W, scheme=4)
    inc = {"value": B * N * N * K / tinc, "unit": "grid-point-steps/s", "batch": B, "steps": K,
           "ms_per_step": 1e3 * tinc / K, "finite": fininc,
           "form": "IMEX-Strang increment form (scheme 4): state kept physical as (u - 1, v), only the diffusion "
                   "increments (H^k - 1) y through the FP32 transforms; 48 B/GPS moved against 32"}

    # roofline of the dominant kernel: the step's compulsory 16 B/GPS (FP32
    # u, v read + written once) is split evenly over its two kernels.
    (ms_gs, n_gs), (ms_col, n_col) = kt[KT_ROW_GS], kt[KT_COL]
    dom = "row_gs" if ms_gs >= ms_col else "col_apply"
    ms_dom, n_dom = (ms_gs, n_gs) if dom == "row_gs" else (ms_col, n_col)
    avg_s = ms_dom / n_dom / 1e3
    algo_bytes = 8.0 * B * N * N
    achieved = algo_bytes / avg_s
    lat = step_latency(kts)
    traffic = measured_traffic(dom, B)

    sweep = {}
    for b in [int(x) for x in args.sweep.split(",") if x]:
        log(f"sweep B={b}")
        t, _, _, _ = time_form(torch.float32, b, K, W)
        sweep[str(b)] = b * N * N * K / t

    anylen = None
    mixed = (96, 100, 192, 200, 250, 384, 500, 768, 1000)
    for n2 in [int(x) for x in args.anylen_n.split(",") if x and int(x) > 0]:
        log(f"any-length IMEX-Strang {n2}^2 B={B}")
        ta, _, fina, _, first = time_form(torch.float32, B, K, W, n=n2, keep_first=True)
        ref = fp64_reference_state(n2, W + K, seed=0)
        leg = {"grid": [n2, n2], "batch": B, "steps": K, "value": B * n2 * n2 * K / ta,
               "kernels": "mixed-radix Stockham" if n2 in mixed else "chirp-z",
               "unit": "grid-point-steps/s", "ms_per_step": 1e3 * ta / K, "finite": fina,
               "vs_power_of_two_headline": (B * n2 * n2 * K / ta) / gps32,
               "rel_l2_fp32_vs_fp64_instance0": rel_l2_pair(first, ref), "rel_l2_after_steps": W + K}
        if anylen is None:
            anylen = leg
        else:
            anylen.setdefault("other_sides", []).append(leg)

    curve = curve_inc = None
    if not args.no_curve:
        log("error curves")
        curve, curve_inc = error_curve([100, 1000, 2000, 3000, 5000, 10000])
        inc["rel_l2_vs_fp64"] = curve_inc

    c4 = c3 = c5 = None
    if args.c4_steps > 0:
        log(f"c4 batched uneven ETDRK4 ({args.c4_instances} instances, {args.c4_steps} steps)")
        c4 = c4_batched_etdrk4(args.c4_instances, args.c4_steps)
    if args.c3_steps > 0:
        log(f"c3 Newton-Krylov 512^2 Dirichlet, {args.c3_steps} BE steps")
        # 512 interior points (DST-I length 513: the chirp-z kernels), and
        # 511 beside it (length 512: the Stockham kernels)
        c3 = {"grid": [512, 512], "h": 1.0 / 513, "dt": 1e-2,
              "fp32_operator": c3_newton_krylov(512, args.c3_steps, torch.float32),
              "fp64_operator": c3_newton_krylov(512, args.c3_steps, torch.float64),
              "grid_511_fp32_operator": c3_newton_krylov(511, args.c3_steps, torch.float32)}
        u32, u64 = c3["fp32_operator"].pop("_u"), c3["fp64_operator"].pop("_u")
        c3["grid_511_fp32_operator"].pop("_u")
        c3["rel_l2_fp32_operator_vs_fp64_operator"] = float(torch.linalg.norm(u32 - u64) / torch.linalg.norm(u64))
    if args.serve_requests > 0:
        log(f"c5 serving {args.serve_requests} requests at {args.serve_rate}/s")
        c5 = c5_serving(args.serve_requests, args.serve_rate)

    kl = None
    if args.kernel_reps > 0:
        log("FD operators, ghost fill, reduction at the headline shape")
        kl = kernel_legs(B, N, args.kernel_reps)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline()

    horizon = None
    if curve:
        ok = [int(c) for c, e in curve.items() if e <= 1e-5]
        horizon = max(ok) if ok else 0
    # which form meets the metric's 1e-5 at config 2's 10 000-step horizon:
    # the fastest of the forms whose measured error there is within it (the
    # FP64 batch is held to the oracle at 1e-12 step for step)
    target = None
    if curve and fp64b:
        e10k, e10k_inc = curve.get("10000"), curve_inc.get("10000")
        cands = [("fp64_batched", fp64b["value"])]
        if e10k is not None and e10k <= 1e-5:
            cands.append(("fp32_batched", gps32))
        if e10k_inc is not None and e10k_inc <= 1e-5:
            cands.append(("fp32_increment_batched", inc["value"]))
        form = max(cands, key=lambda c: c[1])
        target = {"horizon_steps": 10000, "tolerance": 1e-5, "form": form[0], "value": form[1],
                  "unit": "grid-point-steps/s", "ratio_to_fp64_single": form[1] / gps64,
                  "fp32_batched_rel_l2_at_10000": e10k, "fp32_increment_rel_l2_at_10000": e10k_inc,
                  "fp64_rel_l2_vs_oracle": "<= 1e-12 per 20 steps (test suite, step for step)",
                  "note": "the pattern-forming dynamics amplify every rounding after ~2000 steps: the plain FP32 "
                          "chain leaves 1e-5 between 2000 and 3000 steps, the FP32 increment form after 3000 "
                          "(1.2e-5..5.5e-5 at 10 000 over five seeds); the FP64 batch carries the 1e-5 target "
                          "to 10 000 steps"}
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
        "fp64_batched": fp64b,
        "fp32_increment_batched": inc,
        "speedup_fp32_increment_vs_fp64_single": inc["value"] / gps64,
        "speedup_fp64_batched_vs_fp64_single": fp64b["value"] / gps64 if fp64b else None,
        "form_within_1e-5_at_10000_steps": target,
        "rel_l2_fp32_vs_fp64": curve,
        "steps_within_1e-5": horizon,
        "finite": fin32 and fin64,
        "step_latency_ms": lat,
        "p99_step_latency_ms": lat["step_p99_ms"],
        "batch_sweep_gps": sweep or None,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic, "kernel": dom,
                     "algorithmic_bytes_per_launch": algo_bytes, "avg_launch_ms": avg_s * 1e3,
                     "step_frac": 16.0 * gps32 / HBM_PEAK},
        "any_length_imex_strang": anylen,
        "batched_uneven_etdrk4": c4,
        "newton_krylov_512": c3,
        "serving": c5,
        "kernel_rates": kl,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
