"""The Gray-Scott method comparison of benchmarks/benchmark_gray_scott.py on
the MI355X kernels, reported in that script's results-JSON schema.

Conventions are the script's: N = 256 points on [0, 1)^2 (dx = 1/N), the
pseudo-spectral symbol -(kx^2 + ky^2) (:52-57), the initial condition of
np.random.seed(42) with a 20 x 20 perturbation (:60-70), Du = 2e-5,
Dv = 1e-5, F = 0.035, k = 0.065, t_final = 10000 with ceil(t_final / dt)
steps, and the median and IQR of n_reps timed runs after a warm-up.

Served methods:
  'IMEX-Strang'  the script's splitting (:268-289: reaction Euler half-step,
                 exact diffusion, reaction Euler half-step), dt = 0.5: one
                 row and one column pass per step (mjx_gs_imex scheme 3, or
                 batch-plan scheme 2 replayed as a hipGraph);
  'ETDRK4'       the script's ETDRK4 (:336-417, contour coefficients with its
                 f2 as written), dt = 1 (mjx_gs_etdrk4 with
                 mjx_etdrk4_coeffs_form(form 1), or batch-plan scheme 0);
  'RK4 (explicit)'  the script's pseudo-spectral RK4 (:89-127) at 90 % of
                 its stability limit, dt = 0.9 * 2.8 / (max(Du, Dv) lam_max)
                 (mjx_gs_rk4_spectral: three passes per stage);
  'CN (Newton-Krylov)'  the script's preconditioned Crank-Nicolson iteration
                 (:162-229, tol 1e-8 on max|r_u| + max|r_v|, at most 10
                 iterations), dt = 0.5 (mjx_gs_cn_fixed_point: five passes per
                 iteration, converged instances masked on the device).
A single 256^2 instance is launch-bound (each kernel moves ~1-2 MB), so by
default the steps are replayed from hipGraphs of ``steps_per_graph`` steps
(batch plans for the first two, captured call sequences for RK4 and CN).

Schema (benchmark_gray_scott.py:73,141-148,312-319,440-447,458-475):
  {"methods": [{"name", "time_s", "iqr_s", "steps", "dt", "t_final"}, ...],
   "config": {"N", "Du", "Dv", "F", "k", "t_final", "scaling_used",
              "n_reps", "dtype", "device", "laplacian"}}
"""
import json
import math
import time
from pathlib import Path
from typing import Dict, Iterable, Optional, Tuple

import numpy as np
import torch

from . import _hip
from .core._arrays import inst_table, spectral_work

N_DEFAULT = 256
DU, DV, F, K = 2e-5, 1e-5, 0.035, 0.065
T_FINAL = 10000.0
METHODS = ("RK4 (explicit)", "CN (Newton-Krylov)", "IMEX-Strang", "ETDRK4")
CN_TOL, CN_MAX_ITER = 1e-8, 10


def method_dt(name: str, N: int = N_DEFAULT) -> float:
    """The step of each method as the script sets it (:86-88, :156, :264,
    :326)."""
    if name == "RK4 (explicit)":
        lam = spectral_symbol(N)
        lam_max = float(np.max(-(lam[:, None] + lam[None, :])))
        return 0.9 * 2.8 / (max(DU, DV) * lam_max)
    fixed = {"CN (Newton-Krylov)": 0.5, "IMEX-Strang": 0.5, "ETDRK4": 1.0}
    if name not in fixed:
        raise ValueError(f"unknown method {name!r}; served: {list(METHODS)}")
    return fixed[name]

METHOD_KEYS = {"name": str, "time_s": float, "iqr_s": float, "steps": int, "dt": float, "t_final": float}
CONFIG_KEYS = {"N": int, "Du": float, "Dv": float, "F": float, "k": float, "t_final": float, "scaling_used": bool,
               "n_reps": int, "dtype": str, "device": str, "laplacian": str}


def benchmark_ic(N: int = N_DEFAULT) -> Tuple[np.ndarray, np.ndarray]:
    """benchmark_gray_scott.py:60-70: u0 = 1 - 0.05 U, v0 = 0.25 + 0.05 U'
    (U, U' the first two draws of the seed-42 generator), then u0 = 0.5,
    v0 = 0.25 on [N/2-10, N/2+10)^2."""
    rs = np.random.RandomState(42)
    u0 = np.ones((N, N)) - 0.5 * rs.rand(N, N) * 0.1
    v0 = 0.25 * np.ones((N, N)) + 0.5 * rs.rand(N, N) * 0.1
    c, r = N // 2, 10
    u0[c - r:c + r, c - r:c + r] = 0.5
    v0[c - r:c + r, c - r:c + r] = 0.25
    return u0, v0


def spectral_symbol(N: int = N_DEFAULT) -> np.ndarray:
    """-(2 pi fftfreq(N, 1/N))^2 per axis (benchmark_gray_scott.py:52-57)."""
    kk = np.fft.fftfreq(N, 1.0 / N) * 2 * np.pi
    return -(kk ** 2)


class GrayScottMethod:
    """One method of the comparison on B = 1 resident (N, N) state;
    ``steps_per_graph`` > 0 replays hipGraphs of that many steps (a batch
    plan of one instance), 0 launches every kernel directly."""

    def __init__(self, name: str, N: int, dtype: torch.dtype, steps_per_graph: int = 100):
        self.name, self.N, self.dtype, self.dt = name, N, dtype, method_dt(name, N)
        self.plan = None
        self.graph, self.spg = None, int(steps_per_graph)
        self.iters = None
        if name in ("RK4 (explicit)", "CN (Newton-Krylov)"):
            self._init_bench_kernel()
            return
        if steps_per_graph > 0:
            from .batch import GrayScottBatch
            scheme = "benchmark_strang" if name == "IMEX-Strang" else "etdrk4"
            self.plan = GrayScottBatch([N], [F], [K], DU, DV, self.dt, scheme=scheme, dtype=dtype,
                                       steps_per_graph=steps_per_graph, cox_matthews=False)
            self.u, self.v = self.plan.state(0)
            return
        lam = spectral_symbol(N)
        self.u = torch.empty((N, N), dtype=dtype, device="cuda")
        self.v = torch.empty_like(self.u)
        lib = _hip.lib()
        if name == "IMEX-Strang":
            self.lam = torch.from_numpy(lam).to("cuda", dtype)
            self.inst = inst_table([[self.dt * DU, self.dt * DV, 0.0, 0.0, F, K, self.dt]], "cuda")
            self.work = spectral_work(dtype, 1, N, N, "cuda")
        else:
            from .core.fft_integrators import etdrk4_coefficients
            self.coef = etdrk4_coefficients(lam, lam, DU, DV, self.dt, dtype, cox_matthews=False)
            self.inst = inst_table([[0.0, 0.0, 0.0, 0.0, F, K, self.dt]], "cuda")
            nbytes = lib.mjx_etdrk4_workspace_bytes(_hip.dtype_code(dtype), 1, N, N)
            self.work = torch.empty(nbytes, dtype=torch.uint8, device="cuda")

    def _init_bench_kernel(self):
        from .batch import SCHEME_STRANG, autotune
        N, dt = self.N, self.dt
        autotune(self.dtype, SCHEME_STRANG, N, 1)  # row / column tiles for one instance
        self.u = torch.empty((N, N), dtype=self.dtype, device="cuda")
        self.v = torch.empty_like(self.u)
        self.lam = torch.from_numpy(spectral_symbol(N)).to("cuda", self.dtype)
        nbytes = _hip.lib().mjx_gs_bench_workspace_bytes(_hip.dtype_code(self.dtype), 1, N, N)
        self.work = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        if self.name == "RK4 (explicit)":
            self.inst = inst_table([[0.0, 0.0, DU, DV, F, K, dt]], "cuda")
        else:
            self.inst = inst_table([[0.0, 0.0, DU, DV, F, K, dt, 1.0]], "cuda")
            self.inst_pc = inst_table([[0.5 * dt * DU, 0.5 * dt * DV, 0.0, 0.0, F, K, dt, 1.0]], "cuda")
            self.iters = torch.zeros(1, dtype=torch.int32, device="cuda")

    def load(self, u0: torch.Tensor, v0: torch.Tensor):
        self.u.copy_(u0)
        self.v.copy_(v0)

    def _direct(self, n_steps: int) -> int:
        lib, N = _hip.lib(), self.N
        code = _hip.dtype_code(self.dtype)
        if self.name == "RK4 (explicit)":
            return lib.mjx_gs_rk4_spectral(code, 1, N, N, _hip.ptr(self.u), _hip.ptr(self.v), N, N * N,
                                           _hip.ptr(self.lam), _hip.ptr(self.lam), _hip.ptr(self.inst),
                                           int(n_steps), _hip.ptr(self.work), _hip.stream_ptr())
        return lib.mjx_gs_cn_fixed_point(code, 1, N, N, _hip.ptr(self.u), _hip.ptr(self.v), N, N * N,
                                         _hip.ptr(self.lam), _hip.ptr(self.lam), _hip.ptr(self.inst),
                                         _hip.ptr(self.inst_pc), CN_TOL, CN_MAX_ITER, int(n_steps),
                                         _hip.ptr(self.iters), _hip.ptr(self.work), _hip.stream_ptr())

    def advance(self, n_steps: int):
        if self.plan is not None:
            self.plan.run(n_steps)
            return
        if self.name in ("RK4 (explicit)", "CN (Newton-Krylov)"):
            K = self.spg
            if K > 0 and n_steps >= K and self.graph is None:
                _hip.check(self._direct(1), self.name)  # first-call set-up outside the capture
                n_steps -= 1
                torch.cuda.synchronize()
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    rc = self._direct(K)
                _hip.check(rc, self.name)
                self.graph = g
            if self.graph is not None:
                for _ in range(n_steps // K):
                    self.graph.replay()
                n_steps %= K
            if n_steps:
                _hip.check(self._direct(n_steps), self.name)
            return
        lib, N = _hip.lib(), self.N
        code = _hip.dtype_code(self.dtype)
        if self.name == "IMEX-Strang":
            rc = lib.mjx_gs_imex(code, 3, 1, N, N, _hip.ptr(self.u), _hip.ptr(self.v), N, N * N, 0,
                                 _hip.ptr(self.lam), _hip.ptr(self.lam), None, _hip.ptr(self.inst), int(n_steps),
                                 _hip.ptr(self.work), _hip.stream_ptr())
        else:
            rc = lib.mjx_gs_etdrk4(code, 1, N, N, _hip.ptr(self.u), _hip.ptr(self.v), N, N * N, 0,
                                   _hip.ptr(self.coef), _hip.ptr(self.inst), int(n_steps), _hip.ptr(self.work),
                                   _hip.stream_ptr())
        _hip.check(rc, self.name)


def run_benchmark(t_final: float = T_FINAL, n_reps: int = 10, dtype: torch.dtype = torch.float64,
                  methods: Iterable[str] = METHODS, N: int = N_DEFAULT,
                  warmup_steps: int = 10, progress=None, steps_per_graph: int = 100) -> Dict:
    """Time each method over ceil(t_final / dt) steps from the benchmark's
    initial condition (already resident on the device when the clock
    starts), n_reps times after a warm-up, and return the results dict."""
    u0, v0 = benchmark_ic(N)
    U0 = torch.from_numpy(u0).to("cuda", dtype)
    V0 = torch.from_numpy(v0).to("cuda", dtype)
    results: Dict = {"methods": []}
    for name in methods:
        m = GrayScottMethod(name, N, dtype, steps_per_graph)
        n_steps = int(math.ceil(t_final / m.dt))
        m.load(U0, V0)
        m.advance(max(warmup_steps, steps_per_graph + 1))  # captures the graph
        times = []
        for _ in range(n_reps):
            m.load(U0, V0)
            torch.cuda.synchronize()
            start = time.perf_counter()
            m.advance(n_steps)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - start)
        q25, q75 = np.percentile(times, [25, 75])
        results["methods"].append({"name": name, "time_s": float(np.median(times)), "iqr_s": float(q75 - q25),
                                   "steps": n_steps, "dt": float(m.dt), "t_final": float(n_steps * m.dt)})
        if progress:
            progress(results["methods"][-1])
    results["config"] = {"N": N, "Du": DU, "Dv": DV, "F": F, "k": K, "t_final": float(t_final),
                         "scaling_used": False, "n_reps": int(n_reps),
                         "dtype": "float64" if dtype == torch.float64 else "float32",
                         "device": f"cuda:0 ({torch.cuda.get_device_properties(0).gcnArchName})",
                         "laplacian": "pseudo-spectral (-k^2)"}
    return results


def validate_results(obj: Dict) -> None:
    """Raise ValueError unless ``obj`` has the benchmark's results schema."""
    if not isinstance(obj, dict) or set(obj) != {"methods", "config"}:
        raise ValueError("results must be {'methods': [...], 'config': {...}}")
    if not isinstance(obj["methods"], list):
        raise ValueError("'methods' must be a list")
    for m in obj["methods"]:
        if not isinstance(m, dict) or set(m) != set(METHOD_KEYS):
            raise ValueError(f"method entry keys must be {sorted(METHOD_KEYS)}")
        for key, typ in METHOD_KEYS.items():
            if typ is float and not isinstance(m[key], (int, float)) or typ is not float and not isinstance(m[key], typ):
                raise ValueError(f"method field {key!r} must be {typ.__name__}")
    cfg = obj["config"]
    if not isinstance(cfg, dict) or set(cfg) != set(CONFIG_KEYS):
        raise ValueError(f"config keys must be {sorted(CONFIG_KEYS)}")
    for key, typ in CONFIG_KEYS.items():
        val = cfg[key]
        ok = isinstance(val, (int, float)) and not isinstance(val, bool) if typ is float else isinstance(val, typ)
        if typ is int and isinstance(val, bool):
            ok = False
        if not ok:
            raise ValueError(f"config field {key!r} must be {typ.__name__}")


def save_results(results: Dict, path) -> Path:
    """Write ``results`` as the benchmark does (json, indent 2)."""
    validate_results(results)
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "w") as f:
        json.dump(results, f, indent=2)
    return path


def load_results(path) -> Dict:
    with open(path) as f:
        obj = json.load(f)
    validate_results(obj)
    return obj


def compare(ours: Dict, reference: Dict) -> Dict[str, Optional[float]]:
    """Speed-up per method name present in both (reference time / ours)."""
    ref = {m["name"]: m for m in reference["methods"]}
    return {m["name"]: (ref[m["name"]]["time_s"] / m["time_s"] if m["name"] in ref else None)
            for m in ours["methods"]}


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--t-final", type=float, default=T_FINAL)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--fp32", action="store_true")
    ap.add_argument("--out", default="results/gray_scott.json")
    ap.add_argument("--steps-per-graph", type=int, default=100, help="0: direct launches")
    a = ap.parse_args()
    res = run_benchmark(a.t_final, a.reps, torch.float32 if a.fp32 else torch.float64,
                        steps_per_graph=a.steps_per_graph,
                        progress=lambda m: print(f"{m['name']}: {m['time_s']:.3f} s +- {m['iqr_s']:.3f} "
                                                 f"({m['steps']} steps, dt={m['dt']})", flush=True))
    print(f"saved {save_results(res, a.out)}")
