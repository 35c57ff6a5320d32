"""Many Gray-Scott instances of uneven grid sizes advanced together on one
GPU (build-only; the nearest reference code batches fields of one grid,
moljax/core/jit_kernels.py:228-262,356-531 and
moljax/core/fft_integrators.py:473-510).

Instances are grouped by size.  Each group keeps its state resident in one
contiguous (B_g, n, n) pair of device arrays, its parameter rows, its
coefficient or symbol tables and its workspace; mjx_batch_run advances every
group by replaying one hipGraph whose branches are the groups.  Tiles come
from the autotuner (mjx_autotune), run once per (precision, scheme, size) in
a process.  Every instance's result is bitwise that of the same instance
advanced alone.
"""
import ctypes
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _hip
from .core._arrays import inst_table

SCHEME_ETDRK4, SCHEME_STRANG, SCHEME_BENCH_STRANG = 0, 1, 2
SCHEME_HELM_PERIODIC, SCHEME_HELM_DIRICHLET, SCHEME_HELM_NEUMANN = 3, 4, 5
_SCHEMES = {"etdrk4": SCHEME_ETDRK4, "strang": SCHEME_STRANG, "benchmark_strang": SCHEME_BENCH_STRANG}
_HELM = {"periodic": SCHEME_HELM_PERIODIC, "dirichlet": SCHEME_HELM_DIRICHLET, "neumann": SCHEME_HELM_NEUMANN}
_TUNED: Dict[tuple, Tuple[int, int]] = {}


def spectral_symbol_1d(n: int, L: float = 1.0) -> np.ndarray:
    """-(2 pi fftfreq(n, L/n))^2 (benchmarks/benchmark_gray_scott.py:52-57)."""
    k = np.fft.fftfreq(n, L / n) * 2 * np.pi
    return -(k ** 2)


def fd_symbol_1d(n: int, L: float = 1.0) -> np.ndarray:
    """(2 cos(k h) - 2) / h^2 with h = L / n (the 5-point Laplacian)."""
    h = L / n
    k = 2.0 * np.pi * np.fft.fftfreq(n, d=h)
    return (2.0 * np.cos(k * h) - 2.0) / (h * h)


def autotune(dtype: torch.dtype, scheme: int, n: int, B: int, reps: int = 3) -> Tuple[int, int]:
    """Tile lines (row kernels, column kernels) chosen by timing on the
    device; cached per process."""
    key = (dtype, scheme, n)
    hit = _TUNED.get(key)
    if hit is not None:
        return hit
    chosen = (ctypes.c_int * 2)()
    _hip.check(_hip.lib().mjx_autotune(_hip.dtype_code(dtype), scheme, n, int(B), int(reps), chosen,
                                       _hip.stream_ptr()), "mjx_autotune")
    _TUNED[key] = (chosen[0], chosen[1])
    return _TUNED[key]


class GrayScottBatch:
    """Uneven-size Gray-Scott instances advanced by one batched plan.

    sizes[i] is instance i's square grid size; F[i], k[i] its feed and kill
    rates; Du, Dv, dt are shared (one coefficient table per size).
    scheme 'etdrk4' (Cox-Matthews, FP64 contour coefficients; with
    ``cox_matthews=False`` the benchmark's f2 as written), 'strang'
    (IMEX-Strang with implicit diffusion half-steps) or 'benchmark_strang'
    (the Gray-Scott benchmark's reaction / exact diffusion / reaction
    splitting, benchmarks/benchmark_gray_scott.py:268-289).  symbol
    'spectral' (-k^2, the benchmark's) or 'fd' (5-point eigenvalues); the
    domain is [0, L)^2 for every instance."""

    def __init__(self, sizes: Sequence[int], F: Sequence[float], k: Sequence[float], Du: float, Dv: float,
                 dt: float, scheme: str = "etdrk4", dtype: torch.dtype = torch.float32, L: float = 1.0,
                 symbol: str = "spectral", steps_per_graph: int = 16, tune: bool = True,
                 cox_matthews: bool = True):
        if scheme not in _SCHEMES:
            raise ValueError(f"scheme must be one of {sorted(_SCHEMES)}")
        if len(F) != len(sizes) or len(k) != len(sizes):
            raise ValueError("one F and one k per instance")
        self.scheme = _SCHEMES[scheme]
        self.dtype = dtype
        self.sizes = [int(s) for s in sizes]
        self.order = sorted(set(self.sizes))
        sym = spectral_symbol_1d if symbol == "spectral" else fd_symbol_1d
        self._slot: List[Tuple[int, int]] = []  # instance -> (group, index)
        members: Dict[int, List[int]] = {n: [] for n in self.order}
        for i, n in enumerate(self.sizes):
            self._slot.append((self.order.index(n), len(members[n])))
            members[n].append(i)
        self.members = members
        lib = _hip.lib()
        self.u, self.v, self.rows, self.tables, self.work = [], [], [], [], []
        for n in self.order:
            idx = members[n]
            B = len(idx)
            self.u.append(torch.zeros((B, n, n), dtype=dtype, device="cuda"))
            self.v.append(torch.zeros((B, n, n), dtype=dtype, device="cuda"))
            lam = sym(n, L)
            if self.scheme == SCHEME_ETDRK4:
                from .core.fft_integrators import etdrk4_coefficients
                coef = etdrk4_coefficients(lam, lam, Du, Dv, dt, dtype, cox_matthews)
                self.tables.append((coef, None))
                rows = [[0, 0, 0, 0, F[i], k[i], dt] for i in idx]
                nbytes = lib.mjx_etdrk4_workspace_bytes(_hip.dtype_code(dtype), B, n, n)
            else:
                t = torch.from_numpy(lam).to("cuda", dtype)
                self.tables.append((t, t))
                h = 0.5 * dt if self.scheme == SCHEME_STRANG else dt
                rows = [[h * Du, h * Dv, 0, 0, F[i], k[i], dt] for i in idx]
                nbytes = lib.mjx_spectral_workspace_bytes(_hip.dtype_code(dtype), B, n, n)
            self.rows.append(inst_table(rows))
            self.work.append(torch.empty(nbytes, dtype=torch.uint8, device="cuda"))
            if tune:
                autotune(dtype, self.scheme, n, B)
        G = len(self.order)
        arr = lambda ct, vals: (ct * G)(*vals)
        tabs = []
        for t0, t1 in self.tables:
            tabs += [_hip.ptr(t0), _hip.ptr(t1)]
        self._plan = ctypes.c_void_p()
        rc = lib.mjx_batch_create(
            _hip.dtype_code(dtype), self.scheme, G, arr(ctypes.c_int, self.order),
            arr(ctypes.c_int, [len(members[n]) for n in self.order]),
            arr(ctypes.c_void_p, [_hip.ptr(x) for x in self.u]), arr(ctypes.c_void_p, [_hip.ptr(x) for x in self.v]),
            arr(ctypes.c_longlong, self.order), arr(ctypes.c_longlong, [n * n for n in self.order]),
            (ctypes.c_void_p * (2 * G))(*tabs), arr(ctypes.c_void_p, [_hip.ptr(r) for r in self.rows]),
            arr(ctypes.c_void_p, [_hip.ptr(w) for w in self.work]), int(steps_per_graph), ctypes.byref(self._plan))
        _hip.check(rc, "mjx_batch_create")

    @property
    def n_points(self) -> int:
        return sum(n * n for n in self.sizes)

    def state_bytes(self) -> int:
        es = 8 if self.dtype == torch.float64 else 4
        return 2 * es * self.n_points

    def set_state(self, i: int, u, v):
        g, j = self._slot[i]
        self.u[g][j].copy_(torch.as_tensor(u))
        self.v[g][j].copy_(torch.as_tensor(v))

    def state(self, i: int) -> Tuple[torch.Tensor, torch.Tensor]:
        g, j = self._slot[i]
        return self.u[g][j], self.v[g][j]

    def run(self, n_steps: int):
        _hip.check(_hip.lib().mjx_batch_run(self._plan, int(n_steps), 0, _hip.stream_ptr()), "mjx_batch_run")

    def close(self):
        if self._plan:
            _hip.lib().mjx_batch_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def c4_workload(n_instances: int = 1024, seed: int = 0, sizes=(64, 128, 256, 512)):
    """Sizes, F, k of the uneven batched ETDRK4 workload: sizes uniform over
    ``sizes``, F ~ U(0.02, 0.06), k ~ U(0.055, 0.068) from
    numpy.random.default_rng(seed)."""
    rng = np.random.default_rng(seed)
    sz = rng.choice(np.asarray(sizes), size=n_instances)
    F = rng.uniform(0.02, 0.06, size=n_instances)
    k = rng.uniform(0.055, 0.068, size=n_instances)
    return [int(s) for s in sz], F.tolist(), k.tolist()


def dirichlet_symbol_1d(n: int, L: float = 1.0) -> np.ndarray:
    """Eigenvalues -4/h^2 sin^2(pi k / (2(n+1))), k = 1..n, of the 1-D
    5-point Laplacian on n interior vertices of [0, L], h = L/(n+1)
    (moljax/core/fft_nonperiodic.py:140-162)."""
    h = L / (n + 1)
    k = np.arange(1, n + 1, dtype=np.float64)
    return -4.0 / (h * h) * np.sin(np.pi * k / (2 * (n + 1))) ** 2


def neumann_symbol_1d(n: int, L: float = 1.0) -> np.ndarray:
    """Eigenvalues -4/h^2 sin^2(pi k / (2n)), k = 0..n-1, of the 1-D 5-point
    Laplacian on n cells of [0, L] with zero-flux ends, h = L/n
    (moljax/core/fft_nonperiodic.py:165-183)."""
    h = L / n
    k = np.arange(n, dtype=np.float64)
    return -4.0 / (h * h) * np.sin(np.pi * k / (2 * n)) ** 2


class MixedBatch:
    """Gray-Scott and Helmholtz instances of uneven sizes in ONE batched plan
    (mjx_batch_create_mixed): every (kind, size) pair is a group, and one
    hipGraph replay advances all groups, each as a branch of its own.

    add_gray_scott(sizes, F, k, Du, Dv, dt, scheme) adds Gray-Scott
    instances ('etdrk4', 'strang' or 'benchmark_strang', as GrayScottBatch,
    on [0, L)^2 with the spectral or FD symbol); add_helmholtz(kind, sizes,
    D, dt) adds single-field instances whose step is one implicit diffusion
    solve u <- (I - dt D lap_h)^-1 u on the unit square: kind 'periodic' (the
    5-point FD symbol, moljax/core/fft_solvers.py:261-294), 'dirichlet'
    (vertex grid, DST-I) or 'neumann' (cell grid, DCT-II)
    (moljax/core/fft_nonperiodic.py:308-355); D and dt per instance.
    Instances are numbered in the order they are added.  build() creates
    the plan; every instance's result is bitwise that of the instance
    advanced alone."""

    def __init__(self, dtype: torch.dtype = torch.float32, steps_per_graph: int = 16, tune: bool = True):
        self.dtype = dtype
        self.K = int(steps_per_graph)
        self.tune = tune
        self._specs: List[tuple] = []  # (scheme, n, params dict) per instance
        self._plan = None

    def add_gray_scott(self, sizes, F, k, Du: float, Dv: float, dt: float, scheme: str = "etdrk4",
                       L: float = 1.0, symbol: str = "spectral", cox_matthews: bool = True) -> List[int]:
        if scheme not in _SCHEMES:
            raise ValueError(f"scheme must be one of {sorted(_SCHEMES)}")
        if len(F) != len(sizes) or len(k) != len(sizes):
            raise ValueError("one F and one k per instance")
        first = len(self._specs)
        for i, n in enumerate(sizes):
            self._specs.append((_SCHEMES[scheme], int(n), dict(F=float(F[i]), k=float(k[i]), Du=float(Du),
                                                              Dv=float(Dv), dt=float(dt), L=float(L),
                                                              symbol=symbol, cm=bool(cox_matthews))))
        return list(range(first, len(self._specs)))

    def add_helmholtz(self, kind: str, sizes, D, dt) -> List[int]:
        if kind not in _HELM:
            raise ValueError(f"kind must be one of {sorted(_HELM)}")
        D = np.broadcast_to(np.asarray(D, dtype=np.float64), (len(sizes),))
        dt = np.broadcast_to(np.asarray(dt, dtype=np.float64), (len(sizes),))
        first = len(self._specs)
        for i, n in enumerate(sizes):
            self._specs.append((_HELM[kind], int(n), dict(D=float(D[i]), dt=float(dt[i]))))
        return list(range(first, len(self._specs)))

    def build(self):
        if self._plan is not None:
            raise RuntimeError("plan already built")
        lib = _hip.lib()
        dc = _hip.dtype_code(self.dtype)
        keys = []
        members: Dict[tuple, List[int]] = {}
        for i, (sch, n, prm) in enumerate(self._specs):
            key = (sch, n) if sch >= SCHEME_HELM_PERIODIC else (sch, n, prm["Du"], prm["Dv"], prm["dt"], prm["L"],
                                                                  prm["symbol"], prm["cm"])
            if key not in members:
                members[key] = []
                keys.append(key)
            members[key].append(i)
        self._slot = [None] * len(self._specs)
        self.groups = []
        for key in keys:
            sch, n = key[0], key[1]
            idx = members[key]
            B = len(idx)
            for j, i in enumerate(idx):
                self._slot[i] = (len(self.groups), j)
            u = torch.zeros((B, n, n), dtype=self.dtype, device="cuda")
            p0 = self._specs[idx[0]][2]
            if sch >= SCHEME_HELM_PERIODIC:
                v = None
                lam = {SCHEME_HELM_PERIODIC: fd_symbol_1d, SCHEME_HELM_DIRICHLET: dirichlet_symbol_1d,
                       SCHEME_HELM_NEUMANN: neumann_symbol_1d}[sch](n)
                t = torch.from_numpy(lam).to("cuda", self.dtype)
                tables = (t, t)
                rows = [[self._specs[i][2]["dt"] * self._specs[i][2]["D"], 0, 0, 0, 0, 0, self._specs[i][2]["dt"]]
                        for i in idx]
                nbytes = (lib.mjx_spectral_workspace_bytes(dc, B, n, n) if sch == SCHEME_HELM_PERIODIC
                          else B * n * n * (8 if self.dtype == torch.float64 else 4))
            else:
                v = torch.zeros((B, n, n), dtype=self.dtype, device="cuda")
                sym = spectral_symbol_1d if p0["symbol"] == "spectral" else fd_symbol_1d
                lam = sym(n, p0["L"])
                Du, Dv, dt = p0["Du"], p0["Dv"], p0["dt"]
                if sch == SCHEME_ETDRK4:
                    from .core.fft_integrators import etdrk4_coefficients
                    tables = (etdrk4_coefficients(lam, lam, Du, Dv, dt, self.dtype, p0["cm"]), None)
                    rows = [[0, 0, 0, 0, self._specs[i][2]["F"], self._specs[i][2]["k"], dt] for i in idx]
                    nbytes = lib.mjx_etdrk4_workspace_bytes(dc, B, n, n)
                else:
                    t = torch.from_numpy(lam).to("cuda", self.dtype)
                    tables = (t, t)
                    h = 0.5 * dt if sch == SCHEME_STRANG else dt
                    rows = [[h * Du, h * Dv, 0, 0, self._specs[i][2]["F"], self._specs[i][2]["k"], dt] for i in idx]
                    nbytes = lib.mjx_spectral_workspace_bytes(dc, B, n, n)
                if self.tune:
                    autotune(self.dtype, sch, n, B)
            self.groups.append(dict(scheme=sch, n=n, B=B, u=u, v=v, tables=tables, rows=inst_table(rows),
                                    work=torch.empty(nbytes, dtype=torch.uint8, device="cuda")))
        G = len(self.groups)
        arr = lambda ct, vals: (ct * G)(*vals)
        tabs = []
        for g in self.groups:
            tabs += [_hip.ptr(g["tables"][0]), _hip.ptr(g["tables"][1])]
        plan = ctypes.c_void_p()
        rc = lib.mjx_batch_create_mixed(
            dc, G, arr(ctypes.c_int, [g["scheme"] for g in self.groups]), arr(ctypes.c_int, [g["n"] for g in self.groups]),
            arr(ctypes.c_int, [g["B"] for g in self.groups]), arr(ctypes.c_void_p, [_hip.ptr(g["u"]) for g in self.groups]),
            arr(ctypes.c_void_p, [_hip.ptr(g["v"]) for g in self.groups]),
            arr(ctypes.c_longlong, [g["n"] for g in self.groups]),
            arr(ctypes.c_longlong, [g["n"] * g["n"] for g in self.groups]), (ctypes.c_void_p * (2 * G))(*tabs),
            arr(ctypes.c_void_p, [_hip.ptr(g["rows"]) for g in self.groups]),
            arr(ctypes.c_void_p, [_hip.ptr(g["work"]) for g in self.groups]), self.K, ctypes.byref(plan))
        _hip.check(rc, "mjx_batch_create_mixed")
        self._plan = plan
        return self

    def set_state(self, i: int, u, v=None):
        g, j = self._slot[i]
        self.groups[g]["u"][j].copy_(torch.as_tensor(u))
        if v is not None:
            self.groups[g]["v"][j].copy_(torch.as_tensor(v))

    def state(self, i: int):
        g, j = self._slot[i]
        grp = self.groups[g]
        return (grp["u"][j],) if grp["v"] is None else (grp["u"][j], grp["v"][j])

    @property
    def n_points(self) -> int:
        return sum(n * n for _, n, _ in self._specs)

    def run(self, n_steps: int):
        if self._plan is None:
            self.build()
        _hip.check(_hip.lib().mjx_batch_run(self._plan, int(n_steps), 0, _hip.stream_ptr()), "mjx_batch_run")

    def close(self):
        if self._plan:
            _hip.lib().mjx_batch_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
