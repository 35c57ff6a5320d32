"""Continuous-batching request loop for mixed diffusion / Gray-Scott solves
on one GPU (build-only; the reference has no serving path).

Request kinds (all periodic or homogeneous-BC squares, states resident on the
device from admission to retirement):
  'fft_cn'   periodic Crank-Nicolson diffusion, spectral symbol on [0, 1)^2
             (fft_solvers.fft_cn_step semantics);
  'dst'      homogeneous Dirichlet backward-Euler diffusion steps on n - 1
             interior vertices of [0, 1]^2 (DST-I Helmholtz solves);
  'dct'      homogeneous Neumann backward-Euler diffusion steps on n cells of
             [0, 1]^2 (DCT-II Helmholtz solves);
  'imex_gs'  Gray-Scott IMEX-Strang on [0, 2.5]^2 (FD symbol), Du = D,
             Dv = D / 2, F = 0.035, k = 0.065.

Scheduling:
  * admission is FIFO by arrival time and by HBM footprint: a request enters
    when its state and working buffers fit the remaining budget;
  * every request advances in chunks of ``chunk`` steps counted from its own
    first step (the last chunk is the remainder), so its arithmetic does not
    depend on what else is running;
  * requests of one (kind, size) live in the slots of a resident pool; a
    pool's launch is ONE batched launch in which every resident request
    advances min(chunk, its remaining steps): the instance rows carry each
    request's parameters and step count (per-instance spectral powers for
    the linear kinds, a masked Strang chain for Gray-Scott);
  * each pool runs on a stream of its own and is relaunched as soon as its
    previous launch has finished (an event query, no device-wide wait), so
    pools overlap on the device; requests that arrive while others run are
    admitted into free slots at once, and finished ones retire and release
    their footprint when their last launch completes.
Because every kernel is batch-invariant, a request's result is bitwise the
same alone and in any mix.
"""
import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np
import torch

from . import _hip
from .core._arrays import inst_table

KINDS = ("fft_cn", "dst", "dct", "imex_gs")
GS_F, GS_K, GS_L = 0.035, 0.065, 2.5


@dataclass
class Request:
    rid: int
    kind: str
    n: int
    D: float
    dt: float
    steps: int
    seed: int
    arrival: float = 0.0
    # filled by the server
    admitted: float = math.nan
    finished: float = math.nan
    done_steps: int = 0
    state: List[torch.Tensor] = field(default_factory=list)
    ic: Optional[List[np.ndarray]] = None  # host input arriving with the request

    @property
    def grid(self) -> int:
        """Interior points per side (DST-I runs on n - 1 vertices)."""
        return self.n - 1 if self.kind == "dst" else self.n

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def points(self) -> int:
        return self.grid * self.grid * (2 if self.kind == "imex_gs" else 1)


def initial_fields(req: Request) -> List[np.ndarray]:
    """Seeded random-mode initial condition on the host (numpy
    default_rng(seed)), float64."""
    rng = np.random.default_rng(req.seed)
    m = req.grid
    modes = [(rng.normal(0.0, 1.0), int(rng.integers(1, 6)), int(rng.integers(1, 6)), rng.uniform(0, 2 * np.pi))
             for _ in range(4)]
    if req.kind == "dst":
        x = np.arange(1, m + 1) / (m + 1)
        X, Y = np.meshgrid(x, x)
        f = sum(a * np.sin(np.pi * p * X) * np.sin(np.pi * q * Y) for a, p, q, _ in modes)
        fields = [f]
    elif req.kind == "dct":
        x = (np.arange(m) + 0.5) / m
        X, Y = np.meshgrid(x, x)
        f = sum(a * np.cos(np.pi * p * X) * np.cos(np.pi * q * Y) for a, p, q, _ in modes)
        fields = [f]
    else:
        x = np.arange(m) / m
        X, Y = np.meshgrid(x, x)
        f = sum(a * np.cos(2 * np.pi * (p * X + q * Y) + ph) for a, p, q, ph in modes)
        if req.kind == "fft_cn":
            fields = [f]
        else:
            g = (f - f.min()) / max(f.max() - f.min(), 1e-12)
            fields = [1.0 - 0.5 * g, 0.25 * g]
    return [np.ascontiguousarray(a) for a in fields]


def initial_state(req: Request, dtype: torch.dtype) -> List[torch.Tensor]:
    """The request's input on the device: its attached host arrays (``ic``)
    or the seeded random-mode initial condition."""
    fields = req.ic if req.ic is not None else initial_fields(req)
    return [torch.from_numpy(a).to("cuda", dtype) for a in fields]


def attach_inputs(reqs: List[Request], dtype: torch.dtype = torch.float32) -> List[Request]:
    """Generate every request's input on the host ahead of the run (the data
    a client would send with the request), in the serving precision."""
    npd = np.float64 if dtype == torch.float64 else np.float32
    for r in reqs:
        r.ic = [a.astype(npd) for a in initial_fields(r)]
    return reqs


def _spectral(n: int, L: float) -> np.ndarray:
    k = np.fft.fftfreq(n, L / n) * 2 * np.pi
    return -(k ** 2)


def _fd(n: int, L: float) -> np.ndarray:
    h = L / n
    k = 2.0 * np.pi * np.fft.fftfreq(n, d=h)
    return (2.0 * np.cos(k * h) - 2.0) / (h * h)


def _dirichlet(m: int) -> np.ndarray:
    h = 1.0 / (m + 1)
    k = np.arange(1, m + 1, dtype=np.float64)
    return -4.0 / (h * h) * np.sin(np.pi * k / (2 * (m + 1))) ** 2


def _neumann(m: int) -> np.ndarray:
    h = 1.0 / m
    k = np.arange(m, dtype=np.float64)
    return -4.0 / (h * h) * np.sin(np.pi * k / (2 * m)) ** 2


class _Pool:
    """Resident slots of one (kind, grid size): fields (cap, m, m), instance
    rows (cap, 8) with the step count of the next launch in the aux column."""

    def __init__(self, kind: str, m: int, dtype: torch.dtype, cap: int = 8):
        self.kind, self.m, self.dtype = kind, m, dtype
        self.nf = 2 if kind == "imex_gs" else 1
        self.cap = 0
        self.slots: List[Optional[Request]] = []
        self._alloc(cap)

    def _alloc(self, cap: int):
        old = getattr(self, "fields", None)
        self.fields = [torch.zeros((cap, self.m, self.m), dtype=self.dtype, device="cuda") for _ in range(self.nf)]
        if old is not None:
            for f in range(self.nf):
                self.fields[f][:self.cap].copy_(old[f])
        rows = np.zeros((cap, 8))
        if self.cap:
            rows[:self.cap] = self.rows_np
        self.rows_np = rows
        self.rows_host = torch.zeros((cap, 8), dtype=torch.float64).pin_memory()
        self.rows_dev = torch.zeros((cap, 8), dtype=torch.float64, device="cuda")
        lib = _hip.lib()
        dc = _hip.dtype_code(self.dtype)
        nbytes = (lib.mjx_spectral_workspace_bytes(dc, cap, self.m, self.m) if self.kind in ("fft_cn", "imex_gs")
                  else cap * self.m * self.m * (8 if self.dtype == torch.float64 else 4))
        self.work = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        self.slots += [None] * (cap - self.cap)
        self.cap = cap

    def admit(self, r: Request, state: List[torch.Tensor]):
        try:
            i = self.slots.index(None)
        except ValueError:
            i = self.cap
            self._alloc(2 * self.cap)
        self.slots[i] = r
        for f in range(self.nf):
            self.fields[f][i].copy_(state[f])
        self.rows_np[i] = _rows(r)
        return i

    def active(self) -> bool:
        return any(x is not None for x in self.slots)

    def launch(self, chunk: int, st: int) -> int:
        """One batched launch over slots [0, hi): every resident request
        advances min(chunk, its remaining steps); returns the launch's
        maximum step count."""
        hi = max(i for i, x in enumerate(self.slots) if x is not None) + 1
        n = np.zeros(hi)
        for i in range(hi):
            r = self.slots[i]
            if r is not None:
                n[i] = min(chunk, r.steps - r.done_steps)
        self.rows_np[:hi, 7] = n
        self.rows_host[:hi].copy_(torch.from_numpy(self.rows_np[:hi]))
        self.rows_dev[:hi].copy_(self.rows_host[:hi], non_blocking=True)
        run_slots(self.kind, self.fields, hi, self.rows_dev, int(n.max()), self._lam(), self.work, st)
        return int(n.max())

    def _lam(self) -> torch.Tensor:
        if not hasattr(self, "lam"):
            m = self.m
            lam = {"fft_cn": lambda: _spectral(m, 1.0), "imex_gs": lambda: _fd(m, GS_L),
                   "dst": lambda: _dirichlet(m), "dct": lambda: _neumann(m)}[self.kind]()
            self.lam = torch.from_numpy(lam).to("cuda", self.dtype)
        return self.lam


def run_slots(kind: str, fields: List[torch.Tensor], B: int, rows: torch.Tensor, max_steps: int, lam: torch.Tensor,
              work: torch.Tensor, st: int):
    """Advance slots [0, B) of a pool in place; slot b takes rows[b, 7] steps
    (per-instance powers for the linear kinds, a masked Strang chain for
    Gray-Scott)."""
    lib = _hip.lib()
    a = fields[0]
    m = a.shape[-1]
    dc = _hip.dtype_code(a.dtype)
    if max_steps <= 0:
        return
    if kind == "fft_cn":
        rc = lib.mjx_spectral_pair(dc, B, m, m, _hip.ptr(a), None, m, m * m, _hip.ptr(a), None, m, m * m, 0,
                                   _hip.ptr(lam), _hip.ptr(lam), None, _hip.ptr(rows), -3, _hip.ptr(work), st)
        _hip.check(rc, "mjx_spectral_pair")
    elif kind == "imex_gs":
        rc = lib.mjx_gs_imex_strang_var(dc, B, m, m, _hip.ptr(a), _hip.ptr(fields[1]), m, m * m, 0, _hip.ptr(lam),
                                        _hip.ptr(lam), None, _hip.ptr(rows), int(max_steps), _hip.ptr(work), st)
        _hip.check(rc, "mjx_gs_imex_strang_var")
    else:
        rc = lib.mjx_r2r_solve2d(dc, 0 if kind == "dst" else 1, B, m, m, _hip.ptr(a), m, m * m, _hip.ptr(a), m,
                                 m * m, _hip.ptr(lam), _hip.ptr(lam), _hip.ptr(rows), 2, _hip.ptr(work), st)
        _hip.check(rc, "mjx_r2r_solve2d")


def _rows(req: Request) -> List[float]:
    if req.kind == "fft_cn":
        h = 0.5 * req.dt * req.D
        return [h, 0.0, h, 0.0, 0.0, 0.0, req.dt, 0.0]
    if req.kind == "imex_gs":
        return [0.5 * req.dt * req.D, 0.25 * req.dt * req.D, 0.0, 0.0, GS_F, GS_K, req.dt, 0.0]
    return [req.dt * req.D, 0.0, 0.0, 0.0, 0.0, 0.0, req.dt, 0.0]


class Server:
    """The continuous-batching loop.  ``hbm_budget`` bounds the bytes of
    resident request state plus working buffers (footprint = 4x state).

    Every pool runs on a stream of its own and is launched again as soon as
    its previous launch has finished (an event query, no device-wide wait):
    requests are admitted into a pool's free slots at any moment (the input
    copy is ordered after the pool's launch in flight on its stream), pools
    of different kinds and sizes overlap on the device, and a request retires
    -- its result copied out, its slot and footprint freed -- when the launch
    that completes its last chunk has finished."""

    def __init__(self, dtype: torch.dtype = torch.float32, hbm_budget: int = 16 << 30, chunk: int = 32,
                 n_streams: int = 4):
        self.dtype = dtype
        self.budget = int(hbm_budget)
        self.chunk = int(chunk)
        self.streams = [torch.cuda.Stream() for _ in range(n_streams)]
        self.pools: Dict[tuple, _Pool] = {}
        self.used = 0
        self.peak_used = 0
        self.iterations = 0
        self.peak_active = 0
        self.max_group = 0

    def footprint(self, req: Request) -> int:
        es = 8 if self.dtype == torch.float64 else 4
        return 4 * req.points * es

    def _pool(self, r: Request) -> _Pool:
        key = (r.kind, r.grid)
        p = self.pools.get(key)
        if p is None:
            st = self.streams[len(self.pools) % len(self.streams)]
            st.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(st):
                p = self.pools[key] = _Pool(r.kind, r.grid, self.dtype)
            p.stream = st
            p.event = None      # the launch in flight, if any
            p.flight = []       # (slot, request, steps) advanced by it
        return p

    def _finish_launch(self, p: _Pool, t_done: float):
        """The pool's launch has finished: advance its requests and retire
        the completed ones (their results copied on the pool's stream)."""
        with torch.cuda.stream(p.stream):
            for i, r, n in p.flight:
                r.done_steps += n
                if r.done_steps >= r.steps:
                    r.finished = t_done
                    r.state = [f[i].clone() for f in p.fields]
                    p.slots[i] = None
                    self.used -= self.footprint(r)
                    self._n_active -= 1
        p.event = None
        p.flight = []

    def run(self, requests: List[Request], realtime: bool = True) -> List[Request]:
        """Serve ``requests`` (arrival times in seconds from the start);
        returns them with admitted/finished times and final states."""
        for r in requests:
            if self.footprint(r) > self.budget:
                raise ValueError(f"request {r.rid} needs {self.footprint(r)} bytes, above the HBM budget")
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        self._n_active = 0
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: float("inf"))
        while pending or self._n_active:
            progressed = False
            # admission: FIFO by arrival while the footprint budget allows
            now = clock()
            while pending and pending[0].arrival <= now and self.used + self.footprint(pending[0]) <= self.budget:
                r = pending.popleft()
                p = self._pool(r)
                # on the pool's stream: after its launch in flight (a pool that
                # grows copies its old slots there too)
                with torch.cuda.stream(p.stream):
                    p.admit(r, initial_state(r, self.dtype))
                r.admitted = time.perf_counter() - t0
                self.used += self.footprint(r)
                self.peak_used = max(self.peak_used, self.used)
                self._n_active += 1
                progressed = True
            self.peak_active = max(self.peak_active, self._n_active)
            # completions, then new launches of every idle pool with requests
            for p in self.pools.values():
                if p.event is not None and p.event.query():
                    self._finish_launch(p, time.perf_counter() - t0)
                    progressed = True
                if p.event is None and p.active():
                    with torch.cuda.stream(p.stream):
                        hi = max(i for i, x in enumerate(p.slots) if x is not None) + 1
                        p.flight = [(i, p.slots[i], min(self.chunk, p.slots[i].steps - p.slots[i].done_steps))
                                    for i in range(hi) if p.slots[i] is not None]
                        p.launch(self.chunk, p.stream.cuda_stream)
                        p.event = torch.cuda.Event()
                        p.event.record(p.stream)
                    self.iterations += 1
                    self.max_group = max(self.max_group, len(p.flight))
                    progressed = True
            if not progressed:
                if pending and not self._n_active:
                    time.sleep(max(0.0, min(pending[0].arrival - clock(), 0.002)))
                else:
                    time.sleep(2e-5)
        torch.cuda.synchronize()
        return requests


def poisson_stream(n_requests: int, rate: float = 200.0, seed: int = 0, sizes=(64, 128, 256, 512, 1024),
                   kinds=KINDS, steps=(10, 500)) -> List[Request]:
    """Seeded open-loop workload: exponential inter-arrival times at ``rate``
    per second; kind, size uniform; D ~ logU(1e-4, 1e-1); dt ~ logU(1e-3,
    1e-1); steps ~ U{steps[0]..steps[1]}."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for i in range(n_requests):
        t += rng.exponential(1.0 / rate)
        out.append(Request(rid=i, kind=str(rng.choice(kinds)), n=int(rng.choice(sizes)),
                           D=float(10 ** rng.uniform(-4, -1)), dt=float(10 ** rng.uniform(-3, -1)),
                           steps=int(rng.integers(steps[0], steps[1] + 1)), seed=int(rng.integers(1 << 31)),
                           arrival=t))
    return out


def latency_stats(reqs: List[Request]) -> Dict[str, float]:
    lat = np.array([r.latency for r in reqs])
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "max_ms": float(lat.max() * 1e3)}


def sequential(reqs: List[Request], dtype: torch.dtype = torch.float64, keep: bool = False) -> float:
    """Seconds to run the same requests one at a time, each in one launch
    sequence of all its steps (no batching), in ``dtype``; with ``keep`` each
    request's final state is stored in ``state``."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in reqs:
        p = _Pool(r.kind, r.grid, dtype, cap=1)
        p.admit(r, initial_state(r, dtype))
        done = r.done_steps
        p.launch(r.steps, torch.cuda.current_stream().cuda_stream)
        if keep:
            r.state = [f[0] for f in p.fields]
        p.slots[0] = None
        r.done_steps = done
    torch.cuda.synchronize()
    return time.perf_counter() - t0
