"""Scalar reaction-diffusion models u_t = D lap_h u + f(u) served by the
native Newton-Krylov driver, with f a cubic polynomial
f(u) = a3 u^3 + a2 u^2 + a1 u + a0 (Allen-Cahn type f = u - u^3 is
(-1, 0, 1, 0)).

The unknown is an interior array: homogeneous Dirichlet boundaries live on a
vertex grid of n interior points with h = L/(n+1) (zero boundary values, the
grid on which DST-I diagonalises the 5-point Laplacian; the vertex-grid
Dirichlet problem of moljax/core/fft_nonperiodic.py), periodic boundaries on
n cells with h = L/n.
"""
import ctypes
from dataclasses import dataclass, field
from typing import Optional, Tuple

import numpy as np
import torch

from .. import _hip
from .newton_krylov import NKParams, NKStats, create_implicit_residual, newton_krylov_solve
from .state import StateDict, axpby


def _fd_symbol(n: int, d: float) -> np.ndarray:
    k = 2.0 * np.pi * np.fft.fftfreq(n, d=d)
    return (2.0 * np.cos(k * d) - 2.0) / (d * d)


def _dirichlet_symbol(n: int, d: float) -> np.ndarray:
    k = np.arange(1, n + 1, dtype=np.float64)
    return -4.0 / (d * d) * np.sin(np.pi * k / (2 * (n + 1))) ** 2


@dataclass(frozen=True)
class ReactionDiffusionModel:
    ny: int
    nx: int
    dy: float
    dx: float
    bc: str
    D: float
    reaction: Tuple[float, float, float, float]
    params: dict = field(default_factory=dict)
    mjx_kernel: str = "reaction_diffusion"

    @property
    def bc_code(self) -> int:
        return 0 if self.bc == "dirichlet" else 1

    @property
    def grid(self):
        return self

    def as_field(self, u) -> torch.Tensor:
        t = torch.as_tensor(u).to("cuda", torch.float64).contiguous()
        if t.shape != (self.ny, self.nx):
            raise ValueError(f"expected an interior ({self.ny}, {self.nx}) array, got {tuple(t.shape)}")
        return t

    def rhs_field(self, u) -> torch.Tensor:
        """F(u) = D lap_h u + f(u) (the residual kernel with u_old = u,
        dt = theta = 1 returns -F(u) exactly)."""
        u = self.as_field(u)
        r = torch.empty_like(u)
        reac = (ctypes.c_double * 4)(*self.reaction)
        rc = _hip.lib().mjx_rd_residual(self.ny, self.nx, self.bc_code, self.dy, self.dx, self.D, 1.0, 1.0, 1.0, reac,
                                        _hip.ptr(u), _hip.ptr(u), 0, _hip.ptr(r), _hip.stream_ptr())
        _hip.check(rc, "mjx_rd_residual")
        return axpby(-1.0, r)

    def rhs(self, y: StateDict, t: float = 0.0) -> StateDict:
        return {"u": self.rhs_field(y["u"])}

    def precond_tables(self):
        cached = self.params.get("_tables")
        if cached is None:
            if self.bc == "dirichlet":
                lx, ly = _dirichlet_symbol(self.nx, self.dx), _dirichlet_symbol(self.ny, self.dy)
            else:
                lx, ly = _fd_symbol(self.nx, self.dx), _fd_symbol(self.ny, self.dy)
            t64x, t64y = torch.from_numpy(lx).cuda(), torch.from_numpy(ly).cuda()
            cached = (t64x, t64y, t64x.float(), t64y.float())
            self.params["_tables"] = cached
        return cached


def create_reaction_diffusion_model(n: int, L: float = 1.0, D: float = 1e-3,
                                    reaction: Tuple[float, float, float, float] = (-1.0, 0.0, 1.0, 0.0),
                                    bc: str = "dirichlet", ny: Optional[int] = None) -> ReactionDiffusionModel:
    if bc not in ("dirichlet", "periodic"):
        raise ValueError("bc must be 'dirichlet' or 'periodic'")
    ny = n if ny is None else ny
    if bc == "dirichlet":
        dx, dy = L / (n + 1), L / (ny + 1)
    else:
        dx, dy = L / n, L / ny
    return ReactionDiffusionModel(ny=ny, nx=n, dy=dy, dx=dx, bc=bc, D=float(D),
                                  reaction=tuple(float(c) for c in reaction), params={"D": float(D)})


def euler_step(model: ReactionDiffusionModel, y: StateDict, t: float, dt: float) -> StateDict:
    """stepping.py:64-85: y + dt F(y)."""
    u = model.as_field(y["u"])
    return {"u": axpby(1.0, u, float(dt), model.rhs_field(u))}


def be_step(model: ReactionDiffusionModel, y: StateDict, t: float, dt: float, preconditioner=None,
            nk_params: Optional[NKParams] = None) -> Tuple[StateDict, NKStats]:
    """stepping.py:170-214: Newton-Krylov from the explicit Euler guess."""
    res = create_implicit_residual(model, y, t + dt, dt, "be")
    out = newton_krylov_solve(res, euler_step(model, y, t, dt), model, model.params, preconditioner, nk_params, dt)
    return out.solution, out.stats


def cn_step(model: ReactionDiffusionModel, y: StateDict, t: float, dt: float, preconditioner=None,
            nk_params: Optional[NKParams] = None) -> Tuple[StateDict, NKStats]:
    """stepping.py:217-260."""
    res = create_implicit_residual(model, y, t + dt, dt, "cn")
    out = newton_krylov_solve(res, euler_step(model, y, t, dt), model, model.params, preconditioner, nk_params, dt)
    return out.solution, out.stats


def bdf2_step(model: ReactionDiffusionModel, y: StateDict, y_prev: StateDict, t: float, dt: float, dt_prev: float,
              preconditioner=None, nk_params: Optional[NKParams] = None,
              consistent: bool = False) -> Tuple[StateDict, NKStats]:
    """stepping.py:263-310: variable-step BDF2 from the linear extrapolation
    2 y - y_prev (consistent: see bdf2_coefficients)."""
    from .newton_krylov import create_bdf2_residual
    res = create_bdf2_residual(model, y, y_prev, t + dt, dt, dt_prev, consistent)
    guess = {"u": axpby(2.0, model.as_field(y["u"]), -1.0, model.as_field(y_prev["u"]))}
    out = newton_krylov_solve(res, guess, model, model.params, preconditioner, nk_params, dt)
    return out.solution, out.stats


def sine_mode_ic(n: int, seed: int = 0, modes: int = 8, kmax: int = 6) -> np.ndarray:
    """Sum of ``modes`` sine modes a sin(pi p x) sin(pi q y), a ~ N(0, 0.5),
    p, q uniform in 1..kmax, on interior vertices x_i = i/(n+1) (seeded
    numpy default_rng)."""
    rng = np.random.default_rng(seed)
    h = 1.0 / (n + 1)
    x = h * np.arange(1, n + 1)
    X, Y = np.meshgrid(x, x)
    u = np.zeros((n, n))
    for _ in range(modes):
        a = rng.normal(0.0, 0.5)
        p, q = rng.integers(1, kmax + 1, size=2)
        u += a * np.sin(np.pi * p * X) * np.sin(np.pi * q * Y)
    return u


def adaptive_integrate_implicit(model: ReactionDiffusionModel, y0: StateDict, t0: float, t_end: float, dt0: float,
                                method: int = 3, max_steps: int = 10000, cfl_params=None, pid_params=None,
                                preconditioner=None, nk_params: Optional[NKParams] = None, save_every: int = 1,
                                save_states: bool = True):
    """The implicit branch of adaptive_integrate (moljax/core/stepping.py:
    723-1021) for a reaction-diffusion model, every attempt a native
    Newton-Krylov solve:
      BE (3)    err = dt F(y_be);
      CN (4)    err = y_cn - y_be;
      BDF2 (5)  err = y_bdf2 - y_be, one BE step first (the reference sends
                the IMEX values here too; adaptive_integrate routes them to
                the IMEX driver instead);
    accept when the scaled error norm <= 1, the state is finite and Newton
    converged; the next dt from propose_dt with the Newton statistics
    (implicit_robustness_dt), reject -> dt_new / 2.  Histories as the
    reference (t, dt, interior states every save_every accepted steps)."""
    from .dt_policy import (NKStats as PolicyNKStats, PIDParams, compute_error_order,
                            create_initial_controller_state, propose_dt)
    from .state import RED_NONFINITE, RED_SCALED_SQ, _reduce
    from .stepping import AdaptiveResult
    m = int(method)
    if m not in (3, 4, 5):
        raise ValueError("method must be BE (3), CN (4) or BDF2 (5)")
    if save_every < 1:
        raise ValueError("save_every must be >= 1")
    pid = pid_params or PIDParams()
    order = compute_error_order(m)
    u = model.as_field(y0["u"]).clone()
    u_prev, dt_prev = u, float(dt0)
    dt = float(dt0)
    max_saves = max_steps // save_every + 1
    t_hist = torch.zeros(max_saves, dtype=torch.float64)
    dt_hist = torch.zeros(max_saves, dtype=torch.float64)
    t_hist[0] = t0
    hist = [u.clone()] if save_states else []
    ctrl = create_initial_controller_state()
    t, widx, n_acc, n_rej, status, step_count = float(t0), 1, 0, 0, 1, 0
    while status == 1 and t < t_end and widx < max_saves:
        dtc = min(dt, t_end - t)
        y = {"u": u}
        if m == 3 or (m == 5 and step_count < 1):
            y_new, st = be_step(model, y, t, dtc, preconditioner, nk_params)
            err = axpby(dtc, model.rhs_field(y_new["u"]))
        else:
            if m == 4:
                y_new, st = cn_step(model, y, t, dtc, preconditioner, nk_params)
            else:
                y_new, st = bdf2_step(model, y, {"u": u_prev}, t, dtc, dt_prev, preconditioner, nk_params)
            y_be, _ = be_step(model, y, t, dtc, preconditioner, nk_params)
            err = axpby(1.0, y_new["u"], -1.0, y_be["u"])
        un = y_new["u"]
        e = float(torch.sqrt(_reduce(RED_SCALED_SQ, [err], [un], pid.atol, pid.rtol)[0] / un.numel()))
        e = max(e, 1e-10) if e == e else e
        finite = float(_reduce(RED_NONFINITE, [un])[0]) == 0.0
        accept = e <= 1.0 and finite and bool(st.converged)
        nks = PolicyNKStats(bool(st.converged), int(st.newton_iters), int(st.lin_iters), float(st.final_res_norm))
        dt_new, new_ctrl = propose_dt(m, model, model.params, y, t, dtc, e, ctrl, nks, cfl_params, pid, order)
        if accept:
            if n_acc % save_every == 0:
                t_hist[widx] = t + dtc
                dt_hist[widx - 1] = dtc
                if save_states:
                    hist.append(un.clone())
                widx += 1
            t += dtc
            u_prev, u = u, un
            dt_prev = dtc
            dt = dt_new
            ctrl = new_ctrl
            n_acc += 1
            step_count += 1
        else:
            dt = max(dt_new * 0.5, pid.dt_min) if dt_new == dt_new else dt_new
            if dt <= pid.dt_min * 1.1:
                status = 3
            n_rej += 1
        if not finite:
            status = 4
    if t >= t_end:
        status = 0
    elif widx >= max_saves:
        status = 2
    return AdaptiveResult(t_final=t, y_final={"u": u}, t_history=t_hist, y_history={"u": torch.stack(hist)} if hist else {},
                          dt_history=dt_hist, n_steps=widx, n_accepted=n_acc, n_rejected=n_rej, status=status)
