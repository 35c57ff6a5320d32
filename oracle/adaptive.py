"""Adaptive IMEX driver for Gray-Scott, float64 NumPy restatement (test
infrastructure).

Follows adaptive_integrate_imex (moljax/core/stepping.py:1146-1371) with
step-doubling error estimation (stepping.py:566-608), scaled_error_norm
(moljax/core/state.py:312-343), and the PI controller of
moljax/core/dt_policy.py (CFLParams :28-51, PIDParams :54-83,
pid_controller_dt :178-219, handle_rejected_step :222-268, imex_cfl_dt
:430-502, propose_dt_imex :544-610).  Scalars are Python floats (IEEE
float64), so the accept/reject sequence of an FP64 state is reproduced
decision for decision.
"""
import math
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import gray_scott as ogs

SUCCESS, RUNNING, MAX_STEPS_REACHED, DT_TOO_SMALL, NON_FINITE_VALUES = 0, 1, 2, 3, 4


@dataclass(frozen=True)
class CFLParams:
    """dt_policy.py:28-51."""
    safety: float = 0.9
    dt_min: float = 1e-12
    dt_max: float = 1.0
    cfl_advection: float = 0.5
    cfl_diffusion: float = 0.25
    cfl_wave: float = 0.5


@dataclass(frozen=True)
class PIDParams:
    """dt_policy.py:54-83."""
    safety: float = 0.9
    kI: float = 0.7
    kP: float = 0.4
    dt_min: float = 1e-12
    dt_max: float = 1.0
    max_factor: float = 5.0
    min_factor: float = 0.2
    atol: float = 1e-6
    rtol: float = 1e-4


def _nanmax(a, b):
    """jnp.maximum: NaN propagates."""
    if math.isnan(a) or math.isnan(b):
        return math.nan
    return max(a, b)


def _nanmin(a, b):
    if math.isnan(a) or math.isnan(b):
        return math.nan
    return min(a, b)


def _clip(x, lo, hi):
    """jnp.clip = minimum(maximum(x, lo), hi)."""
    return _nanmin(_nanmax(x, lo), hi)


def imex_cfl_dt(params, cfl: CFLParams, min_dx: float) -> float:
    """dt_policy.py:430-502 (advection, wave and reaction limits; diffusion
    is implicit and does not limit)."""
    dt = cfl.dt_max
    v_max = abs(params.get("vx", 0.0)) + abs(params.get("vy", 0.0))
    dt = min(dt, cfl.cfl_advection * min_dx / v_max if v_max > 1e-14 else cfl.dt_max)
    K, rho = params.get("K", 0.0), params.get("rho", 1.0)
    c = max(params.get("wave_speed", 0.0), math.sqrt(K / rho) if K > 1e-14 else 0.0)
    dt = min(dt, cfl.cfl_wave * min_dx / c if c > 1e-14 else cfl.dt_max)
    rate = max(params.get("F", 0.0) + params.get("k", 0.0), params.get("r", 0.0))
    dt = min(dt, 0.5 / rate if rate > 1e-14 else cfl.dt_max)
    return _clip(cfl.safety * dt, cfl.dt_min, cfl.dt_max)


def heisenberg_cfl_dt(params, cfl: CFLParams, min_dx: float) -> float:
    """dt_policy.py:104-175: advection, diffusion (dx^2 / D_max) and wave
    limits."""
    dt = cfl.dt_max
    v_max = abs(params.get("vx", 0.0)) + abs(params.get("vy", 0.0))
    dt = min(dt, cfl.cfl_advection * min_dx / v_max if v_max > 1e-14 else cfl.dt_max)
    D_max = max(params.get("D", 0.0), params.get("Du", 0.0), params.get("Dv", 0.0))
    dt = min(dt, cfl.cfl_diffusion * min_dx ** 2 / D_max if D_max > 1e-14 else cfl.dt_max)
    K, rho = params.get("K", 0.0), params.get("rho", 1.0)
    c = max(params.get("wave_speed", 0.0), math.sqrt(K / rho) if K > 1e-14 else 0.0)
    dt = min(dt, cfl.cfl_wave * min_dx / c if c > 1e-14 else cfl.dt_max)
    return _clip(cfl.safety * dt, cfl.dt_min, cfl.dt_max)


def pid_controller_dt(dt_old, err, prev_err, pid: PIDParams, order):
    """dt_policy.py:178-219; returns (dt_new, new prev_err)."""
    kI, kP = pid.kI / order, pid.kP / order
    e = _nanmax(err, 1e-10)
    pe = _nanmax(prev_err, 1e-10)
    factor = (1.0 / e) ** kI * (pe / e) ** kP
    factor = _clip(factor * pid.safety, pid.min_factor, pid.max_factor)
    return _clip(dt_old * factor, pid.dt_min, pid.dt_max), err


def handle_rejected_step(dt_old, err, pid: PIDParams, order):
    """dt_policy.py:222-268 (never grows)."""
    kI = pid.kI / order
    e = _nanmax(err, 1.0)
    factor = _clip(pid.safety * (1.0 / e) ** kI, pid.min_factor, 1.0)
    return _clip(dt_old * factor, pid.dt_min, pid.dt_max)


def scaled_error_norm(err, y, atol, rtol, ng=1):
    """state.py:312-343: sqrt(mean over interiors of all fields of
    (err / (atol + rtol |y|))^2)."""
    tot, n = 0.0, 0
    for name in err:
        e = err[name][ng:-ng, ng:-ng]
        s = atol + rtol * np.abs(y[name][ng:-ng, ng:-ng])
        tot += float(np.sum((e / s) ** 2))
        n += e.size
    return math.sqrt(tot / n)


@dataclass
class AdaptiveResult:
    t_final: float
    y_final: dict
    t_history: np.ndarray
    dt_history: np.ndarray
    y_saved: List[dict]
    n_steps: int
    n_accepted: int
    n_rejected: int
    status: int
    attempts: List[tuple] = field(default_factory=list)  # (t, dt_clamped, err_ratio, accepted)


def adaptive_integrate_imex(y0, t0, t_end, dt0, F, k, Du, Dv, lam, min_dx, use_strang=True, max_steps=10000,
                            cfl: CFLParams = None, pid: PIDParams = None, save_every=1, ng=1,
                            keep_states=True, max_attempts=None):
    """stepping.py:1146-1371 for the Gray-Scott model (params F, k)."""
    cfl = cfl or CFLParams()
    step = ogs.imex_strang_step if use_strang else ogs.imex_euler_step
    return _adaptive(y0, t0, t_end, dt0, lambda y, h: step(y, h, F, k, Du, Dv, lam, ng),
                     imex_cfl_dt({"F": F, "k": k}, cfl, min_dx), 2 if use_strang else 1, max_steps, pid,
                     save_every, ng, keep_states, max_attempts)


EXPLICIT_ORDER = {"euler": 1, "ssprk3": 3, "rk4": 4}  # dt_policy.py:411-423


def adaptive_integrate_explicit(y0, t0, t_end, dt0, F, k, Du, Dv, dy, dx, method="rk4", max_steps=10000,
                                cfl: CFLParams = None, pid: PIDParams = None, save_every=1, ng=1,
                                keep_states=True, max_attempts=None):
    """The explicit branch of adaptive_integrate (stepping.py:723-1021):
    step doubling with euler/ssprk3/rk4 (stepping.py:502-536), dt limited by
    heisenberg_cfl_dt."""
    cfl = cfl or CFLParams()
    step = {"euler": ogs.euler_step, "ssprk3": ogs.ssprk3_step, "rk4": ogs.rk4_step}[method]
    dt_cfl = heisenberg_cfl_dt({"Du": Du, "Dv": Dv, "F": F, "k": k}, cfl, min(dx, dy))
    return _adaptive(y0, t0, t_end, dt0, lambda y, h: step(y, h, F, k, Du, Dv, dy, dx, ng), dt_cfl,
                     EXPLICIT_ORDER[method], max_steps, pid, save_every, ng, keep_states, max_attempts)


def _adaptive(y0, t0, t_end, dt0, stepper, dt_cfl, order, max_steps, pid, save_every, ng, keep_states,
              max_attempts):
    pid = pid or PIDParams()
    dt = _nanmin(float(dt0), dt_cfl)
    max_saves = max_steps // save_every + 1
    t_hist = np.zeros(max_saves)
    dt_hist = np.zeros(max_saves)
    t_hist[0] = t0
    saved = [{n: v[ng:-ng, ng:-ng].copy() for n, v in y0.items()}] if keep_states else []
    y = {n: v.copy() for n, v in y0.items()}
    t, prev_err, write_idx, n_acc, n_rej, status = float(t0), 1.0, 1, 0, 0, RUNNING
    attempts = []
    while status == RUNNING and t < t_end and write_idx < max_saves:
        if max_attempts is not None and len(attempts) >= max_attempts:
            break
        dtc = _nanmin(dt, t_end - t)
        y_full = stepper(y, dtc)
        y_new = stepper(stepper(y, dtc / 2.0), dtc / 2.0)
        err = {n: y_new[n] - y_full[n] for n in y_new}
        e = _nanmax(scaled_error_norm(err, y_new, pid.atol, pid.rtol, ng), 1e-10)
        finite = all(bool(np.all(np.isfinite(v))) for v in y_new.values())
        accept = (e <= 1.0) and finite
        if e <= 1.0:
            dt_ad, new_prev = pid_controller_dt(dtc, e, prev_err, pid, order)
        else:
            dt_ad, new_prev = handle_rejected_step(dtc, e, pid, order), prev_err
        dt_new = _clip(_nanmin(dt_ad, dt_cfl), pid.dt_min, pid.dt_max)
        attempts.append((t, dtc, e, accept))
        if accept:
            if n_acc % save_every == 0:
                t_hist[write_idx] = t + dtc
                if keep_states:
                    saved.append({n: v[ng:-ng, ng:-ng].copy() for n, v in y_new.items()})
                dt_hist[write_idx - 1] = dtc
                write_idx += 1
            t += dtc
            y = y_new
            dt = dt_new
            prev_err = new_prev
            n_acc += 1
        else:
            dt = _nanmax(dt_new * 0.5, pid.dt_min)
            if dt <= pid.dt_min * 1.1:
                status = DT_TOO_SMALL
            n_rej += 1
        if not finite:
            status = NON_FINITE_VALUES
    if t >= t_end:
        status = SUCCESS
    elif write_idx >= max_saves:
        status = MAX_STEPS_REACHED
    return AdaptiveResult(t_final=t, y_final=y, t_history=t_hist, dt_history=dt_hist, y_saved=saved,
                          n_steps=write_idx, n_accepted=n_acc, n_rejected=n_rej, status=status, attempts=attempts)


def implicit_robustness_dt(dt_old, converged, newton_iters, dt_min=1e-12, dt_max=1.0, growth_factor=1.2,
                           shrink_factor=0.5, fast_iters_threshold=3):
    """dt_policy.py:279-321 (its own dt_max default of 1.0 applies)."""
    dt_new = dt_old if converged else dt_old * shrink_factor
    if converged and newton_iters <= fast_iters_threshold:
        dt_new = _nanmin(dt_new * growth_factor, dt_max)
    return _clip(dt_new, dt_min, dt_max)


IMPLICIT_ORDER = {"be": 1, "cn": 2, "bdf2": 2}  # dt_policy.py:411-423


def adaptive_integrate_implicit(u0, t0, t_end, dt0, D, coef, dy, dx, bc, method="be", precond="dst",
                                max_steps=10000, pid: PIDParams = None, save_every=1, max_attempts=None,
                                nk_kw=None):
    """The implicit branch of adaptive_integrate (stepping.py:723-1021) for a
    scalar reaction-diffusion problem on an interior array (every point is
    interior): BE with err = dt F(y_be); CN and BDF2 with err = y - y_be (BDF2
    starts with one BE step); accept when e <= 1, finite and Newton
    converged; propose_dt for implicit methods (no CFL limit,
    implicit_robustness_dt on the Newton count)."""
    from . import newton_krylov as onk
    pid = pid or PIDParams()
    nk_kw = nk_kw or {}
    order = IMPLICIT_ORDER[method]
    dt = float(dt0)
    max_saves = max_steps // save_every + 1
    t_hist = np.zeros(max_saves)
    dt_hist = np.zeros(max_saves)
    t_hist[0] = t0
    y, y_prev, dt_prev = u0.copy(), u0.copy(), float(dt0)
    t, prev_err, write_idx, n_acc, n_rej, status, step_count = float(t0), 1.0, 1, 0, 0, RUNNING, 0
    attempts = []
    args = (D, coef, dy, dx, bc, precond)
    while status == RUNNING and t < t_end and write_idx < max_saves:
        if max_attempts is not None and len(attempts) >= max_attempts:
            break
        dtc = _nanmin(dt, t_end - t)
        if method == "be" or (method == "bdf2" and step_count < 1):
            st = onk.be_step(y, dtc, *args, **nk_kw)
            y_new = st.u
            err = dtc * onk.rhs(y_new, D, coef, dy, dx, bc)
        else:
            if method == "cn":
                st = onk.cn_step(y, dtc, *args, **nk_kw)
            else:
                st = onk.bdf2_step(y, y_prev, dtc, dt_prev, *args, **nk_kw)
            y_new = st.u
            err = y_new - onk.be_step(y, dtc, *args, **nk_kw).u
        e = _nanmax(math.sqrt(float(np.sum((err / (pid.atol + pid.rtol * np.abs(y_new))) ** 2)) / err.size), 1e-10)
        finite = bool(np.all(np.isfinite(y_new)))
        accept = (e <= 1.0) and finite and bool(st.converged)
        if e <= 1.0:
            dt_ad, new_prev = pid_controller_dt(dtc, e, prev_err, pid, order)
        else:
            dt_ad, new_prev = handle_rejected_step(dtc, e, pid, order), prev_err
        dt_new = _clip(implicit_robustness_dt(dt_ad, bool(st.converged), st.newton_iters), pid.dt_min, pid.dt_max)
        attempts.append((t, dtc, e, accept, st.newton_iters, st.gmres_iters))
        if accept:
            if n_acc % save_every == 0:
                t_hist[write_idx] = t + dtc
                dt_hist[write_idx - 1] = dtc
                write_idx += 1
            t += dtc
            y_prev, y = y, y_new
            dt_prev = dtc
            dt = dt_new
            prev_err = new_prev
            n_acc += 1
            step_count += 1
        else:
            dt = _nanmax(dt_new * 0.5, pid.dt_min)
            if dt <= pid.dt_min * 1.1:
                status = DT_TOO_SMALL
            n_rej += 1
        if not finite:
            status = NON_FINITE_VALUES
    if t >= t_end:
        status = SUCCESS
    elif write_idx >= max_saves:
        status = MAX_STEPS_REACHED
    return AdaptiveResult(t_final=t, y_final={"u": y}, t_history=t_hist, dt_history=dt_hist, y_saved=[],
                          n_steps=write_idx, n_accepted=n_acc, n_rejected=n_rej, status=status, attempts=attempts)
