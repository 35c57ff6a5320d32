"""Time-step policy (moljax/core/dt_policy.py): CFL limits and the PI
controller.

These are the host-side scalar forms, for callers that drive their own loop.
The adaptive drivers run the same controller on the device (per instance,
inside mjx_gs_imex_adaptive), so no host round trip sits between attempts.
Scalars are IEEE float64; NaN propagates through max/min/clip as it does in
jnp.maximum / jnp.minimum / jnp.clip.
"""
import math
from dataclasses import dataclass
from typing import Any, Dict, NamedTuple, Optional, Tuple


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
    """dt_policy.py:54-83 (kI, kP are divided by the method order)."""
    safety: float = 0.9
    kI: float = 0.7
    kP: float = 0.4
    dt_min: float = 1e-12
    dt_max: float = 1.0
    max_factor: float = 5.0
    min_factor: float = 0.2
    atol: float = 1e-6
    rtol: float = 1e-4


class ControllerState(NamedTuple):
    """dt_policy.py:86-101."""
    prev_err_ratio: float
    prevprev_err_ratio: float
    consecutive_rejects: int


class NKStats(NamedTuple):
    """dt_policy.py:270-276."""
    converged: bool
    newton_iters: int
    lin_iters: int
    final_res_norm: float


def create_initial_controller_state(dtype=None) -> ControllerState:
    """dt_policy.py:104-110."""
    return ControllerState(1.0, 1.0, 0)


def _max(a: float, b: float) -> float:
    a, b = float(a), float(b)
    return math.nan if (math.isnan(a) or math.isnan(b)) else max(a, b)


def _min(a: float, b: float) -> float:
    a, b = float(a), float(b)
    return math.nan if (math.isnan(a) or math.isnan(b)) else min(a, b)


def _clip(x: float, lo: float, hi: float) -> float:
    return _min(_max(x, lo), hi)


def _pow(a: float, b: float) -> float:
    if math.isnan(a) or math.isnan(b):
        return math.nan
    if math.isinf(a):
        return math.inf if b > 0 else (0.0 if b < 0 else 1.0)
    return a ** b


def heisenberg_cfl_dt(grid, params: Dict[str, Any], cfl_params: CFLParams) -> float:
    """dt_policy.py:104-175: advection, diffusion and wave limits."""
    dx, dx2 = grid.min_dx, grid.min_dx2
    dt = cfl_params.dt_max
    v_max = abs(params.get("vx", 0.0)) + abs(params.get("vy", 0.0))
    dt = _min(dt, cfl_params.cfl_advection * dx / v_max if v_max > 1e-14 else cfl_params.dt_max)
    D_max = max(params.get("D", 0.0), max(params.get("Du", 0.0), params.get("Dv", 0.0)))
    dt = _min(dt, cfl_params.cfl_diffusion * dx2 / D_max if D_max > 1e-14 else cfl_params.dt_max)
    dt = _min(dt, _wave_limit(params, cfl_params, dx))
    return _clip(cfl_params.safety * dt, cfl_params.dt_min, cfl_params.dt_max)


def _wave_limit(params, cfl_params: CFLParams, dx: float) -> float:
    K, rho = params.get("K", 0.0), params.get("rho", 1.0)
    c = max(params.get("wave_speed", 0.0), math.sqrt(K / rho) if K > 1e-14 else 0.0)
    return cfl_params.cfl_wave * dx / c if c > 1e-14 else cfl_params.dt_max


def pid_controller_dt(dt_old: float, err_ratio: float, controller_state: ControllerState, pid_params: PIDParams,
                      order: int = 4) -> Tuple[float, ControllerState]:
    """dt_policy.py:178-219:
    factor = clip(safety (1/e)^(kI/p) (e_prev/e)^(kP/p), min_factor, max_factor)."""
    kI, kP = pid_params.kI / order, pid_params.kP / order
    e = _max(err_ratio, 1e-10)
    pe = _max(controller_state.prev_err_ratio, 1e-10)
    factor = _pow(1.0 / e, kI) * _pow(pe / e, kP)
    factor = _clip(factor * pid_params.safety, pid_params.min_factor, pid_params.max_factor)
    dt_new = _clip(dt_old * factor, pid_params.dt_min, pid_params.dt_max)
    return dt_new, ControllerState(float(err_ratio), controller_state.prev_err_ratio, 0)


def handle_rejected_step(dt_old: float, err_ratio: float, controller_state: ControllerState,
                         pid_params: PIDParams, order: int = 4) -> Tuple[float, ControllerState]:
    """dt_policy.py:222-268: shrink, never grow."""
    kI = pid_params.kI / order
    e = _max(err_ratio, 1.0)
    factor = _clip(pid_params.safety * _pow(1.0 / e, kI), pid_params.min_factor, 1.0)
    dt_new = _clip(dt_old * factor, pid_params.dt_min, pid_params.dt_max)
    return dt_new, ControllerState(controller_state.prev_err_ratio, controller_state.prevprev_err_ratio,
                                   controller_state.consecutive_rejects + 1)


def implicit_robustness_dt(dt_old: float, nk_stats: NKStats, dt_min: float = 1e-12, dt_max: float = 1.0,
                           growth_factor: float = 1.2, shrink_factor: float = 0.5,
                           fast_iters_threshold: int = 3) -> float:
    """dt_policy.py:279-321."""
    dt_new = dt_old if nk_stats.converged else dt_old * shrink_factor
    if nk_stats.converged and nk_stats.newton_iters <= fast_iters_threshold:
        dt_new = _min(dt_new * growth_factor, dt_max)
    return _clip(dt_new, dt_min, dt_max)


def propose_dt(method: int, grid, params: Dict[str, Any], state, t: float, dt_old: float, err_ratio: float,
               controller_state: ControllerState, nk_stats: Optional[NKStats] = None,
               cfl_params: Optional[CFLParams] = None, pid_params: Optional[PIDParams] = None,
               order: int = 4) -> Tuple[float, ControllerState]:
    """dt_policy.py:324-408."""
    cfl_params = cfl_params or CFLParams()
    pid_params = pid_params or PIDParams()
    if err_ratio <= 1.0:
        dt_ad, new_state = pid_controller_dt(dt_old, err_ratio, controller_state, pid_params, order)
    else:
        dt_ad, new_state = handle_rejected_step(dt_old, err_ratio, controller_state, pid_params, order)
    dt_final = _min(dt_ad, heisenberg_cfl_dt(grid, params, cfl_params)) if method < 3 else dt_ad
    if nk_stats is not None and method >= 3:
        dt_final = implicit_robustness_dt(dt_final, nk_stats)
    return _clip(dt_final, pid_params.dt_min, pid_params.dt_max), new_state


def compute_error_order(method: int) -> int:
    """dt_policy.py:411-423."""
    orders = [1, 3, 4, 1, 2, 2, 1, 2]
    return orders[min(method, len(orders) - 1)]


def imex_cfl_dt(grid, params: Dict[str, Any], cfl_params: CFLParams) -> float:
    """dt_policy.py:430-502: advection, wave and reaction limits (the
    implicit diffusion does not limit)."""
    dx = grid.min_dx
    dt = cfl_params.dt_max
    v_max = abs(params.get("vx", 0.0)) + abs(params.get("vy", 0.0))
    dt = _min(dt, cfl_params.cfl_advection * dx / v_max if v_max > 1e-14 else cfl_params.dt_max)
    dt = _min(dt, _wave_limit(params, cfl_params, dx))
    rate = max(params.get("F", 0.0) + params.get("k", 0.0), params.get("r", 0.0))
    dt = _min(dt, 0.5 / rate if rate > 1e-14 else cfl_params.dt_max)
    return _clip(cfl_params.safety * dt, cfl_params.dt_min, cfl_params.dt_max)


def propose_dt_imex(grid, params: Dict[str, Any], state, t: float, dt_old: float, err_ratio: float,
                    controller_state: ControllerState, cfl_params: Optional[CFLParams] = None,
                    pid_params: Optional[PIDParams] = None, order: int = 2) -> Tuple[float, ControllerState]:
    """dt_policy.py:544-610."""
    cfl_params = cfl_params or CFLParams()
    pid_params = pid_params or PIDParams()
    if err_ratio <= 1.0:
        dt_ad, new_state = pid_controller_dt(dt_old, err_ratio, controller_state, pid_params, order)
    else:
        dt_ad, new_state = handle_rejected_step(dt_old, err_ratio, controller_state, pid_params, order)
    dt_final = _min(dt_ad, imex_cfl_dt(grid, params, cfl_params))
    return _clip(dt_final, pid_params.dt_min, pid_params.dt_max), new_state


def estimate_reaction_stiffness(state, params: Dict[str, Any]) -> float:
    """dt_policy.py:505-541: a bound on |df/du| of the reaction: with F, k
    (Gray-Scott) max(F + k, max|v|^2 + F, 2 max|u| max|v| + F + k), with r
    (Fisher-KPP) at least r; max|.| by the native FP64 reduction."""
    from .state import RED_MAXABS, _device_fields, _reduce
    stiffness = 0.0
    if "F" in params and "k" in params:
        F, k = float(params["F"]), float(params["k"])
        stiffness = max(stiffness, F + k)
        if "u" in state and "v" in state:
            um = float(_reduce(RED_MAXABS, _device_fields([state["u"]])).max())
            vm = float(_reduce(RED_MAXABS, _device_fields([state["v"]])).max())
            stiffness = max(stiffness, vm * vm + F, 2.0 * um * vm + F + k)
    if "r" in params:
        stiffness = max(stiffness, float(params["r"]))
    return stiffness
