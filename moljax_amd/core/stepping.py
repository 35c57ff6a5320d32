"""Time stepping on the MI355X kernels (moljax/core/stepping.py).

The IMEX-Strang Gray-Scott step is one fused kernel chain (mjx_gs_imex_strang):
row pass (x-FFT), column pass (y-FFT, divide, inverse), fused row pass
(inverse x-FFT, Heun reaction, forward x-FFT), column pass, row pass back
with the periodic ghost ring.  Multi-step drivers hand the whole step count to
the kernel chain, which merges the trailing half-solve of one step with the
leading half-solve of the next (one FFT round trip per step).

Models are served by kernel: ``model.metadata['mjx_kernel']`` names the fused
path.  Models without one are rejected with NotImplementedError; there is no
generic per-op fallback.
"""
import ctypes
from enum import IntEnum
from typing import Any, Dict, NamedTuple, Optional, Tuple

import numpy as np
import torch

from .. import _hip
from ._arrays import as_device, field_dtype, inst_table, spectral_work, view2d
from .bc import BCType
from .state import StateDict, tree_sub


class IntegratorType(IntEnum):
    """stepping.py:48-57."""
    EULER = 0
    SSPRK3 = 1
    RK4 = 2
    BE = 3
    CN = 4
    BDF2 = 5
    IMEX_EULER = 6
    IMEX_STRANG = 7
    ETDRK4 = 8  # added: fixed-dt ETDRK4 through adaptive_integrate


def _require_gray_scott(model, diffusivities: Dict[str, float]):
    if model.metadata.get("mjx_kernel") != "gray_scott":
        raise NotImplementedError(
            "the MI355X IMEX path serves Gray-Scott models from create_gray_scott_model / "
            "create_gray_scott_periodic_fft; this model has no fused kernel")
    for name in ("u", "v"):
        if model.bc_spec[name].kind != BCType.PERIODIC:
            raise NotImplementedError("the FFT IMEX path needs periodic boundaries on u and v")
    extra = set(diffusivities) - {"u", "v"}
    if extra:
        raise ValueError(f"diffusivities name unknown fields {sorted(extra)}")


GS_STRANG, GS_EULER, GS_SSPRK2 = 0, 1, 2
# IMEX-Strang in increment form (the same operator; the state stays physical
# and only each diffusion solve's increment (H^k - 1)(y - (1, 0)) goes through
# the transforms, so FP32 transform rounding scales with the increment;
# power-of-two sides).
GS_STRANG_INC = 4


def _gs_coeffs(model, diffusivities: Dict[str, float]):
    """(Du, Dv, F, k) with diffusion switched off for D <= 1e-14 as
    apply_diffusion_inverse_fft does (fft_solvers.py:436-441)."""
    Du = float(diffusivities.get("u", 0.0))
    Dv = float(diffusivities.get("v", 0.0))
    return (Du if Du > 1e-14 else 0.0, Dv if Dv > 1e-14 else 0.0,
            float(model.params.get("F", 0.04)), float(model.params.get("k", 0.06)))


def _gs_rows(model, dt: float, diffusivities: Dict[str, float], B: int, scheme: int = GS_STRANG):
    """Instance rows: Strang solves with dt/2, Euler and SSPRK2 with dt."""
    Du, Dv, F, k = _gs_coeffs(model, diffusivities)
    h = 0.5 * dt if scheme in (GS_STRANG, GS_STRANG_INC) else float(dt)
    return [[h * Du, h * Dv, 0.0, 0.0, F, k, float(dt)]] * B


def gs_imex_strang_inplace(u: torch.Tensor, v: torch.Tensor, lam, rows, n_steps: int,
                           fill_ghosts: bool = True, work: Optional[torch.Tensor] = None):
    """Advance padded (..., ny+2, nx+2) fields u, v in place by n_steps
    IMEX-Strang steps.  ``rows`` is a per-instance list of
    [beta_u, beta_v, 0, 0, F, k, dt] (beta = dt/2 * D) or an already built
    (B, 8) float64 device table; ``lam`` a symbol from laplacian_symbol_2d."""
    gs_imex_inplace(u, v, lam, rows, n_steps, GS_STRANG, fill_ghosts, work)


def gs_imex_inplace(u: torch.Tensor, v: torch.Tensor, lam, rows, n_steps: int, scheme: int = GS_STRANG,
                    fill_ghosts: bool = True, work: Optional[torch.Tensor] = None):
    """As gs_imex_strang_inplace for scheme GS_STRANG, GS_EULER, GS_SSPRK2 or
    GS_STRANG_INC (rows from _gs_rows with the same scheme)."""
    from .fft_solvers import _tables
    tu, pitch, istr, B = view2d(u)
    tv, pitch_v, istr_v, _ = view2d(v)
    if (pitch_v, istr_v) != (pitch, istr) or tu.dtype != tv.dtype:
        raise ValueError("u and v must share dtype and layout")
    ny, nx = tu.shape[-2] - 2, tu.shape[-1] - 2
    tabs = _tables(lam, tu.dtype)
    inst = rows if isinstance(rows, torch.Tensor) else inst_table(rows, tu.device)
    if inst.shape[0] != B:
        raise ValueError("one parameter row per instance is required")
    if work is None:
        work = spectral_work(tu.dtype, B, ny, nx, tu.device)
    rc = _hip.lib().mjx_gs_imex(
        _hip.dtype_code(tu.dtype), int(scheme), B, ny, nx, _hip.ptr(tu[:, 1:, 1:]), _hip.ptr(tv[:, 1:, 1:]), pitch,
        istr, int(bool(fill_ghosts)), _hip.ptr(tabs.lamx), _hip.ptr(tabs.lamy), _hip.ptr(tabs.lam2d),
        _hip.ptr(inst), int(n_steps), _hip.ptr(work), _hip.stream_ptr())
    _hip.check(rc, "mjx_gs_imex")


def _copy_state(y: StateDict):
    u = as_device(y["u"])
    dt = field_dtype(u)
    u = u.to(dt).clone()
    v = as_device(y["v"], dt).clone()
    if u.shape != v.shape:
        raise ValueError("u and v must have the same shape")
    return u, v


def imex_strang_step(model, y: StateDict, t: float, dt: float, fft_cache,
                     diffusivities: Dict[str, float]) -> StateDict:
    """stepping.py:370-435: half implicit diffusion step, Heun reaction step,
    half implicit diffusion step; returns a new state with periodic ghosts."""
    _require_gray_scott(model, diffusivities)
    u, v = _copy_state(y)
    B = view2d(u)[3]
    gs_imex_strang_inplace(u, v, fft_cache.laplacian_symbol, _gs_rows(model, dt, diffusivities, B), 1)
    return {"u": u, "v": v}


def imex_euler_step(model, y: StateDict, t: float, dt: float, fft_cache,
                    diffusivities: Dict[str, float]) -> StateDict:
    """stepping.py:317-367: y_new = (I - dt D lap)^-1 (y + dt N(y))."""
    _require_gray_scott(model, diffusivities)
    u, v = _copy_state(y)
    B = view2d(u)[3]
    gs_imex_inplace(u, v, fft_cache.laplacian_symbol, _gs_rows(model, dt, diffusivities, B, GS_EULER), 1, GS_EULER)
    return {"u": u, "v": v}


def imex_ssprk2_step(model, y: StateDict, t: float, dt: float, fft_cache,
                     diffusivities: Dict[str, float]) -> StateDict:
    """stepping.py:438-495: y_new = y/2 + E(E(y))/2, E the IMEX-Euler map."""
    _require_gray_scott(model, diffusivities)
    u, v = _copy_state(y)
    B = view2d(u)[3]
    gs_imex_inplace(u, v, fft_cache.laplacian_symbol, _gs_rows(model, dt, diffusivities, B, GS_SSPRK2), 1,
                    GS_SSPRK2)
    return {"u": u, "v": v}


def estimate_error_imex_doubling(model, y: StateDict, t: float, dt: float, fft_cache,
                                 diffusivities: Dict[str, float],
                                 use_strang: bool = True) -> Tuple[StateDict, StateDict]:
    """stepping.py:566-608: (two half steps, their difference from one full
    step)."""
    step = imex_strang_step if use_strang else imex_euler_step
    y_full = step(model, y, t, dt, fft_cache, diffusivities)
    y_half = step(model, step(model, y, t, dt / 2.0, fft_cache, diffusivities), t + dt / 2.0, dt / 2.0, fft_cache,
                  diffusivities)
    from .state import axpby
    return y_half, {n: axpby(1.0, y_half[n], -1.0, y_full[n]) for n in y_half}


class AdaptiveResult(NamedTuple):
    """stepping.py:710-720.  For a batch of instances (leading dimension B
    on the state) every field gains that leading dimension."""
    t_final: object
    y_final: StateDict
    t_history: torch.Tensor
    y_history: Dict[str, torch.Tensor]
    dt_history: torch.Tensor
    n_steps: object
    n_accepted: object
    n_rejected: object
    status: object


AD_IMEX_STRANG, AD_IMEX_EULER, AD_EXPL_EULER, AD_EXPL_SSPRK3, AD_EXPL_RK4 = 0, 1, 3, 4, 5


def _adaptive_gs(model, y0: StateDict, t0: float, t_end: float, dt0: float, scheme: int, order: int, dt_cfl: float,
                 coeffs, lam, max_steps: int, pid, save_every: int, save_states: bool,
                 max_attempts: int) -> AdaptiveResult:
    """Shared body of the adaptive drivers: one mjx_gs_imex_adaptive call
    with the per-instance controller state initialised as the reference's
    while_loop carry (stepping.py:1199-1224)."""
    if save_every < 1:
        raise ValueError("save_every must be >= 1")
    u, v = _copy_state(y0)
    batched = u.dim() == 3
    tu, pitch, istr, B = view2d(u)
    tv = view2d(v)[0]
    ny, nx = tu.shape[-2] - 2, tu.shape[-1] - 2
    dev, dt_ = u.device, u.dtype
    dt_init = min(float(dt0), dt_cfl)
    max_saves = max_steps // save_every + 1
    lib = _hip.lib()
    st = np.zeros((B, 16))
    st[:, 0], st[:, 1], st[:, 2], st[:, 3] = t0, dt_init, t_end, dt_cfl
    st[:, 4] = st[:, 5] = 1.0
    st[:, 8], st[:, 9] = 1.0, 1.0
    state = torch.from_numpy(st).to(dev)
    coeffs = torch.tensor([coeffs] * B, dtype=torch.float64, device=dev)
    t_hist = torch.zeros((B, max_saves), dtype=torch.float64, device=dev)
    t_hist[:, 0] = t0
    dt_hist = torch.zeros_like(t_hist)
    y_hist = None
    if save_states:
        y_hist = torch.zeros((B, 2, max_saves, ny, nx), dtype=dt_, device=dev)
        y_hist[:, 0, 0] = tu[:, 1:-1, 1:-1]
        y_hist[:, 1, 0] = tv[:, 1:-1, 1:-1]
    tabs = None
    if lam is not None:
        from .fft_solvers import _tables
        tabs = _tables(lam, dt_)
    work = torch.empty(lib.mjx_gs_adaptive_workspace_bytes(_hip.dtype_code(dt_), scheme, B, ny, nx),
                       dtype=torch.uint8, device=dev)
    g = model.grid
    ctl = (ctypes.c_double * 13)(pid.atol, pid.rtol, pid.safety, pid.kI, pid.kP, pid.dt_min, pid.dt_max,
                                 pid.max_factor, pid.min_factor, float(order), float(save_every), float(g.dy),
                                 float(g.dx))
    done = ctypes.c_longlong(0)
    rc = lib.mjx_gs_imex_adaptive(
        _hip.dtype_code(dt_), scheme, B, ny, nx, _hip.ptr(tu[:, 1:, 1:]), _hip.ptr(tv[:, 1:, 1:]), pitch, istr, 1,
        _hip.ptr(tabs.lamx) if tabs else 0, _hip.ptr(tabs.lamy) if tabs else 0, _hip.ptr(tabs.lam2d) if tabs else 0,
        _hip.ptr(coeffs), ctl, _hip.ptr(state), _hip.ptr(t_hist), _hip.ptr(dt_hist), _hip.ptr(y_hist), max_saves,
        int(max_attempts), ctypes.byref(done), _hip.ptr(work), _hip.stream_ptr())
    _hip.check(rc, "mjx_gs_imex_adaptive")
    s = state.cpu().numpy()
    yh = {"u": y_hist[:, 0], "v": y_hist[:, 1]} if y_hist is not None else {}
    if not batched:
        yh = {k: h[0] for k, h in yh.items()}
        return AdaptiveResult(t_final=float(s[0, 0]), y_final={"u": u, "v": v}, t_history=t_hist[0], y_history=yh,
                              dt_history=dt_hist[0], n_steps=int(s[0, 8]), n_accepted=int(s[0, 6]),
                              n_rejected=int(s[0, 7]), status=int(s[0, 9]))
    return AdaptiveResult(t_final=torch.from_numpy(s[:, 0].copy()), y_final={"u": u, "v": v}, t_history=t_hist,
                          y_history=yh, dt_history=dt_hist, n_steps=torch.from_numpy(s[:, 8].astype(np.int64)),
                          n_accepted=torch.from_numpy(s[:, 6].astype(np.int64)),
                          n_rejected=torch.from_numpy(s[:, 7].astype(np.int64)),
                          status=torch.from_numpy(s[:, 9].astype(np.int64)))


def adaptive_integrate_imex(model, y0: StateDict, t0: float, t_end: float, dt0: float, fft_cache,
                            diffusivities: Dict[str, float], use_strang: bool = True, max_steps: int = 10000,
                            cfl_params=None, pid_params=None, save_every: int = 1, save_states: bool = True,
                            max_attempts: int = 0) -> AdaptiveResult:
    """stepping.py:1146-1371 on the device: step-doubling IMEX (Strang,
    order 2, or Euler, order 1) with the PI controller running per instance
    inside mjx_gs_imex_adaptive.

    y0 holds padded fields (ny+2, nx+2) or a batch (B, ny+2, nx+2) of
    independent instances, each with its own dt sequence.  y_history holds
    interior states (max_steps // save_every + 1 slots) like the reference;
    save_states=False skips it (t_history and dt_history are always kept).
    max_attempts > 0 bounds the number of attempts (status stays RUNNING
    when it cuts the run short)."""
    from .dt_policy import CFLParams, PIDParams, imex_cfl_dt
    _require_gray_scott(model, diffusivities)
    cfl = cfl_params or CFLParams()
    pid = pid_params or PIDParams()
    return _adaptive_gs(model, y0, t0, t_end, dt0, AD_IMEX_STRANG if use_strang else AD_IMEX_EULER,
                        2 if use_strang else 1, imex_cfl_dt(model.grid, model.params, cfl),
                        _gs_coeffs(model, diffusivities), fft_cache.laplacian_symbol, max_steps, pid, save_every,
                        save_states, max_attempts)


_EXPLICIT = {IntegratorType.EULER: 0, IntegratorType.SSPRK3: 1, IntegratorType.RK4: 2}


def _require_gs_explicit(model):
    if getattr(model, "metadata", {}).get("mjx_kernel") != "gray_scott":
        raise NotImplementedError("explicit steps on the MI355X path serve Gray-Scott models "
                                  "(create_gray_scott_model); this model has no fused kernel")
    for name in ("u", "v"):
        if model.bc_spec[name].kind != BCType.PERIODIC:
            raise NotImplementedError("the fused explicit Gray-Scott stages serve periodic boundaries")


def _explicit_coeffs(model):
    p = model.params
    return (float(p.get("Du", 0.16)), float(p.get("Dv", 0.08)), float(p.get("F", 0.04)), float(p.get("k", 0.06)))


def gs_explicit_inplace(model, u: torch.Tensor, v: torch.Tensor, dt: float, n_steps: int, method: int,
                        fill_ghosts: bool = True):
    """n_steps explicit steps (method 0 Euler, 1 SSPRK3, 2 RK4) of the
    periodic Gray-Scott model, in place on padded u, v (mjx_gs_explicit)."""
    tu, pitch, istr, B = view2d(u)
    tv, pv, iv, _ = view2d(v)
    if (pv, iv) != (pitch, istr) or tu.dtype != tv.dtype:
        raise ValueError("u and v must share dtype and layout")
    ny, nx = tu.shape[-2] - 2, tu.shape[-1] - 2
    Du, Dv, F, k = _explicit_coeffs(model)
    rows = inst_table([[0.0, 0.0, Du, Dv, F, k, float(dt)]] * B, u.device)
    lib = _hip.lib()
    work = torch.empty(lib.mjx_gs_explicit_workspace_bytes(_hip.dtype_code(tu.dtype), B, ny, nx), dtype=torch.uint8,
                       device=u.device)
    rc = lib.mjx_gs_explicit(_hip.dtype_code(tu.dtype), int(method), B, ny, nx, _hip.ptr(tu[:, 1:, 1:]),
                             _hip.ptr(tv[:, 1:, 1:]), pitch, istr, int(bool(fill_ghosts)), float(model.grid.dy),
                             float(model.grid.dx), _hip.ptr(rows), int(n_steps), _hip.ptr(work), _hip.stream_ptr())
    _hip.check(rc, "mjx_gs_explicit")


def _explicit_step(model, y: StateDict, dt: float, method: int) -> StateDict:
    if getattr(model, "mjx_kernel", None) == "reaction_diffusion":
        from . import reaction_diffusion as rd
        if method != 0:
            raise NotImplementedError("reaction-diffusion models serve forward Euler explicit steps")
        return rd.euler_step(model, y, 0.0, dt)
    _require_gs_explicit(model)
    u, v = _copy_state(y)
    gs_explicit_inplace(model, u, v, dt, 1, method)
    return {"u": u, "v": v}


def euler_step(model, y: StateDict, t: float, dt: float) -> StateDict:
    """stepping.py:64-85: y + dt F(y)."""
    return _explicit_step(model, y, dt, 0)


def ssprk3_step(model, y: StateDict, t: float, dt: float) -> StateDict:
    """stepping.py:88-124 (Shu-Osher SSPRK3)."""
    return _explicit_step(model, y, dt, 1)


def rk4_step(model, y: StateDict, t: float, dt: float) -> StateDict:
    """stepping.py:127-167 (classical RK4)."""
    return _explicit_step(model, y, dt, 2)


def step_explicit(model, y: StateDict, t: float, dt: float, method: int) -> StateDict:
    """stepping.py:615-643: one explicit step, method 0 Euler, 1 SSPRK3,
    2 RK4."""
    if int(method) not in (0, 1, 2):
        raise ValueError("method must be 0 (Euler), 1 (SSPRK3) or 2 (RK4)")
    return (euler_step, ssprk3_step, rk4_step)[int(method)](model, y, t, dt)


def estimate_error_doubling(model, y: StateDict, t: float, dt: float, step_fn) -> Tuple[StateDict, StateDict]:
    """stepping.py:502-536: one step of dt and two of dt/2; returns the
    two-half-step state and (y_half2 - y_full), unscaled by the order."""
    y_full = step_fn(model, y, t, dt)
    y_half1 = step_fn(model, y, t, dt / 2.0)
    y_half2 = step_fn(model, y_half1, t + dt / 2.0, dt / 2.0)
    return y_half2, tree_sub(y_half2, y_full)


def estimate_error_implicit(model, y: StateDict, t: float, dt: float, y_be: StateDict,
                            y_cn: StateDict) -> StateDict:
    """stepping.py:539-563: y_cn - y_be."""
    return tree_sub(y_cn, y_be)


def step_explicit_with_error(model, y: StateDict, t: float, dt: float, method: int) -> Tuple[StateDict, StateDict]:
    """stepping.py:646-680: step doubling with the method's step."""
    if int(method) not in (0, 1, 2):
        raise ValueError("method must be 0 (Euler), 1 (SSPRK3) or 2 (RK4)")
    return estimate_error_doubling(model, y, t, dt, (euler_step, ssprk3_step, rk4_step)[int(method)])


def be_step(model, y: StateDict, t: float, dt: float, preconditioner=None, nk_params=None):
    """stepping.py:170-214 (reaction-diffusion models, native Newton-Krylov)."""
    from .reaction_diffusion import be_step as _be
    return _be(model, y, t, dt, preconditioner, nk_params)


def cn_step(model, y: StateDict, t: float, dt: float, preconditioner=None, nk_params=None):
    """stepping.py:217-260."""
    from .reaction_diffusion import cn_step as _cn
    return _cn(model, y, t, dt, preconditioner, nk_params)


def bdf2_step(model, y: StateDict, y_prev: StateDict, t: float, dt: float, dt_prev: float, preconditioner=None,
              nk_params=None):
    """stepping.py:263-310."""
    from .reaction_diffusion import bdf2_step as _bdf2
    return _bdf2(model, y, y_prev, t, dt, dt_prev, preconditioner, nk_params)


class AdaptiveState(NamedTuple):
    """stepping.py:683-708: the adaptive loop's carried state (the drivers
    here keep it on the device and in Python locals; the type is kept for
    callers that build or inspect it)."""
    t: Any
    y: StateDict
    dt: Any
    y_prev: StateDict
    dt_prev: Any
    step_count: Any
    controller: Any
    t_history: Any
    y_history: Dict[str, Any]
    dt_history: Any
    write_idx: Any
    n_accepted: Any
    n_rejected: Any
    status: Any


def integrate_fixed_dt(model, y0: StateDict, t0: float, t_end: float, dt: float,
                       method: int = IntegratorType.RK4, save_every: int = 1, preconditioner=None,
                       nk_params=None, n_steps: Optional[int] = None):
    """stepping.py:1024-1143 for explicit methods: int((t_end - t0)/dt) + 1
    steps (the reference's count) unless n_steps is given; returns
    (t_history, y_history, y_final) with the padded state after every
    save_every-th step."""
    m = int(method)
    if m not in _EXPLICIT:
        raise NotImplementedError("integrate_fixed_dt on the MI355X path serves explicit methods; "
                                  "use be_step / cn_step for implicit solves")
    _require_gs_explicit(model)
    if save_every < 1:
        raise ValueError("save_every must be >= 1")
    n = int((t_end - t0) / dt) + 1 if n_steps is None else int(n_steps)
    u, v = _copy_state(y0)
    k = 0
    n_saves = n // save_every
    t_hist = torch.empty(n_saves, dtype=torch.float64)
    hist_u = torch.empty((n_saves,) + tuple(u.shape), dtype=u.dtype, device=u.device)
    hist_v = torch.empty_like(hist_u)
    done = 0
    while done < n:
        chunk = min(save_every, n - done)
        gs_explicit_inplace(model, u, v, dt, chunk, _EXPLICIT[m])
        done += chunk
        if chunk == save_every:
            hist_u[k].copy_(u)
            hist_v[k].copy_(v)
            t_hist[k] = t0 + done * dt
            k += 1
    return t_hist, {"u": hist_u, "v": hist_v}, {"u": u, "v": v}


def adaptive_integrate(model, y0: StateDict, t0: float, t_end: float, dt0: float,
                       method: int = IntegratorType.RK4, max_steps: int = 10000, cfl_params=None, pid_params=None,
                       preconditioner=None, nk_params=None, save_every: int = 1, save_states: bool = True,
                       max_attempts: int = 0) -> AdaptiveResult:
    """stepping.py:723-1021, routed to the device drivers:
      EULER / SSPRK3 / RK4   step doubling with the fused explicit stages,
                             dt limited by heisenberg_cfl_dt, error order
                             compute_error_order(method);
      IMEX_EULER / IMEX_STRANG  adaptive_integrate_imex (the reference sends
                             these values to its BDF2 branch by mistake,
                             stepping.py:864-879);
      ETDRK4                 fixed steps of dt0 (a final shorter step lands
                             on t_end) with the batched ETDRK4 kernel.
      BE / CN / BDF2         reaction-diffusion models: every attempt a native
                             Newton-Krylov solve (adaptive_integrate_implicit)."""
    from .dt_policy import CFLParams, PIDParams, compute_error_order, heisenberg_cfl_dt
    m = int(method)
    cfl = cfl_params or CFLParams()
    pid = pid_params or PIDParams()
    if m in _EXPLICIT:
        _require_gs_explicit(model)
        return _adaptive_gs(model, y0, t0, t_end, dt0, AD_EXPL_EULER + _EXPLICIT[m], compute_error_order(m),
                            heisenberg_cfl_dt(model.grid, model.params, cfl), _explicit_coeffs(model), None,
                            max_steps, pid, save_every, save_states, max_attempts)
    if m in (IntegratorType.IMEX_EULER, IntegratorType.IMEX_STRANG):
        from .model import get_diffusivities, get_fft_cache
        from .fft_solvers import create_fft_cache
        cache = get_fft_cache(model) or create_fft_cache(model.grid, model.dtype)
        return adaptive_integrate_imex(model, y0, t0, t_end, dt0, cache, get_diffusivities(model),
                                       use_strang=(m == IntegratorType.IMEX_STRANG), max_steps=max_steps,
                                       cfl_params=cfl, pid_params=pid, save_every=save_every,
                                       save_states=save_states, max_attempts=max_attempts)
    if m == IntegratorType.ETDRK4:
        return _etdrk4_fixed(model, y0, t0, t_end, dt0, max_steps, save_every)
    if m in (IntegratorType.BE, IntegratorType.CN, IntegratorType.BDF2):
        if getattr(model, "mjx_kernel", None) != "reaction_diffusion":
            raise NotImplementedError("adaptive implicit integration on the MI355X path serves reaction-diffusion "
                                      "models (create_reaction_diffusion_model)")
        from .reaction_diffusion import adaptive_integrate_implicit
        return adaptive_integrate_implicit(model, y0, t0, t_end, dt0, m, max_steps, cfl, pid, preconditioner,
                                           nk_params, save_every, save_states)
    raise ValueError(f"unknown integrator {method}")


def _etdrk4_fixed(model, y0: StateDict, t0: float, t_end: float, dt: float, max_steps: int,
                  save_every: int) -> AdaptiveResult:
    from .fft_integrators import etdrk4_coefficients, gs_etdrk4_inplace, _fd_symbol
    if getattr(model, "metadata", {}).get("mjx_kernel") != "gray_scott":
        raise NotImplementedError("ETDRK4 on the MI355X path serves Gray-Scott models")
    g = model.grid
    Du, Dv, F, k = _explicit_coeffs(model)
    u, v = _copy_state(y0)
    lamx, lamy = _fd_symbol(g.nx, g.dx), _fd_symbol(g.ny, g.dy)
    ui = u[..., 1:-1, 1:-1].contiguous()
    vi = v[..., 1:-1, 1:-1].contiguous()
    B = view2d(ui)[3]
    n_full = int(np.floor((t_end - t0) / dt + 1e-12))
    rem = (t_end - t0) - n_full * dt
    steps = [dt] * n_full + ([rem] if rem > 1e-12 * dt else [])
    max_saves = max_steps // save_every + 1
    t_hist = torch.zeros(max_saves, dtype=torch.float64)
    dt_hist = torch.zeros(max_saves, dtype=torch.float64)
    t_hist[0] = t0
    hist = {"u": [ui.clone()], "v": [vi.clone()]}
    t, w, done = float(t0), 1, 0
    for h in steps:
        if w >= max_saves:
            break
        coef = etdrk4_coefficients(lamx, lamy, Du, Dv, h, ui.dtype)
        gs_etdrk4_inplace(ui, vi, coef, [[0, 0, 0, 0, F, k, h]] * B, 1)
        t += h
        if done % save_every == 0:
            t_hist[w] = t
            dt_hist[w - 1] = h
            hist["u"].append(ui.clone())
            hist["v"].append(vi.clone())
            w += 1
        done += 1
    u[..., 1:-1, 1:-1] = ui
    v[..., 1:-1, 1:-1] = vi
    from .bc_fill import apply_bc
    y = apply_bc({"u": u, "v": v}, g, model.bc_spec, t, model.params)
    status = 0 if done == len(steps) else 2
    return AdaptiveResult(t_final=t, y_final=y, t_history=t_hist,
                          y_history={k_: torch.stack(h) for k_, h in hist.items()}, dt_history=dt_hist, n_steps=w,
                          n_accepted=done, n_rejected=0, status=status)


def integrate_imex_fixed_dt(model, y0: StateDict, t0: float, t_end: float, dt: float, fft_cache,
                            diffusivities: Dict[str, float], use_strang: bool = True, save_every: int = 1,
                            n_steps: Optional[int] = None, increment_form: bool = False
                            ) -> Tuple[torch.Tensor, Dict[str, torch.Tensor], StateDict]:
    """stepping.py:1374-1436.

    Takes int((t_end - t0)/dt) + 1 steps like the reference (its extra step is
    kept) unless ``n_steps`` is given.  Returns (t_history, y_history, y_final)
    where the histories hold the padded state after every ``save_every``-th
    step (the reference stacks every step; save_every=1 gives its shapes).
    ``increment_form`` (Strang, power-of-two sides) runs the same operator
    with only the diffusion increments through the transforms: in float32
    it stays within 1e-5 of float64 over long horizons where the plain form
    drifts (a new option; the reference has one form)."""
    _require_gray_scott(model, diffusivities)
    scheme = GS_STRANG if use_strang else GS_EULER
    if increment_form:
        if not use_strang:
            raise ValueError("increment_form applies to the Strang scheme")
        scheme = GS_STRANG_INC
    if save_every < 1:
        raise ValueError("save_every must be >= 1")
    n = int((t_end - t0) / dt) + 1 if n_steps is None else int(n_steps)
    u, v = _copy_state(y0)
    tu, _, _, B = view2d(u)
    rows = inst_table(_gs_rows(model, dt, diffusivities, B, scheme), u.device)
    work = spectral_work(u.dtype, B, tu.shape[-2] - 2, tu.shape[-1] - 2, u.device)
    n_saves = n // save_every
    t_hist = torch.empty(n_saves, dtype=torch.float64)
    hist_u = torch.empty((n_saves,) + tuple(u.shape), dtype=u.dtype, device=u.device)
    hist_v = torch.empty_like(hist_u)
    done, k = 0, 0
    while done < n:
        chunk = min(save_every, n - done)
        gs_imex_inplace(u, v, fft_cache.laplacian_symbol, rows, chunk, scheme, work=work)
        done += chunk
        if chunk == save_every:
            hist_u[k].copy_(u)
            hist_v[k].copy_(v)
            t_hist[k] = t0 + done * dt
            k += 1
    return t_hist, {"u": hist_u, "v": hist_v}, {"u": u, "v": v}
