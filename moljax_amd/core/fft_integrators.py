"""Exponential time differencing on the MI355X kernels
(moljax/core/fft_integrators.py).

ETDRK4 (Cox-Matthews) for coupled Gray-Scott runs as one batched kernel
chain (mjx_gs_etdrk4) with FP64 contour-integral coefficients computed once
per (grid, dt, Du, Dv) on the device (mjx_etdrk4_coeffs, the Kassam-Trefethen
form of benchmarks/benchmark_gray_scott.py:336-356).  The reference library's
etdrk4_step evaluates the nonlinearity one field at a time (fft_integrators.py
:311,319,327) and cannot couple u and v; here the coupled form is served and
other nonlinearities raise NotImplementedError.

ETD1 (exponential Euler) and ETD2 (exponential Adams-Bashforth-2) run as one
row pass and one column pass per step (mjx_gs_etd_exp) with FP64 tables
e^z, dt phi1(z), dt phi2(z) (mjx_etd_phi_coeffs); ETD2 keeps the previous
step's spectral nonlinear term on the device between steps and calls.
"""
from typing import Callable, Dict, NamedTuple, Optional, Tuple

import numpy as np
import torch

from .. import _hip
from ._arrays import as_device, field_dtype, inst_table, view2d
from .fft_operators import DiffusionOperator
from .state import StateDict


def _fd_symbol(n: int, d: float) -> np.ndarray:
    k = 2.0 * np.pi * np.fft.fftfreq(n, d=d)
    return (2.0 * np.cos(k * d) - 2.0) / (d * d)


def spectral_symbol_1d(n: int, L: float = 1.0) -> np.ndarray:
    """-(2 pi fftfreq(n, L/n))^2, the pseudo-spectral symbol of the ETDRK4
    benchmark (benchmarks/benchmark_gray_scott.py:52-57)."""
    k = np.fft.fftfreq(n, L / n) * 2 * np.pi
    return -(k ** 2)


_COEF_CACHE: Dict[tuple, torch.Tensor] = {}


def etdrk4_coefficients(lamx: np.ndarray, lamy: np.ndarray, Du: float, Dv: float, dt: float,
                        dtype=torch.float32, cox_matthews: bool = True) -> torch.Tensor:
    """(6, ny, nx+2) device tables E, E2, Q, f1, f2, f3 for both fields in the
    separated spectral layout, FP64 contour means rounded once to ``dtype``.
    ``cox_matthews=False`` keeps the benchmark's f2 as written
    (benchmark_gray_scott.py:414-415, without the factor 2).  Cached per
    (symbol, Du, Dv, dt, dtype, form)."""
    lamx = np.ascontiguousarray(lamx, dtype=np.float64)
    lamy = np.ascontiguousarray(lamy, dtype=np.float64)
    key = (lamx.tobytes(), lamy.tobytes(), float(Du), float(Dv), float(dt), dtype, bool(cox_matthews))
    hit = _COEF_CACHE.get(key)
    if hit is not None:
        return hit
    lib = _hip.lib()
    ny, nx = lamy.size, lamx.size
    lx = torch.from_numpy(lamx).cuda()
    ly = torch.from_numpy(lamy).cuda()
    coef = torch.empty((6, ny, nx + 2), dtype=dtype, device="cuda")
    rc = lib.mjx_etdrk4_coeffs_form(_hip.dtype_code(dtype), 0 if cox_matthews else 1, ny, nx, _hip.ptr(lx),
                                    _hip.ptr(ly), float(Du), float(Dv), float(dt), _hip.ptr(coef),
                                    _hip.stream_ptr())
    _hip.check(rc, "mjx_etdrk4_coeffs_form")
    _COEF_CACHE[key] = coef
    return coef


def gs_etdrk4_inplace(u: torch.Tensor, v: torch.Tensor, coef: torch.Tensor, rows, n_steps: int,
                      fill_ghosts: bool = False, work: Optional[torch.Tensor] = None):
    """Advance (..., n, n) interior arrays (or padded ones with
    ``fill_ghosts``: pass the padded tensors) by n_steps ETDRK4 steps.
    ``rows``: per-instance [*, *, *, *, F, k, dt] lists or a (B, 8) table."""
    tu, pitch, istr, B = view2d(u)
    tv, pv, iv, _ = view2d(v)
    if (pv, iv) != (pitch, istr) or tu.dtype != tv.dtype:
        raise ValueError("u and v must share dtype and layout")
    if coef.dtype != tu.dtype:
        raise ValueError("coefficient dtype must match the state dtype")
    off = 1 if fill_ghosts else 0
    ny, nx = tu.shape[-2] - 2 * off, tu.shape[-1] - 2 * off
    inst = rows if isinstance(rows, torch.Tensor) else inst_table(rows, tu.device)
    lib = _hip.lib()
    if work is None:
        nbytes = lib.mjx_etdrk4_workspace_bytes(_hip.dtype_code(tu.dtype), B, ny, nx)
        work = torch.empty(nbytes, dtype=torch.uint8, device=tu.device)
    pu = tu[:, 1:, 1:] if off else tu
    pvv = tv[:, 1:, 1:] if off else tv
    rc = lib.mjx_gs_etdrk4(_hip.dtype_code(tu.dtype), B, ny, nx, _hip.ptr(pu), _hip.ptr(pvv), pitch, istr,
                           int(bool(fill_ghosts)), _hip.ptr(coef), _hip.ptr(inst), int(n_steps), _hip.ptr(work),
                           _hip.stream_ptr())
    _hip.check(rc, "mjx_gs_etdrk4")


_PHI_CACHE: Dict[tuple, torch.Tensor] = {}


def etd_phi_coefficients(lamx: np.ndarray, lamy: np.ndarray, Du: float, Dv: float, dt: float,
                         dtype=torch.float32) -> torch.Tensor:
    """(3, ny, nx+2) device tables e^z, dt phi1(z), dt phi2(z), z = dt D lam,
    for both fields in the separated spectral layout (fft_integrators.py
    :37-64 in FP64, rounded once to ``dtype``).  Cached."""
    lamx = np.ascontiguousarray(lamx, dtype=np.float64)
    lamy = np.ascontiguousarray(lamy, dtype=np.float64)
    key = (lamx.tobytes(), lamy.tobytes(), float(Du), float(Dv), float(dt), dtype)
    hit = _PHI_CACHE.get(key)
    if hit is not None:
        return hit
    lib = _hip.lib()
    ny, nx = lamy.size, lamx.size
    lx = torch.from_numpy(lamx).cuda()
    ly = torch.from_numpy(lamy).cuda()
    coef = torch.empty((3, ny, nx + 2), dtype=dtype, device="cuda")
    rc = lib.mjx_etd_phi_coeffs(_hip.dtype_code(dtype), ny, nx, _hip.ptr(lx), _hip.ptr(ly), float(Du), float(Dv),
                                float(dt), _hip.ptr(coef), _hip.stream_ptr())
    _hip.check(rc, "mjx_etd_phi_coeffs")
    _PHI_CACHE[key] = coef
    return coef


def gs_etd_exp_inplace(u: torch.Tensor, v: torch.Tensor, coef: torch.Tensor, rows, order: int, n_steps: int,
                       prev: int = 0, n_prev: Optional[Tuple[torch.Tensor, torch.Tensor]] = None,
                       fill_ghosts: bool = False, work: Optional[torch.Tensor] = None) -> torch.Tensor:
    """n_steps of ETD1 (order 1) or ETD2 (order 2) in place on (..., n, n)
    arrays.  ETD2's N_{n-1}: prev 0 none (first step ETD1), 1 physical
    ``n_prev`` = (Nu, Nv) laid out like (u, v), 2 the one kept in ``work`` by
    the previous call with the same work.  Returns the workspace."""
    tu, pitch, istr, B = view2d(u)
    tv, pv, iv, _ = view2d(v)
    if (pv, iv) != (pitch, istr) or tu.dtype != tv.dtype:
        raise ValueError("u and v must share dtype and layout")
    if coef.dtype != tu.dtype:
        raise ValueError("coefficient dtype must match the state dtype")
    off = 1 if fill_ghosts else 0
    ny, nx = tu.shape[-2] - 2 * off, tu.shape[-1] - 2 * off
    inst = rows if isinstance(rows, torch.Tensor) else inst_table(rows, tu.device)
    lib = _hip.lib()
    if work is None:
        nbytes = lib.mjx_etdrk4_workspace_bytes(_hip.dtype_code(tu.dtype), B, ny, nx)
        work = torch.empty(nbytes, dtype=torch.uint8, device=tu.device)
    pu = tu[:, 1:, 1:] if off else tu
    pvv = tv[:, 1:, 1:] if off else tv
    nu = nv = None
    if prev == 1:
        if n_prev is None:
            raise ValueError("prev = 1 needs n_prev")
        nu_t, npitch, nistr, _ = view2d(n_prev[0].to(tu.dtype))
        nv_t, _, _, _ = view2d(n_prev[1].to(tu.dtype))
        if (npitch, nistr) != (pitch, istr) or nu_t.shape != tu.shape or nv_t.shape != tu.shape:
            raise ValueError("n_prev must be laid out like the state")
        nu, nv = (nu_t[:, 1:, 1:], nv_t[:, 1:, 1:]) if off else (nu_t, nv_t)
    rc = lib.mjx_gs_etd_exp(_hip.dtype_code(tu.dtype), int(order), B, ny, nx, _hip.ptr(pu), _hip.ptr(pvv), pitch,
                            istr, int(bool(fill_ghosts)), _hip.ptr(coef), _hip.ptr(inst), int(prev),
                            _hip.ptr(nu) if nu is not None else None, _hip.ptr(nv) if nv is not None else None,
                            int(n_steps), _hip.ptr(work), _hip.stream_ptr())
    _hip.check(rc, "mjx_gs_etd_exp")
    return work


def gray_scott_nonlinear_rhs(F: float, k: float) -> Callable[[StateDict, float], StateDict]:
    """N(u, v) = (-u v^2 + F(1-u), u v^2 - (F+k) v) as a nonlinear_rhs callable
    the ETD drivers recognise."""
    from .operators import gray_scott_reaction

    def nonlinear_rhs(state: StateDict, t: float) -> StateDict:
        ru, rv = gray_scott_reaction(state["u"], state["v"], F, k)
        return {"u": ru, "v": rv}

    nonlinear_rhs._mjx_gray_scott = (float(F), float(k))
    return nonlinear_rhs


def _gray_scott_params(nonlinear_rhs) -> Tuple[float, float]:
    tag = getattr(nonlinear_rhs, "_mjx_gray_scott", None)
    if tag is not None:
        return tag
    owner = getattr(nonlinear_rhs, "__self__", None)
    if owner is not None and getattr(owner, "metadata", {}).get("mjx_kernel") == "gray_scott":
        return float(owner.params.get("F", 0.04)), float(owner.params.get("k", 0.06))
    raise NotImplementedError("ETDRK4 on the MI355X path serves the coupled Gray-Scott nonlinearity "
                              "(gray_scott_nonlinear_rhs or a Gray-Scott model's nonlinear_rhs)")


def _linear_symbols(linear_ops) -> Tuple[np.ndarray, np.ndarray, float, float]:
    try:
        ou, ov = linear_ops["u"], linear_ops["v"]
    except KeyError as e:
        raise NotImplementedError("ETDRK4 needs linear operators for both 'u' and 'v'") from e
    for op in (ou, ov):
        if not isinstance(op, DiffusionOperator):
            raise NotImplementedError("ETDRK4 linear operators must be DiffusionOperator instances")
    g = ou.grid
    if ov.grid != g:
        raise ValueError("u and v operators must share the grid")
    return _fd_symbol(g.nx, g.dx), _fd_symbol(g.ny, g.dy), float(ou.D), float(ov.D)


def etdrk4_step(state: StateDict, t: float, dt: float, linear_ops: Dict[str, DiffusionOperator],
                nonlinear_rhs: Callable[[StateDict, float], StateDict]) -> StateDict:
    """fft_integrators.py:252-341, coupled: one ETDRK4 step of (u, v) interior
    arrays with L = D lap_h (FD symbol) per field."""
    F, k = _gray_scott_params(nonlinear_rhs)
    lamx, lamy, Du, Dv = _linear_symbols(linear_ops)
    u = as_device(state["u"])
    dtype = field_dtype(u)
    u = u.to(dtype).clone().contiguous()
    v = as_device(state["v"], dtype).clone().contiguous()
    coef = etdrk4_coefficients(lamx, lamy, Du, Dv, dt, dtype)
    B = view2d(u)[3]
    gs_etdrk4_inplace(u, v, coef, [[0, 0, 0, 0, F, k, dt]] * B, 1)
    return {"u": u, "v": v}


def _exp_setup(state: StateDict, dt: float, linear_ops, nonlinear_rhs):
    F, k = _gray_scott_params(nonlinear_rhs)
    lamx, lamy, Du, Dv = _linear_symbols(linear_ops)
    u = as_device(state["u"])
    dtype = field_dtype(u)
    u = u.to(dtype).clone().contiguous()
    v = as_device(state["v"], dtype).clone().contiguous()
    coef = etd_phi_coefficients(lamx, lamy, Du, Dv, dt, dtype)
    return u, v, coef, [[0, 0, 0, 0, F, k, dt]] * view2d(u)[3], F, k


def etd1_step(state: StateDict, t: float, dt: float, linear_ops: Dict[str, DiffusionOperator],
              nonlinear_rhs: Callable[[StateDict, float], StateDict]) -> StateDict:
    """fft_integrators.py:119-174: u^ <- e^z u^ + dt phi1(z) N^ for the
    coupled Gray-Scott state, z = dt D lam per field."""
    u, v, coef, rows, _, _ = _exp_setup(state, dt, linear_ops, nonlinear_rhs)
    gs_etd_exp_inplace(u, v, coef, rows, 1, 1)
    return {"u": u, "v": v}


def etd2_step(state: StateDict, t: float, dt: float, linear_ops: Dict[str, DiffusionOperator],
              nonlinear_rhs: Callable[[StateDict, float], StateDict],
              N_prev: Optional[StateDict] = None) -> Tuple[StateDict, StateDict]:
    """fft_integrators.py:177-249: exponential Adams-Bashforth-2; returns
    (new state, N of the input state) — ETD1 when N_prev is None."""
    from .operators import gray_scott_reaction
    u, v, coef, rows, F, k = _exp_setup(state, dt, linear_ops, nonlinear_rhs)
    nu, nv = gray_scott_reaction(u, v, F, k)
    if N_prev is None:
        gs_etd_exp_inplace(u, v, coef, rows, 2, 1)
    else:
        npu = as_device(N_prev["u"], u.dtype).contiguous()
        npv = as_device(N_prev["v"], u.dtype).contiguous()
        gs_etd_exp_inplace(u, v, coef, rows, 2, 1, prev=1, n_prev=(npu, npv))
    return {"u": u, "v": v}, {"u": nu, "v": nv}


def etd_integrate(u0: StateDict, t_span: Tuple[float, float], dt: float, linear_ops: Dict[str, DiffusionOperator],
                  nonlinear_rhs: Callable[[StateDict, float], StateDict], method: str = "etd1",
                  save_every: int = 1):
    """fft_integrators.py:344-393: int((t_end - t_start)/dt) steps of 'etd1',
    'etd2' (N_prev carried on the device across saves) or 'etdrk4'; returns
    (t_history, state_history) with the initial state first."""
    if method not in ("etd1", "etd2", "etdrk4"):
        raise ValueError(f"Unknown method: {method}. Use 'etd1', 'etd2', or 'etdrk4'")
    F, k = _gray_scott_params(nonlinear_rhs)
    lamx, lamy, Du, Dv = _linear_symbols(linear_ops)
    t0, t1 = t_span
    n_steps = int((t1 - t0) / dt)
    u = as_device(u0["u"])
    dtype = field_dtype(u)
    u = u.to(dtype).clone().contiguous()
    v = as_device(u0["v"], dtype).clone().contiguous()
    if method == "etdrk4":
        coef = etdrk4_coefficients(lamx, lamy, Du, Dv, dt, dtype)
    else:
        coef = etd_phi_coefficients(lamx, lamy, Du, Dv, dt, dtype)
    rows = inst_table([[0, 0, 0, 0, F, k, dt]] * view2d(u)[3], u.device)
    t_hist = [t0]
    hist = [{"u": u.clone(), "v": v.clone()}]
    done, work = 0, None
    while done < n_steps:
        chunk = min(save_every, n_steps - done)
        if method == "etdrk4":
            gs_etdrk4_inplace(u, v, coef, rows, chunk)
        else:
            order = 1 if method == "etd1" else 2
            work = gs_etd_exp_inplace(u, v, coef, rows, order, chunk, prev=2 if (order == 2 and done) else 0,
                                      work=work)
        done += chunk
        if chunk == save_every:
            t_hist.append(t0 + done * dt)
            hist.append({"u": u.clone(), "v": v.clone()})
    return torch.tensor(t_hist, dtype=torch.float64), hist


def batched_fft_matvec(state: StateDict, linear_ops: Dict[str, DiffusionOperator]) -> StateDict:
    """fft_integrators.py:400-420: L u per field (zeros without an operator)."""
    out = {}
    for name, f in state.items():
        op = linear_ops.get(name)
        out[name] = op.matvec(f) if op is not None else torch.zeros_like(as_device(f))
    return out


def batched_fft_solve(rhs: StateDict, linear_ops: Dict[str, DiffusionOperator], dt: float) -> StateDict:
    """fft_integrators.py:423-445: (I - dt L)^-1 per field (identity without
    an operator)."""
    return {n: (linear_ops[n].solve(f, dt) if n in linear_ops else f) for n, f in rhs.items()}


def batched_fft_exp_matvec(state: StateDict, linear_ops: Dict[str, DiffusionOperator], dt: float) -> StateDict:
    """fft_integrators.py:448-470: exp(dt L) per field."""
    return {n: (linear_ops[n].exp_matvec(f, dt) if n in linear_ops else f) for n, f in state.items()}


def stacked_fft_solve_shared_op(rhs: StateDict, op: DiffusionOperator, dt: float) -> StateDict:
    """fft_integrators.py:473-510: every field solved with one operator in one
    batched launch (the fields stacked along a leading axis)."""
    names = list(rhs.keys())
    if not names:
        return {}
    first = as_device(rhs[names[0]])
    stacked = torch.stack([as_device(rhs[n], field_dtype(first)) for n in names])
    out = op.solve(stacked, dt)
    return {n: out[i] for i, n in enumerate(names)}


def diffusion_only_etd1(u0, t_end: float, dt: float, op: DiffusionOperator) -> torch.Tensor:
    """fft_integrators.py:513-542: max(1, int(t_end/dt)) steps of exp(dt' L)
    with dt' = t_end / n_steps."""
    n = max(1, int(t_end / dt))
    h = t_end / n
    u = as_device(u0)
    for _ in range(n):
        u = op.exp_matvec(u, h)
    return u


class ETDStepResult(NamedTuple):
    """fft_integrators.py:32-36."""
    state: StateDict
    t: float
    diagnostics: dict


def imex_euler_step(state: StateDict, t: float, dt: float, linear_ops: Dict[str, DiffusionOperator],
                    nonlinear_rhs: Callable[[StateDict, float], StateDict]) -> StateDict:
    """fft_integrators.py:545-585: per field, (I - dt L) u_new = u + dt N(u)
    with L's FFT solve, or u + dt N(u) for fields without an operator (the
    a x + b y kernel and the operators' spectral solve)."""
    from .state import axpby
    N = nonlinear_rhs(state, t)
    out = {}
    for name, u in state.items():
        rhs = axpby(1.0, as_device(u), float(dt), as_device(N[name]))
        op = linear_ops.get(name)
        out[name] = op.solve(rhs, dt) if op is not None else rhs
    return out
