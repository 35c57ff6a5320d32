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
from enum import IntEnum
from typing import Dict, Optional, Tuple

import torch

from .. import _hip
from ._arrays import as_device, field_dtype, inst_table, spectral_work, view2d
from .bc import BCType
from .state import StateDict


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


def _gs_rows(model, dt: float, diffusivities: Dict[str, float], B: int):
    Du = float(diffusivities.get("u", 0.0))
    Dv = float(diffusivities.get("v", 0.0))
    bu = 0.5 * dt * Du if Du > 1e-14 else 0.0
    bv = 0.5 * dt * Dv if Dv > 1e-14 else 0.0
    F = float(model.params.get("F", 0.04))
    k = float(model.params.get("k", 0.06))
    return [[bu, bv, 0.0, 0.0, F, k, float(dt)]] * B


def gs_imex_strang_inplace(u: torch.Tensor, v: torch.Tensor, lam, rows, n_steps: int,
                           fill_ghosts: bool = True, work: Optional[torch.Tensor] = None):
    """Advance padded (..., ny+2, nx+2) fields u, v in place by n_steps
    IMEX-Strang steps.  ``rows`` is a per-instance list of
    [beta_u, beta_v, 0, 0, F, k, dt] (beta = dt/2 * D) or an already built
    (B, 8) float64 device table; ``lam`` a symbol from laplacian_symbol_2d."""
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
    rc = _hip.lib().mjx_gs_imex_strang(
        _hip.dtype_code(tu.dtype), B, ny, nx, _hip.ptr(tu[:, 1:, 1:]), _hip.ptr(tv[:, 1:, 1:]), pitch, istr,
        int(bool(fill_ghosts)), _hip.ptr(tabs.lamx), _hip.ptr(tabs.lamy), _hip.ptr(tabs.lam2d), _hip.ptr(inst),
        int(n_steps), _hip.ptr(work), _hip.stream_ptr())
    _hip.check(rc, "mjx_gs_imex_strang")


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


def integrate_imex_fixed_dt(model, y0: StateDict, t0: float, t_end: float, dt: float, fft_cache,
                            diffusivities: Dict[str, float], use_strang: bool = True, save_every: int = 1,
                            n_steps: Optional[int] = None) -> Tuple[torch.Tensor, Dict[str, torch.Tensor], StateDict]:
    """stepping.py:1374-1436.

    Takes int((t_end - t0)/dt) + 1 steps like the reference (its extra step is
    kept) unless ``n_steps`` is given.  Returns (t_history, y_history, y_final)
    where the histories hold the padded state after every ``save_every``-th
    step (the reference stacks every step; save_every=1 gives its shapes)."""
    if not use_strang:
        raise NotImplementedError("IMEX-Euler is not on the MI355X path; use use_strang=True")
    _require_gray_scott(model, diffusivities)
    if save_every < 1:
        raise ValueError("save_every must be >= 1")
    n = int((t_end - t0) / dt) + 1 if n_steps is None else int(n_steps)
    u, v = _copy_state(y0)
    tu, _, _, B = view2d(u)
    rows = inst_table(_gs_rows(model, dt, diffusivities, B), u.device)
    work = spectral_work(u.dtype, B, tu.shape[-2] - 2, tu.shape[-1] - 2, u.device)
    n_saves = n // save_every
    t_hist = torch.empty(n_saves, dtype=torch.float64)
    hist_u = torch.empty((n_saves,) + tuple(u.shape), dtype=u.dtype, device=u.device)
    hist_v = torch.empty_like(hist_u)
    done, k = 0, 0
    while done < n:
        chunk = min(save_every, n - done)
        gs_imex_strang_inplace(u, v, fft_cache.laplacian_symbol, rows, chunk, work=work)
        done += chunk
        if chunk == save_every:
            hist_u[k].copy_(u)
            hist_v[k].copy_(v)
            t_hist[k] = t0 + done * dt
            k += 1
    return t_hist, {"u": hist_u, "v": hist_v}, {"u": u, "v": v}
