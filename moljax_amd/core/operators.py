"""Finite-difference operators and reaction terms (moljax/core/operators.py).

2-D operators take a ghost-padded (ny+2, nx+2) array (or a batch
(..., ny+2, nx+2)) whose ghosts are already filled and return a new padded
array with the stencil on the interior and zeros in the ghost ring, computed
by the mjx_fd_op2d kernel.
"""
from dataclasses import dataclass
from typing import Any, Callable, Dict, List, Optional, Union

import numpy as np
import torch

from .. import _hip
from ._arrays import as_device, field_dtype, view2d
from .grid import Grid2D, GridType
from .state import StateDict, tree_add, tree_zeros_like

FD_LAP, FD_D2, FD_D1C, FD_D1UP = 0, 1, 2, 3


@dataclass(frozen=True)
class LinearOp:
    """operators.py:329-346."""
    name: str
    apply: Callable[[StateDict, GridType, float, Dict[str, Any]], StateDict]
    dt_bound: Optional[Callable[[GridType, Dict[str, Any]], float]] = None


@dataclass(frozen=True)
class NonlinearOp:
    """operators.py:349-361."""
    name: str
    apply: Callable[[StateDict, GridType, float, Dict[str, Any]], StateDict]
    dt_bound: Optional[Callable[[GridType, StateDict, Dict[str, Any]], float]] = None


def op_sum(ops: List[Union[LinearOp, NonlinearOp]]) -> Callable:
    """operators.py:364-380."""
    def combined_apply(state, grid, t, params):
        result = ops[0].apply(state, grid, t, params)
        for op in ops[1:]:
            result = tree_add(result, op.apply(state, grid, t, params))
        return result
    return combined_apply


def state_zero_like(state: StateDict) -> StateDict:
    return tree_zeros_like(state)


def _check_padded(f: torch.Tensor, grid: Grid2D):
    if f.shape[-2:] != (grid.ny_total, grid.nx_total):
        raise ValueError(f"expected padded shape (..., {grid.ny_total}, {grid.nx_total}), got {tuple(f.shape)}")
    if grid.n_ghost != 1:
        raise ValueError("the FD kernels serve n_ghost = 1 grids")


def _fd(f, grid: Grid2D, op: int, axis: int = 1, vel=None) -> torch.Tensor:
    t = as_device(f)
    t = t.to(field_dtype(t))
    _check_padded(t, grid)
    src, pitch, istr, B = view2d(t)
    out = torch.empty((B, grid.ny_total, grid.nx_total), dtype=t.dtype, device=t.device)
    vel_t, vp, vs, vsc = None, 0, 0, 0.0
    if vel is not None:
        if isinstance(vel, (int, float)) or (isinstance(vel, torch.Tensor) and vel.dim() == 0) \
                or (isinstance(vel, np.ndarray) and vel.ndim == 0):
            vsc = float(vel)
        else:
            vt = as_device(vel, t.dtype)
            if vt.shape[-2:] != t.shape[-2:]:
                raise ValueError("velocity array must have the padded field shape")
            if vt.shape != t.shape:
                try:  # broadcast to the field's leading dims (torch raises if it cannot)
                    vt = vt.expand(t.shape).contiguous()
                except RuntimeError as e:
                    raise ValueError(f"velocity shape {tuple(vel.shape)} does not broadcast to the field "
                                     f"shape {tuple(t.shape)}") from e
            vel_t, vp, vs, vB = view2d(vt)
            if vB != B:
                raise ValueError(f"velocity holds {vB} instances, the field {B}")
            vel_t = vel_t[:, 1:, 1:]
    lib = _hip.lib()
    rc = lib.mjx_fd_op2d(_hip.dtype_code(t.dtype), op, B, grid.ny, grid.nx,
                         _hip.ptr(src[:, 1:, 1:]), pitch, istr,
                         _hip.ptr(out[:, 1:, 1:]), out.stride(1), out.stride(0), int(axis),
                         float(grid.dx), float(grid.dy), _hip.ptr(vel_t), vp, vs, vsc, _hip.stream_ptr())
    _hip.check(rc, "mjx_fd_op2d")
    return out.view(t.shape)


def _fd1(f, grid, op: int, vel=None) -> torch.Tensor:
    """1-D padded fields (..., nx + 2) through the 2-D kernel as single rows:
    x-direction stencils read only their own row, and the kernel's zeroed
    ghost rows land in a scratch row above and below."""
    if grid.n_ghost != 1:
        raise ValueError("the FD kernels serve n_ghost = 1 grids")
    t = as_device(f)
    t = t.to(field_dtype(t))
    W = grid.nx + 2
    if t.shape[-1] != W:
        raise ValueError(f"expected padded shape (..., {W}), got {tuple(t.shape)}")
    lead = t.shape[:-1]
    src = t.reshape(-1, W).contiguous()
    B = src.shape[0]
    out = torch.empty((B, 3, W), dtype=t.dtype, device=t.device)
    vel_t, vp, vs, vsc = None, 0, 0, 0.0
    if vel is not None:
        if isinstance(vel, (int, float)) or (isinstance(vel, torch.Tensor) and vel.dim() == 0) \
                or (isinstance(vel, np.ndarray) and vel.ndim == 0):
            vsc = float(vel)
        else:
            vt = as_device(vel, t.dtype)
            if vt.shape[-1] != W:
                raise ValueError("velocity array must have the padded field shape")
            vel_t = vt.expand(t.shape).reshape(-1, W).contiguous()
            vp, vs = W, W
    rc = _hip.lib().mjx_fd_op2d(_hip.dtype_code(t.dtype), op, B, 1, grid.nx, _hip.ptr(src[:, 1:]), W, W,
                                _hip.ptr(out[:, 1, 1:]), W, 3 * W, 1, float(grid.dx), 1.0,
                                _hip.ptr(vel_t[:, 1:]) if vel_t is not None else None, vp, vs, vsc,
                                _hip.stream_ptr())
    _hip.check(rc, "mjx_fd_op2d")
    return out[:, 1, :].reshape(lead + (W,))


def d1_central_1d(f, grid) -> torch.Tensor:
    """operators.py:34-59: (f[i+1] - f[i-1]) / (2 dx) on the interior, zero
    ghosts."""
    return _fd1(f, grid, FD_D1C)


def d1_upwind_1d(f, v, grid) -> torch.Tensor:
    """operators.py:62-111: backward difference where v >= 0, forward where
    v < 0 (v scalar or a padded array)."""
    return _fd1(f, grid, FD_D1UP, v)


def d2_central_1d(f, grid) -> torch.Tensor:
    """operators.py:114-138."""
    return _fd1(f, grid, FD_D2)


def laplacian_1d(f, grid) -> torch.Tensor:
    """operators.py:141-152 (the same as d2_central_1d)."""
    return _fd1(f, grid, FD_D2)


def laplacian_2d(f, grid: Grid2D) -> torch.Tensor:
    """operators.py:311-322: d2/dy2 + d2/dx2 on the interior, zero ghosts."""
    return _fd(f, grid, FD_LAP)


def d2_central_2d(f, grid: Grid2D, axis: int) -> torch.Tensor:
    """operators.py:271-308."""
    return _fd(f, grid, FD_D2, axis)


def d1_central_2d(f, grid: Grid2D, axis: int) -> torch.Tensor:
    """operators.py:159-197."""
    return _fd(f, grid, FD_D1C, axis)


def d1_upwind_2d(f, v, grid: Grid2D, axis: int) -> torch.Tensor:
    """operators.py:200-268: backward difference where v >= 0, forward
    otherwise; v is a scalar or a padded array."""
    return _fd(f, grid, FD_D1UP, axis, vel=v)


def _diffusion_op(field_name: str, D_param: str, lap, cfl: float) -> LinearOp:
    from .state import axpby

    def apply_diffusion(state: StateDict, grid, t: float, params: Dict[str, Any]) -> StateDict:
        D = float(params.get(D_param, 1.0))
        result = tree_zeros_like(state)
        result[field_name] = axpby(D, lap(state[field_name], grid))
        return result

    def dt_bound(grid, params: Dict[str, Any]) -> float:
        return cfl * grid.min_dx2 / (float(params.get(D_param, 1.0)) + 1e-14)

    return LinearOp(name=f"diffusion_{field_name}", apply=apply_diffusion, dt_bound=dt_bound)


def make_diffusion_op_1d(field_name: str, D_param: str = "D") -> LinearOp:
    """operators.py:392-417: D lap on one field (other fields zero); dt
    bound 0.5 dx^2 / D."""
    return _diffusion_op(field_name, D_param, laplacian_1d, 0.5)


def make_diffusion_op_2d(field_name: str, D_param: str = "D") -> LinearOp:
    """operators.py:420-444: D lap_h on one field (other fields zero); dt
    bound 0.25 min(dx, dy)^2 / D."""
    return _diffusion_op(field_name, D_param, laplacian_2d, 0.25)


def make_reaction_op(reactions: Dict[str, Callable[[StateDict, Dict[str, Any]], torch.Tensor]]) -> NonlinearOp:
    """operators.py:623-641: the given per-field reaction callables (other
    fields zero)."""
    def apply_reaction(state: StateDict, grid, t: float, params: Dict[str, Any]) -> StateDict:
        result = tree_zeros_like(state)
        for name, fn in reactions.items():
            result[name] = fn(state, params)
        return result

    return NonlinearOp(name="reactions", apply=apply_reaction, dt_bound=None)


def gray_scott_reaction(u, v, F: float, k: float):
    """(-u v^2 + F(1-u), u v^2 - (F+k) v) point-wise (mjx_gs_reaction)."""
    tu = as_device(u)
    tu = tu.to(field_dtype(tu)).contiguous()
    tv = as_device(v, tu.dtype).contiguous()
    if tu.shape != tv.shape:
        raise ValueError("u and v must have the same shape")
    ru = torch.empty_like(tu)
    rv = torch.empty_like(tv)
    rc = _hip.lib().mjx_gs_reaction(_hip.dtype_code(tu.dtype), tu.numel(), _hip.ptr(tu), _hip.ptr(tv),
                                    _hip.ptr(ru), _hip.ptr(rv), float(F), float(k), _hip.stream_ptr())
    _hip.check(rc, "mjx_gs_reaction")
    return ru, rv


def gray_scott_reaction_op() -> NonlinearOp:
    """operators.py:644-677."""
    def apply_reaction(state, grid, t, params):
        ru, rv = gray_scott_reaction(state["u"], state["v"], params.get("F", 0.04), params.get("k", 0.06))
        return {"u": ru, "v": rv}

    def dt_bound(grid, state, params):
        return 0.5 / (params.get("F", 0.04) + params.get("k", 0.06) + 1e-14)

    return NonlinearOp(name="gray_scott_reaction", apply=apply_reaction, dt_bound=dt_bound)


def fd_laplacian_symbol_2d(ny: int, nx: int, dy: float, dx: float, dtype=torch.float64) -> torch.Tensor:
    """operators.py:734-766 (same as fft_solvers.laplacian_symbol_2d)."""
    from .fft_solvers import laplacian_symbol_2d
    return laplacian_symbol_2d(ny, nx, dy, dx, dtype)


def fd_laplacian_symbol_1d(nx: int, dx: float, dtype=torch.float64) -> torch.Tensor:
    """operators.py:708-731."""
    from .fft_solvers import laplacian_symbol_1d
    return laplacian_symbol_1d(nx, dx, dtype)
