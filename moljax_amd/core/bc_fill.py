"""Ghost-ring fill for padded states (moljax/core/bc.py:310-404), run by the
mjx_bc_fill2d kernel.

Boundary value/flux callables (x, y, t, params) are evaluated on the host
exactly as the reference does: a dict gives per-side values, a scalar is used
on every side, anything else means zeros (bc.py:323-349).
"""
import ctypes
from typing import Any, Dict, Optional

import torch

from .. import _hip
from ._arrays import as_device, field_dtype, view2d
from .bc import BCSpec, BCType, FieldBCSpec
from .grid import Grid2D


def _side_values(fn, grid: Grid2D, t: float, params) -> tuple:
    if fn is None:
        return (0.0, 0.0, 0.0, 0.0)
    X, Y = grid.meshgrid(include_ghost=False)
    val = fn(X, Y, t, params)
    if isinstance(val, dict):
        return tuple(float(val.get(s, 0.0)) for s in ("left", "right", "bottom", "top"))
    if isinstance(val, (int, float)) or (hasattr(val, "ndim") and val.ndim == 0):
        v = float(val)
        return (v, v, v, v)
    return (0.0, 0.0, 0.0, 0.0)


def fill_ghosts(f: torch.Tensor, grid: Grid2D, spec: FieldBCSpec, t: float = 0.0,
                params: Optional[Dict[str, Any]] = None) -> torch.Tensor:
    """In-place ghost fill of one padded field (or a batch of them)."""
    if grid.n_ghost != 1:
        raise ValueError("the BC kernel serves n_ghost = 1 grids")
    kind = int(spec.kind)
    if kind == BCType.DIRICHLET:
        vals = _side_values(spec.value, grid, t, params or {})
    elif kind == BCType.NEUMANN:
        vals = _side_values(spec.flux, grid, t, params or {})
    else:
        vals = (0.0, 0.0, 0.0, 0.0)
    arr = (ctypes.c_double * 4)(*vals)
    tb, pitch, istr, B = view2d(f)
    rc = _hip.lib().mjx_bc_fill2d(_hip.dtype_code(tb.dtype), B, grid.ny, grid.nx, _hip.ptr(tb[:, 1:, 1:]), pitch,
                                  istr, kind, arr, float(grid.dx), float(grid.dy), _hip.stream_ptr())
    _hip.check(rc, "mjx_bc_fill2d")
    return f


def apply_bc(state: Dict[str, torch.Tensor], grid: Grid2D, bc_spec: BCSpec, t: float = 0.0,
             params: Optional[Dict[str, Any]] = None) -> Dict[str, torch.Tensor]:
    """bc.py:358-404: new state with ghost cells filled per field; fields
    without a spec are returned unchanged."""
    if not isinstance(grid, Grid2D):
        raise NotImplementedError("1-D ghost fill is not on the MI355X path")
    out = {}
    for name, f in state.items():
        if name not in bc_spec:
            out[name] = f
            continue
        g = as_device(f)
        g = g.to(field_dtype(g)).clone()
        out[name] = fill_ghosts(g, grid, bc_spec[name], t, params)
    return out
