"""StateDict algebra (moljax/core/state.py).

A state is a dict of field name -> padded array.  The element-wise tree ops
are thin and allocate new arrays (functional semantics of the reference).
Reductions used for accept/reject decisions accumulate in float64.
"""
from typing import Callable, Dict, Tuple

import torch

StateDict = Dict[str, torch.Tensor]


def tree_add(a: StateDict, b: StateDict) -> StateDict:
    """state.py:27-38."""
    return {k: a[k] + b[k] for k in a}


def tree_sub(a: StateDict, b: StateDict) -> StateDict:
    """state.py:41-52."""
    return {k: a[k] - b[k] for k in a}


def tree_scale(a: StateDict, s: float) -> StateDict:
    """state.py:55-66."""
    return {k: v * s for k, v in a.items()}


def tree_axpy(y: StateDict, a: float, x: StateDict) -> StateDict:
    """state.py:69-81: y + a*x."""
    return {k: y[k] + a * x[k] for k in y}


def tree_zeros_like(state: StateDict) -> StateDict:
    return {k: torch.zeros_like(v) for k, v in state.items()}


def tree_ones_like(state: StateDict) -> StateDict:
    return {k: torch.ones_like(v) for k, v in state.items()}


def _interior(f: torch.Tensor, grid, interior_only: bool) -> torch.Tensor:
    if not interior_only:
        return f
    if hasattr(grid, "ny"):
        sy, sx = grid.interior_slice
        return f[..., sy, sx]
    return f[..., grid.interior_slice]


def tree_norm2(state: StateDict, grid, interior_only: bool = True) -> torch.Tensor:
    """state.py:115-137, float64 accumulation."""
    tot = sum(torch.sum(_interior(f, grid, interior_only).double() ** 2) for f in state.values())
    return torch.sqrt(tot)


def tree_norm_inf(state: StateDict, grid, interior_only: bool = True) -> torch.Tensor:
    """state.py:140-160."""
    return torch.stack([torch.max(torch.abs(_interior(f, grid, interior_only))).double()
                        for f in state.values()]).max()


def tree_vdot(a: StateDict, b: StateDict, grid, interior_only: bool = True) -> torch.Tensor:
    """state.py:163-189, float64 accumulation."""
    return sum(torch.sum(_interior(a[k], grid, interior_only).double() * _interior(b[k], grid, interior_only).double())
               for k in a)


def state_interior_size(state: StateDict, grid) -> int:
    n = 0
    for f in state.values():
        n += _interior(f, grid, True).numel()
    return n


def get_interior_state(state: StateDict, grid) -> StateDict:
    return {k: _interior(v, grid, True) for k, v in state.items()}


def scaled_error_norm(error: StateDict, y: StateDict, grid, atol: float, rtol: float) -> torch.Tensor:
    """state.py:312-343: sqrt(mean((err / (atol + rtol |y|))^2)) over interior
    points of all fields, accumulated in float64."""
    tot = None
    for k in error:
        e = _interior(error[k], grid, True).double()
        s = atol + rtol * torch.abs(_interior(y[k], grid, True).double())
        term = torch.sum((e / s) ** 2)
        tot = term if tot is None else tot + term
    return torch.sqrt(tot / state_interior_size(error, grid))


def flatten_to_vec(state: StateDict) -> Tuple[torch.Tensor, Callable[[torch.Tensor], StateDict]]:
    """state.py:192-217: concatenation in sorted-key order (ravel_pytree)."""
    keys = sorted(state)
    shapes = [state[k].shape for k in keys]
    sizes = [state[k].numel() for k in keys]
    vec = torch.cat([state[k].reshape(-1) for k in keys])

    def unravel(v: torch.Tensor) -> StateDict:
        out, o = {}, 0
        for k, shp, n in zip(keys, shapes, sizes):
            out[k] = v[o:o + n].view(shp)
            o += n
        return out

    return vec, unravel
