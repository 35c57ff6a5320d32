"""Status codes, interior access, history buffers and interior
reductions (moljax/core/utils.py).

History buffers are device tensors; save_to_history copies the state's
(interior) fields into slot ``idx`` in place and returns the same dict (the
reference returns updated functional copies, utils.py:126-153).  The
interior reductions run the native FP64 reduction kernel (mjx_reduce2d)."""
from typing import Dict, Tuple

import numpy as np
import torch


class StatusCode:
    """moljax/core/utils.py:265-272."""
    SUCCESS = 0
    RUNNING = 1
    MAX_STEPS_REACHED = 2
    DT_TOO_SMALL = 3
    NON_FINITE_VALUES = 4
    NK_FAILED = 5


def get_interior(field: torch.Tensor, grid) -> torch.Tensor:
    """moljax/core/utils.py get_interior: a view of the interior cells."""
    if hasattr(grid, "ny"):
        sy, sx = grid.interior_slice
        return field[..., sy, sx]
    return field[..., grid.interior_slice]


def is_finite(state) -> bool:
    """moljax/core/utils.py:186-197 (native reduction)."""
    from .state import is_finite as _is_finite
    return _is_finite(state)


def eps_for_dtype(dtype=torch.float64) -> float:
    """utils.py:20-22."""
    return float(torch.finfo(dtype).eps)


def safe_eps(dtype=torch.float64) -> float:
    """utils.py:25-27."""
    return 1e-10 if dtype == torch.float32 else 1e-14


def set_interior(padded: torch.Tensor, interior: torch.Tensor, grid) -> torch.Tensor:
    """utils.py:44-60: a new padded array with the interior replaced."""
    out = padded.clone()
    get_interior(out, grid).copy_(torch.as_tensor(interior, device=out.device, dtype=out.dtype))
    return out


def _dev():
    from .grid import default_device
    return default_device()


def allocate_scalar_history(max_steps: int, dtype=torch.float64) -> torch.Tensor:
    """utils.py:63-74: zeros of shape (max_steps,)."""
    return torch.zeros(int(max_steps), dtype=dtype, device=_dev())


def allocate_field_history(field_shape: Tuple[int, ...], max_steps: int, dtype=torch.float64) -> torch.Tensor:
    """utils.py:77-93: zeros of shape (max_steps,) + field_shape."""
    return torch.zeros((int(max_steps),) + tuple(field_shape), dtype=dtype, device=_dev())


def allocate_state_history(state: Dict[str, torch.Tensor], grid, max_steps: int, interior_only: bool = True,
                           dtype=torch.float64) -> Dict[str, torch.Tensor]:
    """utils.py:96-123: one (max_steps,) + field shape buffer per field."""
    hist = {}
    for name, f in state.items():
        shape = tuple(get_interior(f, grid).shape) if interior_only else tuple(f.shape)
        hist[name] = allocate_field_history(shape, max_steps, dtype)
    return hist


def save_to_history(history: Dict[str, torch.Tensor], state: Dict[str, torch.Tensor], grid, idx: int,
                    interior_only: bool = True) -> Dict[str, torch.Tensor]:
    """utils.py:126-153: history[name][idx] <- the field (interior)."""
    for name, f in state.items():
        data = get_interior(f, grid) if interior_only else f
        history[name][int(idx)].copy_(torch.as_tensor(data, device=history[name].device))
    return history


def create_padded_array(grid, fill_value: float = 0.0, dtype=torch.float64) -> torch.Tensor:
    """utils.py:200-219."""
    if hasattr(grid, "ny"):
        shape = (grid.ny_total, grid.nx_total)
    else:
        shape = (grid.nx_total,)
    return torch.full(shape, float(fill_value), dtype=dtype, device=_dev())


def _interior_reduce(op: int, field, grid) -> float:
    """Over every instance of a (possibly batched) field."""
    from .state import RED_MAXABS, _device_fields, _reduce
    f = _device_fields([get_interior(torch.as_tensor(field), grid)])
    r = _reduce(op, f)
    return float((r.max() if op == RED_MAXABS else r.sum()).item())


def max_abs_interior(field, grid) -> float:
    """utils.py:222-233 (native FP64 reduction)."""
    from .state import RED_MAXABS
    return _interior_reduce(RED_MAXABS, field, grid)


def sum_interior(field, grid) -> float:
    """utils.py:250-262 (native FP64 reduction, fixed order)."""
    from .state import RED_SUM
    return _interior_reduce(RED_SUM, field, grid)


def mean_interior(field, grid) -> float:
    """utils.py:236-247."""
    n = int(np.prod(get_interior(torch.as_tensor(field), grid).shape))
    return sum_interior(field, grid) / n
