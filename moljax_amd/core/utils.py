"""Status codes and small helpers (moljax/core/utils.py)."""
from typing import Dict

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
