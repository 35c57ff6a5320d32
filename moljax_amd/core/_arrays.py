"""Array plumbing shared by the host modules: device placement, strided 2-D
views for the C ABI and per-instance parameter tables.  No arithmetic on
field data happens here."""
import math
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from .. import _hip


def as_device(x, dtype: Optional[torch.dtype] = None) -> torch.Tensor:
    """Any array (torch tensor, NumPy array, DLPack producer, scalar list) as a
    torch tensor on the GPU.  Raises when no GPU is present."""
    _hip.load_library(True)
    if isinstance(x, torch.Tensor):
        t = x
    elif isinstance(x, np.ndarray):
        t = torch.from_numpy(np.ascontiguousarray(x))
    elif hasattr(x, "__dlpack__"):
        t = torch.from_dlpack(x)
    else:
        t = torch.as_tensor(x)
    if t.device.type != "cuda":
        t = t.to("cuda")
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    return t


def field_dtype(t: torch.Tensor) -> torch.dtype:
    if t.dtype in (torch.float32, torch.float64):
        return t.dtype
    return torch.float64


def view2d(t: torch.Tensor) -> Tuple[torch.Tensor, int, int, int]:
    """Return a (B, ny, nx) view, its pitch, instance stride and B for a
    (..., ny, nx) tensor whose last axis is unit-stride; batch axes are
    flattened (a copy is made only when they cannot be expressed with one
    stride)."""
    if t.dim() < 2:
        raise ValueError("expected an array with at least 2 dimensions")
    if t.stride(-1) != 1:
        t = t.contiguous()
    ny, nx = t.shape[-2], t.shape[-1]
    if t.dim() == 2:
        return t.unsqueeze(0), t.stride(0), ny * t.stride(0), 1
    lead = t.shape[:-2]
    B = math.prod(lead)
    try:
        tb = t.view(B, ny, nx)
    except RuntimeError:
        tb = t.reshape(B, ny, nx)
        if tb.data_ptr() != t.data_ptr():
            raise ValueError("batched input must be a strided view; pass a contiguous array")
    return tb, tb.stride(1), tb.stride(0), B


def inst_table(rows: Sequence[Sequence[float]], device=None) -> torch.Tensor:
    """(B, 8) float64 per-instance table:
    [beta_a, beta_b, alpha_a, alpha_b, F, k, dt, 0]."""
    arr = np.zeros((len(rows), 8), dtype=np.float64)
    for i, r in enumerate(rows):
        arr[i, :len(r)] = r
    return torch.from_numpy(arr).to(device or "cuda")


def spectral_work(dtype: torch.dtype, B: int, ny: int, nx: int, device=None) -> torch.Tensor:
    """Complex work array of the periodic spectral entry points, sized by
    mjx_spectral_workspace_bytes (nx columns per row for power-of-two grids,
    2 (nx/2 + 1) for the any-length path)."""
    from .. import _hip
    nbytes = int(_hip.lib().mjx_spectral_workspace_bytes(_hip.dtype_code(dtype), B, ny, nx))
    esize = torch.empty((), dtype=dtype).element_size()
    return torch.empty((nbytes // (2 * esize), 2), dtype=dtype, device=device or "cuda")
