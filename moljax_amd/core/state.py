"""StateDict algebra (moljax/core/state.py).

A state is a dict of field name -> padded array.  The element-wise tree ops
run the native a x + b y + c kernel and allocate new arrays (the reference's
functional semantics).  Reductions used for accept/reject decisions
accumulate in float64.
"""
from typing import Optional, Any, Callable, Dict, Tuple

import torch

StateDict = Dict[str, torch.Tensor]


def axpby(a: float, x, b: float = 0.0, y=None, c: float = 0.0) -> torch.Tensor:
    """A new array a x + b y + c (mjx_axpby), y optional, same shape/dtype."""
    from .. import _hip
    from ._arrays import as_device, field_dtype
    tx = as_device(x)
    tx = tx.to(field_dtype(tx)).contiguous()
    ty = None
    if y is not None:
        ty = as_device(y, tx.dtype).contiguous()
        if ty.shape != tx.shape:
            raise ValueError(f"shapes differ: {tuple(tx.shape)} vs {tuple(ty.shape)}")
    out = torch.empty_like(tx)
    rc = _hip.lib().mjx_axpby(_hip.dtype_code(tx.dtype), tx.numel(), float(a), _hip.ptr(tx), float(b),
                              _hip.ptr(ty) if ty is not None else None, float(c), _hip.ptr(out), _hip.stream_ptr())
    _hip.check(rc, "mjx_axpby")
    return out


def tree_add(a: StateDict, b: StateDict) -> StateDict:
    """state.py:27-38."""
    return {k: axpby(1.0, a[k], 1.0, b[k]) for k in a}


def tree_sub(a: StateDict, b: StateDict) -> StateDict:
    """state.py:41-52."""
    return {k: axpby(1.0, a[k], -1.0, b[k]) for k in a}


def tree_scale(a: StateDict, s: float) -> StateDict:
    """state.py:55-66."""
    return {k: axpby(float(s), v) for k, v in a.items()}


def tree_axpy(y: StateDict, a: float, x: StateDict) -> StateDict:
    """state.py:69-81: y + a*x."""
    return {k: axpby(1.0, y[k], float(a), x[k]) for k in y}


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


RED_SUMSQ, RED_DOT, RED_MAXABS, RED_SCALED_SQ, RED_NONFINITE, RED_SUM = 0, 1, 2, 3, 4, 5


def _rview(f: torch.Tensor):
    """(B, ny, nx) strided view with pitch and instance stride for the
    reduction kernel (1-D fields become one row)."""
    from ._arrays import view2d
    if f.dim() == 1:
        f = f.unsqueeze(0)
    return view2d(f)


_RED_WORK = {}


def _red_work(lib, B: int, device, st: int) -> torch.Tensor:
    """The reduction partials buffer for batch B on a device and stream st,
    kept between calls (work on one stream is ordered; every call rewrites
    it)."""
    key = (B, device.index, st)
    w = _RED_WORK.get(key)
    if w is None:
        w = torch.empty(lib.mjx_reduce_workspace_elems(B), dtype=torch.float64, device=device)
        _RED_WORK[key] = w
    return w


def _reduce(op: int, fields_a, fields_b=None, atol: float = 0.0, rtol: float = 0.0, st: Optional[int] = None
            ) -> torch.Tensor:
    """Per-instance FP64 reduction over several fields (mjx_reduce2d, fixed
    order, bitwise reproducible) on stream st (the current stream when None);
    returns a (B,) float64 device tensor."""
    from .. import _hip
    lib = _hip.lib()
    if st is None:
        st = _hip.stream_ptr()
    out = work = None
    for i, fa in enumerate(fields_a):
        ta, ap, ai, B = _rview(fa)
        tb, bp, bi = None, 0, 0
        if fields_b is not None:
            fb = fields_b[i].to(ta.dtype)
            tb, bp, bi, Bb = _rview(fb)
            if Bb != B or tb.shape != ta.shape:
                raise ValueError("fields of a and b must have the same shape")
        if out is None:
            out = torch.empty(B, dtype=torch.float64, device=ta.device)  # written by the first field
            work = _red_work(lib, B, ta.device, st)
        elif out.numel() != B:
            raise ValueError("all fields must have the same batch shape")
        rc = lib.mjx_reduce2d(_hip.dtype_code(ta.dtype), op, B, ta.shape[-2], ta.shape[-1], _hip.ptr(ta), ap, ai,
                              _hip.ptr(tb), bp, bi, float(atol), float(rtol), _hip.ptr(out), int(i > 0),
                              _hip.ptr(work), st)
        _hip.check(rc, "mjx_reduce2d")
    return out


def _device_fields(fields):
    from ._arrays import as_device, field_dtype
    out = []
    for f in fields:
        t = as_device(f)
        out.append(t.to(field_dtype(t)))
    return out


def _total(op: int, fields_a, fields_b=None, take_sqrt: bool = False) -> torch.Tensor:
    """Batch total (max for RED_MAXABS) of a reduction over several fields,
    in a fixed order on the device (mjx_reduce2d per field, then
    mjx_reduce_finish), optionally its square root: a 0-d float64 tensor."""
    from .. import _hip
    st = _hip.stream_ptr()
    per_inst = _reduce(op, fields_a, fields_b, st=st)
    total = torch.empty((), dtype=torch.float64, device=per_inst.device)
    rc = _hip.lib().mjx_reduce_finish(int(op), per_inst.numel(), per_inst.data_ptr(), int(bool(take_sqrt)),
                                      total.data_ptr(), st)
    _hip.check(rc, "mjx_reduce_finish")
    return total


def tree_norm2(state: StateDict, grid, interior_only: bool = True) -> torch.Tensor:
    """state.py:115-137: sqrt of the sum of squares over all fields, float64
    accumulation."""
    return _total(RED_SUMSQ, _device_fields([_interior(v, grid, interior_only) for v in state.values()]),
                  take_sqrt=True)


def tree_norm_inf(state: StateDict, grid, interior_only: bool = True) -> torch.Tensor:
    """state.py:140-160: max |x| over all fields (NaN propagates)."""
    return _total(RED_MAXABS, _device_fields([_interior(v, grid, interior_only) for v in state.values()]))


def tree_vdot(a: StateDict, b: StateDict, grid, interior_only: bool = True) -> torch.Tensor:
    """state.py:163-189: sum over fields of sum(a*b), float64 accumulation."""
    keys = list(a.keys())
    fa = _device_fields([_interior(a[k], grid, interior_only) for k in keys])
    fb = _device_fields([_interior(b[k], grid, interior_only) for k in keys])
    return _total(RED_DOT, fa, fb)


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
    keys = list(error.keys())
    fe = _device_fields([_interior(error[k], grid, True) for k in keys])
    fy = _device_fields([_interior(y[k], grid, True) for k in keys])
    return torch.sqrt(torch.sum(_reduce(RED_SCALED_SQ, fe, fy, atol, rtol)) / state_interior_size(error, grid))


def is_finite(state: StateDict) -> bool:
    """utils.py:186-197: every value of every field finite (ghosts included)."""
    f = _device_fields(list(state.values()))
    return bool(torch.sum(_reduce(RED_NONFINITE, f)).item() == 0)


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


def unflatten_from_vec(flat_vec: torch.Tensor, template_state: StateDict) -> StateDict:
    """state.py:220-238: the state of template_state's structure and shapes
    holding flat_vec (views into it)."""
    return flatten_to_vec(template_state)[1](flat_vec)


def state_total_size(state: StateDict) -> int:
    """state.py:241-251."""
    return int(sum(f.numel() for f in state.values()))


def create_state_from_interior(interior_arrays: Dict[str, Any], grid, dtype=torch.float64) -> StateDict:
    """state.py:268-298: padded arrays with zero ghosts around the given
    interiors (fill the ghosts with apply_bc)."""
    from ._arrays import as_device
    from .grid import Grid1D, default_device
    out = {}
    for name, interior in interior_arrays.items():
        if isinstance(grid, Grid1D):
            pad = torch.zeros(grid.nx_total, dtype=dtype, device=default_device())
            pad[grid.interior_slice] = as_device(interior, dtype)
        else:
            pad = torch.zeros((grid.ny_total, grid.nx_total), dtype=dtype, device=default_device())
            sy, sx = grid.interior_slice
            pad[sy, sx] = as_device(interior, dtype)
        out[name] = pad
    return out
