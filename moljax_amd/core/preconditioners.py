"""Preconditioners (moljax/core/preconditioners.py) on the MI355X kernels.

The native Newton-Krylov driver takes the preconditioner class as an option
(none / exact diffusion inverse / Jacobi); the ``apply`` methods serve direct
callers with padded StateDicts as in the reference.
"""
from dataclasses import dataclass, field
from typing import Any, Dict, Optional, Protocol

import numpy as np
import torch

from .grid import Grid1D, Grid2D, GridType
from .state import StateDict

PRECOND_NONE, PRECOND_EXACT, PRECOND_JACOBI = 0, 1, 2


@dataclass(frozen=True)
class PrecondContext:
    """preconditioners.py:29-43."""
    grid: GridType
    dt: float
    params: Dict[str, Any]
    extra: Dict[str, Any] = field(default_factory=dict)


class Preconditioner(Protocol):
    """preconditioners.py:46-52: an object with a name and apply(r, ctx)."""
    name: str

    def apply(self, r: StateDict, context: PrecondContext) -> StateDict:
        ...


@dataclass(frozen=True)
class IdentityPreconditioner:
    """preconditioners.py:55-66."""
    name: str = "identity"
    kind: int = PRECOND_NONE

    def apply(self, r: StateDict, context: PrecondContext) -> StateDict:
        return r


@dataclass(frozen=True)
class BlockJacobiPreconditioner:
    """preconditioners.py:69-118: each field divided by
    1 + dt D 4/min_dx^2 (diffusion keys) + dt max|v|/min_dx (advection keys).
    A per-field constant leaves left-preconditioned GMRES unchanged (same
    iterates, scale-free stopping test), so the native Newton-Krylov driver
    runs it as the identity; apply() serves direct callers."""
    name: str = "block_jacobi"
    diffusion_keys: Optional[Dict[str, str]] = None
    advection_keys: Optional[Dict[str, str]] = None
    kind: int = PRECOND_NONE

    def _scale(self, name: str, context: PrecondContext) -> float:
        grid, dt, params = context.grid, float(context.dt), context.params
        scale = 1.0
        if self.diffusion_keys is not None and name in self.diffusion_keys:
            scale += dt * float(params.get(self.diffusion_keys[name], 0.0)) * 4.0 / grid.min_dx2
        if self.advection_keys is not None and name in self.advection_keys:
            v = params.get(self.advection_keys[name], 0.0)
            vmax = float(np.max(np.abs(np.asarray(v.detach().cpu() if isinstance(v, torch.Tensor) else v))))
            scale += dt * vmax / grid.min_dx
        return scale

    def apply(self, r: StateDict, context: PrecondContext) -> StateDict:
        from .state import axpby
        return {name: axpby(1.0 / self._scale(name, context), f) for name, f in r.items()}


@dataclass(frozen=True)
class DiffusionPreconditioner:
    """preconditioners.py:121-198: n_iterations weighted-Jacobi sweeps on
    (I - dt D lap) x = r, x0 = r, diag = 1 + dt D (2/dx^2 + 2/dy^2)."""
    name: str = "diffusion"
    field_diffusivities: Optional[Dict[str, str]] = None
    n_iterations: int = 5
    omega: float = 0.6667
    kind: int = PRECOND_JACOBI

    def apply(self, r: StateDict, context: PrecondContext) -> StateDict:
        """The sweeps run in one native kernel per sweep (mjx_jacobi2d) on the
        padded arrays."""
        from .. import _hip
        from ._arrays import as_device, field_dtype, view2d
        grid, dt, params = context.grid, context.dt, context.params
        out = {}
        for name, rf in r.items():
            key = (self.field_diffusivities or {}).get(name)
            D = float(params.get(key, 0.0)) if key is not None else 0.0
            if D <= 1e-14 or not isinstance(grid, Grid2D):
                out[name] = rf
                continue
            t = as_device(rf)
            t = t.to(field_dtype(t))
            if t.shape[-2:] != (grid.ny_total, grid.nx_total):
                raise ValueError("DiffusionPreconditioner works on padded fields")
            src, sp, ss, B = view2d(t)
            x = torch.empty((B, grid.ny_total, grid.nx_total), dtype=t.dtype, device=t.device)
            work = torch.empty_like(x) if self.n_iterations > 1 else None
            rc = _hip.lib().mjx_jacobi2d(_hip.dtype_code(t.dtype), B, grid.ny, grid.nx, _hip.ptr(src[:, 1:, 1:]), sp,
                                         ss, _hip.ptr(x[:, 1:, 1:]), x.stride(1), x.stride(0),
                                         _hip.ptr(work) if work is not None else None, float(dt) * D, float(grid.dy),
                                         float(grid.dx), float(self.omega), int(self.n_iterations),
                                         _hip.stream_ptr())
            _hip.check(rc, "mjx_jacobi2d")
            out[name] = x.view(t.shape)
        return out


@dataclass(frozen=True)
class FFTDiffusionPreconditioner:
    """preconditioners.py:201-294: exact periodic (I - dt D lap)^-1 per field."""
    name: str = "fft_diffusion"
    field_diffusivity_keys: Optional[Dict[str, str]] = None
    fft_cache: Any = None
    fallback: Optional[Any] = None
    kind: int = PRECOND_EXACT

    def apply(self, r: StateDict, context: PrecondContext) -> StateDict:
        from .fft_solvers import solve_helmholtz_2d
        if self.fft_cache is None:
            return self.fallback.apply(r, context) if self.fallback is not None else r
        grid, dt, params = context.grid, context.dt, context.params
        out = {}
        for name, rf in r.items():
            key = (self.field_diffusivity_keys or {}).get(name)
            D = float(params.get(key, 0.0)) if key else 0.0
            if D <= 1e-14:
                out[name] = rf
                continue
            if rf.shape[-2:] == (grid.ny, grid.nx):
                out[name] = solve_helmholtz_2d(rf, self.fft_cache.laplacian_symbol, dt, D)
            else:
                sy, sx = grid.interior_slice
                x = rf.clone()
                x[..., sy, sx] = solve_helmholtz_2d(rf[..., sy, sx], self.fft_cache.laplacian_symbol, dt, D)
                out[name] = x
        return out


@dataclass(frozen=True)
class DSTDiffusionPreconditioner:
    """Exact (I - dt D lap_h)^-1 for homogeneous Dirichlet boundaries on a
    vertex grid by DST-I (no counterpart in the reference, whose
    DiffusionPreconditioner is Jacobi; the DST form diagonalises the operator)."""
    name: str = "dst_diffusion"
    field_diffusivity_keys: Optional[Dict[str, str]] = None
    kind: int = PRECOND_EXACT

    def apply(self, r: StateDict, context: PrecondContext) -> StateDict:
        from .fft_nonperiodic import laplacian_symbol_dirichlet, solve_helmholtz_dirichlet
        grid, dt, params = context.grid, context.dt, context.params
        out = {}
        for name, rf in r.items():
            key = (self.field_diffusivity_keys or {}).get(name)
            D = float(params.get(key, 0.0)) if key else 0.0
            if D <= 1e-14:
                out[name] = rf
                continue
            ny, nx = rf.shape[-2], rf.shape[-1]
            sym = (laplacian_symbol_dirichlet(ny, grid.dy), laplacian_symbol_dirichlet(nx, grid.dx))
            out[name] = solve_helmholtz_dirichlet(rf, sym, dt, D)
        return out


def create_gray_scott_preconditioner(n_iterations: int = 5, omega: float = 0.6667) -> DiffusionPreconditioner:
    """preconditioners.py:440-460."""
    return DiffusionPreconditioner(field_diffusivities={"u": "Du", "v": "Dv"}, n_iterations=n_iterations,
                                   omega=omega)


def create_gray_scott_fft_preconditioner(fft_cache) -> FFTDiffusionPreconditioner:
    """preconditioners.py:483-488."""
    return FFTDiffusionPreconditioner(field_diffusivity_keys={"u": "Du", "v": "Dv"}, fft_cache=fft_cache)


def create_diffusion_preconditioner(field_diffusivities: Dict[str, str], n_iterations: int = 5
                                    ) -> DiffusionPreconditioner:
    """preconditioners.py:440-457."""
    return DiffusionPreconditioner(field_diffusivities=field_diffusivities, n_iterations=n_iterations)


def create_fft_preconditioner(field_diffusivity_keys: Dict[str, str], fft_cache,
                              fallback: Optional[Any] = None) -> FFTDiffusionPreconditioner:
    """preconditioners.py:460-480."""
    return FFTDiffusionPreconditioner(field_diffusivity_keys=field_diffusivity_keys, fft_cache=fft_cache,
                                      fallback=fallback)
