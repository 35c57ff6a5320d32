"""FFT-diagonalised linear operators (moljax/core/fft_operators.py) on the
MI355X spectral kernels.

DiffusionOperator(grid, D): L = D * lap_h on a periodic 2-D grid with the FD
symbol; matvec (D lap u), solve ((I - dt D lap)^-1) and exp_matvec
(exp(dt D lap) u) each run as one spectral round trip (mjx_spectral_pair).
"""
from dataclasses import dataclass, field
from typing import NamedTuple, Protocol, runtime_checkable

import numpy as np
import torch

from ._arrays import as_device, field_dtype
from .fft_solvers import _tables, create_fft_cache, spectral_pair
from .grid import Grid1D, Grid2D, GridType


class SpectralBounds(NamedTuple):
    """Spectral radius and real/imaginary extents of an operator."""
    rho: float
    re_max: float
    im_max: float
    methods_used: dict


@runtime_checkable
class FFTLinearOperator(Protocol):
    """fft_operators.py:47-78: an FFT-diagonalised operator (eigenvalues,
    matvec, solve, exp_matvec, spectral_bounds); the ETD and implicit
    drivers accept any object with these members."""
    name: str
    grid: GridType

    @property
    def eigenvalues(self) -> torch.Tensor: ...

    def matvec(self, u) -> torch.Tensor: ...

    def solve(self, rhs, dt: float) -> torch.Tensor: ...

    def exp_matvec(self, u, dt: float) -> torch.Tensor: ...

    def spectral_bounds(self) -> SpectralBounds: ...


def _host_eigenvalues(op) -> np.ndarray:
    """The operator's eigenvalues on the host (from the symbol's 1-D tables
    when it has them, else one copy of the device array)."""
    lam = op.eigenvalues
    sep = getattr(getattr(op, "_cache", None), "laplacian_symbol", None)
    sep = getattr(sep, "_mjx_separable", None)
    if sep is not None and hasattr(op, "D"):
        lx, ly = np.asarray(sep[0]), np.asarray(sep[1])
        return op.D * (ly[:, None] + lx[None, :])
    return np.asarray(lam.detach().cpu().numpy() if isinstance(lam, torch.Tensor) else lam)


def exact_cfl_dt(op, method: str = "explicit", safety: float = 0.9) -> float:
    """fft_operators.py:263-300: safety * 2 / max|lambda| for explicit Euler;
    safety for 'imex' and 'etd' (no limit from L)."""
    if method == "explicit":
        rho = float(np.max(np.abs(_host_eigenvalues(op))))
        return float("inf") if rho < 1e-14 else safety * 2.0 / rho
    if method in ("imex", "etd"):
        return safety * 1.0
    raise ValueError(f"Unknown method: {method}")


@dataclass(frozen=True)
class DiffusionOperator:
    """fft_operators.py:80-176."""
    grid: GridType
    D: float
    dtype: torch.dtype = torch.float64
    name: str = field(default="diffusion", repr=False)

    def __post_init__(self):
        if self.D <= 0:
            raise ValueError(f"Diffusion coefficient D must be > 0, got {self.D}")
        if not isinstance(self.grid, Grid2D):
            raise NotImplementedError("1-D periodic operators are not on the MI355X path")
        object.__setattr__(self, "_cache", create_fft_cache(self.grid, self.dtype))

    @property
    def laplacian_symbol(self) -> torch.Tensor:
        return self._cache.laplacian_symbol

    @property
    def eigenvalues(self) -> torch.Tensor:
        """D * (2 cos(k dx) - 2)/dx^2 + ... (fft_operators.py:112-115)."""
        return self.D * self._cache.laplacian_symbol

    def _apply(self, u, alpha: float, beta: float, nrep: int) -> torch.Tensor:
        t = as_device(u)
        t = t.to(field_dtype(t))
        out, _ = spectral_pair(t, None, _tables(self._cache.laplacian_symbol, t.dtype),
                               [[beta, 0.0, alpha, 0.0]], nrep=nrep)
        return out.view(t.shape)

    def matvec(self, u) -> torch.Tensor:
        """D lap u via FFT (fft_operators.py:117-131)."""
        return self._apply(u, float(self.D), 0.0, -1)

    def solve(self, rhs, dt: float) -> torch.Tensor:
        """(I - dt D lap)^-1 rhs (fft_operators.py:133-149)."""
        return self._apply(rhs, 0.0, float(dt) * float(self.D), 1)

    def exp_matvec(self, u, dt: float) -> torch.Tensor:
        """exp(dt D lap) u (fft_operators.py:151-165)."""
        return self._apply(u, 0.0, float(dt) * float(self.D), -2)

    def spectral_bounds(self) -> SpectralBounds:
        lam = _host_eigenvalues(self)
        return SpectralBounds(rho=float(np.max(np.abs(lam))), re_max=float(np.max(lam)), im_max=0.0,
                              methods_used={"exact_fft": f"DiffusionOperator(D={self.D})"})
