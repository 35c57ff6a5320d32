"""Uniform cell-centred grids with ghost cells.

Same contract as moljax/core/grid.py:20-259: arrays are (nx + 2g,) in 1-D and
(ny + 2g, nx + 2g) in 2-D, axis 0 = y (rows), axis 1 = x (columns);
dx = (x_max - x_min) / nx; cell centres at x_min + (i + 1/2) dx.
Coordinates are returned as float64 torch tensors on the default device.
"""
from dataclasses import dataclass
from typing import Tuple, Union

import torch


def default_device() -> torch.device:
    return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")


@dataclass(frozen=True)
class Grid1D:
    """moljax/core/grid.py:20-112."""
    nx: int
    x_min: float
    x_max: float
    n_ghost: int = 1

    @classmethod
    def uniform(cls, nx: int, x_min: float, x_max: float, n_ghost: int = 1) -> "Grid1D":
        return cls(nx=nx, x_min=x_min, x_max=x_max, n_ghost=n_ghost)

    @property
    def dx(self) -> float:
        return (self.x_max - self.x_min) / self.nx

    @property
    def nx_total(self) -> int:
        return self.nx + 2 * self.n_ghost

    @property
    def interior_slice(self) -> slice:
        return slice(self.n_ghost, self.n_ghost + self.nx)

    @property
    def min_dx(self) -> float:
        return self.dx

    @property
    def min_dx2(self) -> float:
        return self.dx ** 2

    def x_coords(self, include_ghost: bool = False) -> torch.Tensor:
        dev = default_device()
        if include_ghost:
            x0 = self.x_min - self.n_ghost * self.dx
            return torch.linspace(x0 + 0.5 * self.dx, x0 + (self.nx_total - 0.5) * self.dx, self.nx_total,
                                  dtype=torch.float64, device=dev)
        return torch.linspace(self.x_min + 0.5 * self.dx, self.x_max - 0.5 * self.dx, self.nx,
                              dtype=torch.float64, device=dev)


@dataclass(frozen=True)
class Grid2D:
    """moljax/core/grid.py:115-259."""
    nx: int
    ny: int
    x_min: float
    x_max: float
    y_min: float
    y_max: float
    n_ghost: int = 1

    @classmethod
    def uniform(cls, nx: int, ny: int, x_min: float, x_max: float, y_min: float, y_max: float,
                n_ghost: int = 1) -> "Grid2D":
        return cls(nx=nx, ny=ny, x_min=x_min, x_max=x_max, y_min=y_min, y_max=y_max, n_ghost=n_ghost)

    @property
    def dx(self) -> float:
        return (self.x_max - self.x_min) / self.nx

    @property
    def dy(self) -> float:
        return (self.y_max - self.y_min) / self.ny

    @property
    def nx_total(self) -> int:
        return self.nx + 2 * self.n_ghost

    @property
    def ny_total(self) -> int:
        return self.ny + 2 * self.n_ghost

    @property
    def interior_slice(self) -> Tuple[slice, slice]:
        return (slice(self.n_ghost, self.n_ghost + self.ny), slice(self.n_ghost, self.n_ghost + self.nx))

    @property
    def min_dx(self) -> float:
        return min(self.dx, self.dy)

    @property
    def min_dx2(self) -> float:
        return self.min_dx ** 2

    def meshgrid(self, include_ghost: bool = False) -> Tuple[torch.Tensor, torch.Tensor]:
        """(X, Y) with X varying along axis 1 and Y along axis 0 (indexing='xy')."""
        dev = default_device()
        if include_ghost:
            x0 = self.x_min - self.n_ghost * self.dx
            y0 = self.y_min - self.n_ghost * self.dy
            x = torch.linspace(x0 + 0.5 * self.dx, x0 + (self.nx_total - 0.5) * self.dx, self.nx_total,
                               dtype=torch.float64, device=dev)
            y = torch.linspace(y0 + 0.5 * self.dy, y0 + (self.ny_total - 0.5) * self.dy, self.ny_total,
                               dtype=torch.float64, device=dev)
        else:
            x = torch.linspace(self.x_min + 0.5 * self.dx, self.x_max - 0.5 * self.dx, self.nx,
                               dtype=torch.float64, device=dev)
            y = torch.linspace(self.y_min + 0.5 * self.dy, self.y_max - 0.5 * self.dy, self.ny,
                               dtype=torch.float64, device=dev)
        return torch.meshgrid(x, y, indexing="xy")


GridType = Union[Grid1D, Grid2D]
