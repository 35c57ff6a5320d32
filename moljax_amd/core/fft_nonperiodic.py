"""DST-I / DCT-II solvers for homogeneous Dirichlet / Neumann boundaries
(moljax/core/fft_nonperiodic.py), on the MI355X r2r kernels.

The reference is 1-D; every solve here also accepts 2-D (..., ny, nx) input
and applies the transform along both axes (axis 0 = y).  For 2-D input the
symbol argument is a pair (lam_y, lam_x) of 1-D symbols, a NonperiodicFFTCache2D,
or one 1-D symbol used on both axes of a square grid.

Sizes: any N in [2, 2048].  Dirichlet with N + 1, and Neumann / cell
Dirichlet with N, a power of two in [16, 1024] run the Stockham kernels; the
other lengths run the chirp-z (Bluestein) kernels; sizes outside [2, 2048]
raise ValueError.

The library's own grids are cell-centred, and moljax/core/bc.py fills a
Dirichlet ghost as 2 g - interior (bc.py:237-258).  That stencil is
diagonalised by the DST-II, not the DST-I above; ``*_dirichlet_cc`` and
BCType.DIRICHLET_CC serve it, and solve_helmholtz_bc /
solve_poisson_bc take a FieldBCSpec with side values or fluxes so that the
solve inverts the FD operator evaluated on bc.py's filled ghosts exactly.
"""
from __future__ import annotations

import ctypes
from enum import Enum
from typing import NamedTuple, Tuple, Union

import numpy as np
import torch

from .. import _hip
from ._arrays import as_device, field_dtype, inst_table, view2d
from .grid import default_device

R2R_DST, R2R_DCT, R2R_DST2 = 0, 1, 2
HELM, POISSON = 0, 1


class BCType(Enum):
    """fft_nonperiodic.py:28-33."""
    DIRICHLET = "dirichlet"
    NEUMANN = "neumann"
    MIXED_DN = "mixed_dn"
    MIXED_ND = "mixed_nd"
    DIRICHLET_CC = "dirichlet_cc"  # cell-centred Dirichlet (bc.py's ghost fill)


class NonperiodicFFTCache(NamedTuple):
    """fft_nonperiodic.py:36-42."""
    N: int
    dx: float
    bc_type: str
    laplacian_symbol: torch.Tensor
    k: torch.Tensor


class NonperiodicFFTCache2D(NamedTuple):
    """2-D cache: per-axis symbols (axis 0 = y)."""
    ny: int
    nx: int
    dy: float
    dx: float
    bc_type: str
    lam_y: torch.Tensor
    lam_x: torch.Tensor


def _dev(a: np.ndarray, dtype) -> torch.Tensor:
    return torch.from_numpy(a).to(default_device(), dtype)


def laplacian_symbol_dirichlet(N: int, dx: float, dtype=torch.float64) -> torch.Tensor:
    """fft_nonperiodic.py:140-162."""
    k = np.arange(1, N + 1, dtype=np.float64)
    return _dev(-4.0 / (dx * dx) * np.sin(np.pi * k / (2 * (N + 1))) ** 2, dtype)


def laplacian_symbol_neumann(N: int, dx: float, dtype=torch.float64) -> torch.Tensor:
    """fft_nonperiodic.py:165-183."""
    k = np.arange(N, dtype=np.float64)
    return _dev(-4.0 / (dx * dx) * np.sin(np.pi * k / (2 * N)) ** 2, dtype)


def laplacian_symbol_dirichlet_cc(N: int, dx: float, dtype=torch.float64) -> torch.Tensor:
    """Cell-centred Dirichlet (ghost = -interior): -4/dx^2 sin^2(pi k/(2N)),
    k = 1..N, the DST-II eigenvalues of the 5-point stencil."""
    k = np.arange(1, N + 1, dtype=np.float64)
    return _dev(-4.0 / (dx * dx) * np.sin(np.pi * k / (2 * N)) ** 2, dtype)


def laplacian_symbol_mixed_dn(N: int, dx: float, dtype=torch.float64) -> torch.Tensor:
    """fft_nonperiodic.py:186-194 (symbol only: no mixed transform kernel)."""
    k = np.arange(1, N + 1, dtype=np.float64)
    return _dev(-4.0 / (dx * dx) * np.sin(np.pi * (2 * k - 1) / (4 * N)) ** 2, dtype)


def create_nonperiodic_fft_cache(N: int, dx: float, bc_type: Union[BCType, str] = BCType.DIRICHLET,
                                 dtype=torch.float64) -> NonperiodicFFTCache:
    """fft_nonperiodic.py:201-240."""
    if isinstance(bc_type, str):
        bc_type = BCType(bc_type.lower())
    if bc_type == BCType.DIRICHLET:
        lam, k = laplacian_symbol_dirichlet(N, dx, dtype), np.arange(1, N + 1)
    elif bc_type == BCType.NEUMANN:
        lam, k = laplacian_symbol_neumann(N, dx, dtype), np.arange(N)
    elif bc_type == BCType.MIXED_DN:
        lam, k = laplacian_symbol_mixed_dn(N, dx, dtype), np.arange(1, N + 1)
    elif bc_type == BCType.DIRICHLET_CC:
        lam, k = laplacian_symbol_dirichlet_cc(N, dx, dtype), np.arange(1, N + 1)
    else:
        raise ValueError(f"Unsupported BC type: {bc_type}")
    return NonperiodicFFTCache(N=N, dx=dx, bc_type=bc_type.value, laplacian_symbol=lam,
                               k=_dev(k.astype(np.float64), dtype))


def create_nonperiodic_fft_cache_2d(ny: int, nx: int, dy: float, dx: float,
                                    bc_type: Union[BCType, str] = BCType.DIRICHLET,
                                    dtype=torch.float64) -> NonperiodicFFTCache2D:
    if isinstance(bc_type, str):
        bc_type = BCType(bc_type.lower())
    if bc_type == BCType.DIRICHLET:
        ly, lx = laplacian_symbol_dirichlet(ny, dy, dtype), laplacian_symbol_dirichlet(nx, dx, dtype)
    elif bc_type == BCType.NEUMANN:
        ly, lx = laplacian_symbol_neumann(ny, dy, dtype), laplacian_symbol_neumann(nx, dx, dtype)
    elif bc_type == BCType.DIRICHLET_CC:
        ly, lx = laplacian_symbol_dirichlet_cc(ny, dy, dtype), laplacian_symbol_dirichlet_cc(nx, dx, dtype)
    else:
        raise ValueError(f"2-D transforms serve Dirichlet, cell-centred Dirichlet or Neumann, not {bc_type}")
    return NonperiodicFFTCache2D(ny=ny, nx=nx, dy=dy, dx=dx, bc_type=bc_type.value, lam_y=ly, lam_x=lx)


# ----------------------------------------------------------- transforms
def _lines(x: torch.Tensor):
    """(nlines, n) view of the last axis with its line stride."""
    if x.stride(-1) != 1:
        x = x.contiguous()
    n = x.shape[-1]
    flat = x.reshape(-1, n)
    if flat.dim() == 2 and flat.stride(1) == 1:
        return flat, flat.stride(0) if flat.shape[0] > 1 else n
    flat = flat.contiguous()
    return flat, n


def _transform(x, kind: int, inverse: bool, scale: float) -> torch.Tensor:
    t = as_device(x)
    t = t.to(field_dtype(t))
    lines, ls = _lines(t)
    out = torch.empty(lines.shape, dtype=t.dtype, device=t.device)
    rc = _hip.lib().mjx_r2r_transform1d(_hip.dtype_code(t.dtype), kind, int(inverse), lines.shape[0],
                                        lines.shape[1], _hip.ptr(lines), ls, _hip.ptr(out), out.stride(0),
                                        float(scale), _hip.stream_ptr())
    _hip.check(rc, "mjx_r2r_transform1d")
    return out.view(t.shape)


def dst_I_fast(x) -> torch.Tensor:
    """fft_nonperiodic.py:97-122: X[k] = sum_n x[n] sin(pi (k+1)(n+1)/(N+1))
    along the last axis."""
    return _transform(x, R2R_DST, False, 1.0)


def idst_I_fast(X) -> torch.Tensor:
    """fft_nonperiodic.py:125-133: dst_I_fast(X) * 2/(N+1)."""
    N = np.shape(X)[-1]
    return _transform(X, R2R_DST, False, 2.0 / (N + 1))


def dst_I(x) -> torch.Tensor:
    """fft_nonperiodic.py:49-77 (the O(N^2) form, which carries an extra
    factor 2 relative to dst_I_fast)."""
    return _transform(x, R2R_DST, False, 2.0)


def idst_I(X) -> torch.Tensor:
    """fft_nonperiodic.py:80-94: dst_I(X) / (N+1)."""
    N = np.shape(X)[-1]
    return _transform(X, R2R_DST, False, 2.0 / (N + 1))


# ---------------------------------------------------------------- solves
def _symbols_2d(symbol, ny: int, nx: int, dtype) -> Tuple[torch.Tensor, torch.Tensor]:
    if isinstance(symbol, NonperiodicFFTCache2D):
        ly, lx = symbol.lam_y, symbol.lam_x
    elif isinstance(symbol, (tuple, list)) and len(symbol) == 2:
        ly, lx = symbol
    else:
        s = as_device(symbol)
        if s.dim() != 1 or ny != nx:
            raise ValueError("2-D solves take (lam_y, lam_x), a NonperiodicFFTCache2D, "
                             "or one 1-D symbol for a square grid")
        ly = lx = s
    ly = as_device(ly, dtype).contiguous()
    lx = as_device(lx, dtype).contiguous()
    if ly.numel() != ny or lx.numel() != nx:
        raise ValueError(f"symbol lengths ({ly.numel()}, {lx.numel()}) do not match grid ({ny}, {nx})")
    return ly, lx


def _solve(kind: int, rhs, symbol, beta: float, mmode: int, sides=None, dy: float = 0.0,
           dx: float = 0.0) -> torch.Tensor:
    """sides: None, or bc.py side values / fluxes (left, right, bottom, top)
    folded into the right-hand side by the kernel (kind 1 or 2)."""
    t = as_device(rhs)
    t = t.to(field_dtype(t))
    lib = _hip.lib()
    inst = inst_table([[beta, 0.0, 0.0, 0.0]], t.device)
    sv = None if sides is None else (ctypes.c_double * 4)(*[float(v) for v in sides])
    if t.dim() == 1:
        lam = as_device(symbol, t.dtype).contiguous()
        if lam.numel() != t.shape[0]:
            raise ValueError("symbol length does not match rhs")
        lines, ls = _lines(t)
        out = torch.empty_like(lines)
        if sv is None:
            rc = lib.mjx_r2r_solve1d(_hip.dtype_code(t.dtype), kind, 1, t.shape[0], _hip.ptr(lines), ls,
                                     _hip.ptr(out), t.shape[0], _hip.ptr(lam), _hip.ptr(inst), mmode,
                                     _hip.stream_ptr())
        else:
            rc = lib.mjx_r2r_solve1d_bc(_hip.dtype_code(t.dtype), kind, 1, t.shape[0], _hip.ptr(lines), ls,
                                        _hip.ptr(out), t.shape[0], _hip.ptr(lam), _hip.ptr(inst), mmode, sv,
                                        float(dx), _hip.stream_ptr())
        _hip.check(rc, "mjx_r2r_solve1d")
        return out.view(t.shape)
    src, pitch, istr, B = view2d(t)
    ny, nx = src.shape[-2], src.shape[-1]
    ly, lx = _symbols_2d(symbol, ny, nx, t.dtype)
    inst = inst_table([[beta, 0.0, 0.0, 0.0]] * B, t.device)
    out = torch.empty((B, ny, nx), dtype=t.dtype, device=t.device)
    work = torch.empty((B, ny, nx), dtype=t.dtype, device=t.device)
    if sv is None:
        rc = lib.mjx_r2r_solve2d(_hip.dtype_code(t.dtype), kind, B, ny, nx, _hip.ptr(src), pitch, istr,
                                 _hip.ptr(out), nx, ny * nx, _hip.ptr(lx), _hip.ptr(ly), _hip.ptr(inst), mmode,
                                 _hip.ptr(work), _hip.stream_ptr())
    else:
        rc = lib.mjx_r2r_solve2d_bc(_hip.dtype_code(t.dtype), kind, B, ny, nx, _hip.ptr(src), pitch, istr,
                                    _hip.ptr(out), nx, ny * nx, _hip.ptr(lx), _hip.ptr(ly), _hip.ptr(inst), mmode,
                                    sv, float(dy), float(dx), _hip.ptr(work), _hip.stream_ptr())
    _hip.check(rc, "mjx_r2r_solve2d")
    return out.view(t.shape)


def solve_helmholtz_dirichlet(rhs, laplacian_symbol, dt: float, D: float) -> torch.Tensor:
    """(I - dt D lap) u = f, u = 0 on the boundary (fft_nonperiodic.py:308-330)."""
    return _solve(R2R_DST, rhs, laplacian_symbol, float(dt) * float(D), HELM)


def solve_helmholtz_neumann(rhs, laplacian_symbol, dt: float, D: float) -> torch.Tensor:
    """(I - dt D lap) u = f, du/dn = 0 (fft_nonperiodic.py:333-355)."""
    return _solve(R2R_DCT, rhs, laplacian_symbol, float(dt) * float(D), HELM)


def solve_poisson_dirichlet(rhs, laplacian_symbol) -> torch.Tensor:
    """lap u = f, u = 0 on the boundary (fft_nonperiodic.py:247-263)."""
    return _solve(R2R_DST, rhs, laplacian_symbol, 0.0, POISSON)


def solve_poisson_neumann(rhs, laplacian_symbol) -> torch.Tensor:
    """lap u = f with du/dn = 0, zero-mean solution (fft_nonperiodic.py:266-305)."""
    return _solve(R2R_DCT, rhs, laplacian_symbol, 0.0, POISSON)


def solve_helmholtz_dirichlet_cc(rhs, laplacian_symbol, dt: float, D: float) -> torch.Tensor:
    """(I - dt D lap) u = f on a cell-centred grid with u's ghosts at
    -interior (homogeneous bc.py Dirichlet), by the DST-II."""
    return _solve(R2R_DST2, rhs, laplacian_symbol, float(dt) * float(D), HELM)


def solve_poisson_dirichlet_cc(rhs, laplacian_symbol) -> torch.Tensor:
    """lap u = f, cell-centred homogeneous Dirichlet, by the DST-II."""
    return _solve(R2R_DST2, rhs, laplacian_symbol, 0.0, POISSON)


_BC_SYMBOLS: dict = {}


def _bc_setup(rhs, grid, spec, t: float, params):
    from .bc import BCType as FieldBC
    from .bc_fill import _side_values
    kind = int(spec.kind)
    if kind == FieldBC.DIRICHLET:
        r2r, sides = R2R_DST2, _side_values(spec.value, grid, t, params or {})
        sym = laplacian_symbol_dirichlet_cc
    elif kind == FieldBC.NEUMANN:
        r2r, sides = R2R_DCT, _side_values(spec.flux, grid, t, params or {})
        sym = laplacian_symbol_neumann
    else:
        raise ValueError("solve_*_bc serves Dirichlet and Neumann specs; periodic fields use fft_solvers")
    tr = as_device(rhs)
    dtype = field_dtype(tr)
    if tr.dim() == 1:
        key = (r2r, tr.shape[-1], grid.dx, dtype)
        if key not in _BC_SYMBOLS:
            _BC_SYMBOLS[key] = sym(tr.shape[-1], grid.dx, dtype)
        return r2r, _BC_SYMBOLS[key], (sides[0], sides[1], 0.0, 0.0)
    ny, nx = tr.shape[-2], tr.shape[-1]
    if (ny, nx) != (grid.ny, grid.nx):
        raise ValueError(f"rhs interior {(ny, nx)} does not match the grid {(grid.ny, grid.nx)}")
    key = (r2r, ny, nx, grid.dy, grid.dx, dtype)
    if key not in _BC_SYMBOLS:
        _BC_SYMBOLS[key] = (sym(ny, grid.dy, dtype), sym(nx, grid.dx, dtype))
    return r2r, _BC_SYMBOLS[key], sides


def solve_helmholtz_bc(rhs, grid, spec, dt: float, D: float, t: float = 0.0, params=None) -> torch.Tensor:
    """u with u - dt D lap_h u = rhs on the interior cells, lap_h the 5-point
    Laplacian (operators.py:311-322) of u padded by bc.py's fill for ``spec``
    (a Dirichlet or Neumann FieldBCSpec, side values / fluxes evaluated as
    bc.py:323-349 does).  rhs: (..., ny, nx) interior arrays (or 1-D lines,
    with the left / right values)."""
    kind, sym, sides = _bc_setup(rhs, grid, spec, t, params)
    return _solve(kind, rhs, sym, float(dt) * float(D), HELM, sides, grid.dy if hasattr(grid, "dy") else grid.dx,
                  grid.dx)


def solve_poisson_bc(rhs, grid, spec, t: float = 0.0, params=None) -> torch.Tensor:
    """lap_h u = rhs with lap_h on u padded by bc.py's fill for ``spec``
    (Neumann: the zero-mean solution; rhs must satisfy the compatibility
    condition with the boundary fluxes)."""
    kind, sym, sides = _bc_setup(rhs, grid, spec, t, params)
    return _solve(kind, rhs, sym, 0.0, POISSON, sides, grid.dy if hasattr(grid, "dy") else grid.dx, grid.dx)


def _etd1(kind: int, u, N_u, eigenvalues, dt: float) -> torch.Tensor:
    tu = as_device(u)
    tu = tu.to(field_dtype(tu))
    tn = as_device(N_u, tu.dtype)
    if tn.shape != tu.shape:
        raise ValueError("u and N_u must have the same shape")
    n = tu.shape[-1]
    eig = as_device(eigenvalues, tu.dtype).reshape(-1).contiguous()
    if eig.numel() != n:
        raise ValueError("eigenvalues length does not match the line length")
    lu, ls = _lines(tu)
    ln, ns = _lines(tn)
    if ns != ls:
        ln, ns = ln.contiguous(), n
        lu, ls = lu.contiguous(), n
    out = torch.empty(lu.shape, dtype=tu.dtype, device=tu.device)
    rc = _hip.lib().mjx_r2r_etd1(_hip.dtype_code(tu.dtype), kind, lu.shape[0], n, _hip.ptr(lu), _hip.ptr(ln), ls,
                                 _hip.ptr(out), out.stride(0), _hip.ptr(eig), float(dt), _hip.stream_ptr())
    _hip.check(rc, "mjx_r2r_etd1")
    return out.view(tu.shape)


def etd1_dirichlet(u, N_u, eigenvalues, dt: float) -> torch.Tensor:
    """fft_nonperiodic.py:363-398: u <- DST^-1[e^z DST u + dt phi1(z) DST N],
    z = dt * eigenvalues (the diffusion eigenvalues D lam of the vertex
    Dirichlet grid), along the last axis."""
    return _etd1(R2R_DST, u, N_u, eigenvalues, dt)


def etd1_neumann(u, N_u, eigenvalues, dt: float) -> torch.Tensor:
    """fft_nonperiodic.py:401-437, with the orthonormal DCT-II."""
    return _etd1(R2R_DCT, u, N_u, eigenvalues, dt)


def etd1_dirichlet_cc(u, N_u, eigenvalues, dt: float) -> torch.Tensor:
    """The same step on a cell-centred Dirichlet grid (bc.py's ghost fill),
    with the DST-II."""
    return _etd1(R2R_DST2, u, N_u, eigenvalues, dt)


def check_compatibility_neumann(rhs, tol: float = 1e-10) -> bool:
    """fft_nonperiodic.py:440-450."""
    from .state import RED_SUM, _reduce
    t = as_device(rhs)
    return abs(float(_reduce(RED_SUM, [t.to(field_dtype(t))]).sum())) < tol


def project_to_compatible(rhs) -> torch.Tensor:
    """fft_nonperiodic.py:453-461."""
    from .state import RED_SUM, _reduce, axpby
    t = as_device(rhs)
    t = t.to(field_dtype(t))
    return axpby(1.0, t, c=-float(_reduce(RED_SUM, [t]).sum()) / t.numel())
