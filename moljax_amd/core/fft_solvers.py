"""FFT-based implicit solvers on periodic grids (MI355X kernels).

Drop-in for moljax/core/fft_solvers.py.  Symbols and caches are built once on
the host in float64; every solve runs through the gfx950 kernels of
libmoljax_hip.so (mjx_spectral_pair): one row pass, one column pass and one
row pass back per solve, two real fields per launch.

Public signatures follow the reference:
  laplacian_symbol_2d(ny, nx, dy, dx, dtype)          fft_solvers.py:124-154
  create_fft_cache(grid, dtype) -> FFTCache2D        fft_solvers.py:161-215
  solve_helmholtz_2d(rhs_interior, symbol, dt, D)    fft_solvers.py:261-294
  solve_helmholtz(rhs_interior, symbol, dt, D)       fft_solvers.py:297-320
  apply_diffusion_inverse_fft(state, grid, dt, diffusivities, fft_cache)
                                                     fft_solvers.py:406-445
New (README.md:58 names it; semantics of benchmarks/benchmark_vs_scipy.py:74-91):
  fft_cn_step(u, D, dx, dt, symbol='spectral', n_steps=1)
"""
from typing import Dict, NamedTuple, Optional, Tuple

import numpy as np
import torch

from .. import _hip
from ._arrays import as_device, field_dtype, inst_table, spectral_work, view2d
from .grid import Grid1D, Grid2D, GridType, default_device


class FFTCache1D(NamedTuple):
    """fft_solvers.py:32-35."""
    laplacian_symbol: torch.Tensor
    k: torch.Tensor


class FFTCache2D(NamedTuple):
    """fft_solvers.py:38-42."""
    laplacian_symbol: torch.Tensor
    kx: torch.Tensor
    ky: torch.Tensor


# --------------------------------------------------------------- symbols
def _np_dtype(dtype):
    if dtype in (torch.float32, np.float32, "float32"):
        return torch.float32
    return torch.float64


def _fd_symbol_1d_np(n: int, d: float) -> np.ndarray:
    k = 2.0 * np.pi * np.fft.fftfreq(n, d=d)
    return (2.0 * np.cos(k * d) - 2.0) / (d * d)


def build_wavenumbers_1d(nx: int, dx: float, dtype=torch.float64) -> torch.Tensor:
    """fft_solvers.py:49-64."""
    k = 2.0 * np.pi * np.fft.fftfreq(nx, d=dx)
    return torch.from_numpy(k).to(device=default_device(), dtype=_np_dtype(dtype))


def build_wavenumbers_2d(ny: int, nx: int, dy: float, dx: float, dtype=torch.float64):
    """fft_solvers.py:67-94: (kx, ky), each (ny, nx); kx varies along axis 1."""
    kx_1d = 2.0 * np.pi * np.fft.fftfreq(nx, d=dx)
    ky_1d = 2.0 * np.pi * np.fft.fftfreq(ny, d=dy)
    kx = np.ascontiguousarray(np.broadcast_to(kx_1d, (ny, nx)))
    ky = np.ascontiguousarray(np.broadcast_to(ky_1d[:, None], (ny, nx)))
    dev, dt = default_device(), _np_dtype(dtype)
    return torch.from_numpy(kx).to(dev, dt), torch.from_numpy(ky).to(dev, dt)


def _tag_separable(lam: torch.Tensor, lamx: np.ndarray, lamy: np.ndarray) -> torch.Tensor:
    # The kernels evaluate lam = lamx[kx] + lamy[ky] from the two 1-D tables
    # instead of streaming the (ny, nx) array; the tag survives only on this
    # exact tensor object, any derived tensor takes the full-array path.
    lam._mjx_separable = (lamx, lamy)
    return lam


def laplacian_symbol_1d(nx: int, dx: float, dtype=torch.float64) -> torch.Tensor:
    """fft_solvers.py:97-121: (2 cos(k dx) - 2) / dx^2."""
    return torch.from_numpy(_fd_symbol_1d_np(nx, dx)).to(default_device(), _np_dtype(dtype))


def laplacian_symbol_2d(ny: int, nx: int, dy: float, dx: float, dtype=torch.float64) -> torch.Tensor:
    """fft_solvers.py:124-154: (2cos(kx dx)-2)/dx^2 + (2cos(ky dy)-2)/dy^2,
    shape (ny, nx)."""
    lx = _fd_symbol_1d_np(nx, dx)
    ly = _fd_symbol_1d_np(ny, dy)
    lam = torch.from_numpy(lx[None, :] + ly[:, None]).to(default_device(), _np_dtype(dtype))
    return _tag_separable(lam, lx, ly)


def create_fft_cache_1d(grid: Grid1D, dtype=torch.float64) -> FFTCache1D:
    """fft_solvers.py:161-178."""
    return FFTCache1D(laplacian_symbol=laplacian_symbol_1d(grid.nx, grid.dx, dtype),
                      k=build_wavenumbers_1d(grid.nx, grid.dx, dtype))


def create_fft_cache_2d(grid: Grid2D, dtype=torch.float64) -> FFTCache2D:
    """fft_solvers.py:181-198."""
    kx, ky = build_wavenumbers_2d(grid.ny, grid.nx, grid.dy, grid.dx, dtype)
    lam = laplacian_symbol_2d(grid.ny, grid.nx, grid.dy, grid.dx, dtype)
    return FFTCache2D(laplacian_symbol=lam, kx=kx, ky=ky)


def create_fft_cache(grid: GridType, dtype=torch.float64):
    """fft_solvers.py:201-215."""
    if isinstance(grid, Grid1D):
        return create_fft_cache_1d(grid, dtype)
    return create_fft_cache_2d(grid, dtype)


# ------------------------------------------------------- kernel plumbing
class _SymbolTables:
    """Device copies of a symbol in the kernel dtype, cached per symbol."""

    def __init__(self, lam: torch.Tensor, dtype: torch.dtype):
        sep = getattr(lam, "_mjx_separable", None)
        if sep is not None:
            self.lamx, self.lamy = (s.to("cuda", dtype).contiguous() if isinstance(s, torch.Tensor) else
                                    torch.from_numpy(np.asarray(s, dtype=np.float64)).to("cuda", dtype)
                                    for s in sep)
            self.lam2d = None
        else:
            self.lamx = self.lamy = None
            self.lam2d = as_device(lam).to(dtype).contiguous()


def _tables(lam, dtype: torch.dtype) -> _SymbolTables:
    if not isinstance(lam, torch.Tensor):
        lam = as_device(lam, torch.float64)
    cache = getattr(lam, "_mjx_tables", None)
    if cache is None:
        cache = {}
        try:
            lam._mjx_tables = cache
        except AttributeError:
            pass
    if dtype not in cache:
        cache[dtype] = _SymbolTables(lam, dtype)
    return cache[dtype]


def spectral_pair(a: torch.Tensor, b: Optional[torch.Tensor], tables: _SymbolTables,
                  rows, nrep: int = 1, a_out: Optional[torch.Tensor] = None,
                  b_out: Optional[torch.Tensor] = None, fill_ghosts: bool = False):
    """Run mjx_spectral_pair on (..., ny, nx) views a (and b).  ``rows`` is the
    per-instance list [beta_a, beta_b, alpha_a, alpha_b]; outputs default to
    fresh contiguous arrays."""
    lib = _hip.lib()
    ta, pitch, istr, B = view2d(a)
    ny, nx = ta.shape[-2], ta.shape[-1]
    dt = ta.dtype
    if b is not None:
        tb, pitch_b, istr_b, Bb = view2d(b)
        if (pitch_b, istr_b, Bb) != (pitch, istr, B) or tb.dtype != dt or tb.shape != ta.shape:
            tb = tb.contiguous().to(dt)
            ta = ta.contiguous()
            _, pitch, istr, B = view2d(ta)
    else:
        tb = None
    if a_out is None:
        a_out = torch.empty((B, ny, nx), dtype=dt, device=ta.device)
    if b is not None and b_out is None:
        b_out = torch.empty((B, ny, nx), dtype=dt, device=ta.device)
    oa, opitch, oistr, OB = view2d(a_out)
    if b_out is not None:
        ob, opitch_b, oistr_b, _ = view2d(b_out)
        if (opitch_b, oistr_b) != (opitch, oistr):
            raise ValueError("outputs a_out and b_out must share one layout")
    else:
        ob = None
    if OB != B:
        raise ValueError("output batch does not match input batch")
    if len(rows) == 1 and B > 1:
        rows = list(rows) * B
    inst = inst_table(rows, ta.device)
    work = spectral_work(dt, B, ny, nx, ta.device)
    rc = lib.mjx_spectral_pair(
        _hip.dtype_code(dt), B, ny, nx, _hip.ptr(ta), _hip.ptr(tb), pitch, istr,
        _hip.ptr(oa), _hip.ptr(ob), opitch, oistr, int(bool(fill_ghosts)),
        _hip.ptr(tables.lamx), _hip.ptr(tables.lamy), _hip.ptr(tables.lam2d),
        _hip.ptr(inst), int(nrep), _hip.ptr(work), _hip.stream_ptr())
    _hip.check(rc, "mjx_spectral_pair")
    return a_out, b_out


# ------------------------------------------------------------ Helmholtz
def solve_helmholtz_2d(rhs_interior, laplacian_symbol, dt: float, D: float) -> torch.Tensor:
    """Solve (I - dt D lap) u = rhs on a periodic grid (fft_solvers.py:261-294).

    rhs_interior: (ny, nx) interior values, or (..., ny, nx) for a batch that
    shares the symbol.  Returns a new array of the same shape and dtype."""
    rhs = as_device(rhs_interior)
    rhs = rhs.to(field_dtype(rhs))
    tabs = _tables(laplacian_symbol, rhs.dtype)
    beta = float(dt) * float(D)
    out, _ = spectral_pair(rhs, None, tabs, [[beta, 0.0, 0.0, 0.0]])
    return out.view(rhs.shape)


_LINE_SYMBOLS: Dict[tuple, torch.Tensor] = {}
_POW2_SIDES = (16, 32, 64, 128, 256, 512, 1024, 2048, 4096)


def _line_symbol_table(sym, rows: int) -> torch.Tensor:
    """The (rows, nx) symbol of stacked lines (y symbol 0), tagged separable:
    host symbols are cached by value, device symbols on the tensor itself."""
    if isinstance(sym, torch.Tensor):
        cache = getattr(sym, "_mjx_lines", None)
        if cache is None:
            cache = {}
            try:
                sym._mjx_lines = cache
            except AttributeError:
                pass
        lam2d = cache.get(rows)
        if lam2d is None:
            lam2d = _tag_separable(sym.reshape(1, -1).expand(rows, -1).contiguous(), sym,
                                   torch.zeros(rows, dtype=torch.float64, device=sym.device))
            cache[rows] = lam2d
        return lam2d
    key = (sym.tobytes(), rows)
    lam2d = _LINE_SYMBOLS.get(key)
    if lam2d is None:
        lam2d = _tag_separable(torch.from_numpy(np.broadcast_to(sym, (rows, sym.shape[-1])).copy()).to(default_device()),
                               sym, np.zeros(rows))
        _LINE_SYMBOLS[key] = lam2d
    return lam2d


def _lines_spectral(t: torch.Tensor, sym, row, nrep: int) -> torch.Tensor:
    """Apply the 1-D spectral multiplier of (sym, row, nrep) along the last
    axis of t (a batch of periodic lines): the lines are stacked as the rows
    of a 2-D array whose y symbol is zero, so the 2-D kernels apply the 1-D
    multiplier row by row.  For power-of-two nx the stacks are padded with
    zero rows to a power of two of at least 16 rows (the Stockham kernels);
    other lengths (2..2048) run the any-length kernels on the stack as it
    is."""
    nx = t.shape[-1]
    lines = t.reshape(-1, nx)
    out = torch.empty_like(lines)
    pow2 = nx in _POW2_SIDES
    step = 4096 if pow2 else 2048
    for i0 in range(0, lines.shape[0], step):
        chunk = lines[i0:i0 + step]
        L = chunk.shape[0]
        rows = max(16, 1 << (L - 1).bit_length()) if pow2 else max(2, L)
        lam2d = _line_symbol_table(sym, rows)
        buf = torch.zeros((rows, nx), dtype=t.dtype, device=t.device)
        buf[:L] = chunk
        res, _ = spectral_pair(buf, None, _tables(lam2d, t.dtype), [row], nrep=nrep)
        out[i0:i0 + L] = res.view(rows, nx)[:L]
    return out.reshape(t.shape)


def solve_helmholtz_1d(rhs_interior, laplacian_symbol, dt: float, D: float) -> torch.Tensor:
    """fft_solvers.py:222-258: (I - dt D lap) u = rhs on periodic lines of nx
    points (a batch along the leading axes)."""
    t = as_device(rhs_interior)
    t = t.to(field_dtype(t))
    nx = t.shape[-1]
    lam = as_device(laplacian_symbol, torch.float64).reshape(-1)
    if lam.numel() != nx:
        raise ValueError(f"symbol length {lam.numel()} does not match the line length {nx}")
    return _lines_spectral(t, lam.contiguous(), [float(dt) * float(D), 0.0, 0.0, 0.0], 1)


def solve_helmholtz(rhs_interior, laplacian_symbol, dt: float, D: float) -> torch.Tensor:
    """fft_solvers.py:297-320: 1-D or 2-D by the symbol's rank."""
    lam = laplacian_symbol
    if getattr(lam, "dim", lambda: np.ndim(lam))() == 1:
        return solve_helmholtz_1d(rhs_interior, lam, dt, D)
    return solve_helmholtz_2d(rhs_interior, lam, dt, D)


def extract_interior(f_padded: torch.Tensor, grid: GridType) -> torch.Tensor:
    """fft_solvers.py:327-342 (a view, no copy)."""
    if isinstance(grid, Grid1D):
        return f_padded[grid.interior_slice]
    sy, sx = grid.interior_slice
    return f_padded[..., sy, sx]


def embed_interior(u_interior: torch.Tensor, grid: GridType, padded_template: torch.Tensor) -> torch.Tensor:
    """fft_solvers.py:345-365: a new padded array with the template's ghosts."""
    out = padded_template.clone()
    if isinstance(grid, Grid1D):
        out[grid.interior_slice] = u_interior
    else:
        sy, sx = grid.interior_slice
        out[..., sy, sx] = u_interior
    return out


def solve_diffusion_fft_field(rhs_padded, grid: Grid2D, dt: float, D: float, fft_cache) -> torch.Tensor:
    """fft_solvers.py:368-399: solve on the interior, ghosts copied from rhs."""
    f = as_device(rhs_padded)
    f = f.to(field_dtype(f))
    out = f.clone()
    tabs = _tables(fft_cache.laplacian_symbol, f.dtype)
    spectral_pair(extract_interior(f, grid), None, tabs, [[float(dt) * float(D), 0.0, 0.0, 0.0]],
                  a_out=extract_interior(out, grid))
    return out


def apply_diffusion_inverse_fft(state_rhs: Dict[str, torch.Tensor], grid: Grid2D, dt: float,
                                diffusivities: Dict[str, float], fft_cache) -> Dict[str, torch.Tensor]:
    """fft_solvers.py:406-445: (I - dt D lap)^-1 per field with D > 1e-14;
    other fields pass through unchanged.  Diffusive fields are solved two per
    kernel launch (packed as a + i b)."""
    result = {}
    solve = []
    for name, f in state_rhs.items():
        D = float(diffusivities.get(name, 0.0))
        if D > 1e-14:
            solve.append((name, D))
        else:
            result[name] = f
    for i in range(0, len(solve), 2):
        pair = solve[i:i + 2]
        fa = as_device(state_rhs[pair[0][0]])
        fa = fa.to(field_dtype(fa))
        outs = [fa.clone()]
        fb = None
        if len(pair) == 2:
            fb = as_device(state_rhs[pair[1][0]]).to(fa.dtype)
            outs.append(fb.clone())
        tabs = _tables(fft_cache.laplacian_symbol, fa.dtype)
        betas = [float(dt) * pair[0][1], float(dt) * pair[1][1] if fb is not None else 0.0]
        spectral_pair(extract_interior(fa, grid), None if fb is None else extract_interior(fb, grid), tabs,
                      [[betas[0], betas[1], 0.0, 0.0]], a_out=extract_interior(outs[0], grid),
                      b_out=None if fb is None else extract_interior(outs[1], grid))
        for (name, _), o in zip(pair, outs):
            result[name] = o
    return {name: result[name] for name in state_rhs}


# ------------------------------------------------------- spectral filters
def exponential_filter_1d(nx: int, strength: float = 36.0, order: int = 36, dtype=torch.float64) -> torch.Tensor:
    """fft_solvers.py:477-503: sigma(k) = exp(-strength (|k|/k_max)^order)
    with |k| folded as min(k, nx - k) and normalised by nx // 2 (clamped to
    1), in FFT order."""
    k = np.arange(nx, dtype=np.float64)
    kn = np.minimum(np.minimum(k, nx - k) / float(nx // 2), 1.0)
    sig = np.exp(-float(strength) * np.power(kn, order))
    return torch.from_numpy(sig).to(default_device(), _np_dtype(dtype))


def exponential_filter_2d(ny: int, nx: int, strength: float = 36.0, order: int = 36,
                          dtype=torch.float64) -> torch.Tensor:
    """fft_solvers.py:506-531: the outer product sigma_y sigma_x, (ny, nx)."""
    sx = exponential_filter_1d(nx, strength, order, torch.float64)
    sy = exponential_filter_1d(ny, strength, order, torch.float64)
    return torch.outer(sy, sx).to(_np_dtype(dtype))


def _symmetrised(sig: np.ndarray) -> np.ndarray:
    """(sigma(k) + sigma(-k)) / 2 (every axis reversed modulo its length):
    the real part of ifft(sigma * fft(u)) for real u and real sigma is the
    transform with this even multiplier, which the half-spectrum kernels
    apply."""
    rev = sig
    for ax in range(sig.ndim):
        rev = np.roll(np.flip(rev, axis=ax), 1, axis=ax)
    return 0.5 * (sig + rev)


def _filter_tables(filter_kernel, dtype: torch.dtype):
    """Kernel tables of a filter kernel (cached on the tensor): the even part
    as a full (ny, nx) multiplier table (2-D) or a line symbol (1-D)."""
    t = filter_kernel if isinstance(filter_kernel, torch.Tensor) else as_device(filter_kernel, torch.float64)
    if t.is_complex():
        raise NotImplementedError("spectral filters take real kernels")
    cache = getattr(t, "_mjx_filter", None)
    if cache is None and getattr(t, "_mjx_even", False):
        # tables made even on the device (mjx_phi_tables): used as they are
        d = t.detach().to(torch.float64).contiguous()
        cache = {"sym": d, "full": d}
    if cache is None:
        sym = _symmetrised(t.detach().double().cpu().numpy())
        cache = {"sym": sym, "full": torch.from_numpy(sym).to(default_device())}
        try:
            t._mjx_filter = cache
        except AttributeError:
            pass
    return cache


def apply_spectral_filter_interior(u_interior, filter_kernel) -> torch.Tensor:
    """fft_solvers.py:534-555: ifft(fft(u) * sigma).real for a 1-D u (or a
    batch of lines) with an (nx,) kernel, ifft2(fft2(u) * sigma).real for
    (..., ny, nx) with an (ny, nx) kernel; one spectral round trip on the
    kernels (multiplier sigma, the matvec form of mjx_spectral_pair)."""
    u = as_device(u_interior)
    u = u.to(field_dtype(u))
    k = filter_kernel if isinstance(filter_kernel, torch.Tensor) else as_device(filter_kernel, torch.float64)
    c = _filter_tables(k, u.dtype)
    if k.dim() == 1:
        if k.numel() != u.shape[-1]:
            raise ValueError(f"filter length {k.numel()} does not match the line length {u.shape[-1]}")
        return _lines_spectral(u, c["sym"], [0.0, 0.0, 1.0, 0.0], -1)
    if tuple(k.shape) != tuple(u.shape[-2:]):
        raise ValueError(f"filter shape {tuple(k.shape)} does not match the grid {tuple(u.shape[-2:])}")
    out, _ = spectral_pair(u, None, _tables(c["full"], u.dtype), [[0.0, 0.0, 1.0, 0.0]], nrep=-1)
    return out.view(u.shape)


def apply_spectral_filter_field(f_padded, grid: GridType, filter_kernel) -> torch.Tensor:
    """fft_solvers.py:558-581: filter the interior, keep the ghosts."""
    f = as_device(f_padded)
    f = f.to(field_dtype(f))
    return embed_interior(apply_spectral_filter_interior(extract_interior(f, grid), filter_kernel), grid, f)


def apply_spectral_filter_state(state: Dict[str, torch.Tensor], grid: GridType, filter_kernel,
                                filter_fields: Optional[Dict[str, bool]] = None) -> Dict[str, torch.Tensor]:
    """fft_solvers.py:584-616: every field (or those filter_fields marks
    True, default True) filtered; the others returned as they are."""
    out = {}
    for name, f in state.items():
        if filter_fields is None or filter_fields.get(name, True):
            out[name] = apply_spectral_filter_field(f, grid, filter_kernel)
        else:
            out[name] = f
    return out


# --------------------------------------------------------------- FFT-CN
def apply_diffusion_fft(state: Dict[str, torch.Tensor], grid: Grid2D, dt: float, diffusivities: Dict[str, float],
                        fft_cache) -> Dict[str, torch.Tensor]:
    """fft_solvers.py:448-474: the implicit diffusion step (the same as
    apply_diffusion_inverse_fft)."""
    return apply_diffusion_inverse_fft(state, grid, dt, diffusivities, fft_cache)


def diffusion_rhs_fft(state: Dict[str, torch.Tensor], grid: Grid2D, diffusivities: Dict[str, float],
                      fft_cache) -> Dict[str, torch.Tensor]:
    """fft_solvers.py:619-662: D lap u by FFT on each field's interior,
    embedded in zero ghosts; fields with D <= 1e-14 (or none) give zeros.
    Two fields with the same layout share one spectral round trip."""
    out = {}
    active = [n for n in state if float(diffusivities.get(n, 0.0)) > 1e-14]
    for n in state:
        if n not in active:
            out[n] = torch.zeros_like(as_device(state[n]))
    for i in range(0, len(active), 2):
        names = active[i:i + 2]
        fs = [as_device(state[n]) for n in names]
        dt_ = field_dtype(fs[0])
        ints = [extract_interior(f.to(dt_), grid) for f in fs]
        tabs = _tables(fft_cache.laplacian_symbol, dt_)
        Ds = [float(diffusivities[n]) for n in names] + [0.0]
        a, b = spectral_pair(ints[0], ints[1] if len(ints) > 1 else None, tabs, [[0.0, 0.0, Ds[0], Ds[1]]], nrep=-1)
        for n, f, r in zip(names, fs, (a, b)):
            pad = torch.zeros_like(f.to(dt_))
            extract_interior(pad, grid).copy_(r.view(extract_interior(pad, grid).shape))
            out[n] = pad
    return out


def _spectral_symbol_np(n: int, dx: float) -> np.ndarray:
    k = np.fft.fftfreq(n, dx) * 2 * np.pi
    return -(k ** 2)


def fft_cn_step(u, D: float, dx: float, dt: float, symbol: str = "spectral", n_steps: int = 1,
                dy: Optional[float] = None) -> torch.Tensor:
    """Crank-Nicolson diffusion step(s) on a periodic grid: u_hat <- G^n u_hat,
    G = (1 + dt D lam / 2) / (1 - dt D lam / 2).

    Reference semantics: benchmarks/benchmark_vs_scipy.py:74-91 (and the same
    form in benchmark_scaling.py:53-79, bench_precision_fp32_fp64.py:47-82):
    lam = -(kx^2 + ky^2) with k = 2 pi fftfreq(n, dx) ('spectral'), or the FD
    symbol of laplacian_symbol_2d ('fd').  n_steps > 1 applies G^n in one FFT
    round trip (the same linear operator as n separate steps).
    u: (..., n0, n1) interior array; axis order does not matter for the
    spectral symbol on equal spacings."""
    v = as_device(u)
    v = v.to(field_dtype(v))
    n0, n1 = v.shape[-2], v.shape[-1]
    dy = dx if dy is None else dy
    if symbol == "spectral":
        lx, ly = _spectral_symbol_np(n1, dx), _spectral_symbol_np(n0, dy)
    elif symbol == "fd":
        lx, ly = _fd_symbol_1d_np(n1, dx), _fd_symbol_1d_np(n0, dy)
    else:
        raise ValueError(f"unknown symbol {symbol!r}: use 'spectral' or 'fd'")
    lam = torch.zeros(1)
    _tag_separable(lam, lx, ly)
    tabs = _tables(lam, v.dtype)
    h = 0.5 * float(dt) * float(D)
    out, _ = spectral_pair(v, None, tabs, [[h, 0.0, h, 0.0]], nrep=int(n_steps))
    return out.view(v.shape)


fft_cn_integrate = fft_cn_step
