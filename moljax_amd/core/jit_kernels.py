"""The reference's jitted kernels (moljax/core/jit_kernels.py) on the MI355X
spectral multiply.

Each kernel is a sum of real spectral multipliers applied to its inputs:
ifft(m * fft(x)).real along the last axis (1-D eigenvalues) or the last two
axes (2-D eigenvalues), which apply_spectral_filter_interior runs as one
native spectral round trip per input (the multiplier's even part, which is
what the real part applies).  The multiplier tables (e^z, dt phi_k(z), the
ETDRK4 weights) are evaluated in FP64 on the device (mjx_phi_tables) from
the eigenvalues with the reference's formulas, series below |z| = 1e-4.
Real eigenvalues only (the complex advection-diffusion solve is not on this
path).
"""
from typing import Callable, Tuple

import numpy as np
import torch

from ._arrays import as_device, field_dtype
from .fft_solvers import apply_spectral_filter_interior, solve_helmholtz_1d, solve_helmholtz_2d
from .state import axpby


def _lam_dev(z) -> torch.Tensor:
    """Eigenvalues as a contiguous FP64 device tensor (device input stays on
    the device)."""
    if (isinstance(z, torch.Tensor) and z.is_complex()) or (not isinstance(z, torch.Tensor) and np.iscomplexobj(z)):
        raise NotImplementedError("complex eigenvalues are not on the MI355X path")
    return as_device(z, torch.float64).contiguous()


def _tables(lam: torch.Tensor, dt: float, mode: int, count: int, even: bool = True):
    """mjx_phi_tables: `count` FP64 multiplier tables shaped like lam (mode 0
    phi1..phi3 of lam; 1 ETD1; 2 ETD2; 3 ETDRK4 weights), on the device.
    even: each table is its even part over lam's last one or two axes (what
    the spectral round trip applies), tagged so that it is used as it is."""
    from .. import _hip
    if lam.dim() > 2:
        raise ValueError("eigenvalues must be one line (n,) or one grid (ny, nx)")
    ny, nx = (1, lam.shape[-1]) if lam.dim() <= 1 else tuple(lam.shape)
    outs = [torch.empty_like(lam) for _ in range(count)]
    ptrs = [_hip.ptr(o) for o in outs] + [None] * (4 - count)
    _hip.check(_hip.lib().mjx_phi_tables(int(ny), int(nx), _hip.ptr(lam), float(dt), int(mode), int(even), *ptrs,
                                         _hip.stream_ptr()), "mjx_phi_tables")
    if even:
        for o in outs:
            o._mjx_even = True
    return outs


def _phi_k(z, k: int) -> torch.Tensor:
    """phi_k of z elementwise (FP64, any shape, on the device)."""
    lam = _lam_dev(z)
    return _tables(lam.reshape(-1), 0.0, 0, 3, even=False)[k - 1].view(lam.shape)


def phi1(z) -> torch.Tensor:
    """jit_kernels.py:29-35: (e^z - 1)/z, 1 + z/2 + z^2/6 + z^3/24 below 1e-4."""
    return _phi_k(z, 1)


def phi2(z) -> torch.Tensor:
    """jit_kernels.py:38-44."""
    return _phi_k(z, 2)


def phi3(z) -> torch.Tensor:
    """jit_kernels.py:47-53."""
    return _phi_k(z, 3)


def _combine(terms) -> torch.Tensor:
    """sum_i ifft(m_i * fft(x_i)).real for (x_i, m_i) pairs, m_i device tables."""
    out = None
    for x, m in terms:
        t = as_device(x)
        t = t.to(field_dtype(t))
        y = apply_spectral_filter_interior(t, m)
        out = y if out is None else axpby(1.0, out, 1.0, y)
    return out


def etd1_kernel_1d(u, N, eigenvalues, dt: float) -> torch.Tensor:
    """jit_kernels.py:60-88: ifft(e^z fft(u) + dt phi1(z) fft(N)).real, z = dt lam."""
    e, p1 = _tables(_lam_dev(eigenvalues), dt, 1, 2)
    return _combine([(u, e), (N, p1)])


etd1_kernel_2d = etd1_kernel_1d  # jit_kernels.py:91-107 (2-D eigenvalues, fft2)


def etd2_kernel_1d(u, N_curr, N_prev, eigenvalues, dt: float) -> torch.Tensor:
    """jit_kernels.py:110-135: e^z u + dt phi1 N_n + dt phi2 (N_n - N_{n-1})."""
    e, a, b = _tables(_lam_dev(eigenvalues), dt, 2, 3)
    return _combine([(u, e), (N_curr, a), (N_prev, b)])


def etdrk4_kernel_1d(u, eigenvalues, dt: float, N_func_vals: Tuple) -> torch.Tensor:
    """jit_kernels.py:138-180: e^z u + dt (b1 N_a + b2 N_b + b3 N_c + b4 N_d),
    b1 = phi1 - 3 phi2 + 4 phi3, b2 = b3 = 2 phi2 - 4 phi3,
    b4 = -phi2 + 4 phi3 (of z = dt lam)."""
    e, b1, b2, b4 = _tables(_lam_dev(eigenvalues), dt, 3, 4)
    return _combine([(u, e)] + [(n, bk) for n, bk in zip(N_func_vals, (b1, b2, b2, b4))])


def helmholtz_solve_1d(rhs, laplacian_symbol, dt: float, D: float) -> torch.Tensor:
    """jit_kernels.py:183-194 (no clamp; 1 - dt D lam >= 1 for lam <= 0)."""
    return solve_helmholtz_1d(rhs, laplacian_symbol, dt, D)


def helmholtz_solve_2d(rhs, laplacian_symbol, dt: float, D: float) -> torch.Tensor:
    """jit_kernels.py:197-208."""
    return solve_helmholtz_2d(rhs, laplacian_symbol, dt, D)


def batched_fft_solve_1d(rhs_stack, laplacian_symbol, dt: float, D: float) -> torch.Tensor:
    """jit_kernels.py:229-250: (n_fields, nx) lines sharing one symbol."""
    return solve_helmholtz_1d(rhs_stack, laplacian_symbol, dt, D)


def batched_exp_matvec_1d(u_stack, eigenvalues, dt: float) -> torch.Tensor:
    """jit_kernels.py:253-266: ifft(e^{dt lam} fft(u)).real per line."""
    return _combine([(u_stack, _tables(_lam_dev(eigenvalues), dt, 1, 2)[0])])


def batched_helmholtz_solve_2d(rhs_stack, laplacian_symbol, dt: float, D: float) -> torch.Tensor:
    """jit_kernels.py:511-531: (n_fields, ny, nx) sharing one symbol."""
    return solve_helmholtz_2d(rhs_stack, laplacian_symbol, dt, D)


batched_etd1_kernel_1d = etd1_kernel_1d  # jit_kernels.py:357-385 (stacked lines)
batched_etd2_kernel_1d = etd2_kernel_1d  # jit_kernels.py:388-419
batched_etd1_kernel_2d = etd1_kernel_1d  # jit_kernels.py:481-508 (stacked 2-D fields)


def multi_operator_etd1_1d(u_stack, N_stack, eigenvalues_stack, dt: float) -> torch.Tensor:
    """jit_kernels.py:445-478: ETD1 with each line's own eigenvalues."""
    lam = _lam_dev(eigenvalues_stack)
    u = as_device(u_stack)
    N = as_device(N_stack, u.dtype)
    return torch.stack([etd1_kernel_1d(u[i], N[i], lam[i], dt) for i in range(lam.shape[0])])


def make_etd1_integrator(eigenvalues, nonlinear_fn: Callable):
    """jit_kernels.py:269-306: integrate(u0, t0, dt, n_steps) ->
    (u_final, t_final) by n_steps ETD1 steps with N = nonlinear_fn(u)."""
    lam = _lam_dev(eigenvalues)

    def integrate(u0, t0: float, dt: float, n_steps: int):
        u = as_device(u0)
        u = u.to(field_dtype(u))
        t = float(t0)
        e, p1 = _tables(lam, dt, 1, 2)  # one table set per call (fixed dt)
        for _ in range(int(n_steps)):
            u = _combine([(u, e), (nonlinear_fn(u), p1)])
            t += float(dt)
        return u, t

    return integrate
