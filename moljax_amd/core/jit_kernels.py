"""The reference's jitted kernels (moljax/core/jit_kernels.py) on the MI355X
spectral multiply.

Each kernel is a sum of real spectral multipliers applied to its inputs:
ifft(m * fft(x)).real along the last axis (1-D eigenvalues) or the last two
axes (2-D eigenvalues), which apply_spectral_filter_interior runs as one
native spectral round trip per input (the multiplier's even part, which is
what the real part applies).  The multiplier tables (e^z, dt phi_k(z), the
ETDRK4 weights) are evaluated in FP64 on the host from the eigenvalues with
the reference's formulas, series below |z| = 1e-4.  Real eigenvalues only
(the complex advection-diffusion solve is not on this path).
"""
from typing import Callable, Tuple

import numpy as np
import torch

from ._arrays import as_device, field_dtype
from .fft_solvers import apply_spectral_filter_interior, solve_helmholtz_1d, solve_helmholtz_2d
from .grid import default_device
from .state import axpby


def _host(z) -> np.ndarray:
    if isinstance(z, torch.Tensor):
        if z.is_complex():
            raise NotImplementedError("complex eigenvalues are not on the MI355X path")
        return z.detach().double().cpu().numpy()
    a = np.asarray(z)
    if np.iscomplexobj(a):
        raise NotImplementedError("complex eigenvalues are not on the MI355X path")
    return a.astype(np.float64)


def _phi_np(k: int, z: np.ndarray) -> np.ndarray:
    with np.errstate(divide="ignore", invalid="ignore"):
        ez = np.exp(z)
        if k == 1:
            taylor, direct = 1.0 + z / 2.0 + z ** 2 / 6.0 + z ** 3 / 24.0, (ez - 1.0) / z
        elif k == 2:
            taylor, direct = 0.5 + z / 6.0 + z ** 2 / 24.0 + z ** 3 / 120.0, (ez - 1.0 - z) / (z * z)
        else:
            taylor = 1.0 / 6.0 + z / 24.0 + z ** 2 / 120.0 + z ** 3 / 720.0
            direct = (ez - 1.0 - z - z ** 2 / 2.0) / z ** 3
    return np.where(np.abs(z) < 1e-4, taylor, direct)


def _dev(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).to(default_device())


def phi1(z) -> torch.Tensor:
    """jit_kernels.py:29-35: (e^z - 1)/z, 1 + z/2 + z^2/6 + z^3/24 below 1e-4."""
    return _dev(_phi_np(1, _host(z)))


def phi2(z) -> torch.Tensor:
    """jit_kernels.py:38-44."""
    return _dev(_phi_np(2, _host(z)))


def phi3(z) -> torch.Tensor:
    """jit_kernels.py:47-53."""
    return _dev(_phi_np(3, _host(z)))


def _combine(terms) -> torch.Tensor:
    """sum_i ifft(m_i * fft(x_i)).real for (x_i, m_i) pairs, m_i host tables."""
    out = None
    for x, m in terms:
        t = as_device(x)
        t = t.to(field_dtype(t))
        y = apply_spectral_filter_interior(t, _dev(m))
        out = y if out is None else axpby(1.0, out, 1.0, y)
    return out


def etd1_kernel_1d(u, N, eigenvalues, dt: float) -> torch.Tensor:
    """jit_kernels.py:60-88: ifft(e^z fft(u) + dt phi1(z) fft(N)).real, z = dt lam."""
    z = float(dt) * _host(eigenvalues)
    return _combine([(u, np.exp(z)), (N, float(dt) * _phi_np(1, z))])


etd1_kernel_2d = etd1_kernel_1d  # jit_kernels.py:91-107 (2-D eigenvalues, fft2)


def etd2_kernel_1d(u, N_curr, N_prev, eigenvalues, dt: float) -> torch.Tensor:
    """jit_kernels.py:110-135: e^z u + dt phi1 N_n + dt phi2 (N_n - N_{n-1})."""
    z = float(dt) * _host(eigenvalues)
    p1, p2 = float(dt) * _phi_np(1, z), float(dt) * _phi_np(2, z)
    return _combine([(u, np.exp(z)), (N_curr, p1 + p2), (N_prev, -p2)])


def etdrk4_kernel_1d(u, eigenvalues, dt: float, N_func_vals: Tuple) -> torch.Tensor:
    """jit_kernels.py:138-180: e^z u + dt (b1 N_a + b2 N_b + b3 N_c + b4 N_d),
    b1 = phi1 - 3 phi2 + 4 phi3, b2 = b3 = 2 phi2 - 4 phi3,
    b4 = -phi2 + 4 phi3 (of z = dt lam)."""
    z = float(dt) * _host(eigenvalues)
    f1, f2, f3 = _phi_np(1, z), _phi_np(2, z), _phi_np(3, z)
    d = float(dt)
    b = [f1 - 3 * f2 + 4 * f3, 2 * f2 - 4 * f3, 2 * f2 - 4 * f3, -f2 + 4 * f3]
    return _combine([(u, np.exp(z))] + [(n, d * bk) for n, bk in zip(N_func_vals, b)])


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
    return _combine([(u_stack, np.exp(float(dt) * _host(eigenvalues)))])


def batched_helmholtz_solve_2d(rhs_stack, laplacian_symbol, dt: float, D: float) -> torch.Tensor:
    """jit_kernels.py:511-531: (n_fields, ny, nx) sharing one symbol."""
    return solve_helmholtz_2d(rhs_stack, laplacian_symbol, dt, D)


batched_etd1_kernel_1d = etd1_kernel_1d  # jit_kernels.py:357-385 (stacked lines)
batched_etd2_kernel_1d = etd2_kernel_1d  # jit_kernels.py:388-419
batched_etd1_kernel_2d = etd1_kernel_1d  # jit_kernels.py:481-508 (stacked 2-D fields)


def multi_operator_etd1_1d(u_stack, N_stack, eigenvalues_stack, dt: float) -> torch.Tensor:
    """jit_kernels.py:445-478: ETD1 with each line's own eigenvalues."""
    lam = _host(eigenvalues_stack)
    u = as_device(u_stack)
    N = as_device(N_stack, u.dtype)
    return torch.stack([etd1_kernel_1d(u[i], N[i], lam[i], dt) for i in range(lam.shape[0])])


def make_etd1_integrator(eigenvalues, nonlinear_fn: Callable):
    """jit_kernels.py:269-306: integrate(u0, t0, dt, n_steps) ->
    (u_final, t_final) by n_steps ETD1 steps with N = nonlinear_fn(u)."""
    lam = _host(eigenvalues)

    def integrate(u0, t0: float, dt: float, n_steps: int):
        u = as_device(u0)
        u = u.to(field_dtype(u))
        t = float(t0)
        for _ in range(int(n_steps)):
            u = etd1_kernel_1d(u, nonlinear_fn(u), lam, dt)
            t += float(dt)
        return u, t

    return integrate
