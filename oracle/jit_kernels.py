"""NumPy float64 restatement of moljax/core/jit_kernels.py (test
infrastructure only; the product path never imports it).  Full complex FFTs
and the real part, as the reference computes them."""
import numpy as np


def _phi(k, z):
    z = np.asarray(z, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        ez = np.exp(z)
        if k == 1:
            t, d = 1.0 + z / 2.0 + z ** 2 / 6.0 + z ** 3 / 24.0, (ez - 1.0) / z
        elif k == 2:
            t, d = 0.5 + z / 6.0 + z ** 2 / 24.0 + z ** 3 / 120.0, (ez - 1.0 - z) / (z * z)
        else:
            t, d = 1.0 / 6.0 + z / 24.0 + z ** 2 / 120.0 + z ** 3 / 720.0, (ez - 1.0 - z - z ** 2 / 2.0) / z ** 3
    return np.where(np.abs(z) < 1e-4, t, d)


def phi1(z):
    """jit_kernels.py:29-35."""
    return _phi(1, z)


def phi2(z):
    """jit_kernels.py:38-44."""
    return _phi(2, z)


def phi3(z):
    """jit_kernels.py:47-53."""
    return _phi(3, z)


def _fft(x, two_d):
    return np.fft.fft2(x) if two_d else np.fft.fft(x, axis=-1)


def _ifft(x, two_d):
    return (np.fft.ifft2(x) if two_d else np.fft.ifft(x, axis=-1)).real


def etd1_kernel(u, N, lam, dt):
    """jit_kernels.py:60-107 (1-D: lam (nx,); 2-D: lam (ny, nx))."""
    two = np.ndim(lam) == 2
    z = dt * lam
    return _ifft(np.exp(z) * _fft(u, two) + dt * phi1(z) * _fft(N, two), two)


def etd2_kernel_1d(u, Nc, Np, lam, dt):
    """jit_kernels.py:110-135."""
    z = dt * lam
    Nch, Nph = _fft(Nc, False), _fft(Np, False)
    return _ifft(np.exp(z) * _fft(u, False) + dt * phi1(z) * Nch + dt * phi2(z) * (Nch - Nph), False)


def etdrk4_kernel_1d(u, lam, dt, Ns):
    """jit_kernels.py:138-180."""
    z = dt * lam
    f1, f2, f3 = phi1(z), phi2(z), phi3(z)
    b = [f1 - 3 * f2 + 4 * f3, 2 * f2 - 4 * f3, 2 * f2 - 4 * f3, -f2 + 4 * f3]
    acc = np.exp(z) * _fft(u, False)
    for bk, n in zip(b, Ns):
        acc = acc + dt * bk * _fft(n, False)
    return _ifft(acc, False)


def helmholtz_solve(rhs, lam, dt, D):
    """jit_kernels.py:183-208 (no clamp)."""
    two = np.ndim(lam) == 2
    return _ifft(_fft(rhs, two) / (1.0 - dt * D * lam), two)


def batched_exp_matvec_1d(u, lam, dt):
    """jit_kernels.py:253-266."""
    return _ifft(np.exp(dt * lam) * _fft(u, False), False)
