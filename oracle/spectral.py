"""Periodic spectral solves, float64 NumPy restatement (test infrastructure).

Follows moljax/core/fft_solvers.py (library conventions: cell-centred grid,
axis 0 = y, FD symbol) and the FFT-CN benchmark family
(benchmarks/benchmark_vs_scipy.py:74-114, benchmarks/sisc/bench_bc_matrix.py:44-95:
vertex grid x_i = i/N, indexing='ij', pseudo-spectral symbol -(kx^2+ky^2)).
"""
import numpy as np


def build_wavenumbers_2d(ny, nx, dy, dx):
    """fft_solvers.py:67-94 -- kx varies along axis 1, ky along axis 0."""
    kx_1d = 2.0 * np.pi * np.fft.fftfreq(nx, d=dx)
    ky_1d = 2.0 * np.pi * np.fft.fftfreq(ny, d=dy)
    kx = np.broadcast_to(kx_1d, (ny, nx)).astype(np.float64)
    ky = np.broadcast_to(ky_1d[:, None], (ny, nx)).astype(np.float64)
    return kx, ky


def laplacian_symbol_1d(nx, dx):
    """fft_solvers.py:97-121."""
    k = 2.0 * np.pi * np.fft.fftfreq(nx, d=dx)
    return (2.0 * np.cos(k * dx) - 2.0) / (dx * dx)


def laplacian_symbol_2d(ny, nx, dy, dx):
    """fft_solvers.py:124-154: (2cos(kx dx)-2)/dx^2 + (2cos(ky dy)-2)/dy^2."""
    kx, ky = build_wavenumbers_2d(ny, nx, dy, dx)
    lam_x = (2.0 * np.cos(kx * dx) - 2.0) / (dx * dx)
    lam_y = (2.0 * np.cos(ky * dy) - 2.0) / (dy * dy)
    return lam_x + lam_y


_FFT_WORKERS = [None]


def set_fft_workers(workers=None):
    """None: numpy.fft complex transforms, the reference's dataflow (the
    default).  An int: scipy.fft real-to-complex transforms with that many
    threads, the same arithmetic up to rounding (a faster CPU baseline)."""
    _FFT_WORKERS[0] = workers


def solve_helmholtz_2d(rhs, lam, dt, D):
    """fft_solvers.py:261-294: full complex fft2, clamped divide, ifft2.real."""
    denom = 1.0 - dt * D * lam
    denom = np.where(np.abs(denom) < 1e-14, 1e-14, denom)
    w = _FFT_WORKERS[0]
    if w is None:
        return np.fft.ifft2(np.fft.fft2(rhs) / denom).real
    import scipy.fft as sf
    h = rhs.shape[-1] // 2 + 1
    return sf.irfft2(sf.rfft2(rhs, workers=w) / denom[..., :h], s=rhs.shape[-2:], workers=w)


def solve_helmholtz_1d(rhs, lam, dt, D):
    """fft_solvers.py:222-258."""
    denom = 1.0 - dt * D * lam
    denom = np.where(np.abs(denom) < 1e-14, 1e-14, denom)
    return np.fft.ifft(np.fft.fft(rhs) / denom).real


def apply_diffusion_inverse_fft(state, ng, dt, diffusivities, lam):
    """fft_solvers.py:406-445 with solve_diffusion_fft_field (:368-399):
    per field with D > 1e-14 solve on the interior and embed into a copy of
    the padded input (ghosts are kept from the input)."""
    out = {}
    for name, f in state.items():
        D = diffusivities.get(name, 0.0)
        if D > 1e-14:
            g = f.copy()
            g[ng:-ng, ng:-ng] = solve_helmholtz_2d(f[ng:-ng, ng:-ng], lam, dt, D)
            out[name] = g
        else:
            out[name] = f
    return out


def spectral_symbol_ij(N, L=1.0):
    """benchmark_vs_scipy.py:74-81: dx = L/N, meshgrid(kx, ky, indexing='ij'),
    lap_eig = -(KX^2 + KY^2)."""
    dx = L / N
    kx = np.fft.fftfreq(N, dx) * 2 * np.pi
    ky = np.fft.fftfreq(N, dx) * 2 * np.pi
    KX, KY = np.meshgrid(kx, ky, indexing="ij")
    return -(KX ** 2 + KY ** 2)


def cn_factor(lap_eig, D, dt):
    """benchmark_vs_scipy.py:81: G = (1 + dt D lam / 2) / (1 - dt D lam / 2)."""
    return (1.0 + 0.5 * dt * D * lap_eig) / (1.0 - 0.5 * dt * D * lap_eig)


def fft_cn_integrate(u0, D, dt, n_steps, L=1.0):
    """benchmark_vs_scipy.py:94-110 (solve_fft_cn_numpy): n_steps of
    fft2 -> * G -> ifft2.real."""
    G = cn_factor(spectral_symbol_ij(u0.shape[0], L), D, dt)
    u = u0.copy()
    for _ in range(n_steps):
        u = np.real(np.fft.ifft2(np.fft.fft2(u) * G))
    return u


def sine_mode_ij(N, L=1.0):
    """benchmark_vs_scipy.py:63-66: x = linspace(0, 1, N, endpoint=False),
    indexing='ij', u0 = sin(2 pi x) sin(2 pi y)."""
    x = np.linspace(0, L, N, endpoint=False)
    X, Y = np.meshgrid(x, x, indexing="ij")
    return np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y)


def periodic_fd_laplacian(u, dy, dx):
    """Periodic 5-point Laplacian on an interior-only array (axis 0 = y)."""
    return ((np.roll(u, -1, 0) - 2 * u + np.roll(u, 1, 0)) / dy ** 2
            + (np.roll(u, -1, 1) - 2 * u + np.roll(u, 1, 1)) / dx ** 2)


def exponential_filter_1d(nx, strength=36.0, order=36):
    """fft_solvers.py:477-503: exp(-strength (min(k, nx-k) / (nx//2))^order),
    the normalised index clamped to 1."""
    k = np.arange(nx, dtype=np.float64)
    kn = np.minimum(np.minimum(k, nx - k) / (nx // 2), 1.0)
    return np.exp(-strength * kn ** order)


def exponential_filter_2d(ny, nx, strength=36.0, order=36):
    """fft_solvers.py:506-531: outer(sigma_y, sigma_x)."""
    return np.outer(exponential_filter_1d(ny, strength, order), exponential_filter_1d(nx, strength, order))


def apply_spectral_filter_interior(u, sigma):
    """fft_solvers.py:534-555: ifft(fft(u) sigma).real (1-D) or
    ifft2(fft2(u) sigma).real (2-D, the last two axes)."""
    if sigma.ndim == 1:
        return np.fft.ifft(np.fft.fft(u, axis=-1) * sigma, axis=-1).real
    return np.fft.ifft2(np.fft.fft2(u) * sigma).real
