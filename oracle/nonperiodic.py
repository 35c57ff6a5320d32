"""DST-I / DCT-II solves for homogeneous Dirichlet / Neumann boundaries,
float64 NumPy restatement (test infrastructure).

Follows moljax/core/fft_nonperiodic.py: dst_I_fast via the odd extension and
a complex FFT (:97-122), idst_I_fast = dst * 2/(N+1) (:125-133), the
eigenvalue formulas (:140-194) and the Poisson/Helmholtz solves (:247-355),
which use the orthonormal DCT-II of jax.scipy.fft (scipy.fft.dct(type=2,
norm='ortho') has the same definition).  The 2-D forms apply the 1-D
transforms along both axes (axis 0 = y).
Pinned by: scipy.fft.dst(type=1) / 2 equals dst_I_fast; sine / cosine modes
are eigenvectors; the 5-point stencil residual of a solve vanishes
(tests/test_oracle_golden.py).

Cell-centred Dirichlet (the ghost = 2 g - interior fill of
moljax/core/bc.py:237-258) is diagonalised by the DST-II (scipy.fft.dst
type 2, norm 'ortho') with eigenvalues -4/h^2 sin^2(pi k/(2N)), k = 1..N.
Inhomogeneous side values / fluxes (bc.py:218-307) enter through the lift b:
the share of the filled ghost ring in the 5-point Laplacian of the boundary
cells (2 g / h^2 for Dirichlet, -2 q / h left / bottom and +2 q / h right /
top for Neumann), so lap_h(fill(u)) = lap0(u) + b.  Pinned by the residual
of lap_h evaluated on oracle.fd.apply_bc's fill (tests/test_oracle_golden.py).
"""
import numpy as np
import scipy.fft as sf


def dst_I_fast(x, axis=-1):
    """X[k] = sum_n x[n] sin(pi (k+1)(n+1)/(N+1)) via the odd extension."""
    x = np.moveaxis(np.asarray(x, dtype=np.float64), axis, -1)
    N = x.shape[-1]
    z = np.zeros(x.shape[:-1] + (1,))
    ext = np.concatenate([z, x, z, -x[..., ::-1]], axis=-1)
    out = -0.5 * np.imag(np.fft.fft(ext, axis=-1)[..., 1:N + 1])
    return np.moveaxis(out, -1, axis)


def idst_I_fast(X, axis=-1):
    N = np.shape(X)[axis]
    return dst_I_fast(X, axis) * 2.0 / (N + 1)


def dct2_ortho(x, axis=-1):
    return sf.dct(np.asarray(x, dtype=np.float64), type=2, norm="ortho", axis=axis)


def idct2_ortho(X, axis=-1):
    return sf.idct(np.asarray(X, dtype=np.float64), type=2, norm="ortho", axis=axis)


def laplacian_symbol_dirichlet(N, dx):
    """fft_nonperiodic.py:140-162: -4/dx^2 sin^2(pi k/(2(N+1))), k=1..N."""
    k = np.arange(1, N + 1, dtype=np.float64)
    return -4.0 / (dx * dx) * np.sin(np.pi * k / (2 * (N + 1))) ** 2


def laplacian_symbol_neumann(N, dx):
    """fft_nonperiodic.py:165-183: -4/dx^2 sin^2(pi k/(2N)), k=0..N-1."""
    k = np.arange(N, dtype=np.float64)
    return -4.0 / (dx * dx) * np.sin(np.pi * k / (2 * N)) ** 2


def solve_helmholtz_dirichlet(rhs, lam, dt, D):
    """fft_nonperiodic.py:308-330 (no clamp on the denominator)."""
    return idst_I_fast(dst_I_fast(rhs) / (1.0 - dt * D * lam))


def solve_helmholtz_neumann(rhs, lam, dt, D):
    """fft_nonperiodic.py:333-355."""
    return idct2_ortho(dct2_ortho(rhs) / (1.0 - dt * D * lam))


def solve_poisson_dirichlet(rhs, lam):
    """fft_nonperiodic.py:247-263."""
    return idst_I_fast(dst_I_fast(rhs) / lam)


def solve_poisson_neumann(rhs, lam):
    """fft_nonperiodic.py:266-305: zero-mean solution (lam_0 = 0 -> 0)."""
    r = dct2_ortho(rhs)
    safe = np.where(np.abs(lam) < 1e-14, 1.0, lam)
    return idct2_ortho(np.where(np.abs(lam) < 1e-14, 0.0, r / safe))


def dst2_ortho(x, axis=-1):
    return sf.dst(np.asarray(x, dtype=np.float64), type=2, norm="ortho", axis=axis)


def idst2_ortho(X, axis=-1):
    return sf.idst(np.asarray(X, dtype=np.float64), type=2, norm="ortho", axis=axis)


def laplacian_symbol_dirichlet_cc(N, dx):
    """Cell-centred Dirichlet: -4/dx^2 sin^2(pi k/(2N)), k = 1..N."""
    k = np.arange(1, N + 1, dtype=np.float64)
    return -4.0 / (dx * dx) * np.sin(np.pi * k / (2 * N)) ** 2


def solve_helmholtz_dirichlet_cc(rhs, lam, dt, D):
    return idst2_ortho(dst2_ortho(rhs) / (1.0 - dt * D * lam))


def bc_lift(shape, kind, values, dy, dx):
    """b with lap_h(fill(u)) = lap0(u) + b on the interior cells; kind
    'dirichlet_cc' or 'neumann', values = (left, right, bottom, top)."""
    b = np.zeros(shape)
    vl, vr, vb, vt = values
    if kind == "dirichlet_cc":
        b[..., 0] += 2.0 * vl / dx ** 2
        b[..., -1] += 2.0 * vr / dx ** 2
        if len(shape) > 1:
            b[..., 0, :] += 2.0 * vb / dy ** 2
            b[..., -1, :] += 2.0 * vt / dy ** 2
    else:
        b[..., 0] += -2.0 * vl / dx
        b[..., -1] += 2.0 * vr / dx
        if len(shape) > 1:
            b[..., 0, :] += -2.0 * vb / dy
            b[..., -1, :] += 2.0 * vt / dy
    return b


def solve_bc_2d(kind, rhs, values, dy, dx, beta=None):
    """u with u - beta lap_h u = rhs (beta given) or lap_h u = rhs (beta
    None, Poisson) where lap_h acts on u padded by bc.py's fill; kind
    'dirichlet_cc' or 'neumann'."""
    ny, nx = rhs.shape[-2:]
    if kind == "dirichlet_cc":
        ly, lx = laplacian_symbol_dirichlet_cc(ny, dy), laplacian_symbol_dirichlet_cc(nx, dx)
    else:
        ly, lx = laplacian_symbol_neumann(ny, dy), laplacian_symbol_neumann(nx, dx)
    b = bc_lift(rhs.shape, kind, values, dy, dx)
    if beta is None:
        return solve_poisson_2d(kind, rhs - b, ly, lx)
    return solve_helmholtz_2d(kind, rhs + beta * b, ly, lx, beta, 1.0)


def _sep(lam_y, lam_x):
    return lam_y[:, None] + lam_x[None, :]


def solve_helmholtz_2d(kind, rhs, lam_y, lam_x, dt, D):
    lam = _sep(lam_y, lam_x)
    if kind == "dirichlet_cc":
        hat = dst2_ortho(dst2_ortho(rhs, -1), -2)
        return idst2_ortho(idst2_ortho(hat / (1.0 - dt * D * lam), -2), -1)
    if kind == "dirichlet":
        hat = dst_I_fast(dst_I_fast(rhs, -1), -2)
        return idst_I_fast(idst_I_fast(hat / (1.0 - dt * D * lam), -2), -1)
    hat = dct2_ortho(dct2_ortho(rhs, -1), -2)
    return idct2_ortho(idct2_ortho(hat / (1.0 - dt * D * lam), -2), -1)


def solve_poisson_2d(kind, rhs, lam_y, lam_x):
    lam = _sep(lam_y, lam_x)
    safe = np.where(np.abs(lam) < 1e-14, 1.0, lam)
    if kind == "dirichlet_cc":
        hat = dst2_ortho(dst2_ortho(rhs, -1), -2)
        return idst2_ortho(idst2_ortho(hat / lam, -2), -1)
    if kind == "dirichlet":
        hat = dst_I_fast(dst_I_fast(rhs, -1), -2)
        return idst_I_fast(idst_I_fast(hat / safe, -2), -1)
    hat = dct2_ortho(dct2_ortho(rhs, -1), -2)
    hat = np.where(np.abs(lam) < 1e-14, 0.0, hat / safe)
    return idct2_ortho(idct2_ortho(hat, -2), -1)


def dirichlet_laplacian_2d(u, dy, dx):
    """5-point Laplacian on interior vertices with zero boundary values."""
    p = np.pad(u, 1)
    return ((p[2:, 1:-1] - 2 * u + p[:-2, 1:-1]) / dy ** 2 + (p[1:-1, 2:] - 2 * u + p[1:-1, :-2]) / dx ** 2)


def neumann_laplacian_2d(u, dy, dx):
    """5-point Laplacian on cells with mirrored (zero-flux) ghosts."""
    p = np.pad(u, 1, mode="edge")
    return ((p[2:, 1:-1] - 2 * u + p[:-2, 1:-1]) / dy ** 2 + (p[1:-1, 2:] - 2 * u + p[1:-1, :-2]) / dx ** 2)


def _phi1_quadratic(z):
    """fft_nonperiodic.py:384-389 (the quadratic series below |z| = 1e-4)."""
    z = np.asarray(z, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        direct = (np.exp(z) - 1.0) / z
    return np.where(np.abs(z) < 1e-4, 1.0 + z / 2.0 + z ** 2 / 6.0, direct)


def etd1_dirichlet(u, N_u, eig, dt):
    """fft_nonperiodic.py:363-398."""
    z = dt * eig
    return idst_I_fast(np.exp(z) * dst_I_fast(u) + dt * _phi1_quadratic(z) * dst_I_fast(N_u))


def etd1_neumann(u, N_u, eig, dt):
    """fft_nonperiodic.py:401-437."""
    z = dt * eig
    return idct2_ortho(np.exp(z) * dct2_ortho(u) + dt * _phi1_quadratic(z) * dct2_ortho(N_u))
