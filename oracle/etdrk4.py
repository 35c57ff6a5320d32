"""Coupled Gray-Scott ETDRK4 with contour-integral coefficients, float64 /
complex128 NumPy restatement (test infrastructure).

Follows benchmarks/benchmark_gray_scott.py:336-417 (Kassam-Trefethen
coefficients with M = 32 points on the upper unit semicircle, four stages,
nine FFTs per field per step), with the symbol passed as 1-D tables
lam(kx, ky) = lamx[kx] + lamy[ky] (axis 0 = y) so that both the benchmark's
pseudo-spectral -k^2 and the library FD symbol can be used.
The benchmark's final combination multiplies (N_a + N_b) by the
Kassam-Trefethen mean f2 = dt(phi2 - 2 phi3) without Kassam-Trefethen's factor
2 (benchmark_gray_scott.py:414-415), so its weights sum to 2/3 dt for a
constant N and the scheme is not consistent.  ``cox_matthews=True`` (the
default) applies the Cox-Matthews weight 2 f2 = dt(2 phi2 - 4 phi3), the same
b2 as the library's etdrk4_step (moljax/core/fft_integrators.py:110-115,
330); ``cox_matthews=False`` restates the benchmark as written.
Pinned by the z -> 0 limits (Q -> dt/2, f1, f2, f3 -> dt/6), the direct
phi-function formulas at moderate |z|, linear exactness (N = 0 gives E u) and
fourth-order convergence of the Cox-Matthews form (tests/test_oracle_golden.py).

Exponential Euler (ETD1) and exponential Adams-Bashforth-2 (ETD2) follow
moljax/core/fft_integrators.py:37-64 (phi1, phi2 with the cubic series for
|z| < 1e-4) and :119-249 (one nonlinear evaluation of the whole state per
step, so u and v couple; ETD2 without N_prev is ETD1).  Pinned by the
series / expm1 forms of phi1, phi2, linear exactness, the reduction to
forward Euler / Adams-Bashforth-2 on spatially constant fields (lam = 0),
and first / second-order convergence (tests/test_oracle_golden.py).

The benchmark's own splitting (benchmark_gray_scott.py:268-289, reaction
half-step / exact diffusion / reaction half-step) is restated as
benchmark_strang; pinned by pure diffusion (F = k = 0, v = 0: e^{n dt D lam}
exactly) and, on spatially constant fields, by two forward-Euler half-steps
per step.  Its RK4 (:89-127) and preconditioned Crank-Nicolson iteration
(:162-229) are benchmark_rk4_step and benchmark_cn_step; pinned by the RK4
and Crank-Nicolson amplification factors of each mode under pure diffusion
(the CN iteration stops after its second update there) and by classical
RK4 of the reaction ODE on constant fields.
"""
import numpy as np


def contour_coeffs(L, dt, M=32):
    """benchmark_gray_scott.py:336-356; returns real (E, E2, Q, f1, f2, f3)."""
    r = np.exp(1j * np.pi * (np.arange(1, M + 1) - 0.5) / M)
    LR = dt * L[..., np.newaxis] + r
    E = np.exp(dt * L)
    E2 = np.exp(dt * L / 2)
    Q = dt * np.mean((np.exp(LR / 2) - 1) / LR, axis=-1)
    f1 = dt * np.mean((-4 - LR + np.exp(LR) * (4 - 3 * LR + LR ** 2)) / LR ** 3, axis=-1)
    f2 = dt * np.mean((2 + LR + np.exp(LR) * (-2 + LR)) / LR ** 3, axis=-1)
    f3 = dt * np.mean((-4 - 3 * LR - LR ** 2 + np.exp(LR) * (4 - LR)) / LR ** 3, axis=-1)
    return E, E2, Q.real, f1.real, f2.real, f3.real


def spectral_symbol_1d(n, L=1.0):
    """-(2 pi fftfreq(n, L/n))^2 (benchmark_gray_scott.py:52-57)."""
    k = np.fft.fftfreq(n, L / n) * 2 * np.pi
    return -(k ** 2)


def gs_coeffs(lamx, lamy, Du, Dv, dt):
    lam = lamy[:, None] + lamx[None, :]
    return contour_coeffs(Du * lam, dt), contour_coeffs(Dv * lam, dt)


def gs_etdrk4_step(u, v, cu, cv, F, k, cox_matthews=True):
    """benchmark_gray_scott.py:369-417 (see the module note on f2)."""
    Eu, E2u, Qu, f1u, f2u, f3u = cu
    Ev, E2v, Qv, f1v, f2v, f3v = cv
    if cox_matthews:
        f2u, f2v = 2 * f2u, 2 * f2v

    def Nu(a, b):
        return -a * b * b + F * (1 - a)

    def Nv(a, b):
        return a * b * b - (F + k) * b

    fft2, ifft2 = np.fft.fft2, np.fft.ifft2
    uh, vh = fft2(u), fft2(v)
    Nnu, Nnv = fft2(Nu(u, v)), fft2(Nv(u, v))
    au = E2u * uh + Qu * Nnu
    av = E2v * vh + Qv * Nnv
    ar, br_ = np.real(ifft2(au)), np.real(ifft2(av))
    Nau, Nav = fft2(Nu(ar, br_)), fft2(Nv(ar, br_))
    bu = E2u * uh + Qu * Nau
    bv = E2v * vh + Qv * Nav
    bur, bvr = np.real(ifft2(bu)), np.real(ifft2(bv))
    Nbu, Nbv = fft2(Nu(bur, bvr)), fft2(Nv(bur, bvr))
    cu_ = E2u * au + Qu * (2 * Nbu - Nnu)
    cv_ = E2v * av + Qv * (2 * Nbv - Nnv)
    cur, cvr = np.real(ifft2(cu_)), np.real(ifft2(cv_))
    Ncu, Ncv = fft2(Nu(cur, cvr)), fft2(Nv(cur, cvr))
    un = Eu * uh + f1u * Nnu + f2u * (Nau + Nbu) + f3u * Ncu
    vn = Ev * vh + f1v * Nnv + f2v * (Nav + Nbv) + f3v * Ncv
    return np.real(ifft2(un)), np.real(ifft2(vn))


def gs_etdrk4(u, v, lamx, lamy, Du, Dv, F, k, dt, n_steps, cox_matthews=True):
    cu, cv = gs_coeffs(lamx, lamy, Du, Dv, dt)
    for _ in range(n_steps):
        u, v = gs_etdrk4_step(u, v, cu, cv, F, k, cox_matthews)
    return u, v


def phi1(z):
    """fft_integrators.py:37-50."""
    z = np.asarray(z, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        direct = (np.exp(z) - 1.0) / z
    return np.where(np.abs(z) < 1e-4, 1.0 + z / 2.0 + z ** 2 / 6.0 + z ** 3 / 24.0, direct)


def phi2(z):
    """fft_integrators.py:53-64."""
    z = np.asarray(z, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        direct = (np.exp(z) - 1.0 - z) / (z * z)
    return np.where(np.abs(z) < 1e-4, 0.5 + z / 6.0 + z ** 2 / 24.0 + z ** 3 / 120.0, direct)


def gs_nonlinear(u, v, F, k):
    return -u * v * v + F * (1 - u), u * v * v - (F + k) * v


def gs_etd1_step(u, v, lamx, lamy, Du, Dv, F, k, dt):
    """fft_integrators.py:119-174 for the coupled Gray-Scott state."""
    lam = lamy[:, None] + lamx[None, :]
    out = []
    Nu, Nv = gs_nonlinear(u, v, F, k)
    for f, N, D in ((u, Nu, Du), (v, Nv, Dv)):
        z = dt * D * lam
        out.append(np.real(np.fft.ifft2(np.exp(z) * np.fft.fft2(f) + dt * phi1(z) * np.fft.fft2(N))))
    return out[0], out[1]


def gs_etd2_step(u, v, lamx, lamy, Du, Dv, F, k, dt, N_prev=None):
    """fft_integrators.py:177-249; returns (u, v, (Nu, Nv) of the input)."""
    N_curr = gs_nonlinear(u, v, F, k)
    if N_prev is None:
        un, vn = gs_etd1_step(u, v, lamx, lamy, Du, Dv, F, k, dt)
        return un, vn, N_curr
    lam = lamy[:, None] + lamx[None, :]
    out = []
    for f, Nn, Np, D in ((u, N_curr[0], N_prev[0], Du), (v, N_curr[1], N_prev[1], Dv)):
        z = dt * D * lam
        Nh, Nph = np.fft.fft2(Nn), np.fft.fft2(Np)
        uh = np.exp(z) * np.fft.fft2(f) + dt * phi1(z) * Nh + dt * phi2(z) * (Nh - Nph)
        out.append(np.real(np.fft.ifft2(uh)))
    return out[0], out[1], N_curr


def gs_etd(u, v, lamx, lamy, Du, Dv, F, k, dt, n_steps, order, N_prev=None):
    """n_steps of ETD1 (order 1) or ETD2 (order 2, N_prev carried as in
    fft_integrators.py:376-387); returns (u, v, N_prev after the last step)."""
    for _ in range(n_steps):
        if order == 1:
            N_prev = gs_nonlinear(u, v, F, k)
            u, v = gs_etd1_step(u, v, lamx, lamy, Du, Dv, F, k, dt)
        else:
            u, v, N_prev = gs_etd2_step(u, v, lamx, lamy, Du, Dv, F, k, dt, N_prev)
    return u, v, N_prev


def benchmark_strang_step(u, v, lamx, lamy, Du, Dv, F, k, dt):
    """benchmark_gray_scott.py:268-289: forward-Euler reaction half-step,
    exact diffusion e^{dt D lam} by FFT, forward-Euler reaction half-step
    (lam(kx, ky) = lamx[kx] + lamy[ky], axis 0 = y)."""
    lam = lamy[:, None] + lamx[None, :]

    def half(a, b):
        abb = a * b * b
        return a + 0.5 * dt * (-abb + F * (1 - a)), b + 0.5 * dt * (abb - (F + k) * b)

    u, v = half(u, v)
    u = np.real(np.fft.ifft2(np.fft.fft2(u) * np.exp(dt * Du * lam)))
    v = np.real(np.fft.ifft2(np.fft.fft2(v) * np.exp(dt * Dv * lam)))
    return half(u, v)


def benchmark_strang(u, v, lamx, lamy, Du, Dv, F, k, dt, n_steps):
    for _ in range(n_steps):
        u, v = benchmark_strang_step(u, v, lamx, lamy, Du, Dv, F, k, dt)
    return u, v


def _bench_lap(f, lam):
    return np.real(np.fft.ifft2(lam * np.fft.fft2(f)))


def benchmark_rk4_step(u, v, lamx, lamy, Du, Dv, F, k, dt):
    """benchmark_gray_scott.py:89-127: classical RK4 of y' = D lap y + R(y)
    with the pseudo-spectral Laplacian."""
    lam = lamy[:, None] + lamx[None, :]

    def rhs(a, b):
        abb = a * b * b
        return (Du * _bench_lap(a, lam) - abb + F * (1 - a), Dv * _bench_lap(b, lam) + abb - (F + k) * b)

    k1u, k1v = rhs(u, v)
    k2u, k2v = rhs(u + 0.5 * dt * k1u, v + 0.5 * dt * k1v)
    k3u, k3v = rhs(u + 0.5 * dt * k2u, v + 0.5 * dt * k2v)
    k4u, k4v = rhs(u + dt * k3u, v + dt * k3v)
    return (u + dt / 6 * (k1u + 2 * k2u + 2 * k3u + k4u), v + dt / 6 * (k1v + 2 * k2v + 2 * k3v + k4v))


def benchmark_cn_step(u_old, v_old, lamx, lamy, Du, Dv, F, k, dt, tol=1e-8, max_iter=10):
    """benchmark_gray_scott.py:162-229: the preconditioned Crank-Nicolson
    iteration; returns (u, v, iterations)."""
    lam = lamy[:, None] + lamx[None, :]
    pu = 1.0 / (1.0 - 0.5 * dt * Du * lam)
    pv = 1.0 / (1.0 - 0.5 * dt * Dv * lam)

    def R(a, b):
        abb = a * b * b
        return -abb + F * (1 - a), abb - (F + k) * b

    Ruo, Rvo = R(u_old, v_old)
    u, v = u_old, v_old
    res_norm, i = 1.0, 0
    while res_norm > tol and i < max_iter:
        Ru, Rv = R(u, v)
        res_u = u - u_old - 0.5 * dt * (Du * (_bench_lap(u, lam) + _bench_lap(u_old, lam)) + Ru + Ruo)
        res_v = v - v_old - 0.5 * dt * (Dv * (_bench_lap(v, lam) + _bench_lap(v_old, lam)) + Rv + Rvo)
        u = u + np.real(np.fft.ifft2(pu * np.fft.fft2(-res_u)))
        v = v + np.real(np.fft.ifft2(pv * np.fft.fft2(-res_v)))
        res_norm = np.max(np.abs(res_u)) + np.max(np.abs(res_v))
        i += 1
    return u, v, i
