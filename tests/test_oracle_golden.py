"""Pins the CPU oracle against the reference's own golden numbers and against
closed forms (no GPU).  A plausible mistake in the oracle (wrong symbol sign
or axis, dropped Heun term, swapped half-steps) fails one of these."""
import json
import math
import os

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from oracle import fd, gray_scott, spectral

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


# ------------------------------------------------------------- FFT-CN
def test_fft_cn_scipy_comparison_golden():
    """benchmarks/results/scipy_comparison.json:18-22 (NumPy FFT-CN, 128^2,
    D=0.01, dt=1e-3, 1000 steps, max error vs exp(-8 pi^2 D t) u0)."""
    g = _golden("fft_cn_golden.json")["scipy_comparison"]
    u0 = spectral.sine_mode_ij(g["N"])
    u = spectral.fft_cn_integrate(u0, g["D"], g["dt"], g["n_steps"])
    exact = np.exp(-8 * np.pi ** 2 * g["D"] * g["dt"] * g["n_steps"]) * u0
    err = np.max(np.abs(u - exact))
    assert err == pytest.approx(g["numpy_error"], rel=1e-9)


@pytest.mark.parametrize("N", [64, 128, 256])
def test_fft_cn_bc_matrix_golden(N):
    """benchmarks/results/bc_matrix.json:8,28,48 (periodic FFT-CN, D=0.1,
    T=0.1, 100 steps; JAX-GPU rounding, hence rel 1e-7)."""
    g = _golden("fft_cn_golden.json")["bc_matrix"]
    u0 = spectral.sine_mode_ij(N)
    u = spectral.fft_cn_integrate(u0, g["D"], g["T"] / g["n_steps"], g["n_steps"])
    exact = np.exp(-8 * np.pi ** 2 * g["D"] * g["T"]) * u0
    assert np.max(np.abs(u - exact)) == pytest.approx(g["errors"][str(N)], rel=1e-7)


def test_fft_cn_config1_closed_form():
    """128^2, D=0.1, dt=0.01, 100 steps: equals G^n u0 to 1e-12 (100 FFT
    round trips of rounding) and the
    analytic decay to the CN time error (4.1e-3 relative L2)."""
    N, D, dt, n = 128, 0.1, 0.01, 100
    u0 = spectral.sine_mode_ij(N)
    u = spectral.fft_cn_integrate(u0, D, dt, n)
    lam = -8 * np.pi ** 2  # continuous eigenvalue of the (1,1) mode
    G = (1 + 0.5 * dt * D * lam) / (1 - 0.5 * dt * D * lam)
    assert gray_scott.rel_l2(u, G ** n * u0) < 1e-12
    r = gray_scott.rel_l2(u, np.exp(lam * D * dt * n) * u0)
    assert 3e-3 < r < 5e-3


# ------------------------------------------------------------- symbols
def test_fd_symbol_is_stencil_eigenvalue():
    """lam(kx, ky) are exactly the eigenvalues of the periodic 5-point
    Laplacian (axis 0 = y): FFT(L_h u) = lam * FFT(u)."""
    rng = np.random.default_rng(0)
    ny, nx, dy, dx = 32, 48, 0.3, 0.7
    u = rng.standard_normal((ny, nx))
    lam = spectral.laplacian_symbol_2d(ny, nx, dy, dx)
    lhs = np.fft.fft2(spectral.periodic_fd_laplacian(u, dy, dx))
    assert np.max(np.abs(lhs - lam * np.fft.fft2(u))) < 1e-10 * np.max(np.abs(lhs))


def test_helmholtz_fd_residual_and_identity():
    rng = np.random.default_rng(1)
    ny, nx, dy, dx, dt, D = 64, 32, 2.5 / 64, 2.5 / 32, 0.7, 0.013
    lam = spectral.laplacian_symbol_2d(ny, nx, dy, dx)
    rhs = rng.standard_normal((ny, nx))
    u = spectral.solve_helmholtz_2d(rhs, lam, dt, D)
    res = u - dt * D * spectral.periodic_fd_laplacian(u, dy, dx) - rhs
    assert np.max(np.abs(res)) < 1e-12
    assert np.max(np.abs(spectral.solve_helmholtz_2d(rhs, lam, dt, 0.0) - rhs)) < 1e-14


# ------------------------------------------------------------- FD / BC
def test_fd_operators_on_sine_modes():
    """Discrete eigenvalues/derivatives of sin modes are exact."""
    ny, nx = 24, 40
    dy, dx = 1.0 / ny, 1.0 / nx
    y = (np.arange(-1, ny + 1) + 0.5) * dy
    x = (np.arange(-1, nx + 1) + 0.5) * dx
    X, Y = np.meshgrid(x, y, indexing="xy")
    f = np.sin(2 * np.pi * X) * np.sin(4 * np.pi * Y)
    lap = fd.laplacian_2d(f, dy, dx)
    ev = (2 * np.cos(2 * np.pi * dx) - 2) / dx ** 2 + (2 * np.cos(4 * np.pi * dy) - 2) / dy ** 2
    assert np.max(np.abs(lap[1:-1, 1:-1] - ev * f[1:-1, 1:-1])) < 1e-9 * abs(ev)
    assert np.all(lap[0] == 0) and np.all(lap[:, 0] == 0)
    d1 = fd.d1_central_2d(f, dy, dx, 1)
    exact = np.sin(2 * np.pi * dx) / dx * np.cos(2 * np.pi * X) * np.sin(4 * np.pi * Y)
    assert np.max(np.abs(d1[1:-1, 1:-1] - exact[1:-1, 1:-1])) < 1e-12


def test_bc_fill_periodic_matches_wrap():
    rng = np.random.default_rng(2)
    f = rng.standard_normal((10, 14))
    g = fd.apply_bc(f, 0)
    inner = g[1:-1, 1:-1]
    assert np.array_equal(g[0, 1:-1], inner[-1]) and np.array_equal(g[-1, 1:-1], inner[0])
    assert np.array_equal(g[1:-1, 0], inner[:, -1]) and g[0, 0] == inner[-1, -1]
    # Dirichlet: ghost + interior = 2 g at every boundary face
    d = fd.apply_bc(f, 1, (1.0, 2.0, 3.0, 4.0))
    assert np.allclose(d[1:-1, 0] + d[1:-1, 1], 2.0) and np.allclose(d[0, 1:-1] + d[1, 1:-1], 6.0)


# ------------------------------------------------------------- Gray-Scott
def _gs_grid(n):
    return 2.5 / n


def test_gs_fixed_point():
    n = 32
    lam = spectral.laplacian_symbol_2d(n, n, _gs_grid(n), _gs_grid(n))
    y = {"u": np.ones((n + 2, n + 2)), "v": np.zeros((n + 2, n + 2))}
    out, _ = gray_scott.integrate(y, 1.0, 5, 0.035, 0.065, 2e-5, 1e-5, lam)
    assert np.max(np.abs(out["u"] - 1)) == 0 and np.max(np.abs(out["v"])) == 0


def test_gs_pure_diffusion_closed_form():
    """F = k = 0 and v = 0: each Strang step is (1 - dt/2 Du lam)^-2 in
    Fourier space, so n steps equal one closed-form multiply."""
    n, dt, Du, steps = 32, 0.8, 3e-3, 7
    h = _gs_grid(n)
    lam = spectral.laplacian_symbol_2d(n, n, h, h)
    rng = np.random.default_rng(3)
    u0 = rng.standard_normal((n, n))
    U = np.zeros((n + 2, n + 2))
    U[1:-1, 1:-1] = u0
    y = {"u": gray_scott.apply_bc_periodic(U), "v": np.zeros((n + 2, n + 2))}
    out, _ = gray_scott.integrate(y, dt, steps, 0.0, 0.0, Du, 1e-5, lam)
    expect = np.fft.ifft2(np.fft.fft2(u0) * (1 - 0.5 * dt * Du * lam) ** (-2 * steps)).real
    assert np.max(np.abs(out["u"][1:-1, 1:-1] - expect)) < 1e-13


def test_gs_reaction_heun_second_order():
    """Spatially uniform state: the Strang step reduces to Heun on the GS
    reaction ODE; the error against a tight ODE solution drops 4x per halving
    of dt (a dropped Heun term would give first order)."""
    F, k, T = 0.035, 0.065, 4.0
    u0, v0 = 0.6, 0.3

    def rhs(t, y):
        u, v = y
        return [-u * v * v + F * (1 - u), u * v * v - (F + k) * v]

    ref = solve_ivp(rhs, (0, T), [u0, v0], rtol=1e-13, atol=1e-14, method="DOP853").y[:, -1]
    n = 16
    lam = spectral.laplacian_symbol_2d(n, n, 0.1, 0.1)
    errs = []
    for steps in (8, 16, 32):
        y = {"u": np.full((n + 2, n + 2), u0), "v": np.full((n + 2, n + 2), v0)}
        out, _ = gray_scott.integrate(y, T / steps, steps, F, k, 1e-3, 1e-3, lam)
        errs.append(abs(out["u"][5, 5] - ref[0]) + abs(out["v"][5, 5] - ref[1]))
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(1.8 < r < 2.2 for r in rates), rates


def test_gs_initial_condition_layout():
    y = gray_scott.initial_condition(64, 64, seed=0)
    u = y["u"]
    assert u.shape == (66, 66)
    inner = u[1:-1, 1:-1]
    assert abs(inner[32, 32] - 0.5) <= 0.01 and abs(inner[0, 0] - 1.0) <= 0.01
    assert np.array_equal(u[0, 1:-1], inner[-1])


def test_fixed_dt_step_count_quirk():
    """stepping.py:1402: int((t_end - t0)/dt) + 1 steps."""
    assert gray_scott.fixed_dt_n_steps(0.0, 10.0, 1.0) == 11
    assert gray_scott.fixed_dt_n_steps(0.0, 1.0, 0.3) == 4


# ------------------------------------------------------- DST-I / DCT-II
def test_dst_matches_library_definition():
    """dst_I_fast restates fft_nonperiodic.py:97-122; scipy's DST-I is the
    same sum times 2."""
    from oracle import nonperiodic as onp
    import scipy.fft as sf
    x = np.random.default_rng(4).standard_normal((3, 31))
    assert np.max(np.abs(onp.dst_I_fast(x) - sf.dst(x, type=1) / 2)) < 1e-13
    assert np.max(np.abs(onp.idst_I_fast(onp.dst_I_fast(x)) - x)) < 1e-13


@pytest.mark.parametrize("kind", ["dirichlet", "neumann"])
def test_nonperiodic_solves_fd_residual(kind):
    """The 2-D DST-I (vertex grid, zero boundary) and DCT-II (cell grid,
    mirrored ghosts) solves invert the 5-point operator exactly; the
    symbols are its eigenvalues."""
    from oracle import nonperiodic as onp
    rng = np.random.default_rng(5)
    if kind == "dirichlet":
        ny, nx = 31, 63
        dy, dx = 1.0 / (ny + 1), 2.0 / (nx + 1)
        ly, lx = onp.laplacian_symbol_dirichlet(ny, dy), onp.laplacian_symbol_dirichlet(nx, dx)
        lap = onp.dirichlet_laplacian_2d
    else:
        ny, nx = 32, 16
        dy, dx = 1.0 / ny, 0.5 / nx
        ly, lx = onp.laplacian_symbol_neumann(ny, dy), onp.laplacian_symbol_neumann(nx, dx)
        lap = onp.neumann_laplacian_2d
    rhs = rng.standard_normal((ny, nx))
    dt, D = 0.3, 0.02
    u = onp.solve_helmholtz_2d(kind, rhs, ly, lx, dt, D)
    assert np.max(np.abs(u - dt * D * lap(u, dy, dx) - rhs)) < 1e-12
    if kind == "dirichlet":
        p = onp.solve_poisson_2d(kind, rhs, ly, lx)
        assert np.max(np.abs(lap(p, dy, dx) - rhs)) < 1e-9
    else:
        r0 = rhs - rhs.mean()
        p = onp.solve_poisson_2d(kind, r0, ly, lx)
        assert np.max(np.abs(lap(p, dy, dx) - r0)) < 1e-9 and abs(p.mean()) < 1e-12


def test_dst2_brute_force_and_cell_dirichlet_eigenmodes():
    """DST-II (ortho) against its O(N^2) sum; the modes sin(pi k (j+1/2)/N)
    are eigenvectors of the 5-point Laplacian on cells padded by bc.py's
    Dirichlet fill (ghost = -interior) with eigenvalue -4/h^2 sin^2(pi k/(2N))."""
    from oracle import fd
    from oracle import nonperiodic as onp
    N, h = 16, 0.1
    x = np.random.default_rng(2).standard_normal(N)
    n = np.arange(N)
    brute = np.array([np.sum(x * np.sin(np.pi * (k + 1) * (2 * n + 1) / (2 * N))) for k in range(N)])
    w = np.where(n == N - 1, np.sqrt(1.0 / N), np.sqrt(2.0 / N))  # ortho weights (last one 1/sqrt(N))
    assert np.allclose(onp.dst2_ortho(x), w * brute, atol=1e-13)
    lam = onp.laplacian_symbol_dirichlet_cc(N, h)
    for k in (1, 5, N):
        mode = np.sin(np.pi * k * (n + 0.5) / N)
        f2 = np.tile(mode, (3, 1))  # constant along y: y-direction Neumann fill keeps it exact
        p = fd.apply_bc(np.pad(f2, 1), 1, dy=h, dx=h)
        p[0, :], p[-1, :] = p[1, :], p[-2, :]
        lap = (p[1:-1, 2:] - 2 * f2 + p[1:-1, :-2]) / h ** 2
        assert np.allclose(lap[1], lam[k - 1] * mode, atol=1e-10)


@pytest.mark.parametrize("kind", ["dirichlet_cc", "neumann"])
def test_bc_solves_invert_fd_operator_on_filled_ghosts(kind):
    """solve_bc_2d with nonzero side values: lap_h evaluated on u padded by
    oracle.fd.apply_bc (bc.py:218-307, pinned by bc_matrix.json) gives
    u - beta lap_h u = rhs and, for Poisson, lap_h u = rhs."""
    from oracle import fd
    from oracle import nonperiodic as onp
    rng = np.random.default_rng(9)
    ny, nx = 16, 32
    dy, dx = 1.0 / ny, 2.0 / nx
    values = (0.3, -0.7, 1.1, 0.25)
    rhs = rng.standard_normal((ny, nx))
    fk = 1 if kind == "dirichlet_cc" else 2

    def lap_h(u):
        return fd.laplacian_2d(fd.apply_bc(np.pad(u, 1), fk, values, dy, dx), dy, dx)[1:-1, 1:-1]

    beta = 0.04
    u = onp.solve_bc_2d(kind, rhs, values, dy, dx, beta)
    assert np.max(np.abs(u - beta * lap_h(u) - rhs)) < 1e-12
    r = rhs
    if kind == "neumann":  # compatibility: sum(rhs) = boundary flux total
        b = onp.bc_lift(rhs.shape, kind, values, dy, dx)
        r = rhs - (rhs - b).mean()
    p = onp.solve_bc_2d(kind, r, values, dy, dx)
    assert np.max(np.abs(lap_h(p) - r)) < 1e-9


def test_poisson_dirichlet_sine_mode():
    """test_fft_nonperiodic.py:90-106 intent: Poisson with f = -pi^2 sin(pi x)
    on a vertex grid gives sin(pi x) to second order."""
    from oracle import nonperiodic as onp
    N = 63
    dx = 1.0 / (N + 1)
    x = dx * np.arange(1, N + 1)
    u = onp.solve_poisson_dirichlet(-np.pi ** 2 * np.sin(np.pi * x), onp.laplacian_symbol_dirichlet(N, dx))
    assert np.max(np.abs(u - np.sin(np.pi * x))) < 5e-2 * 0.1


# ---------------------------------------------------------------- ETDRK4
def test_etdrk4_contour_limits_and_direct_phi():
    """Contour coefficients (Kassam-Trefethen means): z -> 0 limits Q -> dt/2,
    f1 = f2 = f3 -> dt/6; at |z| ~ 1..10 they equal the direct phi-function
    formulas f1 = dt(phi1 - 3 phi2 + 4 phi3), 2 f2 = dt(2 phi2 - 4 phi3) (the
    Cox-Matthews weight of N_a + N_b), f3 = dt(-phi2 + 4 phi3),
    Q = dt/2 phi1(z/2)."""
    from oracle import etdrk4 as oe
    dt = 0.7
    E, E2, Q, f1, f2, f3 = oe.contour_coeffs(np.array([0.0]), dt)
    assert np.allclose([Q[0], f1[0], f2[0], f3[0]], [dt / 2, dt / 6, dt / 6, dt / 6], rtol=1e-13)
    L = np.array([-1.0, -3.0, -7.0, -14.0]) / dt
    z = dt * L
    p1 = (np.exp(z) - 1) / z
    p2 = (np.exp(z) - 1 - z) / z ** 2
    p3 = (np.exp(z) - 1 - z - z ** 2 / 2) / z ** 3
    E, E2, Q, f1, f2, f3 = oe.contour_coeffs(L, dt)
    assert np.allclose(f1, dt * (p1 - 3 * p2 + 4 * p3), rtol=1e-10)
    assert np.allclose(2 * f2, dt * (2 * p2 - 4 * p3), rtol=1e-10)
    assert np.allclose(f3, dt * (-p2 + 4 * p3), rtol=1e-10)
    assert np.allclose(Q, dt / 2 * (np.exp(z / 2) - 1) / (z / 2), rtol=1e-12)


def test_etdrk4_linear_exact_and_fourth_order():
    from oracle import etdrk4 as oe
    n = 32
    lam = oe.spectral_symbol_1d(n)
    rng = np.random.default_rng(6)
    u0 = rng.standard_normal((n, n))
    # N = 0 (F = k = 0, v = 0): exact exponential
    u, v = oe.gs_etdrk4(u0, np.zeros((n, n)), lam, lam, 1e-3, 1e-3, 0.0, 0.0, 0.5, 3)
    expect = np.fft.ifft2(np.fft.fft2(u0) * np.exp(1.5 * 1e-3 * (lam[None, :] + lam[:, None]))).real
    assert np.max(np.abs(u - expect)) < 1e-13
    # smooth coupled problem: error vs a fine run drops ~16x per halving
    x = np.arange(n) / n
    X, Y = np.meshgrid(x, x)
    uu = 0.6 + 0.2 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)
    vv = 0.3 + 0.1 * np.cos(2 * np.pi * X)
    T = 8.0
    ref = oe.gs_etdrk4(uu, vv, lam, lam, 2e-3, 1e-3, 0.035, 0.065, T / 256, 256)
    errs = [np.max(np.abs(oe.gs_etdrk4(uu, vv, lam, lam, 2e-3, 1e-3, 0.035, 0.065, T / s, s)[0] - ref[0]))
            for s in (4, 8, 16)]
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(r > 3.5 for r in rates), rates
    # the benchmark's combination as written (f2 without the factor 2) is not
    # consistent: a constant N is integrated with weight 2/3 dt
    c = oe.contour_coeffs(np.array([0.0]), 0.3)
    assert abs(c[3][0] + 2 * c[4][0] + c[5][0] - 0.3 * 4 / 6) < 1e-14
    bad = [np.max(np.abs(oe.gs_etdrk4(uu, vv, lam, lam, 2e-3, 1e-3, 0.035, 0.065, T / s, s,
                                      cox_matthews=False)[0] - ref[0])) for s in (8, 16)]
    assert math.log2(bad[0] / bad[1]) < 1.5


def test_etd_phi_functions():
    """phi1, phi2 (fft_integrators.py:37-64): the cubic series below
    |z| = 1e-4 and the direct form above agree with expm1(z)/z and with the
    exponential series summed to 40 terms (closed forms independent of the
    restatement); phi(0) = 1, 1/2."""
    from oracle import etdrk4 as oe
    z = np.array([0.0, 3e-5, -9e-5, -2e-3, 0.4, -1.0, -7.5, -40.0])

    def series(z, j):
        return sum(z ** m / math.factorial(m + j) for m in range(40))

    p1, p2 = oe.phi1(z), oe.phi2(z)
    assert p1[0] == 1.0 and p2[0] == 0.5
    nz = z != 0
    assert np.allclose(p1[nz], np.expm1(z[nz]) / z[nz], rtol=1e-12)
    small = np.abs(z) < 1.0
    assert np.allclose(p1[small], series(z[small], 1), rtol=1e-12)
    # the direct phi2 loses digits to cancellation near |z| = 1e-4 (the
    # reference's formula): ~eps/z^2 relative
    assert np.allclose(p2[small], series(z[small], 2), rtol=1e-7)
    big = z[np.abs(z) >= 0.4]
    assert np.allclose(p2[np.abs(z) >= 0.4], (np.expm1(big) - big) / big ** 2, rtol=1e-12)


def test_etd1_etd2_reductions_and_order():
    """ETD1 / ETD2 (fft_integrators.py:119-249): exact for N = 0; on
    spatially constant fields (only lam = 0 is excited) ETD1 is forward Euler
    and ETD2 is Adams-Bashforth-2 on the Gray-Scott ODE; first / second order
    on a smooth coupled problem."""
    from oracle import etdrk4 as oe
    n = 16
    lam = oe.spectral_symbol_1d(n)
    rng = np.random.default_rng(8)
    u0 = rng.standard_normal((n, n))
    for order in (1, 2):
        u, v, _ = oe.gs_etd(u0, np.zeros((n, n)), lam, lam, 1e-3, 1e-3, 0.0, 0.0, 0.5, 3, order)
        expect = np.fft.ifft2(np.fft.fft2(u0) * np.exp(1.5 * 1e-3 * (lam[None, :] + lam[:, None]))).real
        assert np.max(np.abs(u - expect)) < 1e-13 and np.max(np.abs(v)) == 0.0
    F, k, dt, steps = 0.035, 0.065, 0.4, 6
    a, b = 0.7, 0.2

    def f(a, b):
        return -a * b * b + F * (1 - a), a * b * b - (F + k) * b

    ea, eb = a, b
    for _ in range(steps):
        fa, fb = f(ea, eb)
        ea, eb = ea + dt * fa, eb + dt * fb
    u, v, _ = oe.gs_etd(np.full((n, n), a), np.full((n, n), b), lam, lam, 2e-3, 1e-3, F, k, dt, steps, 1)
    assert np.max(np.abs(u - ea)) < 1e-14 and np.max(np.abs(v - eb)) < 1e-14
    qa, qb = f(a, b)
    ab = (a + dt * qa, b + dt * qb)  # the first step is ETD1 = Euler
    for _ in range(steps - 1):
        na, nb = f(*ab)
        ab = (ab[0] + dt * (1.5 * na - 0.5 * qa), ab[1] + dt * (1.5 * nb - 0.5 * qb))
        qa, qb = na, nb
    u, v, _ = oe.gs_etd(np.full((n, n), a), np.full((n, n), b), lam, lam, 2e-3, 1e-3, F, k, dt, steps, 2)
    assert np.max(np.abs(u - ab[0])) < 1e-14 and np.max(np.abs(v - ab[1])) < 1e-14
    x = np.arange(n) / n
    X, Y = np.meshgrid(x, x)
    uu = 0.6 + 0.2 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)
    vv = 0.3 + 0.1 * np.cos(2 * np.pi * X)
    T = 8.0
    ref = oe.gs_etdrk4(uu, vv, lam, lam, 2e-3, 1e-3, F, k, T / 256, 256)
    for order, lo in ((1, 0.85), (2, 1.8)):
        errs = [np.max(np.abs(oe.gs_etd(uu, vv, lam, lam, 2e-3, 1e-3, F, k, T / s, s, order)[0] - ref[0]))
                for s in (16, 32, 64)]
        rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
        assert all(lo < r < order + 0.5 for r in rates), (order, rates)


def test_benchmark_strang_pure_diffusion_and_constant_fields():
    """benchmark_gray_scott.py:268-289 restated: with F = k = 0 and v = 0 it
    is the exact heat semigroup; on constant fields (only lam = 0) each step
    is two forward-Euler half-steps of the reaction ODE."""
    from oracle import etdrk4 as oe
    n = 32
    lam = oe.spectral_symbol_1d(n)
    u0 = np.random.default_rng(3).standard_normal((n, n))
    u, v = oe.benchmark_strang(u0, np.zeros((n, n)), lam, lam, 2e-4, 1e-4, 0.0, 0.0, 0.5, 4)
    expect = np.fft.ifft2(np.fft.fft2(u0) * np.exp(2.0 * 2e-4 * (lam[None, :] + lam[:, None]))).real
    assert np.max(np.abs(u - expect)) < 1e-13 and np.max(np.abs(v)) == 0.0
    F, k, dt = 0.035, 0.065, 0.5
    a, b = 0.6, 0.3
    for _ in range(5):
        for _h in range(2):
            abb = a * b * b
            a, b = a + 0.25 * (-abb + F * (1 - a)), b + 0.25 * (abb - (F + k) * b)
    u, v = oe.benchmark_strang(np.full((n, n), 0.6), np.full((n, n), 0.3), lam, lam, 2e-5, 1e-5, F, k, dt, 5)
    assert np.max(np.abs(u - a)) < 1e-14 and np.max(np.abs(v - b)) < 1e-14


def test_benchmark_rk4_and_cn_closed_forms():
    """benchmark_gray_scott.py:89-127,162-229 restated.  Pure diffusion
    (F = k = 0, v = 0): RK4 multiplies each mode by 1 + z + z^2/2 + z^3/6 +
    z^4/24 and the exactly preconditioned CN iteration reaches
    (1 + z/2)/(1 - z/2) after one update (stopping after the second); on
    constant fields RK4 is classical RK4 of the reaction ODE."""
    from oracle import etdrk4 as oe
    n, D, dt = 32, 2e-4, 0.3
    lam = oe.spectral_symbol_1d(n)
    L = lam[:, None] + lam[None, :]
    z = dt * D * L
    u0 = np.random.default_rng(12).standard_normal((n, n))
    u, v = oe.benchmark_rk4_step(u0, np.zeros((n, n)), lam, lam, D, D, 0.0, 0.0, dt)
    g = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    assert np.max(np.abs(u - np.real(np.fft.ifft2(g * np.fft.fft2(u0))))) < 1e-13 and np.max(np.abs(v)) == 0
    u, v, it = oe.benchmark_cn_step(u0, np.zeros((n, n)), lam, lam, D, D, 0.0, 0.0, dt)
    cn = np.real(np.fft.ifft2((1 + z / 2) / (1 - z / 2) * np.fft.fft2(u0)))
    assert np.max(np.abs(u - cn)) < 1e-13 and it == 2
    F, k, h = 0.035, 0.065, 0.4

    def f(a, b):
        return -a * b * b + F * (1 - a), a * b * b - (F + k) * b

    a, b = 0.6, 0.3
    k1 = f(a, b)
    k2 = f(a + 0.5 * h * k1[0], b + 0.5 * h * k1[1])
    k3 = f(a + 0.5 * h * k2[0], b + 0.5 * h * k2[1])
    k4 = f(a + h * k3[0], b + h * k3[1])
    ea = a + h / 6 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0])
    u, v = oe.benchmark_rk4_step(np.full((n, n), a), np.full((n, n), b), lam, lam, 2e-5, 1e-5, F, k, h)
    assert np.max(np.abs(u - ea)) < 1e-14


def test_cpu_baseline_fft_variant_matches():
    """The scipy.fft R2C variant of the oracle's Helmholtz solve (a CPU
    baseline option) equals the numpy.fft complex dataflow to rounding."""
    from oracle import spectral as osp
    rhs = np.random.default_rng(4).standard_normal((32, 64))
    lam = osp.laplacian_symbol_2d(32, 64, 0.1, 0.05)
    a = osp.solve_helmholtz_2d(rhs, lam, 0.5, 0.02)
    osp.set_fft_workers(2)
    try:
        b = osp.solve_helmholtz_2d(rhs, lam, 0.5, 0.02)
    finally:
        osp.set_fft_workers(None)
    assert np.max(np.abs(a - b)) < 1e-14


def test_etd1_nonperiodic_modes():
    """fft_nonperiodic.py:363-437 restated: on an eigenmode of the stencil
    (DST-I sine on the vertex grid, DCT-II cosine on cells) the step is
    e^{z_k} u_k + dt phi1(z_k) N_k, with phi1 from expm1 (z_k = dt eig_k)."""
    from oracle import nonperiodic as onp
    N, dt, D = 31, 0.4, 0.03
    dx = 1.0 / (N + 1)
    j = np.arange(N)
    eig = D * onp.laplacian_symbol_dirichlet(N, dx)
    for k in (1, 7, N):
        mode = np.sin(np.pi * k * (j + 1) / (N + 1))
        z = dt * eig[k - 1]
        out = onp.etd1_dirichlet(2.0 * mode, -0.5 * mode, eig, dt)
        assert np.allclose(out, (2.0 * np.exp(z) - 0.5 * dt * np.expm1(z) / z) * mode, atol=1e-13)
    M = 32
    eig = D * onp.laplacian_symbol_neumann(M, 1.0 / M)
    jj = np.arange(M)
    for k in (0, 3, M - 1):
        mode = np.cos(np.pi * k * (jj + 0.5) / M)
        z = dt * eig[k]
        phi = 1.0 if k == 0 else np.expm1(z) / z
        out = onp.etd1_neumann(mode, mode, eig, dt)
        assert np.allclose(out, (np.exp(z) + dt * phi) * mode, atol=1e-13)


# --------------------------------------------------------- Newton-Krylov
def test_hand_jvp_matches_golden_ad_norm():
    """benchmarks/results/jvp_vs_fd_sweep.json:3: AD-JVP norm of the BE
    residual u - u_n - dt (D lap u + 10 u^2 (1-u)) on a periodic 64^2 grid
    (dx = 1/64, D = dt = 0.01, numpy legacy seed 42: u0 = |randn| + 0.5,
    v = randn); the hand JVP gives the same vector."""
    from oracle import newton_krylov as onk
    g = _golden("nk_golden.json")["jvp_vs_fd"]
    N = g["N"]
    np.random.seed(g["seed"])
    u0 = np.abs(np.random.randn(N, N)) + 0.5
    v = np.random.randn(N, N)
    dx = 1.0 / N
    K = g["reaction_K"]
    Jv = onk.jvp(u0, v, g["dt"], 1.0, g["D"], (-K, K, 0.0, 0.0), dx, dx, "periodic")
    assert np.linalg.norm(Jv) == pytest.approx(g["ad_jvp_norm"], rel=1e-13)
    # central finite differences at eps = 1e-5 agree to ~1e-10
    R = lambda u: onk.residual(u, u0, g["dt"], 1.0, g["D"], (-K, K, 0.0, 0.0), dx, dx, "periodic")
    eps = 1e-5
    fd = (R(u0 + eps * v) - R(u0 - eps * v)) / (2 * eps)
    assert np.linalg.norm(fd - Jv) / np.linalg.norm(Jv) < 1e-9


def test_gmres_matches_dense_solve_and_precond_exact():
    from oracle import newton_krylov as onk
    n = 15
    h = 1.0 / (n + 1)
    rng = np.random.default_rng(7)
    u = rng.standard_normal((n, n))
    b = rng.standard_normal((n, n))
    args = (0.05, 1.0, 0.02, (-1.0, 0.0, 1.0, 0.0), h, h, "dirichlet")
    A = lambda v: onk.jvp(u, v, *args)
    Ad = np.stack([A(e.reshape(n, n)).ravel() for e in np.eye(n * n)], axis=1)
    xd = np.linalg.solve(Ad, b.ravel()).reshape(n, n)
    M = lambda v: onk.helmholtz_precond(v, 0.05 * 0.02, h, h, "dirichlet")
    x, it, conv, res = onk.gmres(A, b, M, rtol=1e-10, restart=30, maxiter=20)
    assert conv and np.max(np.abs(x - xd)) < 1e-8 * np.max(np.abs(xd))
    x2, it2, conv2, _ = onk.gmres(A, b, None, rtol=1e-10, restart=10, maxiter=200)
    assert conv2 and np.max(np.abs(x2 - xd)) < 1e-7 * np.max(np.abs(xd)) and it2 > it
    r = rng.standard_normal((n, n))
    p = M(r)
    assert np.max(np.abs(p - 0.05 * 0.02 * onk.lap(p, h, h, "dirichlet") - r)) < 1e-12


def test_newton_converges_on_allen_cahn_step():
    from oracle import newton_krylov as onk
    n = 31
    h = 1.0 / (n + 1)
    u_old = onk.sine_mode_ic(n, seed=0)
    res = onk.newton_krylov(u_old, u_old, 1e-2, 1.0, 1e-3, (-1.0, 0.0, 1.0, 0.0), h, h, "dirichlet")
    assert res.converged and res.res_norm < 1e-8
    r = onk.residual(res.u, u_old, 1e-2, 1.0, 1e-3, (-1.0, 0.0, 1.0, 0.0), h, h, "dirichlet")
    assert np.linalg.norm(r) < 1e-8
    assert 0 < res.gmres_iters < 40


def test_bdf2_coefficients_and_order():
    """newton_krylov.py:406-448: at constant steps (omega = 1) the alphas are
    (3/2, -2, 1/2) and beta = 2 dt / 3.  With f(u) = lam u and D = 0 every
    point solves y' = lam y: the consistent form converges to exp(lam T) at
    order 2; the reference form is BDF2 for y' = lam y / alpha0 = 2/3 lam y
    and converges to exp(2/3 lam T), at first order because its
    backward-Euler start step follows the unscaled equation."""
    from oracle import newton_krylov as onk
    a0, a1, a2, beta = onk.bdf2_coefficients(0.1, 0.1)
    assert (a0, a1, a2) == (1.5, -2.0, 0.5) and beta == pytest.approx(0.2 / 3, rel=1e-15)
    assert onk.bdf2_coefficients(0.1, 0.1, consistent=True)[3] == 0.1
    lam, T = -2.0, 1.0

    def run(n, consistent):
        dt = T / n
        u = np.full((3, 3), 1.0)
        prev = u
        u = onk.be_step(u, dt, 0.0, (0.0, 0.0, lam, 0.0), 0.25, 0.25, "dirichlet", precond="none",
                        newton_tol=1e-14).u
        for _ in range(n - 1):
            nxt = onk.bdf2_step(u, prev, dt, dt, 0.0, (0.0, 0.0, lam, 0.0), 0.25, 0.25, "dirichlet",
                                precond="none", consistent=consistent, newton_tol=1e-14).u
            prev, u = u, nxt
        return u[1, 1]
    e1, e2 = abs(run(40, True) - np.exp(lam * T)), abs(run(80, True) - np.exp(lam * T))
    assert 3.0 < e1 / e2 < 5.0, (e1, e2)
    r1, r2 = run(40, False), run(80, False)
    e1, e2 = abs(r1 - np.exp(2.0 / 3.0 * lam * T)), abs(r2 - np.exp(2.0 / 3.0 * lam * T))
    assert 1.7 < e1 / e2 < 2.3 and e2 < 5e-3, (e1, e2)
    assert abs(r2 - np.exp(lam * T)) > 0.1


def test_spectral_filter_closed_forms():
    """Oracle spectral filters (fft_solvers.py:477-555) against closed forms:
    sigma(0) = 1, sigma(nx/2) = e^-strength, even in k, order 2 in closed
    form; the 2-D kernel is the outer product; a real Fourier mode comes back
    times the even part (sigma(k) + sigma(-k))/2 of any real kernel."""
    for nx in (16, 17, 50):
        s = spectral.exponential_filter_1d(nx, strength=3.0, order=2)
        k = np.arange(nx)
        kf = np.minimum(k, nx - k)
        assert s[0] == 1.0
        assert np.allclose(s, np.exp(-3.0 * np.minimum(kf / (nx // 2), 1.0) ** 2), rtol=0, atol=1e-15)
        assert np.array_equal(s[1:], s[1:][::-1])
        if nx % 2 == 0:
            assert s[nx // 2] == math.exp(-3.0)
    s2 = spectral.exponential_filter_2d(12, 20, 5.0, 4)
    assert np.allclose(s2, np.outer(spectral.exponential_filter_1d(12, 5.0, 4), spectral.exponential_filter_1d(20, 5.0, 4)))
    rng = np.random.default_rng(0)
    ny, nx, p, q = 12, 20, 3, 5
    y, x = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    mode = np.cos(2 * np.pi * (p * y / ny + q * x / nx))
    sig = rng.uniform(0.2, 1.0, size=(ny, nx))  # not even
    even = 0.5 * (sig[p, q] + sig[-p, -q])
    assert np.allclose(spectral.apply_spectral_filter_interior(mode, sig), even * mode, atol=1e-14)
    line = np.sin(2 * np.pi * 3 * np.arange(nx) / nx)
    s1 = rng.uniform(0.2, 1.0, size=nx)
    assert np.allclose(spectral.apply_spectral_filter_interior(line, s1), 0.5 * (s1[3] + s1[-3]) * line, atol=1e-14)


def test_jit_kernels_oracle_closed_forms():
    """oracle/jit_kernels.py pinned by closed forms: phi_k(0) = 1/k!, the
    recurrences phi_{k+1} = (phi_k - 1/k!)/z away from 0; ETD1 with N = 0
    multiplies a Fourier mode by e^{dt lam_k}, with lam = 0 gives u + dt N;
    ETD2 with N_prev = N_n is ETD1; ETDRK4 with four equal N is ETD1 (the
    weights sum to phi1); the Helmholtz solve inverts 1 - dt D lam on a
    mode."""
    from oracle import jit_kernels as oj
    z0 = np.array([0.0])
    assert oj.phi1(z0)[0] == 1.0 and oj.phi2(z0)[0] == 0.5 and abs(oj.phi3(z0)[0] - 1 / 6) < 1e-16
    z = np.array([-3.0, -0.5, 0.2, 1.3])
    assert np.allclose(oj.phi2(z), (oj.phi1(z) - 1.0) / z, rtol=1e-13)
    assert np.allclose(oj.phi3(z), (oj.phi2(z) - 0.5) / z, rtol=1e-12)
    nx, dt, q = 32, 0.3, 5
    x = np.arange(nx) / nx
    lam = -(2 * np.pi * np.fft.fftfreq(nx, 1.0 / nx)) ** 2 * 1e-3
    mode = np.cos(2 * np.pi * q * x)
    assert np.allclose(oj.etd1_kernel(mode, 0 * mode, lam, dt), np.exp(dt * lam[q]) * mode, atol=1e-14)
    N = np.sin(2 * np.pi * 3 * x) + 0.3
    assert np.allclose(oj.etd1_kernel(mode, N, 0 * lam, dt), mode + dt * N, atol=1e-14)
    assert np.allclose(oj.etd2_kernel_1d(mode, N, N, lam, dt), oj.etd1_kernel(mode, N, lam, dt), atol=1e-14)
    assert np.allclose(oj.etdrk4_kernel_1d(mode, lam, dt, (N, N, N, N)), oj.etd1_kernel(mode, N, lam, dt),
                       atol=1e-13)
    assert np.allclose(oj.helmholtz_solve(mode, lam, dt, 2.0), mode / (1 - 2.0 * dt * lam[q]), atol=1e-14)
