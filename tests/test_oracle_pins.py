"""Closed-form pins for the oracle functions that the reference's own files
leave open: the GMRES iteration-count convention (the reference reports no
true count, moljax/core/newton_krylov.py:83-132,128,322), the upwind first
derivative (moljax/core/operators.py:200-268) and the implicit dt policy
with the implicit adaptive loop (moljax/core/dt_policy.py:279-321,324-408;
moljax/core/stepping.py:839-990).

Every expected value below is fixed by mathematics (Krylov dimension,
exact difference quotients, saturated controller factors), written as a
literal or a one-line closed form, never as a copy of the oracle's code.
A plausible mistake in the oracle -- an off-by-one in the count, counting
restart cycles or residual checks as iterations, the upwind direction
flipped at v = 0, the robustness growth applied after non-convergence --
fails one of them."""
import math

import numpy as np
import pytest

from oracle import adaptive as oad
from oracle import fd as ofd
from oracle import newton_krylov as onk


# ------------------------------------------------------------ GMRES counts
def _diag_problem(values, mult=3, seed=0, zero=()):
    """A diagonal matrix with the given eigenvalues (each repeated ``mult``
    times) and a positive right-hand side that is zero on the eigenvalues
    listed in ``zero``."""
    rng = np.random.default_rng(seed)
    d = np.repeat(np.asarray(values, dtype=float), mult)
    b = rng.uniform(0.5, 1.5, d.size)
    for z in zero:
        b[d == z] = 0.0
    return d, b


def test_gmres_one_iteration_for_the_identity_and_an_exact_inverse():
    """M A = I: the first Krylov vector spans the solution, so exactly one
    Arnoldi step (one matvec) is counted."""
    rng = np.random.default_rng(1)
    b = rng.standard_normal(17)
    x, it, conv, _ = onk.gmres(lambda v: v, b, rtol=1e-12, restart=30)
    assert (it, conv) == (1, True) and np.allclose(x, b, rtol=1e-14)
    Q = rng.standard_normal((20, 20))
    A = Q @ Q.T + 20 * np.eye(20)
    Ainv = np.linalg.inv(A)
    b = rng.standard_normal(20)
    x, it, conv, _ = onk.gmres(lambda v: A @ v, b, M=lambda v: Ainv @ v, rtol=1e-10, restart=30)
    assert (it, conv) == (1, True) and np.allclose(A @ x, b, rtol=1e-10)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6])
def test_gmres_counts_the_krylov_dimension(k):
    """A diagonal operator with k distinct eigenvalues and b touching all of
    them: the minimal polynomial has degree k, so the residual is nonzero
    after k - 1 steps and zero after k.  The count is exactly k."""
    d, b = _diag_problem(np.arange(1, k + 1), seed=k)
    x, it, conv, ratio = onk.gmres(lambda v: d * v, b, rtol=1e-10, restart=30)
    assert (it, conv) == (k, True)
    assert np.allclose(x, b / d, rtol=1e-10)
    # one step fewer leaves a residual far above the tolerance
    _, it2, conv2, ratio2 = onk.gmres(lambda v: d * v, b, rtol=1e-10, restart=k - 1 if k > 1 else 1, maxiter=1)
    if k > 1:
        assert (it2, conv2) == (k - 1, False) and ratio2 > 1e-6


def test_gmres_counts_only_the_eigenvalues_b_touches():
    d, b = _diag_problem([1.0, 2.0, 3.0, 4.0, 5.0, 6.0], zero=(2.0, 4.0, 6.0), seed=3)
    _, it, conv, _ = onk.gmres(lambda v: d * v, b, rtol=1e-10, restart=30)
    assert (it, conv) == (3, True)


def test_gmres_restart_cycles_count_every_arnoldi_step():
    """rtol = 0 is never met, so every cycle runs its m Arnoldi steps:
    exactly maxiter * m iterations; a restart length equal to the Krylov
    dimension converges inside the first cycle at exactly k (not k + 1)."""
    d, b = _diag_problem(np.arange(1, 9), seed=5)
    for m, cycles in ((3, 4), (2, 5), (1, 7)):
        _, it, conv, _ = onk.gmres(lambda v: d * v, b, rtol=0.0, restart=m, maxiter=cycles)
        assert (it, conv) == (m * cycles, False)
    d, b = _diag_problem([1.0, 3.0, 7.0, 9.0], seed=6)
    _, it, conv, _ = onk.gmres(lambda v: d * v, b, rtol=1e-10, restart=4, maxiter=3)
    assert (it, conv) == (4, True)


def test_gmres_zero_rhs_takes_no_iteration():
    x, it, conv, _ = onk.gmres(lambda v: 2 * v, np.zeros(5), rtol=1e-8)
    assert (it, conv) == (0, True) and np.all(x == 0)


def _mode_sum(n, modes, amps):
    h = 1.0 / (n + 1)
    x = h * np.arange(1, n + 1)
    X, Y = np.meshgrid(x, x)
    return sum(a * np.sin(np.pi * p * X) * np.sin(np.pi * q * Y) for a, (p, q) in zip(amps, modes))


def _dirichlet_eig(n, p, q, h):
    return -4.0 / h ** 2 * (np.sin(np.pi * p / (2 * (n + 1))) ** 2 + np.sin(np.pi * q / (2 * (n + 1))) ** 2)


@pytest.mark.parametrize("modes", [[(1, 1)], [(1, 2), (3, 1)], [(1, 1), (2, 3), (4, 4), (5, 1)]])
def test_newton_krylov_linear_problem_counts_modes(modes):
    """u_t = D lap u + a1 u, backward Euler, exact DST preconditioner: M J =
    I - dt a1 M has one eigenvalue per distinct Laplacian eigenvalue, and the
    first Newton residual lies in the span of the IC's sine modes, so the
    first GMRES solve takes exactly len(modes) iterations.  The Newton update
    is exact: each mode is divided by (1 - dt a1) - dt D lambda."""
    n, D, a1, dt = 31, 0.05, -3.0, 0.1
    h = 1.0 / (n + 1)
    amps = [1.0, -0.7, 0.4, 0.25][:len(modes)]
    u = _mode_sum(n, modes, amps)
    res = onk.be_step(u, dt, D, (0.0, 0.0, a1, 0.0), h, h, "dirichlet", precond="dst", max_newton=1,
                      krylov_tol=1e-10)
    assert res.newton_iters == 1 and res.gmres_iters == len(modes)
    exact = sum(a / ((1 - dt * a1) - dt * D * _dirichlet_eig(n, p, q, h)) *
                _mode_sum(n, [(p, q)], [1.0]) for a, (p, q) in zip(amps, modes))
    assert np.allclose(res.u, exact, rtol=1e-9, atol=1e-11)


def test_newton_krylov_pure_diffusion_one_gmres_per_newton():
    """a1 = 0: the exact preconditioner inverts J, so every GMRES solve takes
    one iteration; Newton tests convergence before its update, so a linear
    problem reports two Newton iterations (solve, then the converged pass)."""
    n, D, dt = 31, 0.05, 0.1
    h = 1.0 / (n + 1)
    u = onk.sine_mode_ic(n, seed=2)
    res = onk.be_step(u, dt, D, (0.0, 0.0, 0.0, 0.0), h, h, "dirichlet", precond="dst")
    assert (res.newton_iters, res.gmres_iters, res.converged) == (2, 2, True)
    assert res.history == [1, 1]


# ------------------------------------------------------------ upwind
def test_upwind_linear_field_exact_slope_any_velocity_sign():
    """f = a x + b y + c: backward and forward quotients both equal the
    slope exactly (up to rounding), whatever the velocity sign."""
    ny, nx, dy, dx = 12, 17, 0.05, 0.03
    yy, xx = np.meshgrid(dy * (np.arange(ny + 2) - 0.5), dx * (np.arange(nx + 2) - 0.5), indexing="ij")
    f = 2.5 * xx - 1.25 * yy + 0.3
    vel = np.sin(7 * xx + 3 * yy)  # changes sign over the grid
    for axis, slope in ((0, -1.25), (1, 2.5)):
        out = ofd.d1_upwind_2d(f, vel, dy, dx, axis)
        assert np.allclose(out[1:-1, 1:-1], slope, rtol=1e-12, atol=1e-12)
        assert np.all(out[0] == 0) and np.all(out[:, 0] == 0) and np.all(out[-1] == 0) and np.all(out[:, -1] == 0)


def test_upwind_quadratic_picks_backward_for_nonnegative_velocity():
    """f = x^2: backward quotient 2x - h, forward 2x + h.  Where v >= 0
    (v = 0 included) the result is 2x - h; where v < 0 it is 2x + h."""
    ny, nx, dy, dx = 6, 9, 0.1, 0.2
    yy, xx = np.meshgrid(dy * np.arange(ny + 2), dx * np.arange(nx + 2), indexing="ij")
    vel = np.where((np.arange(nx + 2) % 3) == 0, 0.0, np.where(np.arange(nx + 2) % 3 == 1, 1.0, -1.0))
    vel = np.broadcast_to(vel, (ny + 2, nx + 2)).copy()
    out = ofd.d1_upwind_2d(xx ** 2, vel, dy, dx, 1)[1:-1, 1:-1]
    x = xx[1:-1, 1:-1]
    v = vel[1:-1, 1:-1]
    want = np.where(v >= 0, 2 * x - dx, 2 * x + dx)
    assert np.allclose(out, want, rtol=1e-12, atol=1e-12)
    # the same along y with scalar velocities of either sign and zero
    g = yy ** 2
    y = yy[1:-1, 1:-1]
    assert np.allclose(ofd.d1_upwind_2d(g, 0.0, dy, dx, 0)[1:-1, 1:-1], 2 * y - dy, atol=1e-12)
    assert np.allclose(ofd.d1_upwind_2d(g, 0.4, dy, dx, 0)[1:-1, 1:-1], 2 * y - dy, atol=1e-12)
    assert np.allclose(ofd.d1_upwind_2d(g, -0.4, dy, dx, 0)[1:-1, 1:-1], 2 * y + dy, atol=1e-12)


# ------------------------------------------------------------ implicit dt policy
def test_implicit_robustness_dt_rules():
    """dt_policy.py:279-321: grow x1.2 (capped at dt_max = 1) when converged
    in <= 3 Newton iterations, keep when converged slower, halve when not
    converged (no growth then), clip to [1e-12, 1]."""
    assert oad.implicit_robustness_dt(0.5, True, 3) == pytest.approx(0.6, rel=1e-15)
    assert oad.implicit_robustness_dt(0.5, True, 0) == pytest.approx(0.6, rel=1e-15)
    assert oad.implicit_robustness_dt(0.5, True, 4) == 0.5
    assert oad.implicit_robustness_dt(0.5, False, 1) == 0.25
    assert oad.implicit_robustness_dt(0.5, False, 9) == 0.25
    assert oad.implicit_robustness_dt(0.9, True, 2) == 1.0
    assert oad.implicit_robustness_dt(3.0, True, 7) == 1.0
    assert oad.implicit_robustness_dt(1e-13, False, 1) == 1e-12
    assert oad.implicit_robustness_dt(2e-12, True, 1, dt_max=5e-12) == pytest.approx(2.4e-12, rel=1e-15)


def _uniform(n, c=1.0):
    return np.full((n, n), c)


@pytest.mark.parametrize("method", ["be", "cn", "bdf2"])
def test_implicit_adaptive_saturated_growth_sequence(method):
    """y' = lam y with lam = -1e-12 and D = 0: the error ratio is tiny, so
    the PI factor saturates at max_factor 5 and the Newton-count rule adds
    x1.2 -- dt_k = min(dt0 * 6^k, 1) with the last step landing on t_end,
    and no rejections."""
    n, dt0, t_end = 8, 1e-3, 3.0
    h = 1.0 / (n + 1)
    res = oad.adaptive_integrate_implicit(_uniform(n), 0.0, t_end, dt0, 0.0, (0.0, 0.0, -1e-12, 0.0), h, h,
                                          "dirichlet", method=method, precond="dst", max_steps=100)
    want, t = [], 0.0
    dt = dt0
    while t < t_end:
        step = min(dt, t_end - t)
        want.append(step)
        t += step
        dt = min(dt * 6.0, 1.0)
    got = [a[1] for a in res.attempts]
    assert res.n_rejected == 0 and res.status == oad.SUCCESS
    assert len(got) == len(want) == 7  # 1e-3, 6e-3, 0.036, 0.216, 1, 1, remainder 0.741
    assert np.allclose(got, want, rtol=1e-12)
    assert res.t_final == pytest.approx(t_end, rel=1e-15)
    assert np.allclose(res.dt_history[:len(want)], want, rtol=1e-12)


def test_implicit_adaptive_saturated_rejection():
    """A stiff decay lam = -1e6 from dt0 = 1 with backward Euler: the error
    ratio is ~1e6, so handle_rejected_step clips at min_factor 0.2, the
    Newton-count rule still grows x1.2 (Newton converged fast) and the reject
    halves: the second attempt uses 0.2 * 1.2 * 0.5 = 0.12."""
    n = 4
    h = 1.0 / (n + 1)
    res = oad.adaptive_integrate_implicit(_uniform(n), 0.0, 10.0, 1.0, 0.0, (0.0, 0.0, -1e6, 0.0), h, h,
                                          "dirichlet", method="be", precond="dst", max_steps=50, max_attempts=3)
    assert [a[3] for a in res.attempts[:2]] == [False, False]
    assert res.attempts[0][1] == 1.0
    assert res.attempts[1][1] == pytest.approx(0.12, rel=1e-14)
    assert res.attempts[2][1] == pytest.approx(0.12 * 0.12, rel=1e-14)
    assert math.isclose(res.attempts[0][2], 1e6 / (1 + 1e6) / (1e-6 + 1e-4 / (1 + 1e6)), rel_tol=1e-9)
