"""Newton-Krylov for implicit reaction-diffusion steps, float64 NumPy
restatement (test infrastructure).

Problem class: u_t = D lap_h u + f(u), f(u) = a3 u^3 + a2 u^2 + a1 u + a0, on
an interior (ny, nx) array with either homogeneous Dirichlet boundaries on a
vertex grid (zero outside) or periodic wrap.  Backward Euler / Crank-Nicolson
residual as moljax/core/newton_krylov.py:362-403 (create_implicit_residual):
    R(u) = u - u_old - theta dt F(u) - (1 - theta) dt F(u_old),  F = D lap + f,
with its Jacobian-vector product written out by hand:
    J v = v - theta dt (D lap v + f'(u) v).

Newton loop as newton_krylov.py:171-359: the convergence test uses the
residual norm before the update, three backtracking residual evaluations
(accept when ||r_new|| < (1 - 1e-4 alpha) ||r||, alpha halves otherwise),
fallback to the damped full step when none is accepted.

Krylov solve: left-preconditioned restarted GMRES(m) (the reference solves
M A x = M b with jax.scipy.sparse.linalg.gmres, newton_krylov.py:111-125).
Counting convention (fixed here because the reference reports no true
count): one iteration = one Arnoldi step (one matvec); the recursive Givens
residual |g_{j+1}| is compared with rtol * ||M b|| after every step;
maxiter counts restart cycles like JAX's gmres.  Orthogonalisation is
classical Gram-Schmidt applied twice (CGS2).

Pinned by: the hand JVP against central finite differences and against the
golden AD-JVP norm of benchmarks/results/jvp_vs_fd_sweep.json:3
(205.99961307732647, periodic roll Laplacian, f = 10 u^2 (1 - u)); GMRES
against a direct dense solve on small systems; Newton on problems with known
roots (tests/test_oracle_golden.py).
"""
import numpy as np

from . import nonperiodic as onp
from . import spectral as osp


def lap(u, dy, dx, bc):
    if bc == "dirichlet":
        return onp.dirichlet_laplacian_2d(u, dy, dx)
    return osp.periodic_fd_laplacian(u, dy, dx)


def reaction(u, coef):
    a3, a2, a1, a0 = coef
    return ((a3 * u + a2) * u + a1) * u + a0


def reaction_prime(u, coef):
    a3, a2, a1, _ = coef
    return (3 * a3 * u + 2 * a2) * u + a1


def rhs(u, D, coef, dy, dx, bc):
    return D * lap(u, dy, dx, bc) + reaction(u, coef)


def residual(u, u_old, dt, theta, D, coef, dy, dx, bc, F_old=None, lead=1.0):
    """lead * u - u_old - theta dt F(u) [- (1 - theta) dt F_old]."""
    r = lead * u - u_old - theta * dt * rhs(u, D, coef, dy, dx, bc)
    if theta < 1.0:
        r = r - (1.0 - theta) * dt * F_old
    return r


def jvp(u, v, dt, theta, D, coef, dy, dx, bc, lead=1.0):
    return lead * v - theta * dt * (D * lap(v, dy, dx, bc) + reaction_prime(u, coef) * v)


def bdf2_coefficients(dt, dt_prev, consistent=False):
    """newton_krylov.py:436-440 (variable-step BDF2, omega = dt/dt_prev):
    alpha0 = (1+2w)/(1+w), alpha1 = -(1+w), alpha2 = w^2/(1+w) and
    beta = dt (1+w)/(1+2w).  That beta is the one of the alpha0-normalised
    form while the alphas are not normalised, so the reference scheme
    integrates y' = F/alpha0 (2/3 F at constant steps); consistent=True
    gives beta = dt, the variable-step BDF2 of the docstring at :418."""
    omega = dt / dt_prev
    beta = dt if consistent else dt * (1.0 + omega) / (1.0 + 2.0 * omega)
    return (1.0 + 2.0 * omega) / (1.0 + omega), -(1.0 + omega), omega ** 2 / (1.0 + omega), beta


def helmholtz_precond(r, beta, dy, dx, bc):
    """(I - beta lap_h)^-1 exactly: DST-I (Dirichlet) or FFT (periodic)."""
    ny, nx = r.shape
    if bc == "dirichlet":
        ly, lx = onp.laplacian_symbol_dirichlet(ny, dy), onp.laplacian_symbol_dirichlet(nx, dx)
        return onp.solve_helmholtz_2d("dirichlet", r, ly, lx, 1.0, beta)
    lam = osp.laplacian_symbol_2d(ny, nx, dy, dx)
    return osp.solve_helmholtz_2d(r, lam, 1.0, beta)


def jacobi_precond(r, beta, dy, dx, bc, n_iter=5, omega=0.6667):
    """preconditioners.py:166-198 on interior arrays: x0 = r,
    x <- omega (r + beta lap x)/diag + (1 - omega) x, diag = 1 + beta(2/dx^2 + 2/dy^2)."""
    diag = 1.0 + beta * (2.0 / dx ** 2 + 2.0 / dy ** 2)
    x = r.copy()
    for _ in range(n_iter):
        x = omega * (r + beta * lap(x, dy, dx, bc)) / diag + (1.0 - omega) * x
    return x


def gmres(A, b, M=None, rtol=1e-8, restart=30, maxiter=50):
    """Returns (x, iterations, converged, recursive residual ratio)."""
    shape = np.shape(b)
    x, it, conv, ratio = _gmres_flat(lambda v: np.ravel(A(v.reshape(shape))), np.ravel(b),
                                     None if M is None else (lambda v: np.ravel(M(v.reshape(shape)))),
                                     rtol, restart, maxiter)
    return x.reshape(shape), it, conv, ratio


def _gmres_flat(A, b, M, rtol, restart, maxiter):
    Mf = (lambda v: v) if M is None else M
    x = np.zeros_like(b)
    bb = Mf(b)
    bnorm = np.linalg.norm(bb)
    if bnorm == 0.0:
        return x, 0, True, 0.0
    target = rtol * bnorm
    iters = 0
    res = bnorm
    for _ in range(maxiter):
        r = Mf(b - A(x)) if iters > 0 else bb
        beta = np.linalg.norm(r)
        if beta <= target:
            return x, iters, True, beta / bnorm
        V = [r / beta]
        H = np.zeros((restart + 1, restart))
        cs = np.zeros(restart)
        sn = np.zeros(restart)
        g = np.zeros(restart + 1)
        g[0] = beta
        k = 0
        for j in range(restart):
            w = Mf(A(V[j]))
            iters += 1
            Vm = np.array(V)
            h = Vm @ w
            w = w - Vm.T @ h
            h2 = Vm @ w
            w = w - Vm.T @ h2
            h = h + h2
            hn = np.linalg.norm(w)
            H[:j + 1, j] = h
            H[j + 1, j] = hn
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            den = np.hypot(H[j, j], H[j + 1, j])
            cs[j] = H[j, j] / den
            sn[j] = H[j + 1, j] / den
            H[j, j] = den
            H[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            res = abs(g[j + 1])
            k = j + 1
            if res <= target or hn == 0.0:
                break
            V.append(w / hn)
        y = np.zeros(k)
        for i in range(k - 1, -1, -1):
            y[i] = (g[i] - H[i, i + 1:k] @ y[i + 1:k]) / H[i, i]
        x = x + np.array(V[:k]).T @ y
        if res <= target:
            return x, iters, True, res / bnorm
    return x, iters, False, res / bnorm


class NKResult:
    def __init__(self, u, converged, newton_iters, gmres_iters, res_norm, history):
        self.u, self.converged, self.newton_iters = u, converged, newton_iters
        self.gmres_iters, self.res_norm, self.history = gmres_iters, res_norm, history


def newton_krylov(u0, u_old, dt, theta, D, coef, dy, dx, bc, precond="dst", newton_tol=1e-8,
                  krylov_tol=1e-8, max_newton=10, max_krylov=50, restart=30, damping=1.0,
                  backtrack_factor=0.5, max_backtrack=3, min_decrease=1e-4, F_old=None, lead=1.0, precond_dt=None):
    """newton_krylov.py:171-359 with the GMRES above.  The preconditioner
    approximates (I - precond_dt D lap)^-1 with precond_dt the dt of the
    preconditioner context, the step dt (newton_krylov.py:214, passed by
    stepping.py:211,257,307); by default it is dt."""
    def R(u):
        return residual(u, u_old, dt, theta, D, coef, dy, dx, bc, F_old, lead)

    beta = (dt if precond_dt is None else precond_dt) * D
    if precond == "dst":
        M = lambda v: helmholtz_precond(v, beta, dy, dx, bc)
    elif precond == "jacobi":
        M = lambda v: jacobi_precond(v, beta, dy, dx, bc)
    else:
        M = None
    x = u0.copy()
    r0 = np.linalg.norm(R(x))
    converged = r0 < newton_tol
    res_norm = r0
    it, lin_total, history = 0, 0, []
    while it < max_newton and not converged:
        r = R(x)
        rn = np.linalg.norm(r)
        converged = rn < newton_tol
        A = lambda v, x=x: jvp(x, v, dt, theta, D, coef, dy, dx, bc, lead)
        dxv, lin, _, _ = gmres(A, -r, M, krylov_tol, restart, max_krylov)
        lin_total += lin
        history.append(lin)
        alpha, x_bt, r_bt, accepted = damping, x, rn, False
        for _ in range(max_backtrack):
            x_new = x_bt + alpha * dxv
            r_new = np.linalg.norm(R(x_new))
            accept = r_new < (1.0 - min_decrease * alpha) * r_bt
            if not (accept or accepted):
                alpha = alpha * backtrack_factor
            if accept and not accepted:
                x_bt, r_bt = x_new, r_new
            accepted = accepted or accept
        if np.allclose(x_bt, x):
            x_bt = x + damping * dxv
        x = x_bt
        res_norm = r_bt
        it += 1
    return NKResult(x, converged, it, lin_total, res_norm, history)


def sine_mode_ic(n, seed=0, modes=8, kmax=6):
    """Sum of ``modes`` sine modes sin(pi p x) sin(pi q y) with amplitudes
    N(0, 0.5) and p, q uniform in 1..kmax on interior vertices x_i = i/(n+1)."""
    rng = np.random.default_rng(seed)
    h = 1.0 / (n + 1)
    x = h * np.arange(1, n + 1)
    X, Y = np.meshgrid(x, x)
    u = np.zeros((n, n))
    for _ in range(modes):
        a = rng.normal(0.0, 0.5)
        p, q = rng.integers(1, kmax + 1, size=2)
        u += a * np.sin(np.pi * p * X) * np.sin(np.pi * q * Y)
    return u


def be_step(u, dt, D, coef, dy, dx, bc, precond="dst", **kw):
    """stepping.py:170-214: Newton-Krylov from the explicit Euler guess."""
    return newton_krylov(u + dt * rhs(u, D, coef, dy, dx, bc), u, dt, 1.0, D, coef, dy, dx, bc, precond, **kw)


def cn_step(u, dt, D, coef, dy, dx, bc, precond="dst", **kw):
    """stepping.py:217-260."""
    F0 = rhs(u, D, coef, dy, dx, bc)
    return newton_krylov(u + dt * F0, u, dt, 0.5, D, coef, dy, dx, bc, precond, F_old=F0, **kw)


def bdf2_step(u, u_prev, dt, dt_prev, D, coef, dy, dx, bc, precond="dst", consistent=False, **kw):
    """stepping.py:263-310 with the residual of newton_krylov.py:406-448."""
    a0, a1, a2, beta = bdf2_coefficients(dt, dt_prev, consistent)
    g = -(a1 * u + a2 * u_prev)
    return newton_krylov(2.0 * u - u_prev, g, beta, 1.0, D, coef, dy, dx, bc, precond, lead=a0, precond_dt=dt, **kw)
