"""Gray-Scott IMEX-Strang on the library data layout, float64 NumPy
restatement (test infrastructure).

Follows the reference dataflow step for step: ghost-padded (ny+2, nx+2)
arrays, periodic ghost fill (moljax/core/bc.py:202-215), the Gray-Scott
reaction on padded arrays (moljax/core/operators.py:644-667),
imex_strang_step (moljax/core/stepping.py:370-435) with
apply_diffusion_inverse_fft (moljax/core/fft_solvers.py:406-445), and the
fixed-dt driver step count int((t_end - t0)/dt) + 1 (stepping.py:1402).
"""
import numpy as np

from . import spectral


def apply_bc_periodic(f, ng=1):
    """bc.py:202-215: x ghosts first, then y ghosts (full rows, so the corners
    come from the opposite corner)."""
    f = f.copy()
    ny, nx = f.shape[0] - 2 * ng, f.shape[1] - 2 * ng
    f[:, :ng] = f[:, nx:nx + ng]
    f[:, nx + ng:] = f[:, ng:2 * ng]
    f[:ng, :] = f[ny:ny + ng, :]
    f[ny + ng:, :] = f[ng:2 * ng, :]
    return f


def apply_bcs(state, ng=1):
    return {k: apply_bc_periodic(v, ng) for k, v in state.items()}


def reaction(state, F, k):
    """operators.py:655-667."""
    u, v = state["u"], state["v"]
    uvv = u * v * v
    return {"u": -uvv + F * (1.0 - u), "v": uvv - (F + k) * v}


def imex_strang_step(y, dt, F, k, Du, Dv, lam, ng=1):
    """stepping.py:370-435 (model.apply_bcs = periodic ghost fill)."""
    diff = {"u": Du, "v": Dv}
    dt_half = dt / 2.0
    y = apply_bcs(y, ng)
    y_star = spectral.apply_diffusion_inverse_fft(y, ng, dt_half, diff, lam)
    y_star = apply_bcs(y_star, ng)
    R1 = reaction(y_star, F, k)
    y_tilde = {n: y_star[n] + dt * R1[n] for n in y_star}
    y_tilde = apply_bcs(y_tilde, ng)
    R2 = reaction(y_tilde, F, k)
    R_avg = {n: R1[n] + R2[n] for n in R1}
    y_double = {n: y_star[n] + 0.5 * dt * R_avg[n] for n in y_star}
    y_double = apply_bcs(y_double, ng)
    y_new = spectral.apply_diffusion_inverse_fft(y_double, ng, dt_half, diff, lam)
    return apply_bcs(y_new, ng)


def imex_euler_step(y, dt, F, k, Du, Dv, lam, ng=1):
    """stepping.py:317-367: y_new = (I - dt D lap)^-1 (y + dt N(y))."""
    diff = {"u": Du, "v": Dv}
    y = apply_bcs(y, ng)
    N = reaction(y, F, k)
    rhs = {n: y[n] + dt * N[n] for n in y}
    return apply_bcs(spectral.apply_diffusion_inverse_fft(rhs, ng, dt, diff, lam), ng)


def imex_ssprk2_step(y, dt, F, k, Du, Dv, lam, ng=1):
    """stepping.py:438-495: y1 = E(y), y2 = E(y1), y_new = y/2 + y2/2 with E
    the IMEX-Euler map above."""
    y = apply_bcs(y, ng)
    y1 = imex_euler_step(y, dt, F, k, Du, Dv, lam, ng)
    y2 = imex_euler_step(y1, dt, F, k, Du, Dv, lam, ng)
    return apply_bcs({n: 0.5 * y[n] + 0.5 * y2[n] for n in y}, ng)


def laplacian_padded(f, dy, dx, ng=1):
    """operators.py:159-322: d2_central axis 0 + axis 1 on the interior of a
    padded array, zeros in the ghost ring."""
    out = np.zeros_like(f)
    c = f[ng:-ng, ng:-ng]
    d2y = (f[ng + 1:f.shape[0] - ng + 1, ng:-ng] - 2.0 * c + f[ng - 1:f.shape[0] - ng - 1, ng:-ng]) / dy ** 2
    d2x = (f[ng:-ng, ng + 1:f.shape[1] - ng + 1] - 2.0 * c + f[ng:-ng, ng - 1:f.shape[1] - ng - 1]) / dx ** 2
    out[ng:-ng, ng:-ng] = d2y + d2x
    return out


def rhs(y, F, k, Du, Dv, dy, dx, ng=1):
    """model.py:100-118 for the periodic Gray-Scott model: apply_bcs, then
    D lap (interior) + reaction (padded)."""
    y = apply_bcs(y, ng)
    R = reaction(y, F, k)
    return {"u": Du * laplacian_padded(y["u"], dy, dx, ng) + R["u"],
            "v": Dv * laplacian_padded(y["v"], dy, dx, ng) + R["v"]}


def euler_step(y, dt, F, k, Du, Dv, dy, dx, ng=1):
    """stepping.py:64-85."""
    f = rhs(y, F, k, Du, Dv, dy, dx, ng)
    return {n: y[n] + dt * f[n] for n in y}


def ssprk3_step(y, dt, F, k, Du, Dv, dy, dx, ng=1):
    """stepping.py:88-124."""
    f0 = rhs(y, F, k, Du, Dv, dy, dx, ng)
    y1 = {n: y[n] + dt * f0[n] for n in y}
    f1 = rhs(y1, F, k, Du, Dv, dy, dx, ng)
    y2 = {n: (0.75 * y[n] + 0.25 * y1[n]) + (0.25 * dt) * f1[n] for n in y}
    f2 = rhs(y2, F, k, Du, Dv, dy, dx, ng)
    return {n: ((1.0 / 3.0) * y[n] + (2.0 / 3.0) * y2[n]) + (2.0 / 3.0 * dt) * f2[n] for n in y}


def rk4_step(y, dt, F, k, Du, Dv, dy, dx, ng=1):
    """stepping.py:127-167."""
    k1 = rhs(y, F, k, Du, Dv, dy, dx, ng)
    k2 = rhs({n: y[n] + 0.5 * dt * k1[n] for n in y}, F, k, Du, Dv, dy, dx, ng)
    k3 = rhs({n: y[n] + 0.5 * dt * k2[n] for n in y}, F, k, Du, Dv, dy, dx, ng)
    k4 = rhs({n: y[n] + dt * k3[n] for n in y}, F, k, Du, Dv, dy, dx, ng)
    return {n: y[n] + dt / 6.0 * (((k1[n] + 2.0 * k2[n]) + 2.0 * k3[n]) + k4[n]) for n in y}


def integrate(y0, dt, n_steps, F, k, Du, Dv, lam, ng=1, record=None):
    """Fixed-dt loop (stepping.py:1414-1434 body).  ``record`` is an optional
    set of step counts at which the state is saved."""
    y = {n: v.copy() for n, v in y0.items()}
    saved = {}
    for s in range(1, n_steps + 1):
        y = imex_strang_step(y, dt, F, k, Du, Dv, lam, ng)
        if record is not None and s in record:
            saved[s] = {n: v.copy() for n, v in y.items()}
    return y, saved


def fixed_dt_n_steps(t0, t_end, dt):
    """stepping.py:1402 -- note the extra step."""
    return int((t_end - t0) / dt) + 1


def initial_condition(ny, nx, seed, square=20, noise=0.01, ng=1):
    """Gray-Scott IC on interior cells (ny, nx): u = 1, v = 0, a centred
    square of square x square cells with u = 0.5, v = 0.25, plus uniform noise
    in [-noise, noise) on u then v from numpy default_rng(seed); returned as
    ghost-padded arrays with the periodic ring filled (domain convention of
    examples/gray_scott_2d.py:42-73, noise source made reproducible)."""
    rng = np.random.default_rng(seed)
    u = np.ones((ny, nx))
    v = np.zeros((ny, nx))
    cy, cx = ny // 2, nx // 2
    h = square // 2
    u[cy - h:cy - h + square, cx - h:cx - h + square] = 0.5
    v[cy - h:cy - h + square, cx - h:cx - h + square] = 0.25
    u = u + rng.uniform(-noise, noise, size=(ny, nx))
    v = v + rng.uniform(-noise, noise, size=(ny, nx))
    U = np.zeros((ny + 2 * ng, nx + 2 * ng))
    V = np.zeros_like(U)
    U[ng:-ng, ng:-ng] = u
    V[ng:-ng, ng:-ng] = v
    return {"u": apply_bc_periodic(U, ng), "v": apply_bc_periodic(V, ng)}


def rel_l2(a, b):
    """Relative L2 error of a against b (b is the reference)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
