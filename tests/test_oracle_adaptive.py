"""Pinning the adaptive-IMEX oracle and the host controller (CPU).

What fixes the restatement independently of itself:
  * the reference's own controller tests (tests/test_dt_policy.py:100-148)
    run against both the oracle controller and moljax_amd.core.dt_policy;
  * a PI factor that reduces to the safety factor (e = e_prev = 1) and the
    clip limits;
  * the Gray-Scott fixed point u = 1, v = 0: zero error, so dt grows by the
    max factor until the reaction CFL cap, a closed-form dt sequence that
    pins the loop (clamping to t_end, histories, counters, status);
  * pure diffusion of one Fourier mode: the first attempt's error norm in
    closed form from the per-step amplification factors of Strang,
    (1 - dt/2 D lam)^-2, and of IMEX-Euler, (1 - dt D lam)^-1;
  * IMEX-Euler and SSPRK2 steps against their one-mode closed forms.
"""
import math

import numpy as np
import pytest

from oracle import adaptive as oad
from oracle import gray_scott as ogs
from oracle import spectral as osp
from moljax_amd.core import dt_policy as mdp


def _controllers():
    yield "oracle", lambda dt, e, prev, pid, order: oad.pid_controller_dt(dt, e, prev, pid, order)[0]
    yield "host", lambda dt, e, prev, pid, order: mdp.pid_controller_dt(
        dt, e, mdp.ControllerState(prev, 1.0, 0), mdp.PIDParams(**pid.__dict__), order)[0]


@pytest.mark.parametrize("which", ["oracle", "host"])
def test_reference_controller_behaviour(which):
    """tests/test_dt_policy.py:103-148."""
    pid_fn = dict(_controllers())[which]
    P = oad.PIDParams
    assert pid_fn(0.1, 0.1, 1.0, P(safety=0.9, max_factor=2.0), 4) > 0.1
    assert pid_fn(0.1, 2.0, 1.0, P(safety=0.9, min_factor=0.2), 4) < 0.1
    p = P(dt_min=0.001, dt_max=1.0, max_factor=100.0)
    assert pid_fn(0.5, 1e-6, 1.0, p, 4) <= p.dt_max
    assert pid_fn(0.5, 1e6, 1.0, p, 4) >= p.dt_min


@pytest.mark.parametrize("which", ["oracle", "host"])
def test_controller_limits(which):
    pid_fn = dict(_controllers())[which]
    P = oad.PIDParams()
    assert pid_fn(0.1, 1.0, 1.0, P, 2) == pytest.approx(0.09, rel=1e-15)  # safety only
    assert pid_fn(0.1, 1e-9, 1.0, P, 2) == pytest.approx(0.5, rel=1e-15)  # max_factor
    assert pid_fn(0.1, 1e9, 1.0, P, 2) == pytest.approx(0.02, rel=1e-15)  # min_factor
    # rejected step: never grows, min_factor floor
    assert oad.handle_rejected_step(0.1, 1.0, P, 2) == pytest.approx(0.09, rel=1e-15)
    assert oad.handle_rejected_step(0.1, 1e9, P, 2) == pytest.approx(0.02, rel=1e-15)
    h = mdp.handle_rejected_step(0.1, 1.0, mdp.create_initial_controller_state(), mdp.PIDParams(), 2)
    assert h[0] == pytest.approx(0.09, rel=1e-15) and h[1].consecutive_rejects == 1


def test_imex_cfl_dt_gray_scott():
    """dt_policy.py:430-502 for F + k = 0.1: 0.9 * min(1, 0.5/0.1) = 0.9;
    an advection speed caps it at cfl_adv dx / v."""
    c = oad.CFLParams()
    assert oad.imex_cfl_dt({"F": 0.04, "k": 0.06}, c, 0.1) == pytest.approx(0.9)
    assert oad.imex_cfl_dt({"F": 0.4, "k": 0.6}, c, 0.1) == pytest.approx(0.45)

    class G:
        min_dx = 0.1
        min_dx2 = 0.01
    assert mdp.imex_cfl_dt(G, {"F": 0.04, "k": 0.06, "vx": 1.0}, mdp.CFLParams()) == pytest.approx(0.9 * 0.05)
    assert mdp.heisenberg_cfl_dt(G, {"Du": 0.16}, mdp.CFLParams()) == pytest.approx(0.9 * 0.25 * 0.01 / 0.16)
    assert mdp.compute_error_order(7) == 2 and mdp.compute_error_order(6) == 1


def test_fixed_point_dt_sequence_closed_form():
    N, L = 16, 2.5
    h = L / N
    lam = osp.laplacian_symbol_2d(N, N, h, h)
    U = np.ones((N + 2, N + 2))
    V = np.zeros((N + 2, N + 2))
    r = oad.adaptive_integrate_imex({"u": U, "v": V}, 0.0, 5.0, 0.01, 0.04, 0.06, 0.16, 0.08, lam, h,
                                    max_steps=100)
    expect, t = [], 0.0
    dt = 0.01
    while t < 5.0:
        dc = min(dt, 5.0 - t)
        expect.append(dc)
        t += dc
        dt = min(dt * 5.0, 0.9)
    assert r.status == oad.SUCCESS and r.n_rejected == 0
    assert r.n_accepted == len(expect) and r.n_steps == len(expect) + 1
    assert np.allclose(r.dt_history[:len(expect)], expect, rtol=1e-14, atol=0)
    assert np.allclose(r.t_history[1:len(expect) + 1], np.cumsum(expect), rtol=1e-14)
    assert r.t_final == pytest.approx(5.0, abs=1e-14)
    assert np.max(np.abs(r.y_final["u"] - 1.0)) < 1e-14 and np.max(np.abs(r.y_final["v"])) == 0.0


def _mode_state(N, L, m, a):
    h = L / N
    x = h * np.arange(N)
    u = 1.0 + a * np.cos(2 * np.pi * m * x / L)[None, :] * np.ones((N, 1))
    U = np.zeros((N + 2, N + 2))
    U[1:-1, 1:-1] = u
    return {"u": ogs.apply_bc_periodic(U), "v": np.zeros((N + 2, N + 2))}, x


@pytest.mark.parametrize("use_strang", [True, False])
def test_first_attempt_error_closed_form(use_strang):
    N, L, m, a, Du, dt = 32, 2.5, 3, 0.3, 0.16, 0.2
    h = L / N
    lam = osp.laplacian_symbol_2d(N, N, h, h)
    y0, x = _mode_state(N, L, m, a)
    lm = (2.0 * np.cos(2 * np.pi * m / N) - 2.0) / h ** 2
    if use_strang:
        g = lambda s: (1.0 - s / 2 * Du * lm) ** -2
    else:
        g = lambda s: 1.0 / (1.0 - s * Du * lm)
    pid = oad.PIDParams(atol=1e-4, rtol=1e-3)
    r = oad.adaptive_integrate_imex(y0, 0.0, 10.0, dt, 0.0, 0.0, Du, 0.0, lam, h, use_strang=use_strang,
                                    pid=pid, max_attempts=1, cfl=oad.CFLParams(dt_max=10.0))
    c = np.cos(2 * np.pi * m * x / L)
    y_new = 1.0 + a * g(dt / 2) ** 2 * c
    err = a * (g(dt / 2) ** 2 - g(dt)) * c
    e = math.sqrt(np.sum((err / (pid.atol + pid.rtol * np.abs(y_new))) ** 2) * N / (2 * N * N))
    assert r.attempts[0][2] == pytest.approx(e, rel=1e-9)


def test_euler_and_ssprk2_steps_closed_form():
    N, L, m, a, Du, dt = 32, 2.5, 2, 0.2, 0.16, 0.3
    h = L / N
    lam = osp.laplacian_symbol_2d(N, N, h, h)
    y0, x = _mode_state(N, L, m, a)
    lm = (2.0 * np.cos(2 * np.pi * m / N) - 2.0) / h ** 2
    g = 1.0 / (1.0 - dt * Du * lm)
    c = np.cos(2 * np.pi * m * x / L)[None, :]
    ye = ogs.imex_euler_step(y0, dt, 0.0, 0.0, Du, 0.0, lam)
    assert np.max(np.abs(ye["u"][1:-1, 1:-1] - (1.0 + a * g * c))) < 1e-14
    ys = ogs.imex_ssprk2_step(y0, dt, 0.0, 0.0, Du, 0.0, lam)
    assert np.max(np.abs(ys["u"][1:-1, 1:-1] - (1.0 + a * 0.5 * (1 + g * g) * c))) < 1e-14
    # with reaction: the Euler step is (I - dt D lam)^-1 (y + dt R(y)) on the fixed-point-free state
    y1 = {"u": np.full((N + 2, N + 2), 0.7), "v": np.full((N + 2, N + 2), 0.2)}
    r = ogs.reaction(y1, 0.04, 0.06)
    ye = ogs.imex_euler_step(y1, dt, 0.04, 0.06, Du, 0.08, lam)
    assert np.max(np.abs(ye["u"] - (0.7 + dt * r["u"]))) < 1e-14
    assert np.max(np.abs(ye["v"] - (0.2 + dt * r["v"]))) < 1e-14


def test_status_codes():
    N, L = 16, 2.5
    h = L / N
    lam = osp.laplacian_symbol_2d(N, N, h, h)
    y0 = ogs.initial_condition(N, N, seed=0, square=6)
    r = oad.adaptive_integrate_imex(y0, 1.0, 1.0, 0.1, 0.04, 0.06, 0.16, 0.08, lam, h)
    assert r.status == oad.SUCCESS and r.n_accepted == 0 and not r.attempts
    r = oad.adaptive_integrate_imex(y0, 0.0, 100.0, 0.1, 0.04, 0.06, 0.16, 0.08, lam, h, max_steps=3)
    assert r.status == oad.MAX_STEPS_REACHED and r.n_steps == 4
    pid = oad.PIDParams(atol=1e-12, rtol=1e-12, dt_min=0.05)
    r = oad.adaptive_integrate_imex(y0, 0.0, 1.0, 0.5, 0.04, 0.06, 0.16, 0.08, lam, h, pid=pid)
    assert r.status == oad.DT_TOO_SMALL and r.n_rejected >= 1
    bad = {k: v.copy() for k, v in y0.items()}
    bad["u"][5, 5] = np.nan
    r = oad.adaptive_integrate_imex(bad, 0.0, 1.0, 0.1, 0.04, 0.06, 0.16, 0.08, lam, h)
    assert r.status == oad.NON_FINITE_VALUES and r.n_accepted == 0


@pytest.mark.parametrize("method,poly", [
    ("euler", lambda z: 1 + z),
    ("ssprk3", lambda z: 1 + z + z * z / 2 + z ** 3 / 6),
    ("rk4", lambda z: 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24)])
def test_explicit_steps_mode_amplification(method, poly):
    """Pure diffusion of one Fourier mode: an explicit RK step multiplies the
    mode by its stability polynomial at z = dt D lam_m (lam_m the 5-point
    eigenvalue), independently of the stage arithmetic."""
    N, L, m, a, Du, dt = 32, 2.5, 3, 0.3, 0.16, 0.02
    h = L / N
    y0, x = _mode_state(N, L, m, a)
    lm = (2.0 * np.cos(2 * np.pi * m / N) - 2.0) / h ** 2
    step = {"euler": ogs.euler_step, "ssprk3": ogs.ssprk3_step, "rk4": ogs.rk4_step}[method]
    y = step(y0, dt, 0.0, 0.0, Du, 0.0, h, h)
    c = np.cos(2 * np.pi * m * x / L)[None, :]
    assert np.max(np.abs(y["u"][1:-1, 1:-1] - (1.0 + a * poly(dt * Du * lm) * c))) < 1e-14
    # reaction only (uniform state): the stage recurrences on the scalar ODE
    y1 = {"u": np.full((N + 2, N + 2), 0.7), "v": np.full((N + 2, N + 2), 0.2)}
    yr = step(y1, dt, 0.04, 0.06, Du, 0.08, h, h)

    def f(u, v):
        return -u * v * v + 0.04 * (1 - u), u * v * v - 0.1 * v
    u, v = 0.7, 0.2
    if method == "euler":
        fu, fv = f(u, v)
        ue, ve = u + dt * fu, v + dt * fv
    elif method == "ssprk3":
        f0 = f(u, v)
        u1, v1 = u + dt * f0[0], v + dt * f0[1]
        f1 = f(u1, v1)
        u2, v2 = 0.75 * u + 0.25 * (u1 + dt * f1[0]), 0.75 * v + 0.25 * (v1 + dt * f1[1])
        f2 = f(u2, v2)
        ue, ve = u / 3 + 2 / 3 * (u2 + dt * f2[0]), v / 3 + 2 / 3 * (v2 + dt * f2[1])
    else:
        k1 = f(u, v)
        k2 = f(u + dt / 2 * k1[0], v + dt / 2 * k1[1])
        k3 = f(u + dt / 2 * k2[0], v + dt / 2 * k2[1])
        k4 = f(u + dt * k3[0], v + dt * k3[1])
        ue = u + dt / 6 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0])
        ve = v + dt / 6 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1])
    assert np.max(np.abs(yr["u"][1:-1, 1:-1] - ue)) < 1e-15
    assert np.max(np.abs(yr["v"][1:-1, 1:-1] - ve)) < 1e-15


def test_heisenberg_cfl_and_explicit_adaptive_loop():
    N, L = 32, 2.5
    h = L / N
    c = oad.CFLParams()
    assert oad.heisenberg_cfl_dt({"Du": 0.16, "Dv": 0.08}, c, h) == pytest.approx(0.9 * 0.25 * h * h / 0.16)
    y0 = ogs.initial_condition(N, N, seed=2, square=8)
    r = oad.adaptive_integrate_explicit(y0, 0.0, 2.0, 0.5, 0.04, 0.06, 0.16, 0.08, h, h, method="rk4",
                                        pid=oad.PIDParams(atol=1e-5, rtol=1e-4))
    assert r.status == oad.SUCCESS and r.t_final == pytest.approx(2.0)
    # every accepted dt respects the diffusion CFL cap
    assert np.all(r.dt_history[:r.n_steps - 1] <= oad.heisenberg_cfl_dt({"Du": 0.16, "Dv": 0.08}, c, h) + 1e-15)
