"""Explicit Gray-Scott steps (fused FD stencil + reaction + RK stage kernels)
and adaptive_integrate routing on the GPU vs the float64 oracle.

Tolerances: float64 interiors <= 1e-12 relative L2 after several steps,
float32 <= 1e-6; the adaptive explicit driver reproduces the oracle's
accept/reject sequence (FP64) with dt histories to 1e-10."""
import numpy as np
import pytest
import torch

from oracle import adaptive as oad
from oracle import gray_scott as ogs
from oracle import spectral as osp

pytestmark = pytest.mark.gpu

L = 2.5
GS = dict(Du=0.16, Dv=0.08, F=0.04, k=0.06)


def _np(t):
    return t.detach().double().cpu().numpy()


def _model(ny, nx, dtype):
    from moljax_amd.core import Grid2D, create_gray_scott_model
    grid = Grid2D.uniform(nx, ny, 0.0, L, 0.0, L)
    return create_gray_scott_model(grid, dtype=dtype, **GS), grid


def _dev(y0, dtype):
    return {k: torch.from_numpy(v).to("cuda", dtype) for k, v in y0.items()}


@pytest.mark.parametrize("method", ["euler", "ssprk3", "rk4"])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-6)])
@pytest.mark.parametrize("shape", [(64, 64), (48, 80), (37, 45), (70, 300), (1, 7), (20, 1), (130, 256)])
def test_explicit_steps_match_oracle(gpu, method, dtype, tol, shape):
    from moljax_amd.core import stepping
    ny, nx = shape
    model, grid = _model(ny, nx, dtype)
    y0 = ogs.initial_condition(ny, nx, seed=1, square=12)
    dt = min(1e-3, 0.2 * min(grid.dx, grid.dy) ** 2 / GS["Du"])  # inside the explicit stability limit
    ostep = {"euler": ogs.euler_step, "ssprk3": ogs.ssprk3_step, "rk4": ogs.rk4_step}[method]
    gstep = {"euler": stepping.euler_step, "ssprk3": stepping.ssprk3_step, "rk4": stepping.rk4_step}[method]
    ref, y = y0, _dev(y0, dtype)
    for _ in range(5):
        ref = ostep(ref, dt, GS["F"], GS["k"], GS["Du"], GS["Dv"], grid.dy, grid.dx)
        y = gstep(model, y, 0.0, dt)
    for f in ("u", "v"):
        assert ogs.rel_l2(_np(y[f])[1:-1, 1:-1], ref[f][1:-1, 1:-1]) <= tol, f


def test_integrate_fixed_dt_rk4(gpu):
    from moljax_amd.core.stepping import IntegratorType, integrate_fixed_dt
    model, grid = _model(32, 32, torch.float64)
    y0 = ogs.initial_condition(32, 32, seed=2, square=8)
    dt = 2e-3
    t_hist, hist, fin = integrate_fixed_dt(model, _dev(y0, torch.float64), 0.0, 0.02, dt, IntegratorType.RK4,
                                           save_every=4)
    n = int(0.02 / dt) + 1
    ref = y0
    for _ in range(n):
        ref = ogs.rk4_step(ref, dt, GS["F"], GS["k"], GS["Du"], GS["Dv"], grid.dy, grid.dx)
    assert ogs.rel_l2(_np(fin["u"])[1:-1, 1:-1], ref["u"][1:-1, 1:-1]) < 1e-12
    assert hist["u"].shape[0] == n // 4 and float(t_hist[-1]) == pytest.approx(4 * (n // 4) * dt)


@pytest.mark.parametrize("method", ["rk4", "ssprk3"])
def test_adaptive_explicit_matches_oracle(gpu, method):
    from moljax_amd.core.dt_policy import PIDParams
    from moljax_amd.core.stepping import IntegratorType, adaptive_integrate
    model, grid = _model(32, 32, torch.float64)
    y0 = ogs.initial_condition(32, 32, seed=4, square=8)
    pid = dict(atol=1e-6, rtol=1e-5)
    ref = oad.adaptive_integrate_explicit(y0, 0.0, 0.2, 0.05, GS["F"], GS["k"], GS["Du"], GS["Dv"], grid.dy,
                                          grid.dx, method=method, pid=oad.PIDParams(**pid), max_steps=2000)
    m = IntegratorType.RK4 if method == "rk4" else IntegratorType.SSPRK3
    res = adaptive_integrate(model, _dev(y0, torch.float64), 0.0, 0.2, 0.05, method=m, max_steps=2000,
                             pid_params=PIDParams(**pid))
    print(f"{method}: oracle acc={ref.n_accepted} rej={ref.n_rejected}; gpu acc={res.n_accepted} "
          f"rej={res.n_rejected}", flush=True)
    assert (res.n_accepted, res.n_rejected, res.status, res.n_steps) == \
        (ref.n_accepted, ref.n_rejected, ref.status, ref.n_steps)
    w = ref.n_steps
    assert np.allclose(_np(res.dt_history)[:w], ref.dt_history[:w], rtol=1e-10, atol=0)
    assert ogs.rel_l2(_np(res.y_final["u"])[1:-1, 1:-1], ref.y_final["u"][1:-1, 1:-1]) < 1e-10


def test_adaptive_integrate_routes_imex_and_etdrk4(gpu):
    from moljax_amd.core import Grid2D, create_gray_scott_periodic_fft
    from moljax_amd.core.dt_policy import PIDParams
    from moljax_amd.core.fft_integrators import etdrk4_coefficients, gs_etdrk4_inplace, _fd_symbol
    from moljax_amd.core.stepping import IntegratorType, adaptive_integrate, adaptive_integrate_imex
    n = 32
    grid = Grid2D.uniform(n, n, 0.0, L, 0.0, L)
    model, cache, diff = create_gray_scott_periodic_fft(grid, dtype=torch.float64, **GS)
    y0 = ogs.initial_condition(n, n, seed=6, square=8)
    pid = PIDParams(atol=1e-4, rtol=1e-3)
    a = adaptive_integrate(model, _dev(y0, torch.float64), 0.0, 3.0, 0.5, method=IntegratorType.IMEX_STRANG,
                           pid_params=pid, max_steps=200)
    b = adaptive_integrate_imex(model, _dev(y0, torch.float64), 0.0, 3.0, 0.5, cache, diff, pid_params=pid,
                                max_steps=200)
    assert (a.n_accepted, a.n_rejected, a.status) == (b.n_accepted, b.n_rejected, b.status)
    assert torch.equal(a.y_final["u"], b.y_final["u"])
    # ETDRK4: fixed dt0 steps, the last one shortened onto t_end
    r = adaptive_integrate(model, _dev(y0, torch.float64), 0.0, 1.25, 0.5, method=IntegratorType.ETDRK4)
    assert r.status == 0 and r.n_accepted == 3 and r.t_final == pytest.approx(1.25)
    ui = torch.from_numpy(y0["u"][1:-1, 1:-1].copy()).cuda()
    vi = torch.from_numpy(y0["v"][1:-1, 1:-1].copy()).cuda()
    lx, ly = _fd_symbol(n, grid.dx), _fd_symbol(n, grid.dy)
    for h in (0.5, 0.5, 0.25):
        coef = etdrk4_coefficients(lx, ly, GS["Du"], GS["Dv"], h, torch.float64)
        gs_etdrk4_inplace(ui, vi, coef, [[0, 0, 0, 0, GS["F"], GS["k"], h]], 1)
    assert torch.equal(r.y_final["u"][1:-1, 1:-1], ui)
    with pytest.raises(NotImplementedError):
        adaptive_integrate(model, _dev(y0, torch.float64), 0.0, 1.0, 0.1, method=IntegratorType.BE)
