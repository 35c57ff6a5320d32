"""Coupled Gray-Scott ETDRK4 and FFT operators on the GPU vs the float64
oracle.  Tolerances (relative L2): coefficients 1e-12; float64 trajectories
1e-12; float32 trajectories 1e-5 over 10 steps."""
import numpy as np
import pytest
import torch

from oracle import etdrk4 as oe
from oracle import gray_scott as ogs
from oracle import spectral as osp

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().double().cpu().numpy()


def _ic(n, seed):
    u, v = ogs.initial_condition(n, n, seed, square=max(2, n * 20 // 256))["u"], None
    y = ogs.initial_condition(n, n, seed, square=max(2, n * 20 // 256))
    return y["u"][1:-1, 1:-1].copy(), y["v"][1:-1, 1:-1].copy()


@pytest.mark.parametrize("n", [32, 256])
def test_coefficients_match_oracle(gpu, n):
    from moljax_amd.core.fft_integrators import etdrk4_coefficients, spectral_symbol_1d
    lam = spectral_symbol_1d(n)
    coef = _np(etdrk4_coefficients(lam, lam, 2e-5, 1e-5, 1.0, torch.float64))
    cu, cv = oe.gs_coeffs(lam, lam, 2e-5, 1e-5, 1.0)
    h = n // 2
    for t in range(6):
        w = 2.0 if t == 4 else 1.0  # f2 holds the Cox-Matthews weight 2 f2
        assert ogs.rel_l2(coef[t][:, :h + 1], w * np.real(cu[t])[:, :h + 1]) < 1e-12
        assert ogs.rel_l2(coef[t][:, h + 1:], w * np.real(cv[t])[:, :h + 1]) < 1e-12


@pytest.mark.parametrize("n", [64, 256])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_gs_etdrk4_matches_oracle(gpu, n, dtype, tol):
    from moljax_amd.core.fft_integrators import etdrk4_coefficients, gs_etdrk4_inplace, spectral_symbol_1d
    lam = spectral_symbol_1d(n)
    F, k, Du, Dv, dt, steps = 0.035, 0.065, 2e-5, 1e-5, 1.0, 10
    u0, v0 = _ic(n, 0)
    U = torch.from_numpy(u0).to("cuda", dtype)
    V = torch.from_numpy(v0).to("cuda", dtype)
    coef = etdrk4_coefficients(lam, lam, Du, Dv, dt, dtype)
    gs_etdrk4_inplace(U, V, coef, [[0, 0, 0, 0, F, k, dt]], steps)
    ru, rv = oe.gs_etdrk4(u0, v0, lam, lam, Du, Dv, F, k, dt, steps)
    eu, ev = ogs.rel_l2(_np(U), ru), ogs.rel_l2(_np(V), rv)
    print(f"n={n} {dtype} rel_l2 u={eu:.2e} v={ev:.2e}", flush=True)
    assert eu <= tol and ev <= tol


def test_batched_per_instance_params_and_invariance(gpu):
    from moljax_amd.core.fft_integrators import etdrk4_coefficients, gs_etdrk4_inplace, spectral_symbol_1d
    n, steps = 64, 6
    lam = spectral_symbol_1d(n)
    params = [(0.03, 0.06), (0.05, 0.062), (0.04, 0.058)]
    ics = [_ic(n, s) for s in range(3)]
    U = torch.from_numpy(np.stack([a for a, _ in ics])).cuda()
    V = torch.from_numpy(np.stack([b for _, b in ics])).cuda()
    coef = etdrk4_coefficients(lam, lam, 2e-5, 1e-5, 1.0, torch.float64)
    rows = [[0, 0, 0, 0, F, k, 1.0] for F, k in params]
    u1, v1 = U[1].clone(), V[1].clone()
    gs_etdrk4_inplace(U, V, coef, rows, steps)
    gs_etdrk4_inplace(u1, v1, coef, rows[1:2], steps)
    assert torch.equal(U[1], u1) and torch.equal(V[1], v1)
    for i, (F, k) in enumerate(params):
        ru, rv = oe.gs_etdrk4(ics[i][0], ics[i][1], lam, lam, 2e-5, 1e-5, F, k, 1.0, steps)
        assert ogs.rel_l2(_np(U[i]), ru) < 1e-12 and ogs.rel_l2(_np(V[i]), rv) < 1e-12


@pytest.mark.parametrize("n,B", [(256, 64), (512, 32), (512, 64)])
def test_row_tiles_per_workgroup_do_not_change_results(gpu, n, B):
    """FP32 batches this large run the row pass with two row tiles per
    workgroup, or four at 512^2 when the grid keeps >= 2048 workgroups (the
    loads of the later tiles issued before the first is transformed); a
    lone instance runs one tile per workgroup.  Each row's arithmetic is the
    same, so the results are bitwise equal."""
    from moljax_amd.core.fft_integrators import etdrk4_coefficients, gs_etdrk4_inplace, spectral_symbol_1d
    steps = 3
    lam = spectral_symbol_1d(n)
    ics = [_ic(n, s % 4) for s in range(B)]
    U = torch.from_numpy(np.stack([a for a, _ in ics])).to("cuda", torch.float32)
    V = torch.from_numpy(np.stack([b for _, b in ics])).to("cuda", torch.float32)
    coef = etdrk4_coefficients(lam, lam, 2e-5, 1e-5, 1.0, torch.float32)
    rows = [[0, 0, 0, 0, 0.03 + 0.0003 * i, 0.06, 1.0] for i in range(B)]
    pick = [0, B // 2 + 1, B - 1]
    alone = []
    for i in pick:
        u1, v1 = U[i].clone(), V[i].clone()
        gs_etdrk4_inplace(u1, v1, coef, rows[i:i + 1], steps)
        alone.append((u1, v1))
    gs_etdrk4_inplace(U, V, coef, rows, steps)
    for i, (u1, v1) in zip(pick, alone):
        print(f"n={n} B={B} instance {i}: max |diff| u={float((U[i] - u1).abs().max()):.3e}", flush=True)
        assert torch.equal(U[i], u1) and torch.equal(V[i], v1), i


def test_etdrk4_step_api_fd_symbol(gpu):
    """etdrk4_step / etd_integrate with DiffusionOperator (library FD symbol)."""
    from moljax_amd.core import Grid2D
    from moljax_amd.core.fft_integrators import etd_integrate, etdrk4_step, gray_scott_nonlinear_rhs
    from moljax_amd.core.fft_operators import DiffusionOperator
    n = 64
    grid = Grid2D.uniform(n, n, 0.0, 2.5, 0.0, 2.5)
    ops = {"u": DiffusionOperator(grid, 2e-5), "v": DiffusionOperator(grid, 1e-5)}
    nl = gray_scott_nonlinear_rhs(0.035, 0.065)
    u0, v0 = _ic(n, 1)
    st = {"u": torch.from_numpy(u0).cuda(), "v": torch.from_numpy(v0).cuda()}
    one = etdrk4_step(st, 0.0, 1.0, ops, nl)
    lam = osp.laplacian_symbol_1d(n, grid.dx)
    ru, rv = oe.gs_etdrk4(u0, v0, lam, lam, 2e-5, 1e-5, 0.035, 0.065, 1.0, 1)
    assert ogs.rel_l2(_np(one["u"]), ru) < 1e-12
    t, hist = etd_integrate(st, (0.0, 4.0), 1.0, ops, nl, method="etdrk4", save_every=2)
    assert len(hist) == 3 and np.allclose(t.numpy(), [0, 2, 4])
    ru4, _ = oe.gs_etdrk4(u0, v0, lam, lam, 2e-5, 1e-5, 0.035, 0.065, 1.0, 4)
    assert ogs.rel_l2(_np(hist[-1]["u"]), ru4) < 1e-12


def test_diffusion_operator(gpu):
    from moljax_amd.core import Grid2D
    from moljax_amd.core.fft_operators import DiffusionOperator
    grid = Grid2D.uniform(64, 32, 0.0, 1.0, 0.0, 2.0)
    op = DiffusionOperator(grid, 0.01)
    u = np.random.default_rng(8).standard_normal((32, 64))
    lam = osp.laplacian_symbol_2d(32, 64, grid.dy, grid.dx)
    t = torch.from_numpy(u).cuda()
    assert ogs.rel_l2(_np(op.matvec(t)), 0.01 * osp.periodic_fd_laplacian(u, grid.dy, grid.dx)) < 1e-12
    assert ogs.rel_l2(_np(op.solve(t, 0.3)), osp.solve_helmholtz_2d(u, lam, 0.3, 0.01)) < 1e-13
    ex = np.fft.ifft2(np.fft.fft2(u) * np.exp(0.3 * 0.01 * lam)).real
    assert ogs.rel_l2(_np(op.exp_matvec(t, 0.3)), ex) < 1e-13
    with pytest.raises(ValueError):
        DiffusionOperator(grid, 0.0)


def test_operator_wrappers(gpu):
    """fft_integrators.py:400-542 and fft_solvers.py:448-474,619-662 wrappers
    over the operator kernels, and exact_cfl_dt (fft_operators.py:263-300)
    against the closed form rho = D (4/dx^2 + 4/dy^2) of even grids."""
    from moljax_amd.core import Grid2D, create_fft_cache
    from moljax_amd.core.fft_integrators import (batched_fft_exp_matvec, batched_fft_matvec, batched_fft_solve,
                                                 diffusion_only_etd1, stacked_fft_solve_shared_op)
    from moljax_amd.core.fft_operators import DiffusionOperator, FFTLinearOperator, exact_cfl_dt
    from moljax_amd.core.fft_solvers import apply_diffusion_fft, diffusion_rhs_fft
    grid = Grid2D.uniform(64, 32, 0.0, 1.0, 0.0, 2.0)
    opu, opv = DiffusionOperator(grid, 0.01), DiffusionOperator(grid, 0.002)
    assert isinstance(opu, FFTLinearOperator)
    lam = osp.laplacian_symbol_2d(32, 64, grid.dy, grid.dx)
    rng = np.random.default_rng(9)
    a, b = rng.standard_normal((32, 64)), rng.standard_normal((32, 64))
    st = {"u": torch.from_numpy(a).cuda(), "v": torch.from_numpy(b).cuda(), "w": torch.from_numpy(a).cuda()}
    ops = {"u": opu, "v": opv}
    mv = batched_fft_matvec(st, ops)
    assert ogs.rel_l2(_np(mv["u"]), 0.01 * osp.periodic_fd_laplacian(a, grid.dy, grid.dx)) < 1e-12
    assert float(mv["w"].abs().max()) == 0.0
    sv = batched_fft_solve(st, ops, 0.3)
    assert ogs.rel_l2(_np(sv["v"]), osp.solve_helmholtz_2d(b, lam, 0.3, 0.002)) < 1e-13
    assert torch.equal(sv["w"], st["w"])
    ex = batched_fft_exp_matvec(st, ops, 0.3)
    assert ogs.rel_l2(_np(ex["u"]), np.fft.ifft2(np.fft.fft2(a) * np.exp(0.3 * 0.01 * lam)).real) < 1e-13
    sh = stacked_fft_solve_shared_op({"u": st["u"], "v": st["v"]}, opu, 0.3)
    assert ogs.rel_l2(_np(sh["v"]), osp.solve_helmholtz_2d(b, lam, 0.3, 0.01)) < 1e-13
    d1 = diffusion_only_etd1(st["u"], 1.0, 0.3, opu)
    assert ogs.rel_l2(_np(d1), np.fft.ifft2(np.fft.fft2(a) * np.exp(1.0 * 0.01 * lam)).real) < 1e-12
    assert exact_cfl_dt(opu) == pytest.approx(0.9 * 2.0 / (0.01 * (4 / grid.dx ** 2 + 4 / grid.dy ** 2)), rel=1e-12)
    assert exact_cfl_dt(opu, "etd") == 0.9
    with pytest.raises(ValueError):
        exact_cfl_dt(opu, "rk9")
    cache = create_fft_cache(grid)
    pad = {"u": torch.from_numpy(np.pad(a, 1)).cuda(), "v": torch.from_numpy(np.pad(b, 1)).cuda(),
           "w": torch.from_numpy(np.pad(a, 1)).cuda()}
    rhs = diffusion_rhs_fft(pad, grid, {"u": 0.01, "v": 0.002}, cache)
    assert ogs.rel_l2(_np(rhs["u"])[1:-1, 1:-1], 0.01 * osp.periodic_fd_laplacian(a, grid.dy, grid.dx)) < 1e-12
    assert ogs.rel_l2(_np(rhs["v"])[1:-1, 1:-1], 0.002 * osp.periodic_fd_laplacian(b, grid.dy, grid.dx)) < 1e-12
    assert float(rhs["u"][0].abs().max()) == 0.0 and float(rhs["w"].abs().max()) == 0.0
    imp = apply_diffusion_fft(pad, grid, 0.3, {"u": 0.01, "v": 0.002}, cache)
    assert ogs.rel_l2(_np(imp["u"])[1:-1, 1:-1], osp.solve_helmholtz_2d(a, lam, 0.3, 0.01)) < 1e-13
