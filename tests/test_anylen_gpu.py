"""Grid sizes outside the power-of-two rule (the chirp-z kernels) on the GPU
vs the float64 oracle and library routines.

Tolerances (relative L2): float64 1e-12, float32 5e-6 for transforms and
solves; Gray-Scott IMEX over 20 steps float64 1e-12, float32 3e-5.  The raw
transforms are checked against scipy.fft (DST-I = dst type 1 / 2, DCT-II =
dct type 2 / 2, DST-II = dst type 2 / 2; the unnormalised DCT-III / DST-III
from scipy's type 3 with the first / last term restored), the periodic
solves against numpy FFTs and the FD-stencil residual."""
import numpy as np
import pytest
import scipy.fft as sfft
import torch

from oracle import etdrk4 as oe
from oracle import gray_scott as ogs
from oracle import nonperiodic as onp
from oracle import spectral as osp

pytestmark = pytest.mark.gpu
TOL = {torch.float64: 1e-12, torch.float32: 5e-6}
DTYPES = [torch.float64, torch.float32]


def _np(t):
    return t.detach().double().cpu().numpy()


def _in(x, dtype):
    if dtype == torch.float32:
        x = x.astype(np.float32).astype(np.float64)
    return x, torch.from_numpy(np.ascontiguousarray(x)).to("cuda", dtype)


def _raw(kind, inverse, x):
    """Unnormalised transforms along the last axis, from scipy.fft."""
    if kind == 0:
        return sfft.dst(x, type=1, axis=-1) / 2
    if kind == 1:
        if not inverse:
            return sfft.dct(x, type=2, axis=-1) / 2
        return (sfft.dct(x, type=3, axis=-1) + x[..., :1]) / 2
    if not inverse:
        return sfft.dst(x, type=2, axis=-1) / 2
    sign = (-1.0) ** np.arange(x.shape[-1])
    return (sfft.dst(x, type=3, axis=-1) + sign * x[..., -1:]) / 2


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n", [2, 3, 20, 50, 126, 513, 1000, 2048])
def test_raw_transforms(gpu, n, dtype):
    from moljax_amd.core import fft_nonperiodic as fn
    x, t = _in(np.random.default_rng(n).standard_normal((3, n)), dtype)
    for kind in (0, 1, 2):
        for inverse in (False, True):
            if kind == 0 and inverse:
                continue
            out = _np(fn._transform(t, kind, inverse, 1.0))
            r = ogs.rel_l2(out, _raw(kind, inverse, x))
            assert r <= TOL[dtype], (kind, inverse, r)
    # DST-I pair of the reference API
    assert ogs.rel_l2(_np(fn.dst_I_fast(t)), onp.dst_I_fast(x)) <= TOL[dtype]
    assert ogs.rel_l2(_np(fn.idst_I_fast(fn.dst_I_fast(t))), x) <= 4 * TOL[dtype]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n", [5, 50, 100, 300])
def test_helmholtz_1d_all_kinds(gpu, n, dtype):
    from moljax_amd.core import fft_nonperiodic as fn
    x, t = _in(np.random.default_rng(n + 1).standard_normal(n), dtype)
    dx = 1.0 / (n + 1)
    out = fn.solve_helmholtz_dirichlet(t, fn.laplacian_symbol_dirichlet(n, dx), 0.1, 0.5)
    ref = onp.solve_helmholtz_dirichlet(x, onp.laplacian_symbol_dirichlet(n, dx), 0.1, 0.5)
    assert ogs.rel_l2(_np(out), ref) <= TOL[dtype]
    dx = 2.0 / n
    out = fn.solve_helmholtz_neumann(t, fn.laplacian_symbol_neumann(n, dx), 0.1, 0.5)
    ref = onp.solve_helmholtz_neumann(x, onp.laplacian_symbol_neumann(n, dx), 0.1, 0.5)
    assert ogs.rel_l2(_np(out), ref) <= TOL[dtype]
    out = fn.solve_helmholtz_dirichlet_cc(t, fn.laplacian_symbol_dirichlet_cc(n, dx), 0.1, 0.5)
    ref = onp.solve_helmholtz_dirichlet_cc(x, onp.laplacian_symbol_dirichlet_cc(n, dx), 0.1, 0.5)
    assert ogs.rel_l2(_np(out), ref) <= TOL[dtype]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("ny,nx", [(20, 30), (50, 37), (100, 127), (3, 2)])
def test_helmholtz_and_poisson_2d(gpu, ny, nx, dtype):
    from moljax_amd.core import fft_nonperiodic as fn
    rhs, t = _in(np.random.default_rng(ny * nx).standard_normal((ny, nx)), dtype)
    dy, dx = 1.0 / (ny + 1), 1.5 / (nx + 1)
    ly, lx = onp.laplacian_symbol_dirichlet(ny, dy), onp.laplacian_symbol_dirichlet(nx, dx)
    u = fn.solve_helmholtz_dirichlet(t, (fn.laplacian_symbol_dirichlet(ny, dy), fn.laplacian_symbol_dirichlet(nx, dx)),
                                     0.05, 0.3)
    assert ogs.rel_l2(_np(u), onp.solve_helmholtz_2d("dirichlet", rhs, ly, lx, 0.05, 0.3)) <= TOL[dtype]
    if dtype == torch.float64:
        # (I - dt D lap_h) u = rhs with the vertex Dirichlet stencil
        res = _np(u) - 0.015 * onp.dirichlet_laplacian_2d(_np(u), dy, dx)
        assert ogs.rel_l2(res, rhs) < 1e-12
    dy, dx = 1.0 / ny, 1.5 / nx
    ly, lx = onp.laplacian_symbol_neumann(ny, dy), onp.laplacian_symbol_neumann(nx, dx)
    z = rhs - rhs.mean()
    p = fn.solve_poisson_neumann(torch.from_numpy(z).to("cuda", dtype),
                                 (fn.laplacian_symbol_neumann(ny, dy), fn.laplacian_symbol_neumann(nx, dx)))
    assert ogs.rel_l2(_np(p), onp.solve_poisson_2d("neumann", z, ly, lx)) <= 4 * TOL[dtype]
    ly, lx = onp.laplacian_symbol_dirichlet_cc(ny, dy), onp.laplacian_symbol_dirichlet_cc(nx, dx)
    u = fn.solve_helmholtz_dirichlet_cc(t, (fn.laplacian_symbol_dirichlet_cc(ny, dy),
                                            fn.laplacian_symbol_dirichlet_cc(nx, dx)), 0.05, 0.3)
    assert ogs.rel_l2(_np(u), onp.solve_helmholtz_2d("dirichlet_cc", rhs, ly, lx, 0.05, 0.3)) <= TOL[dtype]


@pytest.mark.parametrize("kind", ["dirichlet", "neumann"])
def test_etd1_any_length(gpu, kind):
    from moljax_amd.core import fft_nonperiodic as fn
    n, dt = 50, 0.3
    rng = np.random.default_rng(11)
    u, N = rng.standard_normal((3, n)), rng.standard_normal((3, n))
    if kind == "dirichlet":
        eig = 0.2 * onp.laplacian_symbol_dirichlet(n, 1.0 / (n + 1))
        out = fn.etd1_dirichlet(torch.from_numpy(u).cuda(), torch.from_numpy(N).cuda(), torch.from_numpy(eig).cuda(), dt)
        ref = onp.etd1_dirichlet(u, N, eig, dt)
    else:
        eig = 0.2 * onp.laplacian_symbol_neumann(n, 1.0 / n)
        out = fn.etd1_neumann(torch.from_numpy(u).cuda(), torch.from_numpy(N).cuda(), torch.from_numpy(eig).cuda(), dt)
        ref = onp.etd1_neumann(u, N, eig, dt)
    assert ogs.rel_l2(_np(out), ref) < 1e-12


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("ny,nx", [(24, 32), (50, 50), (100, 60), (7, 9), (33, 2)])
def test_periodic_helmholtz(gpu, ny, nx, dtype):
    from moljax_amd.core import fft_solvers as fs
    rhs, t = _in(np.random.default_rng(ny + nx).standard_normal((2, ny, nx)), dtype)
    dy, dx = 1.0 / ny, 2.0 / nx
    lam = fs.laplacian_symbol_2d(ny, nx, dy, dx)
    out = _np(fs.solve_helmholtz_2d(t, lam, 0.1, 0.4))
    lam_np = osp.laplacian_symbol_2d(ny, nx, dy, dx)
    for b in range(2):
        assert ogs.rel_l2(out[b], osp.solve_helmholtz_2d(rhs[b], lam_np, 0.1, 0.4)) <= TOL[dtype]
    if dtype == torch.float64:
        res = out[0] - 0.04 * osp.periodic_fd_laplacian(out[0], dy, dx)
        assert ogs.rel_l2(res, rhs[0]) < 1e-12


@pytest.mark.parametrize("nrep", [1, 2, 5, -1, -2])
def test_periodic_pair_multipliers(gpu, nrep):
    """Two fields with their own (alpha, beta) through every multiplier form:
    ((1 + a lam)/(1 - b lam))^nrep, the matvec a lam, and e^{b lam}."""
    from moljax_amd.core import fft_solvers as fs
    ny, nx = 45, 30
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal((ny, nx)), rng.standard_normal((ny, nx))
    lam = fs.laplacian_symbol_2d(ny, nx, 1.0 / ny, 1.0 / nx)
    tabs = fs._tables(lam, torch.float64)
    row = [0.01, 0.02, 0.005, -0.003]
    oa, ob = fs.spectral_pair(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), tabs, [row], nrep=nrep)
    L = osp.laplacian_symbol_2d(ny, nx, 1.0 / ny, 1.0 / nx)
    for f, out, (beta, alpha) in ((a, oa, (row[0], row[2])), (b, ob, (row[1], row[3]))):
        if nrep == -1:
            m = alpha * L
        elif nrep == -2:
            m = np.exp(beta * L)
        else:
            m = ((1 + alpha * L) / (1 - beta * L)) ** nrep
        ref = np.fft.ifft2(m * np.fft.fft2(f)).real
        assert ogs.rel_l2(_np(out[0]), ref) < 1e-12, nrep


def test_fft_cn_step_any_length(gpu):
    from moljax_amd.core.fft_solvers import fft_cn_step
    N, D, dt, n = 100, 0.1, 0.01, 10
    u0 = osp.sine_mode_ij(N)
    u = fft_cn_step(torch.from_numpy(u0).cuda(), D, 1.0 / N, dt, n_steps=n)
    ref = osp.fft_cn_integrate(u0, D, dt, n)
    assert ogs.rel_l2(_np(u), ref) < 1e-13


P = dict(Du=2e-5, Dv=1e-5, F=0.035, k=0.065)


@pytest.mark.parametrize("ny,nx", [(40, 60), (100, 100)])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 3e-5)])
def test_gray_scott_strang_any_length(gpu, ny, nx, dtype, tol):
    """20 IMEX-Strang steps.  A chirp-z transform runs two radix-2 FFTs of
    length L >= 2n, so its float32 rounding is about twice the Stockham
    path's; v, whose norm sits in the 20 x 20 square, shows it most
    (1.4e-5 at 100 x 100)."""
    from moljax_amd.core import Grid2D, create_gray_scott_periodic_fft
    from moljax_amd.core.stepping import integrate_imex_fixed_dt
    steps = 20
    grid = Grid2D.uniform(nx, ny, 0.0, 2.5 * nx / 100, 0.0, 2.5 * ny / 100)
    model, cache, diff = create_gray_scott_periodic_fft(grid, dtype=dtype, **P)
    y0 = ogs.initial_condition(ny, nx, 1)
    lam = osp.laplacian_symbol_2d(ny, nx, grid.dy, grid.dx)
    ref, _ = ogs.integrate(y0, 1.0, steps, P["F"], P["k"], P["Du"], P["Dv"], lam)
    y = {k: torch.from_numpy(v).to("cuda", dtype) for k, v in y0.items()}
    _, _, fin = integrate_imex_fixed_dt(model, y, 0.0, 0.0, 1.0, cache, diff, n_steps=steps, save_every=steps)
    for k in ("u", "v"):
        r = ogs.rel_l2(_np(fin[k]), ref[k])
        print(f"{ny}x{nx} {dtype} {k} rel_l2={r:.3e}", flush=True)
        assert r <= tol, k
        f = _np(fin[k])
        assert np.array_equal(f, ogs.apply_bc_periodic(f))


@pytest.mark.parametrize("scheme", ["euler", "ssprk2"])
def test_gray_scott_euler_ssprk2_any_length(gpu, scheme):
    from moljax_amd.core import Grid2D, create_gray_scott_periodic_fft
    from moljax_amd.core.stepping import imex_euler_step, imex_ssprk2_step
    ny, nx = 30, 50
    grid = Grid2D.uniform(nx, ny, 0.0, 0.75, 0.0, 0.45)
    model, cache, diff = create_gray_scott_periodic_fft(grid, dtype=torch.float64, **P)
    y0 = ogs.initial_condition(ny, nx, 2)
    lam = osp.laplacian_symbol_2d(ny, nx, grid.dy, grid.dx)
    y = {k: torch.from_numpy(v).cuda() for k, v in y0.items()}
    if scheme == "euler":
        out = imex_euler_step(model, y, 0.0, 1.0, cache, diff)
        ref = ogs.imex_euler_step(y0, 1.0, P["F"], P["k"], P["Du"], P["Dv"], lam)
    else:
        out = imex_ssprk2_step(model, y, 0.0, 1.0, cache, diff)
        ref = ogs.imex_ssprk2_step(y0, 1.0, P["F"], P["k"], P["Du"], P["Dv"], lam)
    for k in ("u", "v"):
        assert ogs.rel_l2(_np(out[k]), ref[k]) < 1e-12, k


def _rect_ic(ny, nx, seed):
    y = ogs.initial_condition(ny, nx, seed, square=10)
    return y["u"][1:-1, 1:-1].copy(), y["v"][1:-1, 1:-1].copy()


@pytest.mark.parametrize("ny,nx", [(48, 48), (40, 50), (37, 33)])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 3e-5)])
def test_etdrk4_any_length(gpu, ny, nx, dtype, tol):
    """10 coupled ETDRK4 steps with the FP64 contour coefficients on grids the
    power-of-two chain does not take (rectangular, odd) vs the oracle."""
    from moljax_amd.core.fft_integrators import etdrk4_coefficients, gs_etdrk4_inplace, spectral_symbol_1d
    lx, ly = spectral_symbol_1d(nx), spectral_symbol_1d(ny)
    F, k, Du, Dv, dt, steps = 0.035, 0.065, 2e-5, 1e-5, 1.0, 10
    u0, v0 = _rect_ic(ny, nx, 0)
    U = torch.from_numpy(u0).to("cuda", dtype)
    V = torch.from_numpy(v0).to("cuda", dtype)
    coef = etdrk4_coefficients(lx, ly, Du, Dv, dt, dtype)
    gs_etdrk4_inplace(U, V, coef, [[0, 0, 0, 0, F, k, dt]], steps)
    ru, rv = oe.gs_etdrk4(u0, v0, lx, ly, Du, Dv, F, k, dt, steps)
    eu, ev = ogs.rel_l2(_np(U), ru), ogs.rel_l2(_np(V), rv)
    print(f"ETDRK4 {ny}x{nx} {dtype}: rel_l2 u={eu:.2e} v={ev:.2e}", flush=True)
    assert eu <= tol and ev <= tol


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("ny,nx", [(40, 50), (37, 33)])
def test_etd_exp_any_length(gpu, order, ny, nx):
    from moljax_amd.core.fft_integrators import etd_phi_coefficients, gs_etd_exp_inplace, spectral_symbol_1d
    lx, ly = spectral_symbol_1d(nx), spectral_symbol_1d(ny)
    F, k, dt, steps = 0.035, 0.065, 1.0, 9
    u0, v0 = _rect_ic(ny, nx, 1)
    U, V = torch.from_numpy(u0).cuda(), torch.from_numpy(v0).cuda()
    coef = etd_phi_coefficients(lx, ly, 2e-5, 1e-5, dt, torch.float64)
    gs_etd_exp_inplace(U, V, coef, [[0, 0, 0, 0, F, k, dt]], order, steps)
    ru, rv, _ = oe.gs_etd(u0, v0, lx, ly, 2e-5, 1e-5, F, k, dt, steps, order)
    assert ogs.rel_l2(_np(U), ru) < 1e-12 and ogs.rel_l2(_np(V), rv) < 1e-12


@pytest.mark.parametrize("nx,lines", [(50, 3), (100, 1), (30, 2100)])
def test_periodic_helmholtz_1d_any_length(gpu, nx, lines):
    """fft_solvers.solve_helmholtz_1d on line lengths outside the
    power-of-two rule (a stack of lines, chunked at 2048 rows)."""
    from moljax_amd.core import fft_solvers as fs
    rhs = np.random.default_rng(nx + lines).standard_normal((lines, nx))
    dx = 1.0 / nx
    out = _np(fs.solve_helmholtz_1d(torch.from_numpy(rhs).cuda(), fs.laplacian_symbol_1d(nx, dx), 0.1, 0.3))
    ref = osp.solve_helmholtz_1d(rhs, osp.laplacian_symbol_1d(nx, dx), 0.1, 0.3)
    assert ogs.rel_l2(out, ref) < 1e-12


def test_any_length_api_inputs(gpu):
    """Non-contiguous views, extra batch dimensions and NumPy inputs on grids
    the chirp-z kernels serve, against the oracle."""
    from moljax_amd.core import fft_nonperiodic as fn
    from moljax_amd.core import fft_solvers as fs
    rng = np.random.default_rng(21)
    base = rng.standard_normal((50, 30))
    lam = fs.laplacian_symbol_2d(30, 50, 1.0 / 30, 1.0 / 50)
    lam_np = osp.laplacian_symbol_2d(30, 50, 1.0 / 30, 1.0 / 50)
    t = torch.from_numpy(base).cuda().t()  # (30, 50), non-contiguous
    assert not t.is_contiguous()
    out = _np(fs.solve_helmholtz_2d(t, lam, 0.1, 0.2))
    assert ogs.rel_l2(out, osp.solve_helmholtz_2d(base.T, lam_np, 0.1, 0.2)) < 1e-12
    stack = rng.standard_normal((2, 3, 30, 50))
    out = _np(fs.solve_helmholtz_2d(stack, lam, 0.1, 0.2))  # NumPy input, batch dims
    assert out.shape == stack.shape
    for i in range(2):
        for j in range(3):
            assert ogs.rel_l2(out[i, j], osp.solve_helmholtz_2d(stack[i, j], lam_np, 0.1, 0.2)) < 1e-12
    x = rng.standard_normal((7, 45)).astype(np.float32)
    d = fn.solve_helmholtz_neumann(x[2], fn.laplacian_symbol_neumann(45, 1.0 / 45), 0.1, 0.5)
    assert d.dtype == torch.float32
    ref = onp.solve_helmholtz_neumann(x[2].astype(np.float64), onp.laplacian_symbol_neumann(45, 1.0 / 45), 0.1, 0.5)
    assert ogs.rel_l2(_np(d), ref) < 5e-6
