"""Cell-centred Dirichlet (DST-II) and bc.py boundary values / fluxes in the
non-periodic solves, on the GPU vs the float64 oracle.

The solves must invert the library's 5-point Laplacian evaluated on the
ghost ring that moljax/core/bc.py:218-307 fills (Dirichlet ghost =
2 g - interior, Neumann ghost = interior -/+ 2 h q).  Tolerances (relative
L2): float64 1e-13, float32 2e-6; operator residuals 1e-11 (float64)."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import fd
from oracle import gray_scott as ogs
from oracle import nonperiodic as onp

pytestmark = pytest.mark.gpu
TOL = {torch.float64: 1e-13, torch.float32: 2e-6}


def _np(t):
    return t.detach().double().cpu().numpy()


def _in(x, dtype):
    if dtype == torch.float32:
        x = x.astype(np.float32).astype(np.float64)
    return x, torch.from_numpy(x).to("cuda", dtype)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("N", [16, 128, 1024])
def test_dst2_raw_transform_brute_force(gpu, N, dtype):
    """mjx_r2r_transform1d kind 2: DST-II and DST-III against their sums."""
    from moljax_amd import _hip
    x, t = _in(np.random.default_rng(N).standard_normal((3, N)), dtype)
    n = np.arange(N)
    S = np.sin(np.pi * np.outer(n + 1, 2 * n + 1) / (2 * N))  # S[k, m]
    lib = _hip.lib()
    for inverse, want in ((0, x @ S.T), (1, x @ S)):
        out = torch.empty_like(t)
        rc = lib.mjx_r2r_transform1d(_hip.dtype_code(dtype), 2, inverse, 3, N, _hip.ptr(t), N, _hip.ptr(out), N, 1.0,
                                     _hip.stream_ptr())
        _hip.check(rc, "transform1d")
        assert ogs.rel_l2(_np(out), want) <= 10 * TOL[dtype], inverse


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("shape", [(16, 16), (32, 128), (256, 64), (1024, 16)])
def test_cell_dirichlet_2d_matches_oracle(gpu, shape, dtype):
    from moljax_amd.core import fft_nonperiodic as fn
    ny, nx = shape
    dy, dx = 1.0 / ny, 2.0 / nx
    x, t = _in(np.random.default_rng(ny + nx).standard_normal((2, ny, nx)), dtype)
    cache = fn.create_nonperiodic_fft_cache_2d(ny, nx, dy, dx, "dirichlet_cc", dtype)
    out = fn.solve_helmholtz_dirichlet_cc(t, cache, 0.2, 0.01)
    ly, lx = onp.laplacian_symbol_dirichlet_cc(ny, dy), onp.laplacian_symbol_dirichlet_cc(nx, dx)
    for b in range(2):
        ref = onp.solve_helmholtz_2d("dirichlet_cc", x[b], ly, lx, 0.2, 0.01)
        err = ogs.rel_l2(_np(out[b]), ref)
        print(f"{shape} {dtype} instance {b}: rel_l2 {err:.2e}", flush=True)
        assert err <= TOL[dtype]
    p = fn.solve_poisson_dirichlet_cc(t[0], cache)
    assert ogs.rel_l2(_np(p), onp.solve_poisson_2d("dirichlet_cc", x[0], ly, lx)) <= 10 * TOL[dtype]


@pytest.mark.parametrize("N", [16, 256])
def test_cell_dirichlet_1d(gpu, N):
    from moljax_amd.core import fft_nonperiodic as fn
    dx = 1.0 / N
    x, t = _in(np.random.default_rng(3).standard_normal(N), torch.float64)
    lam = fn.create_nonperiodic_fft_cache(N, dx, "dirichlet_cc").laplacian_symbol
    out = fn.solve_helmholtz_dirichlet_cc(t, lam, 0.1, 0.5)
    ref = onp.solve_helmholtz_dirichlet_cc(x, onp.laplacian_symbol_dirichlet_cc(N, dx), 0.1, 0.5)
    assert ogs.rel_l2(_np(out), ref) <= 1e-13


@pytest.mark.parametrize("kind", ["dirichlet", "neumann"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("ny,nx", [(32, 64), (30, 45)])
def test_bc_values_invert_fd_operator(gpu, kind, dtype, ny, nx):
    """solve_helmholtz_bc / solve_poisson_bc with per-side values: equal to
    the oracle's lift + solve, and u - beta lap_h(fill(u)) = rhs with the
    fill of oracle.fd.apply_bc (bc.py) and the native fill + FD Laplacian."""
    from moljax_amd.core import FieldBCSpec, Grid2D
    from moljax_amd.core.bc_fill import fill_ghosts
    from moljax_amd.core.fft_nonperiodic import solve_helmholtz_bc, solve_poisson_bc
    from moljax_amd.core.operators import laplacian_2d
    grid = Grid2D.uniform(nx, ny, 0.0, 2.0, 0.0, 1.0)
    vals = {"left": 0.3, "right": -0.7, "bottom": 1.1, "top": 0.25}
    sides = (vals["left"], vals["right"], vals["bottom"], vals["top"])
    fn_ = lambda X, Y, t, p: dict(vals)
    spec = FieldBCSpec.dirichlet(fn_) if kind == "dirichlet" else FieldBCSpec.neumann(fn_)
    okind = "dirichlet_cc" if kind == "dirichlet" else "neumann"
    fk = 1 if kind == "dirichlet" else 2
    x, t = _in(np.random.default_rng(4).standard_normal((ny, nx)), dtype)
    dt, D = 0.3, 0.05
    u = solve_helmholtz_bc(t, grid, spec, dt, D)
    ref = onp.solve_bc_2d(okind, x, sides, grid.dy, grid.dx, dt * D)
    err = ogs.rel_l2(_np(u), ref)
    print(f"{kind} {dtype}: helmholtz rel_l2 {err:.2e}", flush=True)
    assert err <= TOL[dtype]
    if dtype == torch.float64:
        un = _np(u)
        lap = fd.laplacian_2d(fd.apply_bc(np.pad(un, 1), fk, sides, grid.dy, grid.dx), grid.dy, grid.dx)[1:-1, 1:-1]
        assert np.max(np.abs(un - dt * D * lap - x)) < 1e-11
        # the same with the native ghost fill and FD Laplacian
        pad = torch.zeros((ny + 2, nx + 2), dtype=dtype, device="cuda")
        pad[1:-1, 1:-1] = u
        fill_ghosts(pad, grid, spec)
        lap_n = _np(laplacian_2d(pad, grid))[1:-1, 1:-1]
        assert np.max(np.abs(un - dt * D * lap_n - x)) < 1e-11
    r = x
    if kind == "neumann":
        b = onp.bc_lift(x.shape, okind, sides, grid.dy, grid.dx)
        r = x - (x - b).mean()
    p = solve_poisson_bc(torch.from_numpy(r).to("cuda", dtype), grid, spec)
    pref = onp.solve_bc_2d(okind, r, sides, grid.dy, grid.dx)
    assert ogs.rel_l2(_np(p), pref) <= 50 * TOL[dtype]


def test_bc_values_1d_and_batch(gpu):
    """1-D lines take the left / right values; a batch of 2-D fields shares
    the side values, each instance equal to its lone solve."""
    from moljax_amd import _hip
    from moljax_amd.core import FieldBCSpec, Grid2D
    from moljax_amd.core.fft_nonperiodic import solve_helmholtz_bc
    from oracle.nonperiodic import bc_lift, laplacian_symbol_dirichlet_cc, solve_helmholtz_dirichlet_cc
    N, dx = 64, 1.0 / 64
    x = np.random.default_rng(6).standard_normal(N)
    lam = torch.from_numpy(laplacian_symbol_dirichlet_cc(N, dx)).cuda()
    inst = torch.tensor([[0.02, 0, 0, 0, 0, 0, 0, 0]], dtype=torch.float64, device="cuda")
    t = torch.from_numpy(x).cuda()
    out = torch.empty_like(t)
    sv = (ctypes.c_double * 2)(0.5, -0.25)
    rc = _hip.lib().mjx_r2r_solve1d_bc(1, 2, 1, N, _hip.ptr(t), N, _hip.ptr(out), N, _hip.ptr(lam), _hip.ptr(inst), 0,
                                       sv, dx, _hip.stream_ptr())
    _hip.check(rc, "solve1d_bc")
    b = bc_lift((N,), "dirichlet_cc", (0.5, -0.25, 0.0, 0.0), dx, dx)
    ref = solve_helmholtz_dirichlet_cc(x + 0.02 * b, laplacian_symbol_dirichlet_cc(N, dx), 0.02, 1.0)
    assert ogs.rel_l2(_np(out), ref) <= 1e-13
    grid = Grid2D.uniform(32, 16, 0.0, 1.0, 0.0, 1.0)
    spec = FieldBCSpec.dirichlet(lambda X, Y, t, p: 0.75)
    X = torch.from_numpy(np.random.default_rng(7).standard_normal((3, 16, 32))).cuda()
    ub = solve_helmholtz_bc(X, grid, spec, 0.1, 0.2)
    for i in range(3):
        assert torch.equal(ub[i], solve_helmholtz_bc(X[i], grid, spec, 0.1, 0.2))
    # Helmholtz power mode and DST-I do not take boundary values
    sv4 = (ctypes.c_double * 4)(0, 0, 0, 0)
    w = torch.empty_like(X)
    lamx = torch.from_numpy(laplacian_symbol_dirichlet_cc(32, 1 / 32)).cuda()
    lamy = torch.from_numpy(laplacian_symbol_dirichlet_cc(16, 1 / 16)).cuda()
    inst3 = inst.repeat(3, 1)
    for kind, mmode in ((0, 0), (2, 2)):
        rc = _hip.lib().mjx_r2r_solve2d_bc(1, kind, 3, 16, 32, _hip.ptr(X), 32, 512, _hip.ptr(w), 32, 512,
                                           _hip.ptr(lamx), _hip.ptr(lamy), _hip.ptr(inst3), mmode, sv4, 1 / 16, 1 / 32,
                                           _hip.ptr(torch.empty_like(X)), _hip.stream_ptr())
        assert rc != 0


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_etd1_nonperiodic(gpu, dtype):
    """etd1_dirichlet / etd1_neumann (fft_nonperiodic.py:363-437) and the
    cell-centred variant vs the oracle, batched lines."""
    from moljax_amd.core import fft_nonperiodic as fn
    tol = 1e-13 if dtype == torch.float64 else 2e-6
    rng = np.random.default_rng(11)
    dt, D = 0.2, 0.05
    for kind, N in (("dirichlet", 63), ("neumann", 64), ("dirichlet_cc", 64)):
        x, tu = _in(rng.standard_normal((3, N)), dtype)
        y, tn = _in(rng.standard_normal((3, N)), dtype)
        if kind == "dirichlet":
            eig = D * onp.laplacian_symbol_dirichlet(N, 1.0 / (N + 1))
            got = fn.etd1_dirichlet(tu, tn, torch.from_numpy(eig), dt)
            want = np.stack([onp.etd1_dirichlet(x[i], y[i], eig, dt) for i in range(3)])
        elif kind == "neumann":
            eig = D * onp.laplacian_symbol_neumann(N, 1.0 / N)
            got = fn.etd1_neumann(tu, tn, torch.from_numpy(eig), dt)
            want = np.stack([onp.etd1_neumann(x[i], y[i], eig, dt) for i in range(3)])
        else:
            eig = D * onp.laplacian_symbol_dirichlet_cc(N, 1.0 / N)
            z = dt * eig
            ph = onp._phi1_quadratic(z)
            want = np.stack([onp.idst2_ortho(np.exp(z) * onp.dst2_ortho(x[i]) + dt * ph * onp.dst2_ortho(y[i]))
                             for i in range(3)])
            got = fn.etd1_dirichlet_cc(tu, tn, torch.from_numpy(eig), dt)
        err = ogs.rel_l2(_np(got), want)
        print(f"etd1 {kind} {dtype}: rel_l2 {err:.2e}", flush=True)
        assert err <= tol, kind


def test_periodic_helmholtz_1d_and_fd_1d(gpu):
    """solve_helmholtz_1d (fft_solvers.py:222-258) vs the oracle; the 1-D FD
    operators (operators.py:34-152) on a sine mode against their exact
    discrete values, ghosts zero."""
    from moljax_amd.core import Grid1D
    from moljax_amd.core.fft_solvers import laplacian_symbol_1d, solve_helmholtz, solve_helmholtz_1d
    from moljax_amd.core.operators import d1_central_1d, d1_upwind_1d, d2_central_1d, laplacian_1d
    from oracle import spectral as osp
    n, dx = 64, 0.05
    lam = laplacian_symbol_1d(n, dx)
    x = np.random.default_rng(2).standard_normal((5, n))
    got = solve_helmholtz_1d(torch.from_numpy(x).cuda(), lam, 0.3, 0.02)
    want = np.stack([osp.solve_helmholtz_1d(x[i], osp.laplacian_symbol_1d(n, dx), 0.3, 0.02) for i in range(5)])
    assert ogs.rel_l2(_np(got), want) < 1e-13
    one = solve_helmholtz(torch.from_numpy(x[2]).cuda(), lam, 0.3, 0.02)
    assert ogs.rel_l2(_np(one), want[2]) < 1e-13
    g = Grid1D.uniform(n, 0.0, n * dx)
    k = 2 * np.pi * 3 / (n * dx)
    xc = (np.arange(n + 2) - 0.5) * dx
    f = torch.from_numpy(np.sin(k * xc)).cuda()
    interior = np.sin(k * xc[1:-1])
    d2 = _np(d2_central_1d(f, g))
    assert np.allclose(d2[1:-1], (2 * np.cos(k * dx) - 2) / dx ** 2 * interior, atol=1e-9) and d2[0] == d2[-1] == 0
    assert torch.equal(laplacian_1d(f, g), d2_central_1d(f, g))
    d1 = _np(d1_central_1d(f, g))
    assert np.allclose(d1[1:-1], np.sin(k * dx) / dx * np.cos(k * xc[1:-1]), atol=1e-10)
    lin = torch.from_numpy(3.0 * xc).cuda()
    vel = torch.from_numpy(np.where(np.arange(n + 2) % 2 == 0, 1.0, -1.0)).cuda()
    up = _np(d1_upwind_1d(lin, vel, g))
    assert np.allclose(up[1:-1], 3.0, atol=1e-12) and up[0] == 0
    upb = _np(d1_upwind_1d(f, -2.0, g))
    assert np.allclose(upb[1:-1], (np.sin(k * xc[2:]) - interior) / dx, atol=1e-12)
