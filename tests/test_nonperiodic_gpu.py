"""DST-I / DCT-II transforms and Dirichlet / Neumann solves on the GPU vs
the float64 oracle.  Tolerances (relative L2): float64 1e-13, float32 2e-6."""
import numpy as np
import pytest
import torch

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
@pytest.mark.parametrize("N", [15, 31, 127, 511, 1023])
def test_dst_transform(gpu, N, dtype):
    from moljax_amd.core import fft_nonperiodic as fn
    x, t = _in(np.random.default_rng(N).standard_normal((3, N)), dtype)
    assert ogs.rel_l2(_np(fn.dst_I_fast(t)), onp.dst_I_fast(x)) <= TOL[dtype]
    assert ogs.rel_l2(_np(fn.idst_I_fast(fn.dst_I_fast(t))), x) <= TOL[dtype]
    assert ogs.rel_l2(_np(fn.dst_I(t[0])), 2 * onp.dst_I_fast(x[0])) <= TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("N", [15, 63, 255])
def test_helmholtz_dirichlet_1d(gpu, N, dtype):
    from moljax_amd.core import fft_nonperiodic as fn
    dx = 1.0 / (N + 1)
    x, t = _in(np.random.default_rng(1).standard_normal(N), dtype)
    cache = fn.create_nonperiodic_fft_cache(N, dx, "dirichlet")
    out = fn.solve_helmholtz_dirichlet(t, cache.laplacian_symbol, 0.1, 0.5)
    ref = onp.solve_helmholtz_dirichlet(x, onp.laplacian_symbol_dirichlet(N, dx), 0.1, 0.5)
    assert ogs.rel_l2(_np(out), ref) <= TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("N", [16, 64, 1024])
def test_helmholtz_neumann_1d(gpu, N, dtype):
    from moljax_amd.core import fft_nonperiodic as fn
    dx = 2.0 / N
    x, t = _in(np.random.default_rng(2).standard_normal(N), dtype)
    lam = fn.laplacian_symbol_neumann(N, dx)
    out = fn.solve_helmholtz_neumann(t, lam, 0.1, 0.5)
    ref = onp.solve_helmholtz_neumann(x, onp.laplacian_symbol_neumann(N, dx), 0.1, 0.5)
    assert ogs.rel_l2(_np(out), ref) <= TOL[dtype]


def test_poisson_1d(gpu):
    from moljax_amd.core import fft_nonperiodic as fn
    N = 63
    dx = 1.0 / (N + 1)
    xg = dx * np.arange(1, N + 1)
    u = fn.solve_poisson_dirichlet(torch.from_numpy(-np.pi ** 2 * np.sin(np.pi * xg)).cuda(),
                                   fn.laplacian_symbol_dirichlet(N, dx))
    assert np.max(np.abs(_np(u) - np.sin(np.pi * xg))) < 5e-3
    r = np.random.default_rng(3).standard_normal(32)
    r -= r.mean()
    un = fn.solve_poisson_neumann(torch.from_numpy(r).cuda(), fn.laplacian_symbol_neumann(32, 0.1))
    assert ogs.rel_l2(_np(un), onp.solve_poisson_neumann(r, onp.laplacian_symbol_neumann(32, 0.1))) < 1e-12


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("kind,ny,nx", [("dirichlet", 15, 15), ("dirichlet", 63, 255), ("dirichlet", 511, 511),
                                        ("neumann", 16, 16), ("neumann", 128, 32), ("neumann", 512, 1024)])
def test_helmholtz_2d(gpu, kind, ny, nx, dtype):
    from moljax_amd.core import fft_nonperiodic as fn
    print(f"{kind} {ny}x{nx} {dtype}", flush=True)
    if kind == "dirichlet":
        dy, dx = 1.0 / (ny + 1), 1.0 / (nx + 1)
    else:
        dy, dx = 1.0 / ny, 1.0 / nx
    cache = fn.create_nonperiodic_fft_cache_2d(ny, nx, dy, dx, kind)
    x, t = _in(np.random.default_rng(ny + nx).standard_normal((ny, nx)), dtype)
    fn_solve = fn.solve_helmholtz_dirichlet if kind == "dirichlet" else fn.solve_helmholtz_neumann
    out = fn_solve(t, cache, 0.5, 1e-3)
    ly = _np(cache.lam_y)
    lx = _np(cache.lam_x)
    ref = onp.solve_helmholtz_2d(kind, x, ly, lx, 0.5, 1e-3)
    assert ogs.rel_l2(_np(out), ref) <= TOL[dtype]
    if dtype == torch.float64:
        lap = onp.dirichlet_laplacian_2d if kind == "dirichlet" else onp.neumann_laplacian_2d
        u = _np(out)
        assert np.max(np.abs(u - 0.5e-3 * lap(u, dy, dx) - x)) < 1e-11


def test_helmholtz_2d_batched_and_poisson(gpu):
    from moljax_amd.core import fft_nonperiodic as fn
    ny = nx = 63
    h = 1.0 / 64
    cache = fn.create_nonperiodic_fft_cache_2d(ny, nx, h, h, "dirichlet")
    rhs = np.random.default_rng(7).standard_normal((4, ny, nx))
    out = fn.solve_helmholtz_dirichlet(torch.from_numpy(rhs).cuda(), cache, 1.0, 0.01)
    for b in range(4):
        single = fn.solve_helmholtz_dirichlet(torch.from_numpy(rhs[b]).cuda(), cache, 1.0, 0.01)
        assert torch.equal(out[b], single)
    p = fn.solve_poisson_dirichlet(torch.from_numpy(rhs[0]).cuda(), (cache.lam_y, cache.lam_x))
    ref = onp.solve_poisson_2d("dirichlet", rhs[0], _np(cache.lam_y), _np(cache.lam_x))
    assert ogs.rel_l2(_np(p), ref) < 1e-12


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("kind,n", [("dirichlet", 511), ("dirichlet", 512), ("neumann", 512)])
def test_line_tiles_do_not_change_results(gpu, kind, n, dtype):
    # A lone instance runs its line passes on narrow workgroups (a quarter of
    # the default tile, so that the launch spreads over the CUs); a batch of
    # eight runs the default tile.  The per-line arithmetic is the same, so
    # each instance must match its lone solve bit for bit.  Dirichlet 511
    # and Neumann 512 use the Stockham kernels, Dirichlet 512 the chirp-z ones.
    from moljax_amd.core import fft_nonperiodic as fn
    h = 1.0 / (n + 1) if kind == "dirichlet" else 1.0 / n
    cache = fn.create_nonperiodic_fft_cache_2d(n, n, h, h, kind)
    solve = fn.solve_helmholtz_dirichlet if kind == "dirichlet" else fn.solve_helmholtz_neumann
    rhs = np.random.default_rng(n).standard_normal((8, n, n))
    out = solve(torch.from_numpy(rhs).to("cuda", dtype), cache, 1.0, 0.01)
    for b in (0, 5):
        single = solve(torch.from_numpy(rhs[b]).to("cuda", dtype), cache, 1.0, 0.01)
        assert torch.equal(out[b], single), b
    if dtype == torch.float64:
        ref = onp.solve_helmholtz_2d(kind, rhs[0], _np(cache.lam_y), _np(cache.lam_x), 1.0, 0.01)
        assert ogs.rel_l2(_np(out[0]), ref) < 1e-12


def test_unsupported_nonperiodic_size(gpu):
    from moljax_amd.core import fft_nonperiodic as fn
    with pytest.raises(ValueError):
        fn.solve_helmholtz_dirichlet(torch.zeros(4097, dtype=torch.float64, device="cuda"),
                                     fn.laplacian_symbol_dirichlet(4097, 0.1), 0.1, 0.1)
