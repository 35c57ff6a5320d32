"""Mixed-radix Stockham lengths (96, 100, 192, 200, 250, 384, 500, 768, 1000:
radix-3 and radix-5 stages beside the power-of-two ones) on the GPU.

The periodic spectral pair is checked against numpy's FFT (a library routine
fixed by the mathematics) for every multiplier form, every mixed length in the
row and in the column role, beside power-of-two sides, with separable and full
2-D symbols; Gray-Scott IMEX-Strang against the oracle.  Tolerances (relative
L2): float64 1e-12, float32 3e-6 for one spectral multiply, Gray-Scott over 20
steps float64 1e-12, float32 1e-5."""
import numpy as np
import pytest
import torch

from oracle import gray_scott as ogs
from oracle import spectral as osp

pytestmark = pytest.mark.gpu

MIXED = [96, 100, 192, 200, 250, 384, 500, 768, 1000]
# every mixed length as ny and as nx, next to another mixed or a power-of-two side
SHAPES = [(96, 100), (100, 96), (192, 200), (200, 192), (250, 250), (384, 500), (500, 384), (768, 1000),
          (1000, 768), (250, 256), (256, 250), (16, 1000), (1000, 16), (96, 4096)]


def _np(t):
    return t.detach().double().cpu().numpy()


def _mult(nrep, L, alpha, beta):
    if nrep == -1:
        return alpha * L
    if nrep == -2:
        return np.exp(beta * L)
    return ((1 + alpha * L) / (1 - beta * L)) ** nrep


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 3e-6)])
@pytest.mark.parametrize("ny,nx", SHAPES)
def test_spectral_pair_mixed_lengths(gpu, ny, nx, dtype, tol):
    from moljax_amd.core import fft_solvers as fs
    rng = np.random.default_rng(ny * 7 + nx)
    a, b = rng.standard_normal((2, ny, nx))
    if dtype == torch.float32:
        a, b = a.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64)
    dy, dx = 1.0 / ny, 2.0 / nx
    lam = fs.laplacian_symbol_2d(ny, nx, dy, dx)
    L = osp.laplacian_symbol_2d(ny, nx, dy, dx)
    row = [0.01 * dy * dy, 0.02 * dy * dy, 0.005 * dy * dy, -0.003 * dy * dy]
    ta = torch.from_numpy(a).to("cuda", dtype)
    tb = torch.from_numpy(b).to("cuda", dtype)
    for nrep in (1, 2, 3, -1, -2):
        oa, ob = fs.spectral_pair(ta, tb, fs._tables(lam, dtype), [row], nrep=nrep)
        for f, out, (beta, alpha) in ((a, oa, (row[0], row[2])), (b, ob, (row[1], row[3]))):
            ref = np.fft.ifft2(_mult(nrep, L, alpha, beta) * np.fft.fft2(f)).real
            r = ogs.rel_l2(_np(out[0]), ref)
            assert r <= tol, (nrep, r)


@pytest.mark.parametrize("ny,nx", [(250, 250), (96, 200), (256, 250)])
def test_spectral_pair_full_2d_symbol(gpu, ny, nx):
    """A symbol that is not a sum of 1-D symbols (the lam2d kernels); real
    fields need a symbol even in (kx, ky) -> (-kx, -ky)."""
    from moljax_amd.core import fft_solvers as fs
    rng = np.random.default_rng(5)
    a = rng.standard_normal((ny, nx))
    L = osp.laplacian_symbol_2d(ny, nx, 1.0 / ny, 1.0 / nx)
    g = np.cos(2 * np.pi * np.arange(nx) / nx)[None, :] * np.cos(4 * np.pi * np.arange(ny) / ny)[:, None]
    L2 = L * (1.0 + 0.1 * g)
    lam = torch.from_numpy(L2).cuda()
    row = [1e-5, 2e-5, 3e-6, 1e-6]
    for nrep in (1, 2, 4):
        oa, _ = fs.spectral_pair(torch.from_numpy(a).cuda(), None, fs._tables(lam, torch.float64), [row], nrep=nrep)
        ref = np.fft.ifft2(_mult(nrep, L2, row[2], row[0]) * np.fft.fft2(a)).real
        assert ogs.rel_l2(_np(oa[0]), ref) < 1e-12, nrep


@pytest.mark.parametrize("n", MIXED)
def test_helmholtz_inverse_round_trip(gpu, n):
    """(I - c lam) applied by the matvec form after the solve returns the
    right-hand side (a check that needs no reference transform)."""
    from moljax_amd.core import fft_solvers as fs
    rng = np.random.default_rng(n)
    rhs = rng.standard_normal((3, n, n))
    lam = fs.laplacian_symbol_2d(n, n, 1.0 / n, 1.0 / n)
    c = 0.3 / n ** 2
    u = fs.solve_helmholtz_2d(torch.from_numpy(rhs).cuda(), lam, c, 1.0)
    lap_u, _ = fs.spectral_pair(u, None, fs._tables(lam, torch.float64), [[0, 0, 1.0, 0]] * 3, nrep=-1)
    back = _np(u) - c * _np(lap_u)
    assert ogs.rel_l2(back, rhs) < 1e-12


P = dict(Du=2e-5, Dv=1e-5, F=0.035, k=0.065)


@pytest.mark.parametrize("ny,nx", [(250, 250), (100, 100), (96, 200), (256, 250)])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_gray_scott_strang_mixed(gpu, ny, nx, dtype, tol):
    """20 IMEX-Strang steps on mixed-radix grids vs the oracle, ghost ring
    periodic."""
    from moljax_amd.core import Grid2D, create_gray_scott_periodic_fft
    from moljax_amd.core.stepping import integrate_imex_fixed_dt
    steps = 20
    grid = Grid2D.uniform(nx, ny, 0.0, 2.5 * nx / 256, 0.0, 2.5 * ny / 256)
    model, cache, diff = create_gray_scott_periodic_fft(grid, dtype=dtype, **P)
    y0 = ogs.initial_condition(ny, nx, 3)
    lam = osp.laplacian_symbol_2d(ny, nx, grid.dy, grid.dx)
    ref, _ = ogs.integrate(y0, 1.0, steps, P["F"], P["k"], P["Du"], P["Dv"], lam)
    y = {k: torch.from_numpy(v).to("cuda", dtype) for k, v in y0.items()}
    _, _, fin = integrate_imex_fixed_dt(model, y, 0.0, 0.0, 1.0, cache, diff, n_steps=steps, save_every=steps)
    for k in ("u", "v"):
        f = _np(fin[k])
        r = ogs.rel_l2(f, ref[k])
        print(f"{ny}x{nx} {dtype} {k} rel_l2={r:.3e}", flush=True)
        assert r <= tol, k
        assert np.array_equal(f, ogs.apply_bc_periodic(f))


def test_strang_batch_mixed_equals_alone(gpu):
    """A batch of 250^2 instances with their own parameters: each instance
    equals the same instance run alone, bit for bit."""
    from moljax_amd.core import Grid2D, create_gray_scott_periodic_fft
    from moljax_amd.core._arrays import inst_table, spectral_work
    from moljax_amd.core.stepping import _gs_rows, gs_imex_strang_inplace
    from moljax_amd.workloads import gray_scott_batch
    n, B = 250, 5
    grid = Grid2D.uniform(n, n, 0.0, 2.5, 0.0, 2.5)
    model, cache, diff = create_gray_scott_periodic_fft(grid, dtype=torch.float32, **P)
    U, V = gray_scott_batch(n, n, list(range(B)), dtype=torch.float32)
    rows = [list(r) for r in _gs_rows(model, 1.0, diff, B)]
    for i, r in enumerate(rows):
        r[4] += 0.001 * i
    lam = cache.laplacian_symbol
    Ub, Vb = U.clone(), V.clone()
    gs_imex_strang_inplace(Ub, Vb, lam, inst_table(rows), 7, work=spectral_work(torch.float32, B, n, n))
    for i in (0, 3):
        Ui, Vi = U[i:i + 1].clone(), V[i:i + 1].clone()
        gs_imex_strang_inplace(Ui, Vi, lam, inst_table([rows[i]]), 7, work=spectral_work(torch.float32, 1, n, n))
        assert torch.equal(Ui[0], Ub[i]) and torch.equal(Vi[0], Vb[i]), i
