"""moljax_amd.core.jit_kernels (the reference's jit_kernels.py) against the
float64 oracle: ETD1/ETD2/ETDRK4 kernels, phi1-3, the Helmholtz and batched
variants and the exponential matvec, on power-of-two and other lengths."""
import numpy as np
import pytest
import torch

from oracle import jit_kernels as oj
from oracle import spectral as osp

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().double().cpu().numpy()


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _lam1(n, dx, D=1.0, skew=False):
    lam = D * osp.laplacian_symbol_1d(n, dx)
    if skew:  # a real multiplier that is not even in k: only its even part acts on real data
        lam = lam - 0.05 * np.arange(n)
    return lam


@pytest.mark.parametrize("n", [64, 50])
def test_phi_functions(gpu, n):
    from moljax_amd.core import jit_kernels as jk
    z = np.concatenate([np.linspace(-40, 2, n - 4), [0.0, 5e-5, -5e-5, 1e-3]])
    for k, f in ((1, jk.phi1), (2, jk.phi2), (3, jk.phi3)):
        got = f(torch.from_numpy(z))
        assert got.is_cuda
        ref = getattr(oj, f"phi{k}")(z)
        assert np.max(np.abs(_np(got) - ref) / np.maximum(np.abs(ref), 1e-300)) < 1e-15


@pytest.mark.parametrize("n,skew", [(128, False), (90, False), (64, True)])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-13), (torch.float32, 2e-5)])
def test_etd_kernels_1d(gpu, n, skew, dtype, tol):
    from moljax_amd.core import jit_kernels as jk
    rng = np.random.default_rng(n)
    dx, dt = 1.0 / n, 0.02
    lam = _lam1(n, dx, 1e-3, skew)
    u, N0, N1, N2, N3 = (rng.standard_normal((3, n)) for _ in range(5))
    d = lambda a: torch.from_numpy(a).cuda().to(dtype)  # noqa: E731
    got = jk.etd1_kernel_1d(d(u), d(N0), lam, dt)
    assert got.dtype == dtype and got.shape == (3, n)
    assert _rel(_np(got), oj.etd1_kernel(u, N0, lam, dt)) < tol
    got = jk.etd2_kernel_1d(d(u), d(N0), d(N1), lam, dt)
    assert _rel(_np(got), oj.etd2_kernel_1d(u, N0, N1, lam, dt)) < tol
    got = jk.etdrk4_kernel_1d(d(u), lam, dt, (d(N0), d(N1), d(N2), d(N3)))
    assert _rel(_np(got), oj.etdrk4_kernel_1d(u, lam, dt, (N0, N1, N2, N3))) < tol
    got = jk.batched_exp_matvec_1d(d(u), torch.from_numpy(lam), dt)
    assert _rel(_np(got), oj.batched_exp_matvec_1d(u, lam, dt)) < tol
    assert jk.batched_etd1_kernel_1d is jk.etd1_kernel_1d and jk.batched_etd2_kernel_1d is jk.etd2_kernel_1d


@pytest.mark.parametrize("ny,nx", [(64, 128), (48, 40)])
def test_etd1_and_helmholtz_2d(gpu, ny, nx):
    from moljax_amd.core import jit_kernels as jk
    rng = np.random.default_rng(ny + nx)
    lam = osp.laplacian_symbol_2d(ny, nx, 1.0 / ny, 1.0 / nx)
    u, N = rng.standard_normal((2, ny, nx)), rng.standard_normal((2, ny, nx))
    got = jk.etd1_kernel_2d(torch.from_numpy(u[0]).cuda(), torch.from_numpy(N[0]).cuda(), 1e-3 * lam, 0.05)
    assert _rel(_np(got), oj.etd1_kernel(u[0], N[0], 1e-3 * lam, 0.05)) < 1e-13
    got = jk.batched_etd1_kernel_2d(torch.from_numpy(u).cuda(), torch.from_numpy(N).cuda(), 1e-3 * lam, 0.05)
    assert _rel(_np(got), oj.etd1_kernel(u, N, 1e-3 * lam, 0.05)) < 1e-13
    got = jk.helmholtz_solve_2d(torch.from_numpy(u[1]).cuda(), lam, 0.1, 2e-3)
    assert _rel(_np(got), oj.helmholtz_solve(u[1], lam, 0.1, 2e-3)) < 1e-13
    got = jk.batched_helmholtz_solve_2d(torch.from_numpy(u).cuda(), lam, 0.1, 2e-3)
    assert _rel(_np(got), oj.helmholtz_solve(u, lam, 0.1, 2e-3)) < 1e-13


@pytest.mark.parametrize("n", [256, 100])
def test_helmholtz_1d_and_batched(gpu, n):
    from moljax_amd.core import jit_kernels as jk
    rng = np.random.default_rng(7)
    lam = osp.laplacian_symbol_1d(n, 1.0 / n)
    r = rng.standard_normal((4, n))
    got = jk.helmholtz_solve_1d(torch.from_numpy(r[0]).cuda(), lam, 0.1, 1e-3)
    assert _rel(_np(got), oj.helmholtz_solve(r[0], lam, 0.1, 1e-3)) < 1e-13
    got = jk.batched_fft_solve_1d(torch.from_numpy(r).cuda(), lam, 0.1, 1e-3)
    assert _rel(_np(got), oj.helmholtz_solve(r, lam, 0.1, 1e-3)) < 1e-13


def test_multi_operator_and_integrator(gpu):
    from moljax_amd.core import jit_kernels as jk
    n, dt = 64, 0.01
    rng = np.random.default_rng(3)
    lam = np.stack([D * osp.laplacian_symbol_1d(n, 1.0 / n) for D in (1e-3, 5e-3, 2e-2)])
    u, N = rng.standard_normal((3, n)), rng.standard_normal((3, n))
    got = jk.multi_operator_etd1_1d(torch.from_numpy(u).cuda(), torch.from_numpy(N).cuda(), lam, dt)
    ref = np.stack([oj.etd1_kernel(u[i], N[i], lam[i], dt) for i in range(3)])
    assert _rel(_np(got), ref) < 1e-13
    integ = jk.make_etd1_integrator(lam[0], lambda v: v - v ** 3)
    uf, tf = integ(torch.from_numpy(0.5 * u[0]).cuda(), 0.25, dt, 20)
    ref = 0.5 * u[0]
    for _ in range(20):
        ref = oj.etd1_kernel(ref, ref - ref ** 3, lam[0], dt)
    assert _rel(_np(uf), ref) < 1e-12 and tf == pytest.approx(0.25 + 20 * dt, rel=1e-15)
    with pytest.raises(NotImplementedError):
        jk.etd1_kernel_1d(torch.from_numpy(u).cuda(), torch.from_numpy(N).cuda(), lam[0] + 1j, dt)


def test_reference_known_answers(gpu):
    """The reference's own checks (tests/test_jit_kernels.py:149-196,467-490):
    ETD1 with N = 0 decays sin(2 pi x) by e^{-4 pi^2 D dt} (1-D, rel 1e-6) and
    sin(2 pi x) sin(2 pi y) by e^{-8 pi^2 D dt} (2-D, rel 1e-5); a stacked
    call equals the single-line calls."""
    from moljax_amd.core import jit_kernels as jk
    nx, D, dt = 128, 0.01, 0.01
    x = np.arange(nx) / nx
    u0 = np.sin(2 * np.pi * x)
    lam = D * osp.laplacian_symbol_1d(nx, 1.0 / nx)
    u1 = _np(jk.etd1_kernel_1d(torch.from_numpy(u0).cuda(), torch.zeros(nx, dtype=torch.float64, device="cuda"),
                               lam, dt))
    assert _rel(u1, np.exp(-4 * np.pi ** 2 * D * dt) * u0) < 1e-6
    n2 = 64
    X, Y = np.meshgrid(np.arange(n2) / n2, np.arange(n2) / n2, indexing="ij")
    v0 = np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y)
    lam2 = D * osp.laplacian_symbol_2d(n2, n2, 1.0 / n2, 1.0 / n2)
    v1 = _np(jk.etd1_kernel_2d(torch.from_numpy(v0).cuda(), torch.zeros_like(torch.from_numpy(v0)).cuda(), lam2, dt))
    assert _rel(v1, np.exp(-8 * np.pi ** 2 * D * dt) * v0) < 1e-5
    rng = np.random.default_rng(11)
    U, N = rng.standard_normal((4, nx)), rng.standard_normal((4, nx))
    stacked = _np(jk.batched_etd1_kernel_1d(torch.from_numpy(U).cuda(), torch.from_numpy(N).cuda(), lam, dt))
    for i in range(4):
        one = _np(jk.etd1_kernel_1d(torch.from_numpy(U[i]).cuda(), torch.from_numpy(N[i]).cuda(), lam, dt))
        assert np.max(np.abs(one - stacked[i])) < 1e-14
