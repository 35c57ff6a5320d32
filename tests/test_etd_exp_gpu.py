"""Coupled Gray-Scott ETD1 (exponential Euler) and ETD2 (exponential
Adams-Bashforth-2) on the GPU vs the float64 oracle
(moljax/core/fft_integrators.py:119-249).  Tolerances (relative L2): tables
1e-13; float64 trajectories 1e-12; float32 trajectories 1e-5."""
import numpy as np
import pytest
import torch

from oracle import etdrk4 as oe
from oracle import gray_scott as ogs
from oracle import spectral as osp

pytestmark = pytest.mark.gpu

DU, DV = 2e-5, 1e-5


def _np(t):
    return t.detach().double().cpu().numpy()


def _ic(n, seed):
    y = ogs.initial_condition(n, n, seed, square=max(2, n * 20 // 256))
    return y["u"][1:-1, 1:-1].copy(), y["v"][1:-1, 1:-1].copy()


@pytest.mark.parametrize("n", [32, 256])
def test_phi_tables_match_oracle(gpu, n):
    from moljax_amd.core.fft_integrators import etd_phi_coefficients, spectral_symbol_1d
    lam = spectral_symbol_1d(n)
    dt = 1.0
    coef = _np(etd_phi_coefficients(lam, lam, DU, DV, dt, torch.float64))
    h = n // 2
    L = lam[:, None] + lam[None, :]
    for half, D in ((slice(0, h + 1), DU), (slice(h + 1, n + 2), DV)):
        z = dt * D * L[:, :h + 1]
        assert ogs.rel_l2(coef[0][:, half], np.exp(z)) < 1e-13
        assert ogs.rel_l2(coef[1][:, half], dt * oe.phi1(z)) < 1e-13
        # the direct phi2 = (e^z - 1 - z)/z^2 of the reference cancels: one
        # rounding of e^z becomes ~eps/z^2, so compare within that bound
        bound = 4 * np.finfo(float).eps * (1 + np.abs(z) + np.exp(z)) / np.maximum(z * z, 1e-8) + 1e-15
        assert np.all(np.abs(coef[2][:, half] - dt * oe.phi2(z)) <= dt * bound), D


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("n", [64, 256])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_gs_etd_matches_oracle(gpu, order, n, dtype, tol):
    from moljax_amd.core.fft_integrators import etd_phi_coefficients, gs_etd_exp_inplace, spectral_symbol_1d
    lam = spectral_symbol_1d(n)
    F, k, dt, steps = 0.035, 0.065, 1.0, 9
    u0, v0 = _ic(n, 0)
    U = torch.from_numpy(u0).to("cuda", dtype)
    V = torch.from_numpy(v0).to("cuda", dtype)
    coef = etd_phi_coefficients(lam, lam, DU, DV, dt, dtype)
    gs_etd_exp_inplace(U, V, coef, [[0, 0, 0, 0, F, k, dt]], order, steps)
    ru, rv, _ = oe.gs_etd(u0, v0, lam, lam, DU, DV, F, k, dt, steps, order)
    eu, ev = ogs.rel_l2(_np(U), ru), ogs.rel_l2(_np(V), rv)
    print(f"ETD{order} n={n} {dtype}: rel_l2 u={eu:.2e} v={ev:.2e}", flush=True)
    assert eu <= tol and ev <= tol


@pytest.mark.parametrize("order", [1, 2])
def test_batched_per_instance_and_alone(gpu, order):
    from moljax_amd.core.fft_integrators import etd_phi_coefficients, gs_etd_exp_inplace, spectral_symbol_1d
    n, steps = 128, 7
    lam = spectral_symbol_1d(n)
    params = [(0.03, 0.06), (0.05, 0.062), (0.04, 0.058)]
    ics = [_ic(n, s) for s in range(3)]
    U = torch.from_numpy(np.stack([a for a, _ in ics])).cuda()
    V = torch.from_numpy(np.stack([b for _, b in ics])).cuda()
    coef = etd_phi_coefficients(lam, lam, DU, DV, 1.0, torch.float64)
    rows = [[0, 0, 0, 0, F, k, 1.0] for F, k in params]
    u1, v1 = U[1].clone(), V[1].clone()
    gs_etd_exp_inplace(U, V, coef, rows, order, steps)
    gs_etd_exp_inplace(u1, v1, coef, rows[1:2], order, steps)
    assert torch.equal(U[1], u1) and torch.equal(V[1], v1)
    for i, (F, k) in enumerate(params):
        ru, rv, _ = oe.gs_etd(ics[i][0], ics[i][1], lam, lam, DU, DV, F, k, 1.0, steps, order)
        assert ogs.rel_l2(_np(U[i]), ru) < 1e-12 and ogs.rel_l2(_np(V[i]), rv) < 1e-12


@pytest.mark.parametrize("n", [64, 48])
def test_etd2_history_modes_agree(gpu, n):
    """ETD2's N_{n-1} three ways: one call of 8 steps; 3 + 5 steps with the
    workspace's N^ (prev 2) — to rounding; 2 + 1 steps (prev 2), then 5 steps
    from the physical N(u_2) (prev 1) in a fresh workspace — vs the oracle."""
    from moljax_amd.core.fft_integrators import etd_phi_coefficients, gs_etd_exp_inplace, spectral_symbol_1d
    from moljax_amd.core.operators import gray_scott_reaction
    F, k = 0.04, 0.06
    lam = spectral_symbol_1d(n)
    u0, v0 = _ic(n, 2)
    coef = etd_phi_coefficients(lam, lam, DU, DV, 1.0, torch.float64)
    rows = [[0, 0, 0, 0, F, k, 1.0]]
    A = [torch.from_numpy(u0).cuda(), torch.from_numpy(v0).cuda()]
    gs_etd_exp_inplace(A[0], A[1], coef, rows, 2, 8)
    Bs = [torch.from_numpy(u0).cuda(), torch.from_numpy(v0).cuda()]
    w = gs_etd_exp_inplace(Bs[0], Bs[1], coef, rows, 2, 3)
    gs_etd_exp_inplace(Bs[0], Bs[1], coef, rows, 2, 5, prev=2, work=w)
    # a call boundary round-trips the state through physical space: rounding
    # only (the chirp-z transforms of n = 48 run four FFTs of 128 per
    # round trip against one FFT of 64, hence the wider bound)
    rb = 1e-14 if n & (n - 1) == 0 else 1e-13
    assert ogs.rel_l2(_np(Bs[0]), _np(A[0])) < rb and ogs.rel_l2(_np(Bs[1]), _np(A[1])) < rb
    C = [torch.from_numpy(u0).cuda(), torch.from_numpy(v0).cuda()]
    w = gs_etd_exp_inplace(C[0], C[1], coef, rows, 2, 2)
    n2 = gray_scott_reaction(C[0], C[1], F, k)  # N(u_2)
    gs_etd_exp_inplace(C[0], C[1], coef, rows, 2, 1, prev=2, work=w)  # u_3
    # continue from u_3 with N(u_2) given physically in a fresh workspace
    gs_etd_exp_inplace(C[0], C[1], coef, rows, 2, 5, prev=1, n_prev=n2)
    ru, rv, _ = oe.gs_etd(u0, v0, lam, lam, DU, DV, F, k, 1.0, 8, 2)
    for got in (A, C):
        assert ogs.rel_l2(_np(got[0]), ru) < 1e-12 and ogs.rel_l2(_np(got[1]), rv) < 1e-12


def test_step_api_and_integrate(gpu):
    """etd1_step / etd2_step / etd_integrate('etd1' | 'etd2') with
    DiffusionOperator (library FD symbol) against the oracle."""
    from moljax_amd.core import Grid2D
    from moljax_amd.core.fft_integrators import etd1_step, etd2_step, etd_integrate, gray_scott_nonlinear_rhs
    from moljax_amd.core.fft_operators import DiffusionOperator
    n = 64
    grid = Grid2D.uniform(n, n, 0.0, 2.5, 0.0, 2.5)
    ops = {"u": DiffusionOperator(grid, DU), "v": DiffusionOperator(grid, DV)}
    F, k = 0.035, 0.065
    nl = gray_scott_nonlinear_rhs(F, k)
    u0, v0 = _ic(n, 1)
    lam = osp.laplacian_symbol_1d(n, grid.dx)
    st = {"u": torch.from_numpy(u0).cuda(), "v": torch.from_numpy(v0).cuda()}
    one = etd1_step(st, 0.0, 1.0, ops, nl)
    r1u, r1v = oe.gs_etd1_step(u0, v0, lam, lam, DU, DV, F, k, 1.0)
    assert ogs.rel_l2(_np(one["u"]), r1u) < 1e-12 and ogs.rel_l2(_np(one["v"]), r1v) < 1e-12
    s, N = etd2_step(st, 0.0, 1.0, ops, nl)
    assert torch.equal(s["u"], one["u"])
    nu, nv = oe.gs_nonlinear(u0, v0, F, k)
    assert ogs.rel_l2(_np(N["u"]), nu) < 1e-14 and ogs.rel_l2(_np(N["v"]), nv) < 1e-14
    for _ in range(3):
        s, N = etd2_step(s, 0.0, 1.0, ops, nl, N)
    ru, rv, _ = oe.gs_etd(u0, v0, lam, lam, DU, DV, F, k, 1.0, 4, 2)
    assert ogs.rel_l2(_np(s["u"]), ru) < 1e-12 and ogs.rel_l2(_np(s["v"]), rv) < 1e-12
    for method, order in (("etd1", 1), ("etd2", 2)):
        t, hist = etd_integrate(st, (0.0, 6.0), 1.0, ops, nl, method=method, save_every=2)
        assert len(hist) == 4 and np.allclose(t.numpy(), [0, 2, 4, 6])
        for j, steps in ((1, 2), (3, 6)):
            ru, rv, _ = oe.gs_etd(u0, v0, lam, lam, DU, DV, F, k, 1.0, steps, order)
            assert ogs.rel_l2(_np(hist[j]["u"]), ru) < 1e-12, (method, j)
            assert ogs.rel_l2(_np(hist[j]["v"]), rv) < 1e-12, (method, j)
    t, hist = etd_integrate(st, (0.0, 2.0), 1.0, ops, nl)  # the reference's default method is ETD1
    assert ogs.rel_l2(_np(hist[-1]["u"]), oe.gs_etd(u0, v0, lam, lam, DU, DV, F, k, 1.0, 2, 1)[0]) < 1e-12
    with pytest.raises(ValueError):
        etd_integrate(st, (0.0, 2.0), 1.0, ops, nl, method="rk4")
