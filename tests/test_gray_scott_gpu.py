"""Gray-Scott IMEX-Strang on the GPU vs the float64 oracle.

Tolerances (relative L2 against the oracle's reference dataflow):
float64 <= 1e-12 over 20 steps; float32 <= 1e-5 over 20 steps at the
config-2 parameters (Du=2e-5, Dv=1e-5, F=0.035, k=0.065, dt=1, [0,2.5]^2)."""
import numpy as np
import pytest
import torch

from oracle import gray_scott as ogs
from oracle import spectral as osp

pytestmark = pytest.mark.gpu

P = dict(Du=2e-5, Dv=1e-5, F=0.035, k=0.065)


def _np(t):
    return t.detach().double().cpu().numpy()


def _model(n, dtype):
    from moljax_amd.core import Grid2D, create_gray_scott_periodic_fft
    grid = Grid2D.uniform(n, n, 0.0, 2.5, 0.0, 2.5)
    return grid, create_gray_scott_periodic_fft(grid, dtype=dtype, **P)


def _oracle_run(n, steps, seed=0, dt=1.0):
    y0 = ogs.initial_condition(n, n, seed)
    lam = osp.laplacian_symbol_2d(n, n, 2.5 / n, 2.5 / n)
    out, _ = ogs.integrate(y0, dt, steps, P["F"], P["k"], P["Du"], P["Dv"], lam)
    return y0, out


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-13), (torch.float32, 2e-6)])
def test_single_step_matches_oracle(gpu, dtype, tol):
    from moljax_amd.core.stepping import imex_strang_step
    n = 64
    grid, (model, cache, diff) = _model(n, dtype)
    y0, ref = _oracle_run(n, 1)
    y = {k: torch.from_numpy(v).to("cuda", dtype) for k, v in y0.items()}
    out = imex_strang_step(model, y, 0.0, 1.0, cache, diff)
    for k in ("u", "v"):
        assert ogs.rel_l2(_np(out[k]), ref[k]) <= tol, k
    # inputs untouched (functional semantics)
    assert np.array_equal(_np(y["u"]), y0["u"].astype(np.float32) if dtype == torch.float32 else y0["u"])


@pytest.mark.parametrize("n", [64, 256])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_multi_step_matches_oracle(gpu, n, dtype, tol):
    from moljax_amd.core.stepping import integrate_imex_fixed_dt
    steps = 20
    grid, (model, cache, diff) = _model(n, dtype)
    y0, ref = _oracle_run(n, steps)
    y = {k: torch.from_numpy(v).to("cuda", dtype) for k, v in y0.items()}
    _, _, fin = integrate_imex_fixed_dt(model, y, 0.0, 0.0, 1.0, cache, diff, n_steps=steps, save_every=steps)
    for k in ("u", "v"):
        r = ogs.rel_l2(_np(fin[k]), ref[k])
        print(f"n={n} {dtype} {k} rel_l2={r:.3e}", flush=True)
        assert r <= tol, k


def test_ghost_ring_is_periodic(gpu):
    from moljax_amd.core.stepping import imex_strang_step
    n = 32
    grid, (model, cache, diff) = _model(n, torch.float64)
    y0, _ = _oracle_run(n, 0)
    out = imex_strang_step(model, {k: torch.from_numpy(v).cuda() for k, v in y0.items()}, 0.0, 1.0, cache, diff)
    for k in ("u", "v"):
        f = _np(out[k])
        assert np.array_equal(f, ogs.apply_bc_periodic(f))


def test_fixed_point(gpu):
    from moljax_amd.core.stepping import integrate_imex_fixed_dt
    n = 128
    grid, (model, cache, diff) = _model(n, torch.float32)
    y = {"u": torch.ones((n + 2, n + 2), device="cuda"), "v": torch.zeros((n + 2, n + 2), device="cuda")}
    _, _, fin = integrate_imex_fixed_dt(model, y, 0.0, 0.0, 1.0, cache, diff, n_steps=50, save_every=50)
    assert float((fin["u"] - 1).abs().max()) < 1e-6 and float(fin["v"].abs().max()) < 1e-6


def test_batch_equals_single_bitwise(gpu):
    """Batch invariance: each instance of a batch equals the same instance run
    alone, bit for bit (no cross-instance reductions)."""
    from moljax_amd.core.stepping import _gs_rows, gs_imex_strang_inplace
    n, B, steps = 64, 6, 15
    grid, (model, cache, diff) = _model(n, torch.float32)
    ys = [ogs.initial_condition(n, n, seed) for seed in range(B)]
    U = torch.from_numpy(np.stack([y["u"] for y in ys])).to("cuda", torch.float32)
    V = torch.from_numpy(np.stack([y["v"] for y in ys])).to("cuda", torch.float32)
    u1, v1 = U[2].clone(), V[2].clone()
    gs_imex_strang_inplace(U, V, cache.laplacian_symbol, _gs_rows(model, 1.0, diff, B), steps)
    gs_imex_strang_inplace(u1, v1, cache.laplacian_symbol, _gs_rows(model, 1.0, diff, 1), steps)
    assert torch.equal(U[2], u1) and torch.equal(V[2], v1)


@pytest.mark.parametrize("n,B", [(512, 32), (256, 1024)])
def test_two_row_tiles_per_workgroup_bitwise(gpu, n, B):
    """A 512^2 FP32 batch of 32, and a 256^2 batch of 1024 (its 512 MiB
    packed plane outgrows the Infinity Cache), run the row pass with two row
    tiles per workgroup (the second tile's loads in flight during the
    first); a lone instance runs one.  Each row's arithmetic is the same:
    bitwise equal."""
    from moljax_amd.core.stepping import _gs_rows, gs_imex_strang_inplace
    steps = 4
    grid, (model, cache, diff) = _model(n, torch.float32)
    ys = [ogs.initial_condition(n, n, seed % 3) for seed in range(B)]
    U = torch.from_numpy(np.stack([y["u"] for y in ys])).to("cuda", torch.float32)
    V = torch.from_numpy(np.stack([y["v"] for y in ys])).to("cuda", torch.float32)
    pick = [0, 17, B - 1]
    alone = []
    for i in pick:
        u1, v1 = U[i].clone(), V[i].clone()
        gs_imex_strang_inplace(u1, v1, cache.laplacian_symbol, _gs_rows(model, 1.0, diff, 1), steps)
        alone.append((u1, v1))
    gs_imex_strang_inplace(U, V, cache.laplacian_symbol, _gs_rows(model, 1.0, diff, B), steps)
    for i, (u1, v1) in zip(pick, alone):
        assert torch.equal(U[i], u1) and torch.equal(V[i], v1), i


def test_per_instance_parameters(gpu):
    """Instances with different (F, k, D, dt) in one launch each match the
    oracle run with their own parameters."""
    from moljax_amd.core.stepping import gs_imex_strang_inplace
    n, steps = 32, 10
    from moljax_amd.core import laplacian_symbol_2d
    lam = laplacian_symbol_2d(n, n, 2.5 / n, 2.5 / n)
    params = [(0.03, 0.06, 3e-5, 1e-5, 1.0), (0.05, 0.062, 1e-5, 2e-5, 0.5)]
    ys = [ogs.initial_condition(n, n, s) for s in range(2)]
    U = torch.from_numpy(np.stack([y["u"] for y in ys])).cuda()
    V = torch.from_numpy(np.stack([y["v"] for y in ys])).cuda()
    rows = [[0.5 * dt * Du, 0.5 * dt * Dv, 0, 0, F, k, dt] for (F, k, Du, Dv, dt) in params]
    gs_imex_strang_inplace(U, V, lam, rows, steps)
    lam_np = osp.laplacian_symbol_2d(n, n, 2.5 / n, 2.5 / n)
    for i, (F, k, Du, Dv, dt) in enumerate(params):
        ref, _ = ogs.integrate(ys[i], dt, steps, F, k, Du, Dv, lam_np)
        assert ogs.rel_l2(_np(U[i]), ref["u"]) < 1e-12
        assert ogs.rel_l2(_np(V[i]), ref["v"]) < 1e-12


def test_history_saves(gpu):
    from moljax_amd.core.stepping import integrate_imex_fixed_dt
    n = 32
    grid, (model, cache, diff) = _model(n, torch.float64)
    y0, _ = _oracle_run(n, 0)
    y = {k: torch.from_numpy(v).cuda() for k, v in y0.items()}
    t, hist, fin = integrate_imex_fixed_dt(model, y, 0.0, 5.0, 1.0, cache, diff, save_every=2)
    assert hist["u"].shape == (3, n + 2, n + 2)  # 6 steps (quirk: +1), saves after 2, 4, 6
    assert np.allclose(t.numpy(), [2.0, 4.0, 6.0])
    _, ref = _oracle_run(n, 4)
    assert ogs.rel_l2(_np(hist["u"][1]), ref["u"]) < 1e-12
    assert torch.equal(hist["u"][2], fin["u"])
