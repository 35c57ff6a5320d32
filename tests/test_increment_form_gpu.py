"""Increment-form IMEX-Strang (mjx_gs_imex scheme 4) on the GPU.

The same operator as the plain Strang chain: the state stays physical and
each diffusion solve is applied as y + (H^k - 1)(y - (1, 0)), so only the
increment passes through the FP32 transforms.

Tolerances (relative L2 over both fields):
- against the float64 oracle over 20 steps: float64 <= 1e-12, float32 <= 1e-5;
- float64 increment form against the float64 plain chain over 200 steps
  <= 1e-12 (two roundings of one operator);
- config 2 (256^2, Du=2e-5, Dv=1e-5, F=0.035, k=0.065, dt=1): float32
  increment form within 1e-5 of the float64 chain to 3,000 steps, and at
  10,000 steps at most a third of the plain float32 chain's drift.  The
  float64 chain stands in for the oracle here (it is held to the oracle at
  1e-12 per 20 steps; 10,000 oracle steps on the CPU would take minutes).
"""
import numpy as np
import pytest
import torch

from oracle import gray_scott as ogs
from oracle import spectral as osp

pytestmark = pytest.mark.gpu

P = dict(Du=2e-5, Dv=1e-5, F=0.035, k=0.065)


def _np(t):
    return t.detach().double().cpu().numpy()


def _model(ny, nx, dtype):
    from moljax_amd.core import Grid2D, create_gray_scott_periodic_fft
    grid = Grid2D.uniform(nx, ny, 0.0, 2.5, 0.0, 2.5 * ny / nx)
    return create_gray_scott_periodic_fft(grid, dtype=dtype, **P)


def _run(y0, ny, nx, dtype, steps, increment_form, ghosts=True):
    from moljax_amd.core.stepping import integrate_imex_fixed_dt
    model, cache, diff = _model(ny, nx, dtype)
    y = {k: torch.from_numpy(v).to("cuda", dtype) for k, v in y0.items()}
    _, _, fin = integrate_imex_fixed_dt(model, y, 0.0, 0.0, 1.0, cache, diff, n_steps=steps, save_every=steps,
                                        increment_form=increment_form)
    return fin


def _rel_pair(a, b):
    num = np.sqrt(sum(np.sum((a[k] - b[k]) ** 2) for k in ("u", "v")))
    return num / np.sqrt(sum(np.sum(b[k] ** 2) for k in ("u", "v")))


@pytest.mark.parametrize("ny,nx", [(64, 64), (256, 256), (64, 128)])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_matches_oracle(gpu, ny, nx, dtype, tol):
    steps = 20
    y0 = ogs.initial_condition(ny, nx, 0)
    lam = osp.laplacian_symbol_2d(ny, nx, 2.5 / nx, 2.5 / nx)
    ref, _ = ogs.integrate(y0, 1.0, steps, P["F"], P["k"], P["Du"], P["Dv"], lam)
    fin = _run(y0, ny, nx, dtype, steps, True)
    for k in ("u", "v"):
        r = ogs.rel_l2(_np(fin[k])[1:-1, 1:-1], ref[k][1:-1, 1:-1])
        print(f"{ny}x{nx} {dtype} {k} rel_l2={r:.3e}", flush=True)
        assert r <= tol, k


@pytest.mark.parametrize("steps", [1, 2, 200])
def test_fp64_equals_plain_chain(gpu, steps):
    n = 128
    y0 = ogs.initial_condition(n, n, 3)
    a = _run(y0, n, n, torch.float64, steps, True)
    b = _run(y0, n, n, torch.float64, steps, False)
    r = _rel_pair({k: _np(a[k]) for k in a}, {k: _np(b[k]) for k in b})
    print(f"steps={steps} fp64 increment vs plain rel_l2={r:.3e}", flush=True)
    assert r <= 1e-12


def test_ghost_ring_is_periodic(gpu):
    n = 32
    y0 = ogs.initial_condition(n, n, 1)
    fin = _run(y0, n, n, torch.float64, 3, True)
    for k in ("u", "v"):
        f = _np(fin[k])
        assert np.array_equal(f, ogs.apply_bc_periodic(f))


def test_batch_equals_alone_bitwise(gpu):
    from moljax_amd.core import laplacian_symbol_2d
    from moljax_amd.core._arrays import inst_table
    from moljax_amd.core.stepping import GS_STRANG_INC, gs_imex_inplace
    from moljax_amd.workloads import gray_scott_batch
    n, B, steps = 64, 5, 30
    lam = laplacian_symbol_2d(n, n, 2.5 / n, 2.5 / n)
    rows = [[0.5 * P["Du"], 0.5 * P["Dv"], 0.0, 0.0, 0.02 + 0.01 * i, 0.06, 1.0] for i in range(B)]
    U, V = gray_scott_batch(n, n, list(range(B)), dtype=torch.float32)
    gs_imex_inplace(U, V, lam, rows, steps, GS_STRANG_INC)
    for i in range(B):
        u1, v1 = gray_scott_batch(n, n, [i], dtype=torch.float32)
        gs_imex_inplace(u1, v1, lam, inst_table([rows[i]]), steps, GS_STRANG_INC)
        assert torch.equal(u1[0], U[i]) and torch.equal(v1[0], V[i]), i


def test_config2_long_horizon(gpu):
    """Config 2 over 10,000 steps: the float32 increment form stays within
    1e-5 of float64 to 3,000 steps (the plain float32 chain leaves 1e-5
    between 2,000 and 3,000) and drifts at least 3x less at 10,000 steps.
    It does not hold 1e-5 to 10,000 steps (measured 1.2e-5 to 5.5e-5 over
    five seeds), which is why the float64 batch carries that horizon."""
    from moljax_amd.core.stepping import GS_STRANG, GS_STRANG_INC, gs_imex_inplace
    from moljax_amd.core import laplacian_symbol_2d
    from moljax_amd.core._arrays import inst_table
    from moljax_amd.workloads import gray_scott_batch
    n = 256
    lam = laplacian_symbol_2d(n, n, 2.5 / n, 2.5 / n)
    row = inst_table([[0.5 * P["Du"], 0.5 * P["Dv"], 0.0, 0.0, P["F"], P["k"], 1.0]])
    st = {}
    for name, dt, sc in (("ref", torch.float64, GS_STRANG), ("inc", torch.float32, GS_STRANG_INC),
                         ("plain", torch.float32, GS_STRANG)):
        st[name] = (gray_scott_batch(n, n, [0], dtype=dt), sc)
    errs = {}
    done = 0
    for c in (3000, 10000):
        for (U, V), sc in st.values():
            gs_imex_inplace(U, V, lam, row, c - done, sc)
        done = c
        (U64, V64), _ = st["ref"]
        den = torch.sqrt((U64[0, 1:-1, 1:-1] ** 2).sum() + (V64[0, 1:-1, 1:-1] ** 2).sum())
        for name in ("inc", "plain"):
            (U, V), _ = st[name]
            num = torch.sqrt(((U[0, 1:-1, 1:-1].double() - U64[0, 1:-1, 1:-1]) ** 2).sum()
                             + ((V[0, 1:-1, 1:-1].double() - V64[0, 1:-1, 1:-1]) ** 2).sum())
            errs[(name, c)] = float(num / den)
    print("config 2 rel_l2 vs float64: " + ", ".join(f"{k[0]}@{k[1]}={v:.3e}" for k, v in errs.items()), flush=True)
    assert errs[("inc", 3000)] <= 1e-5
    assert errs[("inc", 10000)] <= errs[("plain", 10000)] / 3


def test_rejects_non_power_of_two(gpu):
    from moljax_amd.core import laplacian_symbol_2d
    from moljax_amd.core.stepping import GS_STRANG_INC, gs_imex_inplace
    from moljax_amd.workloads import gray_scott_batch
    n = 96
    U, V = gray_scott_batch(n, n, [0], dtype=torch.float32)
    with pytest.raises(ValueError):
        gs_imex_inplace(U, V, laplacian_symbol_2d(n, n, 0.1, 0.1), [[1e-5, 5e-6, 0, 0, 0.035, 0.065, 1.0]], 2,
                        GS_STRANG_INC)
