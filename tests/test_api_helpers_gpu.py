"""Reference helpers kept on the drop-in surface (stepping, state,
operators, preconditioners, fft_integrators, dt_policy) against the float64
oracle or their defining formulas."""
import numpy as np
import pytest
import torch

from oracle import fd as ofd
from oracle import gray_scott as ogs
from oracle import spectral as osp

pytestmark = pytest.mark.gpu

GS = dict(Du=0.16, Dv=0.08, F=0.04, k=0.06)


def _np(t):
    return t.detach().double().cpu().numpy()


def _gs(n):
    from moljax_amd.core import Grid2D, create_gray_scott_periodic_fft
    grid = Grid2D.uniform(n, n, 0.0, 2.5, 0.0, 2.5)
    model, cache, diff = create_gray_scott_periodic_fft(grid, **GS)
    return grid, model, cache, diff


def test_step_explicit_and_error_doubling(gpu):
    """stepping.py:502-680: step_explicit, estimate_error_doubling and
    step_explicit_with_error against the oracle's explicit steps (interiors;
    the native steps refill the periodic ghosts)."""
    from moljax_amd.core import stepping as st
    n, dt = 32, 0.05
    grid, model, _, _ = _gs(n)
    y0 = ogs.initial_condition(n, n, 3)
    y = {k: torch.from_numpy(v).cuda() for k, v in y0.items()}
    steps = (ogs.euler_step, ogs.ssprk3_step, ogs.rk4_step)
    args = (GS["F"], GS["k"], GS["Du"], GS["Dv"], grid.dy, grid.dx)
    inner = (slice(1, -1), slice(1, -1))
    for m in range(3):
        got = st.step_explicit(model, y, 0.0, dt, m)
        ref = steps[m](y0, dt, *args)
        for f in ("u", "v"):
            assert ogs.rel_l2(_np(got[f])[inner], ref[f][inner]) < 1e-13, (m, f)
        yh, err = st.step_explicit_with_error(model, y, 0.0, dt, m)
        full = steps[m](y0, dt, *args)
        h1 = ogs.apply_bcs(steps[m](y0, dt / 2, *args))
        h2 = steps[m](h1, dt / 2, *args)
        for f in ("u", "v"):
            assert ogs.rel_l2(_np(yh[f])[inner], h2[f][inner]) < 1e-13
            assert np.max(np.abs(_np(err[f])[inner] - (h2[f] - full[f])[inner])) < 1e-12
    e = st.estimate_error_implicit(model, y, 0.0, dt, {"u": y["u"]}, {"u": 2 * y["u"]})
    assert torch.equal(e["u"], y["u"])
    with pytest.raises(ValueError):
        st.step_explicit(model, y, 0.0, dt, 3)


def test_implicit_steps_in_stepping_are_the_native_ones(gpu):
    from moljax_amd.core import stepping as st
    from moljax_amd.core.preconditioners import DSTDiffusionPreconditioner
    from moljax_amd.core.reaction_diffusion import be_step, create_reaction_diffusion_model, sine_mode_ic
    model = create_reaction_diffusion_model(31, D=1e-3)
    y = {"u": torch.from_numpy(sine_mode_ic(31, seed=1)).cuda()}
    P = DSTDiffusionPreconditioner(field_diffusivity_keys={"u": "D"})
    a, sa = st.be_step(model, y, 0.0, 1e-2, P)
    b, sb = be_step(model, y, 0.0, 1e-2, P)
    assert torch.equal(a["u"], b["u"]) and sa.lin_iters == sb.lin_iters


def test_diffusion_and_reaction_op_factories(gpu):
    """operators.py:392-444,623-641."""
    from moljax_amd.core import Grid2D, make_diffusion_op_2d, make_reaction_op
    grid = Grid2D.uniform(40, 24, 0.0, 2.0, 0.0, 1.0)
    rng = np.random.default_rng(2)
    u = rng.standard_normal((26, 42))
    v = rng.standard_normal((26, 42))
    state = {"u": torch.from_numpy(u).cuda(), "v": torch.from_numpy(v).cuda()}
    op = make_diffusion_op_2d("u", "Du")
    out = op.apply(state, grid, 0.0, {"Du": 0.3})
    ref = 0.3 * ofd.laplacian_2d(u, grid.dy, grid.dx)
    assert np.max(np.abs(_np(out["u"]) - ref)) < 1e-12 * np.max(np.abs(ref))
    assert torch.count_nonzero(out["v"]) == 0
    assert op.dt_bound(grid, {"Du": 0.3}) == pytest.approx(0.25 * grid.min_dx2 / (0.3 + 1e-14), rel=1e-15)
    rop = make_reaction_op({"v": lambda s, p: s["u"] * p["a"]})
    r = rop.apply(state, grid, 0.0, {"a": 2.0})
    assert torch.equal(r["v"], state["u"] * 2.0) and torch.count_nonzero(r["u"]) == 0


def test_block_jacobi(gpu):
    """preconditioners.py:69-118: r / (1 + dt D 4/min_dx^2 + dt max|v|/min_dx);
    in the native Newton-Krylov driver it runs as the identity (a constant
    left preconditioner leaves GMRES unchanged)."""
    from moljax_amd.core import Grid2D
    from moljax_amd.core.newton_krylov import NKParams
    from moljax_amd.core.preconditioners import BlockJacobiPreconditioner, IdentityPreconditioner, PrecondContext
    from moljax_amd.core.reaction_diffusion import be_step, create_reaction_diffusion_model, sine_mode_ic
    grid = Grid2D.uniform(16, 16, 0.0, 1.0, 0.0, 1.0)
    P = BlockJacobiPreconditioner(diffusion_keys={"u": "D"}, advection_keys={"u": "vel"})
    r = {"u": torch.ones((18, 18), dtype=torch.float64, device="cuda"), "w": torch.ones(3, device="cuda",
                                                                                        dtype=torch.float64)}
    out = P.apply(r, PrecondContext(grid=grid, dt=0.1, params={"D": 0.2, "vel": -3.0}))
    scale = 1.0 + 0.1 * 0.2 * 4.0 / grid.min_dx2 + 0.1 * 3.0 / grid.min_dx
    assert np.allclose(_np(out["u"]), 1.0 / scale, rtol=1e-15)
    assert torch.equal(out["w"], r["w"])
    model = create_reaction_diffusion_model(31, D=1e-3)
    y = {"u": torch.from_numpy(sine_mode_ic(31, seed=2)).cuda()}
    nkp = NKParams(krylov_tol=1e-8)
    a, sa = be_step(model, y, 0.0, 1e-2, BlockJacobiPreconditioner(diffusion_keys={"u": "D"}), nkp)
    b, sb = be_step(model, y, 0.0, 1e-2, IdentityPreconditioner(), nkp)
    assert (sa.newton_iters, sa.lin_iters) == (sb.newton_iters, sb.lin_iters)
    assert torch.equal(a["u"], b["u"])


def test_fft_integrators_imex_euler_step(gpu):
    """fft_integrators.py:545-585 on interior fields with the coupled
    Gray-Scott nonlinearity vs the oracle (FD symbol solve)."""
    from moljax_amd.core import Grid2D
    from moljax_amd.core.fft_integrators import gray_scott_nonlinear_rhs, imex_euler_step
    from moljax_amd.core.fft_operators import DiffusionOperator
    ny, nx, dt = 32, 48, 0.5
    grid = Grid2D.uniform(nx, ny, 0.0, 2.5, 0.0, 1.7)
    y0 = ogs.initial_condition(ny, nx, 4)
    s0 = {k: v[1:-1, 1:-1].copy() for k, v in y0.items()}
    ops = {"u": DiffusionOperator(grid, 2e-3)}  # v has no operator: explicit Euler
    out = imex_euler_step({k: torch.from_numpy(v).cuda() for k, v in s0.items()}, 0.0, dt, ops,
                          gray_scott_nonlinear_rhs(GS["F"], GS["k"]))
    N = ogs.reaction(s0, GS["F"], GS["k"])
    lam = osp.laplacian_symbol_2d(ny, nx, grid.dy, grid.dx)
    ref_u = osp.solve_helmholtz_2d(s0["u"] + dt * N["u"], lam, dt, 2e-3)
    assert ogs.rel_l2(_np(out["u"]), ref_u) < 1e-13
    assert ogs.rel_l2(_np(out["v"]), s0["v"] + dt * N["v"]) < 1e-15


def test_reaction_stiffness_and_state_helpers(gpu):
    from moljax_amd.core import (Grid2D, create_state_from_interior, estimate_reaction_stiffness,
                                 state_total_size, unflatten_from_vec, flatten_to_vec)
    rng = np.random.default_rng(5)
    u, v = rng.uniform(-1, 1, (10, 12)), rng.uniform(-2, 2, (10, 12))
    st = {"u": torch.from_numpy(u).cuda(), "v": torch.from_numpy(v).cuda()}
    um, vm = np.abs(u).max(), np.abs(v).max()
    want = max(0.1 + 0.05, vm ** 2 + 0.1, 2 * um * vm + 0.1 + 0.05, 0.7)
    assert estimate_reaction_stiffness(st, {"F": 0.1, "k": 0.05, "r": 0.7}) == pytest.approx(want, rel=1e-15)
    assert estimate_reaction_stiffness(st, {"r": 0.7}) == 0.7
    grid = Grid2D.uniform(12, 10, 0.0, 1.0, 0.0, 1.0)
    pad = create_state_from_interior({"u": u}, grid)
    assert pad["u"].shape == (12, 14) and torch.equal(pad["u"][1:-1, 1:-1].cpu(), torch.from_numpy(u))
    assert float(pad["u"][0].abs().sum()) == 0.0
    assert state_total_size(st) == 240
    vec, _ = flatten_to_vec(st)
    back = unflatten_from_vec(vec, st)
    assert torch.equal(back["u"], st["u"]) and torch.equal(back["v"], st["v"])
