"""Newton-Krylov (residual, hand JVP, GMRES, preconditioners) on the GPU vs
the float64 oracle.  With a float64 Krylov operator the Newton and GMRES
iteration counts equal the oracle's exactly and solutions agree to 1e-10
(relative L2); with a float32 operator the Newton residual still converges
below newton_tol (it is evaluated in float64)."""
import math

import numpy as np
import pytest
import torch

from oracle import gray_scott as ogs
from oracle import newton_krylov as onk

pytestmark = pytest.mark.gpu

AC = (-1.0, 0.0, 1.0, 0.0)  # f(u) = u - u^3


def _np(t):
    return t.detach().double().cpu().numpy()


def test_jvp_golden_norm_on_gpu(gpu):
    """benchmarks/results/jvp_vs_fd_sweep.json:3 through the JVP kernel."""
    import json, os
    from moljax_amd.core.newton_krylov import RDResidual
    from moljax_amd.core.reaction_diffusion import create_reaction_diffusion_model
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "nk_golden.json")))["jvp_vs_fd"]
    N, K = g["N"], g["reaction_K"]
    np.random.seed(g["seed"])
    u0 = np.abs(np.random.randn(N, N)) + 0.5
    v = np.random.randn(N, N)
    model = create_reaction_diffusion_model(N, L=1.0, D=g["D"], reaction=(-K, K, 0.0, 0.0), bc="periodic")
    res = RDResidual(model, {"u": torch.from_numpy(u0)}, g["dt"], 1.0)
    Jv = res.jvp(torch.from_numpy(u0), torch.from_numpy(v))
    assert float(torch.linalg.norm(Jv)) == pytest.approx(g["ad_jvp_norm"], rel=1e-12)
    Jv32 = res.jvp(torch.from_numpy(u0), torch.from_numpy(v).float())
    assert float(torch.linalg.norm(Jv32.double())) == pytest.approx(g["ad_jvp_norm"], rel=1e-5)


@pytest.mark.parametrize("bc,n", [("dirichlet", 31), ("periodic", 32)])
def test_residual_kernel(gpu, bc, n):
    from moljax_amd.core.newton_krylov import create_implicit_residual
    from moljax_amd.core.reaction_diffusion import create_reaction_diffusion_model
    model = create_reaction_diffusion_model(n, D=2e-3, reaction=(-1.0, 0.3, 1.0, 0.1), bc=bc)
    rng = np.random.default_rng(1)
    u, uo = rng.standard_normal((n, n)), rng.standard_normal((n, n))
    for method, theta in (("be", 1.0), ("cn", 0.5)):
        res = create_implicit_residual(model, {"u": torch.from_numpy(uo)}, 0.1, 0.05, method)
        r = _np(res({"u": torch.from_numpy(u)})["u"])
        F_old = onk.rhs(uo, 2e-3, model.reaction, model.dy, model.dx, bc)
        ref = onk.residual(u, uo, 0.05, theta, 2e-3, model.reaction, model.dy, model.dx, bc, F_old)
        assert np.max(np.abs(r - ref)) < 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("precond", ["dst", "jacobi", "none"])
@pytest.mark.parametrize("bc,n", [("dirichlet", 63), ("periodic", 64), ("dirichlet", 50), ("periodic", 40)])
def test_be_step_counts_match_oracle(gpu, precond, bc, n):
    from moljax_amd.core.newton_krylov import NKParams
    from moljax_amd.core.preconditioners import (DSTDiffusionPreconditioner, DiffusionPreconditioner,
                                                 FFTDiffusionPreconditioner, IdentityPreconditioner)
    from moljax_amd.core.reaction_diffusion import be_step, create_reaction_diffusion_model, sine_mode_ic
    D, dt = 1e-3, 1e-2
    model = create_reaction_diffusion_model(n, D=D, reaction=AC, bc=bc)
    u0 = sine_mode_ic(n, seed=0)
    if precond == "dst":
        P = DSTDiffusionPreconditioner(field_diffusivity_keys={"u": "D"}) if bc == "dirichlet" else \
            FFTDiffusionPreconditioner(field_diffusivity_keys={"u": "D"}, fft_cache=True)
    elif precond == "jacobi":
        P = DiffusionPreconditioner(field_diffusivities={"u": "D"})
    else:
        P = IdentityPreconditioner()
    nkp = NKParams(krylov_tol=1e-8, gmres_restart=30)
    y, st = be_step(model, {"u": torch.from_numpy(u0)}, 0.0, dt, P, nkp)
    guess = u0 + dt * onk.rhs(u0, D, AC, model.dy, model.dx, bc)
    ref = onk.newton_krylov(guess, u0, dt, 1.0, D, AC, model.dy, model.dx, bc, precond=precond,
                            krylov_tol=1e-8)
    print(f"{bc} {precond}: newton {st.newton_iters}/{ref.newton_iters} gmres {st.lin_iters}/{ref.gmres_iters}",
          flush=True)
    assert st.converged == ref.converged and st.newton_iters == ref.newton_iters
    assert st.lin_iters == ref.gmres_iters
    assert ogs.rel_l2(_np(y["u"]), ref.u) < 1e-10
    assert st.final_res_norm < 1e-8


def test_mixed_precision_operator_converges(gpu):
    from moljax_amd.core.newton_krylov import NKParams
    from moljax_amd.core.preconditioners import DSTDiffusionPreconditioner
    from moljax_amd.core.reaction_diffusion import be_step, create_reaction_diffusion_model, sine_mode_ic
    n, D, dt = 127, 1e-3, 1e-2
    model = create_reaction_diffusion_model(n, D=D, reaction=AC)
    u0 = sine_mode_ic(n, seed=3)
    P = DSTDiffusionPreconditioner(field_diffusivity_keys={"u": "D"})
    y64, s64 = be_step(model, {"u": torch.from_numpy(u0)}, 0.0, dt, P, NKParams(krylov_tol=1e-8))
    y32, s32 = be_step(model, {"u": torch.from_numpy(u0)}, 0.0, dt, P,
                       NKParams(krylov_tol=1e-8, operator_dtype=torch.float32))
    print(f"fp64 op: newton {s64.newton_iters} gmres {s64.lin_iters}; fp32 op: newton {s32.newton_iters} "
          f"gmres {s32.lin_iters} res {s32.final_res_norm:.2e}", flush=True)
    assert s32.converged and s32.final_res_norm < 1e-8
    assert ogs.rel_l2(_np(y32["u"]), _np(y64["u"])) < 1e-9


def test_cn_step_matches_oracle(gpu):
    from moljax_amd.core.newton_krylov import NKParams
    from moljax_amd.core.preconditioners import DSTDiffusionPreconditioner
    from moljax_amd.core.reaction_diffusion import cn_step, create_reaction_diffusion_model, sine_mode_ic
    n, D, dt = 31, 5e-3, 2e-2
    model = create_reaction_diffusion_model(n, D=D, reaction=AC)
    u0 = sine_mode_ic(n, seed=5)
    y, st = cn_step(model, {"u": torch.from_numpy(u0)}, 0.0, dt,
                    DSTDiffusionPreconditioner(field_diffusivity_keys={"u": "D"}), NKParams(krylov_tol=1e-8))
    F0 = onk.rhs(u0, D, AC, model.dy, model.dx, "dirichlet")
    ref = onk.newton_krylov(u0 + dt * F0, u0, dt, 0.5, D, AC, model.dy, model.dx, "dirichlet", precond="dst",
                            F_old=F0)
    assert st.newton_iters == ref.newton_iters and st.lin_iters == ref.gmres_iters
    assert ogs.rel_l2(_np(y["u"]), ref.u) < 1e-10


def test_dot_kernel_deterministic_and_exact(gpu):
    import ctypes
    from moljax_amd import _hip
    rng = np.random.default_rng(11)
    n, k = 100003, 11
    V = rng.standard_normal((k, n))
    w = rng.standard_normal(n)
    Vt, wt = torch.from_numpy(V).cuda(), torch.from_numpy(w).cuda()
    part = torch.empty(_hip.lib().mjx_dot_partial_elems(k), dtype=torch.float64, device="cuda")
    outs = []
    for _ in range(2):
        out = torch.empty(k, dtype=torch.float64, device="cuda")
        _hip.check(_hip.lib().mjx_dot_f64(n, k, _hip.ptr(Vt), _hip.ptr(wt), _hip.ptr(out), _hip.ptr(part),
                                          _hip.stream_ptr()), "dot")
        outs.append(out.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])
    exact = np.array([math.fsum(V[i] * w) for i in range(k)])
    assert np.max(np.abs(outs[0] - exact)) < 1e-12 * np.sqrt(n)


@pytest.mark.parametrize("n_iter", [0, 1, 5])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-14), (torch.float32, 1e-6)])
def test_diffusion_preconditioner_apply(gpu, n_iter, dtype, tol):
    """DiffusionPreconditioner.apply (preconditioners.py:121-198, native
    mjx_jacobi2d) against the recurrence on whole padded arrays with the
    oracle's 5-point Laplacian (zero on the ghost ring); the FFT factory's
    preconditioner inverts the periodic operator exactly."""
    from oracle import fd
    from oracle import spectral as osp
    from moljax_amd.core import Grid2D, create_fft_cache
    from moljax_amd.core.preconditioners import (PrecondContext, create_diffusion_preconditioner,
                                                 create_fft_preconditioner)
    grid = Grid2D.uniform(64, 32, 0.0, 1.2, 0.0, 2.0)
    rng = np.random.default_rng(21)
    r = {"u": rng.standard_normal((2, 34, 66)), "v": rng.standard_normal((34, 66))}
    params = {"Du": 0.03, "Dv": 0.0}
    dt = 0.7
    P = create_diffusion_preconditioner({"u": "Du", "v": "Dv"}, n_iterations=n_iter)
    ctx = PrecondContext(grid=grid, dt=dt, params=params)
    out = P.apply({k: torch.from_numpy(v).to("cuda", dtype) for k, v in r.items()}, ctx)
    D = params["Du"]
    diag = 1.0 + dt * D * (2.0 / grid.dx ** 2 + 2.0 / grid.dy ** 2)
    for b in range(2):
        rf = r["u"][b].astype(np.float32).astype(np.float64) if dtype == torch.float32 else r["u"][b]
        x = rf
        for _ in range(n_iter):
            x = P.omega * (rf + dt * D * fd.laplacian_2d(x, grid.dy, grid.dx)) / diag + (1.0 - P.omega) * x
        err = ogs.rel_l2(out["u"][b].double().cpu().numpy(), x)
        assert err <= tol, (b, err)
    assert torch.equal(out["v"].cpu(), torch.from_numpy(r["v"]).to(dtype))  # D = 0: unchanged
    if n_iter == 1 and dtype == torch.float64:
        F = create_fft_preconditioner({"u": "Du"}, create_fft_cache(grid))
        y = F.apply({"u": torch.from_numpy(r["u"][0][1:-1, 1:-1]).cuda()}, ctx)["u"]
        lam = osp.laplacian_symbol_2d(32, 64, grid.dy, grid.dx)
        assert ogs.rel_l2(y.cpu().numpy(), osp.solve_helmholtz_2d(r["u"][0][1:-1, 1:-1], lam, dt, D)) < 1e-13


def test_c3_512_be_steps_match_oracle(gpu):
    """The Newton-Krylov workload at its own size: 512^2 interior vertices,
    homogeneous Dirichlet (DST-I length 513, the chirp-z kernels), u_t = D
    lap u - u^3 + u with D = 1e-3, backward Euler dt = 1e-2, GMRES(30) rtol
    1e-8, DST preconditioner, two steps.  FP64 operator: Newton and GMRES
    counts equal to the oracle's and the state within 1e-10; FP32 operator:
    converged below newton_tol with the state within 1e-6."""
    from moljax_amd.core.newton_krylov import NKParams
    from moljax_amd.core.preconditioners import DSTDiffusionPreconditioner
    from moljax_amd.core.reaction_diffusion import be_step, create_reaction_diffusion_model, sine_mode_ic
    n, D, dt = 512, 1e-3, 1e-2
    model = create_reaction_diffusion_model(n, D=D, reaction=AC, bc="dirichlet")
    u0 = sine_mode_ic(n, seed=0)
    P = DSTDiffusionPreconditioner(field_diffusivity_keys={"u": "D"})
    ref_u = u0
    refs = []
    for s in range(2):
        r = onk.be_step(ref_u, dt, D, AC, model.dy, model.dx, "dirichlet", precond="dst", krylov_tol=1e-8)
        refs.append(r)
        ref_u = r.u
    for op_dtype in (torch.float64, torch.float32):
        y = {"u": torch.from_numpy(u0).cuda()}
        for s in range(2):
            y, st = be_step(model, y, s * dt, dt, P, NKParams(krylov_tol=1e-8, gmres_restart=30,
                                                                 operator_dtype=op_dtype))
            r = refs[s]
            err = ogs.rel_l2(_np(y["u"]), r.u)
            print(f"c3 512^2 step {s} {op_dtype}: newton {st.newton_iters}/{r.newton_iters} "
                  f"gmres {st.lin_iters}/{r.gmres_iters} rel_l2 {err:.2e} res {st.final_res_norm:.2e}", flush=True)
            assert st.converged and st.final_res_norm < 1e-8
            if op_dtype == torch.float64:
                assert (st.newton_iters, st.lin_iters) == (r.newton_iters, r.gmres_iters)
                assert err < 1e-10
            else:
                assert err < 1e-6


@pytest.mark.parametrize("max_backtrack", [0, 1, 3])
def test_rejected_backtracking_trials_match_oracle(gpu, max_backtrack):
    """Every backtracking trial rejected (min_residual_decrease = 1.2 asks
    ||R(x_t)|| < (1 - 1.2 alpha) r_bt, which the trials at alpha = 1, 0.5,
    0.25 miss by 20 % or more), so each Newton iteration takes the damped
    fallback x + damping dx (newton_krylov.py:272-316), decided on the device
    by the launch that follows each trial; and no trials at all."""
    from moljax_amd.core.newton_krylov import NKParams
    from moljax_amd.core.preconditioners import DSTDiffusionPreconditioner
    from moljax_amd.core.reaction_diffusion import be_step, create_reaction_diffusion_model, sine_mode_ic
    n, D, dt, bc = 63, 1e-3, 1e-2, "dirichlet"
    model = create_reaction_diffusion_model(n, D=D, reaction=AC, bc=bc)
    u0 = sine_mode_ic(n, seed=5)
    P = DSTDiffusionPreconditioner(field_diffusivity_keys={"u": "D"})
    nkp = NKParams(krylov_tol=1e-8, gmres_restart=30, min_residual_decrease=1.2, max_backtrack=max_backtrack)
    y, st = be_step(model, {"u": torch.from_numpy(u0)}, 0.0, dt, P, nkp)
    guess = u0 + dt * onk.rhs(u0, D, AC, model.dy, model.dx, bc)
    ref = onk.newton_krylov(guess, u0, dt, 1.0, D, AC, model.dy, model.dx, bc, precond="dst", krylov_tol=1e-8,
                            min_decrease=1.2, max_backtrack=max_backtrack)
    print(f"max_backtrack {max_backtrack}: newton {st.newton_iters}/{ref.newton_iters} "
          f"gmres {st.lin_iters}/{ref.gmres_iters}", flush=True)
    assert st.converged == ref.converged and st.newton_iters == ref.newton_iters
    assert st.lin_iters == ref.gmres_iters
    assert ogs.rel_l2(_np(y["u"]), ref.u) < 1e-10
    assert abs(st.final_res_norm - ref.res_norm) <= 1e-6 * ref.res_norm + 1e-14
