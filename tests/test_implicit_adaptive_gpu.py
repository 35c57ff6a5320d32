"""BDF2 Newton-Krylov steps and the implicit branch of adaptive_integrate on
the GPU vs the float64 oracle: Newton/GMRES counts equal, solutions to
1e-10; the adaptive BE / CN / BDF2 loops take the oracle's accept/reject
decisions with dt histories to 1e-10."""
import numpy as np
import pytest
import torch

from oracle import adaptive as oad
from oracle import gray_scott as ogs
from oracle import newton_krylov as onk

pytestmark = pytest.mark.gpu

AC = (-1.0, 0.0, 1.0, 0.0)


def _np(t):
    return t.detach().double().cpu().numpy()


@pytest.mark.parametrize("consistent", [False, True])
@pytest.mark.parametrize("bc,n,precond", [("dirichlet", 63, "dst"), ("periodic", 64, "dst"), ("dirichlet", 31, "jacobi")])
def test_bdf2_step_matches_oracle(gpu, bc, n, precond, consistent):
    from moljax_amd.core.newton_krylov import NKParams
    from moljax_amd.core.preconditioners import (DSTDiffusionPreconditioner, DiffusionPreconditioner,
                                                 FFTDiffusionPreconditioner)
    from moljax_amd.core.reaction_diffusion import bdf2_step, create_reaction_diffusion_model, sine_mode_ic
    D, dt, dt_prev = 2e-3, 1.5e-2, 1e-2
    model = create_reaction_diffusion_model(n, D=D, reaction=AC, bc=bc)
    u_prev = sine_mode_ic(n, seed=1)
    u = u_prev + dt_prev * onk.rhs(u_prev, D, AC, model.dy, model.dx, bc)
    if precond == "dst":
        P = DSTDiffusionPreconditioner(field_diffusivity_keys={"u": "D"}) if bc == "dirichlet" else \
            FFTDiffusionPreconditioner(field_diffusivity_keys={"u": "D"}, fft_cache=True)
    else:
        P = DiffusionPreconditioner(field_diffusivities={"u": "D"})
    y, st = bdf2_step(model, {"u": torch.from_numpy(u)}, {"u": torch.from_numpy(u_prev)}, 0.0, dt, dt_prev, P,
                      NKParams(krylov_tol=1e-8), consistent=consistent)
    ref = onk.bdf2_step(u, u_prev, dt, dt_prev, D, AC, model.dy, model.dx, bc, precond=precond,
                        consistent=consistent)
    print(f"bdf2 {bc} {precond} consistent={consistent}: newton {st.newton_iters}/{ref.newton_iters} "
          f"gmres {st.lin_iters}/{ref.gmres_iters}", flush=True)
    assert st.converged and (st.newton_iters, st.lin_iters) == (ref.newton_iters, ref.gmres_iters)
    assert ogs.rel_l2(_np(y["u"]), ref.u) < 1e-10


@pytest.mark.parametrize("method,t_end", [("be", 0.01), ("cn", 0.2), ("bdf2", 0.03)])
def test_adaptive_implicit_matches_oracle(gpu, method, t_end):
    from moljax_amd.core.dt_policy import PIDParams
    from moljax_amd.core.newton_krylov import NKParams
    from moljax_amd.core.preconditioners import DSTDiffusionPreconditioner
    from moljax_amd.core.reaction_diffusion import create_reaction_diffusion_model, sine_mode_ic
    from moljax_amd.core.stepping import IntegratorType, adaptive_integrate
    n, D = 31, 1e-2
    model = create_reaction_diffusion_model(n, D=D, reaction=AC, bc="dirichlet")
    u0 = sine_mode_ic(n, seed=0)
    pid = dict(atol=1e-4, rtol=1e-3)
    ref = oad.adaptive_integrate_implicit(u0, 0.0, t_end, 0.05, D, AC, model.dy, model.dx, "dirichlet",
                                          method=method, pid=oad.PIDParams(**pid), nk_kw=dict(krylov_tol=1e-8))
    m = {"be": IntegratorType.BE, "cn": IntegratorType.CN, "bdf2": IntegratorType.BDF2}[method]
    res = adaptive_integrate(model, {"u": torch.from_numpy(u0)}, 0.0, t_end, 0.05, method=m,
                             pid_params=PIDParams(**pid),
                             preconditioner=DSTDiffusionPreconditioner(field_diffusivity_keys={"u": "D"}),
                             nk_params=NKParams(krylov_tol=1e-8))
    print(f"{method}: oracle acc={ref.n_accepted} rej={ref.n_rejected}; gpu acc={res.n_accepted} "
          f"rej={res.n_rejected}", flush=True)
    assert ref.n_rejected > 0
    assert (res.n_accepted, res.n_rejected, res.status, res.n_steps) == \
        (ref.n_accepted, ref.n_rejected, ref.status, ref.n_steps)
    w = ref.n_steps
    assert np.allclose(res.dt_history.numpy()[:w], ref.dt_history[:w], rtol=1e-10, atol=0)
    assert ogs.rel_l2(_np(res.y_final["u"]), ref.y_final["u"]) < 1e-9
