"""Newton-Krylov implicit solves (moljax/core/newton_krylov.py) on the
native MI355X driver.

newton_krylov_solve keeps the reference's signature and Newton semantics
(convergence tested before the update, three backtracking residual trials,
fallback damped step, newton_krylov.py:171-359).  The Krylov solve is the
library's left-preconditioned restarted GMRES(m) with FP64 CGS2 and Givens;
NKStats.lin_iters reports the true GMRES iteration count (one per Arnoldi
step), where the reference reports max_iters per Newton iteration
(newton_krylov.py:128).

Residuals are served by kernel: create_implicit_residual of a
reaction-diffusion model returns a residual object the driver recognises;
other residual callables raise NotImplementedError.
"""
import ctypes
from typing import Any, Callable, Dict, NamedTuple, Optional

import numpy as np
import torch

from .. import _hip
from .preconditioners import PRECOND_EXACT, PRECOND_JACOBI, PRECOND_NONE, IdentityPreconditioner, PrecondContext
from .state import StateDict, axpby


class NKParams(NamedTuple):
    """newton_krylov.py:29-38, plus the GMRES restart length and the
    precision of the Krylov operator (float64, or float32 for a mixed
    FP32-matvec / FP64-Arnoldi solve)."""
    max_newton_iters: int = 10
    max_krylov_iters: int = 50
    newton_tol: float = 1e-8
    krylov_tol: float = 1e-6
    damping: float = 1.0
    backtrack_factor: float = 0.5
    max_backtrack: int = 3
    min_residual_decrease: float = 1e-4
    gmres_restart: int = 30
    operator_dtype: torch.dtype = torch.float64


class NKStats(NamedTuple):
    """newton_krylov.py:41-46 (lin_iters: true total GMRES iterations)."""
    converged: bool
    newton_iters: int
    lin_iters: int
    final_res_norm: float


class NKResult(NamedTuple):
    """newton_krylov.py:49-52."""
    solution: StateDict
    stats: NKStats


class RDResidual:
    """R(u) = c u - g - b F(u) - fold for a ReactionDiffusionModel, evaluated
    by the mjx_rd_residual kernel; callable on a state {'u': array}.
      BE / CN (newton_krylov.py:362-403): c = 1, g = u_old, b = theta dt,
        fold = (1 - theta) dt F(u_old)  (theta = 1 / 1/2);
      BDF2 (newton_krylov.py:406-448): c = alpha0, g = -(alpha1 y_n +
        alpha2 y_{n-1}), b = beta, no fold (see create_bdf2_residual)."""

    def __init__(self, model, y_old: StateDict, dt: float, theta: float, lead: float = 1.0, g=None):
        self.model = model
        self.dt = float(dt)
        self.theta = float(theta)
        self.lead = float(lead)
        self.u_old = model.as_field(y_old["u"]) if g is None else model.as_field(g)
        self.fold = None
        if g is None and theta < 1.0:
            self.fold = axpby((1.0 - theta) * self.dt, model.rhs_field(self.u_old))

    def field(self, u: torch.Tensor) -> torch.Tensor:
        m = self.model
        u = m.as_field(u)
        r = torch.empty_like(u)
        reac = (ctypes.c_double * 4)(*m.reaction)
        rc = _hip.lib().mjx_rd_residual(m.ny, m.nx, m.bc_code, m.dy, m.dx, m.D, self.dt, self.theta, self.lead, reac,
                                        _hip.ptr(u), _hip.ptr(self.u_old), _hip.ptr(self.fold), _hip.ptr(r),
                                        _hip.stream_ptr())
        _hip.check(rc, "mjx_rd_residual")
        return r

    def __call__(self, y: StateDict) -> StateDict:
        return {"u": self.field(y["u"])}

    def jvp(self, u: torch.Tensor, v: torch.Tensor) -> torch.Tensor:
        """J(u) v by the hand-written JVP kernel (v's dtype: float32 or float64)."""
        m = self.model
        u = m.as_field(u)
        v = torch.as_tensor(v, device="cuda")
        if v.dtype not in (torch.float32, torch.float64):
            v = v.double()
        v = v.contiguous()
        w = torch.empty_like(v)
        reac = (ctypes.c_double * 4)(*m.reaction)
        rc = _hip.lib().mjx_rd_jvp(_hip.dtype_code(v.dtype), m.ny, m.nx, m.bc_code, m.dy, m.dx, m.D, self.dt,
                                   self.theta, self.lead, reac, _hip.ptr(u), _hip.ptr(v), _hip.ptr(w), _hip.stream_ptr())
        _hip.check(rc, "mjx_rd_jvp")
        return w


def create_implicit_residual(model, y_old: StateDict, t_new: float, dt: float, method: str = "be"):
    """newton_krylov.py:362-403."""
    if method not in ("be", "cn"):
        raise ValueError(f"Unknown method: {method}")
    if getattr(model, "mjx_kernel", None) != "reaction_diffusion":
        raise NotImplementedError("implicit residuals on the MI355X path serve reaction-diffusion models "
                                  "(create_reaction_diffusion_model)")
    return RDResidual(model, y_old, dt, 1.0 if method == "be" else 0.5)


def bdf2_coefficients(dt: float, dt_prev: float, consistent: bool = False):
    """newton_krylov.py:436-440: (alpha0, alpha1, alpha2, beta) of the
    variable-step BDF2 residual, omega = dt / dt_prev.  The reference pairs
    unnormalised alphas with the normalised beta = dt (1+w)/(1+2w), which
    integrates y' = F / alpha0; consistent=True uses beta = dt (the scheme
    its docstring states)."""
    omega = dt / dt_prev
    alpha0 = (1.0 + 2.0 * omega) / (1.0 + omega)
    alpha1 = -(1.0 + omega)
    alpha2 = omega ** 2 / (1.0 + omega)
    beta = dt if consistent else dt * (1.0 + omega) / (1.0 + 2.0 * omega)
    return alpha0, alpha1, alpha2, beta


def create_bdf2_residual(model, y_n: StateDict, y_nm1: StateDict, t_new: float, dt: float, dt_prev: float,
                         consistent: bool = False):
    """newton_krylov.py:406-448:
    R = alpha0 y + alpha1 y_n + alpha2 y_{n-1} - beta F(y)."""
    if getattr(model, "mjx_kernel", None) != "reaction_diffusion":
        raise NotImplementedError("implicit residuals on the MI355X path serve reaction-diffusion models "
                                  "(create_reaction_diffusion_model)")
    a0, a1, a2, beta = bdf2_coefficients(dt, dt_prev, consistent)
    g = axpby(-a1, model.as_field(y_n["u"]), -a2, model.as_field(y_nm1["u"]))
    return RDResidual(model, y_n, beta, 1.0, lead=a0, g=g)


def _precond_kind(preconditioner) -> int:
    if preconditioner is None or isinstance(preconditioner, IdentityPreconditioner):
        return PRECOND_NONE
    kind = getattr(preconditioner, "kind", None)
    if kind is None:
        raise NotImplementedError("preconditioner class not served by the native Newton-Krylov driver")
    return int(kind)


def newton_krylov_solve(residual_fn: Callable[[StateDict], StateDict], x0: StateDict, grid, params: Dict[str, Any],
                        preconditioner=None, nk_params: Optional[NKParams] = None, dt: float = 1.0) -> NKResult:
    """newton_krylov.py:171-359: solve R(x) = 0 from x0."""
    if nk_params is None:
        nk_params = NKParams()
    if not isinstance(residual_fn, RDResidual):
        raise NotImplementedError("newton_krylov_solve on the MI355X path takes residuals from "
                                  "create_implicit_residual of a reaction-diffusion model")
    m = residual_fn.model
    u = m.as_field(x0["u"]).clone()
    kind = _precond_kind(preconditioner)
    jac_iters = int(getattr(preconditioner, "n_iterations", 5))
    jac_omega = float(getattr(preconditioner, "omega", 0.6667))
    restart = int(nk_params.gmres_restart)
    lib = _hip.lib()
    work = torch.empty(lib.mjx_nk_workspace_bytes(m.ny, m.nx, restart), dtype=torch.uint8, device="cuda")
    tabs = m.precond_tables()
    lam = (ctypes.c_void_p * 4)(*[_hip.ptr(t) for t in tabs])
    fp = (ctypes.c_double * 8)(nk_params.newton_tol, nk_params.krylov_tol, nk_params.damping,
                               nk_params.backtrack_factor, nk_params.min_residual_decrease, jac_omega,
                               residual_fn.lead, float(dt))
    ip = (ctypes.c_int * 7)(_hip.dtype_code(nk_params.operator_dtype), kind, jac_iters,
                            int(nk_params.max_newton_iters), int(nk_params.max_krylov_iters), restart,
                            int(nk_params.max_backtrack))
    stats = (ctypes.c_double * 4)()
    reac = (ctypes.c_double * 4)(*m.reaction)
    rc = lib.mjx_rd_newton_krylov(m.ny, m.nx, m.bc_code, m.dy, m.dx, m.D, residual_fn.dt, residual_fn.theta, reac,
                                  _hip.ptr(u), _hip.ptr(residual_fn.u_old), _hip.ptr(residual_fn.fold),
                                  ctypes.cast(lam, ctypes.c_void_p), fp, ip, stats, _hip.ptr(work),
                                  _hip.stream_ptr())
    _hip.check(rc, "mjx_rd_newton_krylov")
    st = NKStats(converged=bool(stats[0]), newton_iters=int(stats[1]), lin_iters=int(stats[2]),
                 final_res_norm=float(stats[3]))
    return NKResult(solution={"u": u}, stats=st)
