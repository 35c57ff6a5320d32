"""moljax_amd.core: drop-in host API of moljax/core over the MI355X kernels."""
from .grid import Grid1D, Grid2D, GridType
from .bc import BCType, FieldBCSpec, BCSpec, create_default_bc_spec
from .bc_fill import apply_bc
from .state import (StateDict, tree_add, tree_sub, tree_scale, tree_axpy, tree_zeros_like, tree_ones_like,
                    tree_norm2, tree_norm_inf, tree_vdot, scaled_error_norm, flatten_to_vec, unflatten_from_vec,
                    state_total_size, create_state_from_interior)
from .utils import (StatusCode, get_interior, set_interior, eps_for_dtype, safe_eps, allocate_scalar_history,
                    allocate_field_history, allocate_state_history, save_to_history, create_padded_array,
                    max_abs_interior, mean_interior, sum_interior, is_finite)
from .operators import (LinearOp, NonlinearOp, laplacian_2d, d2_central_2d, d1_central_2d, d1_upwind_2d,
                        d1_central_1d, d1_upwind_1d, d2_central_1d, laplacian_1d, gray_scott_reaction_op,
                        fd_laplacian_symbol_1d, fd_laplacian_symbol_2d, make_diffusion_op_1d, make_diffusion_op_2d,
                        make_reaction_op)
from .fft_solvers import (FFTCache1D, FFTCache2D, build_wavenumbers_1d, build_wavenumbers_2d,
                          laplacian_symbol_1d, laplacian_symbol_2d, create_fft_cache, create_fft_cache_1d,
                          create_fft_cache_2d, solve_helmholtz_1d, solve_helmholtz_2d, solve_helmholtz,
                          extract_interior,
                          embed_interior, solve_diffusion_fft_field, apply_diffusion_inverse_fft, fft_cn_step,
                          apply_diffusion_fft, diffusion_rhs_fft, exponential_filter_1d, exponential_filter_2d,
                          apply_spectral_filter_interior, apply_spectral_filter_field, apply_spectral_filter_state)
from .model import (MOLModel, create_gray_scott_model, create_gray_scott_periodic_fft, get_fft_cache,
                    get_diffusivities)
from .stepping import (IntegratorType, AdaptiveResult, imex_strang_step, imex_euler_step, imex_ssprk2_step,
                       estimate_error_imex_doubling, integrate_imex_fixed_dt, adaptive_integrate_imex,
                       adaptive_integrate, integrate_fixed_dt, euler_step, ssprk3_step, rk4_step, step_explicit,
                       step_explicit_with_error, estimate_error_doubling, estimate_error_implicit, AdaptiveState)
from .dt_policy import (CFLParams, PIDParams, ControllerState, create_initial_controller_state, pid_controller_dt,
                        handle_rejected_step, heisenberg_cfl_dt, imex_cfl_dt, implicit_robustness_dt, propose_dt,
                        propose_dt_imex, compute_error_order, estimate_reaction_stiffness)
from .fft_nonperiodic import (BCType as NonperiodicBCType, dst_I_fast, idst_I_fast, solve_helmholtz_dirichlet,
                              solve_helmholtz_neumann, solve_poisson_dirichlet, solve_poisson_neumann,
                              laplacian_symbol_dirichlet_cc, solve_helmholtz_dirichlet_cc, solve_poisson_dirichlet_cc,
                              solve_helmholtz_bc, solve_poisson_bc, NonperiodicFFTCache, NonperiodicFFTCache2D,
                              laplacian_symbol_dirichlet, laplacian_symbol_neumann, laplacian_symbol_mixed_dn,
                              create_nonperiodic_fft_cache, create_nonperiodic_fft_cache_2d, etd1_dirichlet,
                              etd1_neumann, etd1_dirichlet_cc)
from .fft_operators import DiffusionOperator, FFTLinearOperator, SpectralBounds, exact_cfl_dt
from .fft_integrators import (etd1_step, etd2_step, etdrk4_step, etd_integrate, batched_fft_matvec,
                              batched_fft_solve, batched_fft_exp_matvec, stacked_fft_solve_shared_op,
                              diffusion_only_etd1, ETDStepResult)
from .newton_krylov import NKParams, NKStats, NKResult, newton_krylov_solve, create_implicit_residual
from .preconditioners import (PrecondContext, Preconditioner, IdentityPreconditioner, BlockJacobiPreconditioner,
                              DiffusionPreconditioner,
                              FFTDiffusionPreconditioner, DSTDiffusionPreconditioner,
                              create_gray_scott_preconditioner, create_gray_scott_fft_preconditioner,
                              create_diffusion_preconditioner, create_fft_preconditioner)
from .reaction_diffusion import (ReactionDiffusionModel, create_reaction_diffusion_model, be_step, cn_step,
                                 bdf2_step)
from .jit_kernels import (phi1, phi2, phi3, etd1_kernel_1d, etd1_kernel_2d, etd2_kernel_1d, etdrk4_kernel_1d,
                          helmholtz_solve_1d, helmholtz_solve_2d, batched_fft_solve_1d, batched_exp_matvec_1d,
                          make_etd1_integrator, batched_etd1_kernel_1d, batched_etd2_kernel_1d, multi_operator_etd1_1d,
                          batched_etd1_kernel_2d, batched_helmholtz_solve_2d)
