"""CPU oracle for the moljax stiff-PDE hot path -- TEST INFRASTRUCTURE ONLY.

This package is a NumPy/SciPy float64 restatement of the reference moljax
algorithms (each function cites the reference file:line it follows).  It is
what the HIP path is compared against in the tests, and the CPU baseline
timed by bench.py.  Only ``tests/``, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg may import it; the product package ``moljax_amd`` never
does, and has no CPU fallback.

Pinning (what fixes each function independently of this restatement):
  spectral.fft_cn_*            reference-held golden numbers
                               (benchmarks/results/scipy_comparison.json:21,
                               benchmarks/results/bc_matrix.json:8,28,48) and the
                               closed form G^n u0.
  spectral.laplacian_symbol_2d the periodic 5-point stencil applied to Fourier
                               modes (exact eigenvalues).
  spectral.solve_helmholtz_2d  FD residual: (I - dt D lap_h) u = rhs to 1e-13.
  gray_scott.*                 fixed point (u=1, v=0), pure-diffusion closed
                               form (1 - dt/2 D lam)^-2 per step, reaction
                               formula on scalars.
  newton_krylov.*              golden AD-JVP norm
                               (benchmarks/results/jvp_vs_fd_sweep.json:3),
                               central finite differences, GMRES against a
                               dense solve, the DST preconditioner's exact
                               inverse, BDF2 coefficients and order.
  newton_krylov.gmres          its iteration count (tests/test_oracle_pins.py):
                               1 for M A = I, k for a diagonal operator with k
                               distinct eigenvalues that b touches, m per
                               restart cycle when the tolerance is never met,
                               0 for b = 0; planted off-by-one mistakes fail.
  newton_krylov.newton_krylov  Newton / GMRES counts of the linear problem
                               (one GMRES step per sine mode of the residual)
                               and of pure diffusion (one per Newton step).
  fd.*                         sine modes (exact discrete eigenvalues) and the
                               periodic wrap of the ghost fill.
  fd.d1_upwind_2d              exact slope of a linear field for either sign
                               of the velocity, and the one-sided difference
                               it picks on a quadratic.
  nonperiodic.*                scipy.fft.dst / dct of the same definitions,
                               brute-force DST-II, FD residual on filled ghosts.
  etdrk4.*                     contour limits (Q -> dt/2, f1..f3 -> dt/6), the
                               linear problem (N = 0) exactly E u, fourth order
                               (ETDRK4), first and second order (ETD1, ETD2).
  adaptive.*                   controller closed forms, fixed-point dt
                               sequence, first-attempt error in closed form,
                               status codes.
  adaptive.implicit_robustness_dt and adaptive_integrate_implicit: the dt
                               rules case by case, and the saturated-growth
                               dt sequence (and rejection) of y' = lam y with
                               D = 0 in closed form.
  jit_kernels.*                phi_k(0) = 1/k! and the phi recurrences, ETD1 on
                               a Fourier mode, ETD2 / ETDRK4 with equal N
                               reduce to ETD1, the Helmholtz inverse on a mode.
  spectral.exponential_filter_* and apply_spectral_filter_*: closed forms of
                               the filter values and their action on modes.
"""
